"""MI355X (gfx950) RigidBodyIntegrator step: HIP kernels behind a C ABI.

The compute path is ``librbhip.so`` (include/rbhip.h), bound by
:mod:`openmm_rigidbody_plugin_amd.rbhip` (the library itself is loaded on
first use); the C++ plugin layer that mirrors the reference's
RigidBodyIntegrator / IntegrateRigidBodyStepKernel API lives in ``host/`` and
builds into ``librigidbody_hip.so``, bound by
:mod:`openmm_rigidbody_plugin_amd.plugin`.
"""
from . import plugin, rbhip, systems  # noqa: F401

__all__ = ["plugin", "rbhip", "systems"]
