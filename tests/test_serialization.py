"""Serialization of RigidBodyIntegrator (host code, no device): the
reference's round-trip test (serialization/tests/TestSerializeRigidBodyIntegrator.cpp)
and the extra cases of tests/cpp/TestSerializeRigidBodyIntegrator.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "_bin", "TestSerializeRigidBodyIntegrator")


def test_serialize_rigid_body_integrator():
    from openmm_rigidbody_plugin_amd import build
    if not os.path.exists(build.PLUGIN):
        pytest.skip("C++ plugin layer not built")
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("Done")
