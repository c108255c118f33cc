"""The reference's Python names for the integrator (python/rigidbodyplugin.i:51-72)
before it is bound to a context: no GPU is touched.  Settings follow the
reference's rules (openmmapi/src/RigidBodyIntegrator.cpp:26-38): a negative
rotation mode is refused, and queries and steps need a bound context."""
import pytest

from openmm_rigidbody_plugin_amd.integrator import RigidBodyIntegrator


def test_settings_before_binding():
    it = RigidBodyIntegrator(0.002, [0, 1, 1, 1, 2, 2, 2])
    assert it.getRotationMode() == 0 and it.getComputeRefinedEnergies() is False
    assert it.getStepSize() == 0.002
    it.setRotationMode(10)
    it.setComputeRefinedEnergies(True)
    it.setStepSize(0.001)
    assert it.getRotationMode() == 10 and it.getComputeRefinedEnergies() is True
    assert it.getStepSize() == 0.001
    with pytest.raises(RuntimeError, match="negative"):
        it.setRotationMode(-1)


def test_unbound_integrator_refuses_steps_and_queries():
    it = RigidBodyIntegrator(0.002, [1, 1, 1])
    for call in (lambda: it.step(1), it.getKineticEnergies, it.getRefinedKineticEnergies,
                 it.getPotentialEnergyRefinement, it.getRigidBodySystem):
        with pytest.raises(RuntimeError, match="not bound"):
            call()
