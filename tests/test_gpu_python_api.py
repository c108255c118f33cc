"""The reference's Python names for the integrator (python/rigidbodyplugin.i:51-72)
bound to a system on the GPU: the same results as the plugin layer driven
directly (bit for bit), the FP64 oracle to the mixed tolerances, the
reference's refusals once bound (openmmapi/src/RigidBodyIntegrator.cpp:26-38),
and RigidBodySystem's DOF and energies."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from openmm_rigidbody_plugin_amd.integrator import Context, RigidBodyIntegrator

pytestmark = pytest.mark.gpu


def bind(spec, refined=False):
    it = RigidBodyIntegrator(spec.dt, spec.body_indices)
    it.setRotationMode(spec.rotation_mode)
    it.setComputeRefinedEnergies(refined)
    ctx = Context(spec.masses, it, tether_k=spec.tether_k, tether_r0=spec.tether_r0)
    ctx.setPositions(spec.positions)
    ctx.setVelocities(spec.velocities)
    return it, ctx


@pytest.mark.parametrize("mode", [10, 0])
def test_python_api_matches_plugin_and_oracle(hip, oracle, mode):
    from openmm_rigidbody_plugin_amd import plugin
    spec = systems.water_c1(rotation_mode=mode)
    it, ctx = bind(spec)
    sim = plugin.Simulation.from_spec(spec, precision="mixed")
    ref = oracle.OracleSim.from_spec(spec)
    it.step(10)
    sim.step(10)
    ref.step(10)
    st = ctx.getState(getPositions=True, getVelocities=True, getEnergy=True)
    np.testing.assert_array_equal(st["positions"], sim.state()["positions"])
    assert np.abs(st["positions"] - ref.positions()).max() <= 1e-6
    ke = it.getKineticEnergies()
    np.testing.assert_array_equal(ke, sim.energies()["kinetic"])
    np.testing.assert_allclose(ke, ref.kinetic_energies(), rtol=1e-5)
    rbs = it.getRigidBodySystem()
    assert rbs.getNumDOF() == ref.counts()["numDOF"]
    assert rbs.getKineticEnergy() == pytest.approx(sum(ke), rel=1e-15)
    assert (rbs.getTranslationalEnergy(), rbs.getRotationalEnergy()) == tuple(ke)
    assert it.getRefinedKineticEnergies() == ke and it.getPotentialEnergyRefinement() == 0.0


def test_python_api_refusals_once_bound(hip):
    spec = systems.water_c1(rotation_mode=10)
    it, ctx = bind(spec)
    with pytest.raises(RuntimeError, match="already bound"):
        it.setRotationMode(0)
    with pytest.raises(RuntimeError, match="already bound"):
        it.setComputeRefinedEnergies(True)
    with pytest.raises(RuntimeError, match="already bound"):
        Context(spec.masses, it)


def test_python_api_refined_energies_and_step_size(hip):
    from openmm_rigidbody_plugin_amd import plugin
    spec = systems.water_c1(rotation_mode=10)
    it, ctx = bind(spec, refined=True)
    sim = plugin.Simulation.from_spec(spec, precision="mixed", refined_energies=True)
    for s in (it, sim):
        s.step(5)
    it.setStepSize(0.5 * spec.dt)
    sim.set_step_size(0.5 * spec.dt)
    for s in (it, sim):
        s.step(5)
    e = sim.energies()
    np.testing.assert_array_equal(it.getRefinedKineticEnergies(), e["refined_kinetic"])
    assert it.getPotentialEnergyRefinement() == e["potential_refinement"]
    assert it.getRefinedKineticEnergies() != it.getKineticEnergies()
    np.testing.assert_array_equal(ctx.getState(getPositions=True)["positions"], sim.state()["positions"])


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_python_force_field_through_the_slot(hip, oracle, mode, precision):
    """A force field written in Python (the tether, in numpy) passed as the
    Context's `force` runs between part 1 and part 2 of every step and gives
    the oracle's trajectory (C1 tolerances, tests/tolerances.py)."""
    from tolerances import TOL, check_state
    spec = systems.water_c1(rotation_mode=mode)
    calls = []

    def tether(x):
        calls.append(1)
        d = x - spec.tether_r0
        return -spec.tether_k[:, None] * d, 0.5 * float((spec.tether_k[:, None] * d * d).sum())

    it = RigidBodyIntegrator(spec.dt, spec.body_indices)
    it.setRotationMode(mode)
    ctx = Context(spec.masses, it, precision=precision, force=tether)
    ctx.setPositions(spec.positions)
    ctx.setVelocities(spec.velocities)
    n0 = len(calls)
    it.step(10)
    assert len(calls) - n0 == 10          # one force evaluation per step
    ref = oracle.OracleSim.from_spec(spec)
    ref.step(10)
    st = ctx.getState(getPositions=True, getVelocities=True, getEnergy=True)
    check_state(spec, ref.positions(), ref.velocities(), st["positions"], st["velocities"], precision, 10)
    np.testing.assert_allclose(it.getKineticEnergies(), ref.kinetic_energies(), rtol=TOL[precision]["ke"])
    assert st["potential"] == pytest.approx(ref.potential_energy(), rel=1e-6)


def test_python_force_field_errors_are_raised(hip):
    from openmm_rigidbody_plugin_amd import plugin
    spec = systems.water_box(8, 0.3125, seed=3)

    def broken(x):
        raise ValueError("no forces today")

    it = RigidBodyIntegrator(spec.dt, spec.body_indices)
    ctx = Context(spec.masses, it, force=broken)
    with pytest.raises(plugin.PluginError, match="no forces today"):
        ctx.setPositions(spec.positions)


def test_state_data_reporter_on_the_gpu(hip, oracle):
    """The reference's StateDataReporter columns (python/statedatareporter.py)
    from a bound integrator stepped by app.Simulation: the step and time, the
    kinetic energy split, the temperature over the rigid DOF, and the refined
    energies, each against the oracle at the report steps."""
    import io
    from openmm_rigidbody_plugin_amd import app
    spec = systems.water_c1(rotation_mode=10)
    it, ctx = bind(spec, refined=True)
    ref = oracle.OracleSim(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode)
    ref.set_tether(spec.tether_k, spec.tether_r0)
    ref.set_refined(True)
    ref.set_positions(spec.positions)
    ref.set_velocities(spec.velocities)
    sim = app.Simulation(it, ctx)
    out = io.StringIO()
    sim.reporters.append(app.StateDataReporter(out, 5, step=True, time=True, kineticEnergy=True, temperature=True,
                                               translationalEnergy=True, rotationalEnergy=True,
                                               refinedKineticEnergy=True, refinedTemperature=True))
    sim.step(10)
    lines = out.getvalue().splitlines()
    assert lines[0].startswith('#"Step","Time (ps)","Kinetic Energy (kJ/mole)","Temperature (K)"')
    assert len(lines) == 3
    dof = ref.counts()["numDOF"]
    for row, steps in zip(lines[1:], (5, 10)):
        ref.step(5)
        v = [float(x) for x in row.split(",")]
        kt, kr = ref.kinetic_energies()
        rkt, rkr = ref.refined_kinetic_energies()
        R = app.MOLAR_GAS_CONSTANT_R
        assert v[0] == steps and v[1] == pytest.approx(steps * spec.dt, rel=1e-12)
        # mixed precision: kinetic energies to 1e-5, refined ones (differences
        # of fp32 quaternions) to 2e-5
        np.testing.assert_allclose(v[2:6], [kt + kr, 2 * (kt + kr) / (dof * R), kt, kr], rtol=1e-5)
        np.testing.assert_allclose(v[6:], [rkt + rkr, 2 * (rkt + rkr) / (dof * R)], rtol=2e-5)
