"""HIP path vs the FP64 oracle on the same seeded inputs, through the C ABI.

Tolerances (after 10 steps) are those of tests/tolerances.py: in mixed
precision positions <= 1e-6 nm, each body's momentum and angular momentum
within 1e-5 of its own magnitude above the floor the force field's fp32
rounding sets, free-atom velocities likewise, kinetic energies <= 1e-5; in
double precision 1e-10 nm, 1e-9 relative (floor 1e-10 of the RMS), 1e-11.
Integer bookkeeping (body numbering, atom order, counts, DOF) is bit-exact.
"""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from tolerances import TOL, check_state

pytestmark = pytest.mark.gpu


def configs():
    return {
        "water512+ar64/nosquish": lambda: systems.water_c1(rotation_mode=10),
        "water512+ar64/exact": lambda: systems.water_c1(rotation_mode=0),
        "hetero2k+free5k/exact": lambda: systems.heterogeneous(2000, 5000, seed=5, rotation_mode=0),
        "hetero2k+free5k/nosquish": lambda: systems.heterogeneous(2000, 5000, seed=5, rotation_mode=10),
    }


def compare_states(spec, o, g, precision, steps=10):
    check_state(spec, o.positions(), o.velocities(), g.positions(), g.velocities(), precision, steps)
    ko, kg = o.kinetic_energies(), g.kinetic_energies()
    assert np.allclose(kg, ko, rtol=TOL[precision]["ke"], atol=0)


@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("name", list(configs()))
def test_ten_steps_match_oracle(hip, oracle, name, precision):
    spec = configs()[name]()
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision=precision)
    tol = TOL[precision]
    # state after setPositions/setVelocities
    ko, kg = o.kinetic_energies(), g.kinetic_energies()
    assert np.allclose(kg, ko, rtol=tol["ke"], atol=0)
    o.step(10)
    g.step(10)
    compare_states(spec, o, g, precision)
    t, nsteps = g.time()
    assert t == pytest.approx(o.time(), rel=1e-15) and nsteps == 10


@pytest.mark.parametrize("name", list(configs()))
def test_bookkeeping_bit_exact(hip, oracle, name):
    spec = configs()[name]()
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec)
    assert g.counts() == o.counts()
    assert np.array_equal(g.body_index(), o.body_index())
    assert np.array_equal(g.atom_index(), o.atom_index())


def test_body_state_after_setup_double(hip, oracle):
    """FP64 device setup = the host build of the reference: same frames,
    quaternions, body-fixed coordinates and momenta."""
    spec = systems.heterogeneous(1500, 100, seed=8, rotation_mode=0)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision="double")
    bo, bg = o.body_states(), g.body_states()
    for a, b in zip(bo, bg):
        assert a["N"] == b["N"] and a["dof"] == b["dof"] and a["loc"] == b["loc"]
        for key, tol in (("mass", 1e-14), ("I", 1e-12), ("invI", 1e-9), ("rcm", 1e-13), ("pcm", 1e-12),
                         ("q", 1e-12), ("pi", 1e-11), ("force", 1e-11), ("torque", 1e-11)):
            scale = max(1.0, np.abs(a[key]).max())
            assert np.abs(np.asarray(a[key]) - np.asarray(b[key])).max() <= tol * scale, (key, a[key], b[key])
    assert np.abs(o.body_fixed_positions() - g.body_fixed_positions()).max() < 1e-13


def test_single_bond_known_answer_on_gpu(hip):
    """tests/TestRigidBodyIntegrator.h:49-83 through the HIP path (free atoms, bond force)."""
    for precision in ("mixed", "double"):
        g = hip.Context([2.0, 2.0], [0, 0], 0.01, rotation_mode=0, precision=precision)
        g.add_bonds([[0, 1]], 1.5, 1.0)
        g.set_positions(np.array([[-1.0, 0, 0], [1.0, 0, 0]]))
        g.set_velocities(np.zeros((2, 3)))
        e0 = g.kinetic_energies().sum() + g.compute_forces(energy=True)
        for i in range(0, 1000, 10):
            t = g.time()[0]
            x, v = g.positions(), g.velocities()
            d = 1.5 + 0.5 * np.cos(t)
            assert np.allclose(x[0], [-0.5 * d, 0, 0], atol=0.02) and np.allclose(x[1], [0.5 * d, 0, 0], atol=0.02)
            sp = -0.5 * np.sin(t)
            assert np.allclose(v[1], [0.5 * sp, 0, 0], atol=0.02)
            assert g.kinetic_energies().sum() + g.potential_energy() == pytest.approx(e0, abs=0.01)
            g.step(10)
        assert g.time()[0] == pytest.approx(10.0, abs=1e-5)


@pytest.mark.parametrize("mode", [10, 0])
def test_energy_drift_matches_reference(hip, oracle, mode):
    """10^4 steps of the 512-water + 64-argon system (NO-SQUISH n = 10 and
    the exact rotation).

    The tethered rigid bodies are chaotic: after ~10^3 steps any two runs
    (FP64 GPU included) decorrelate, so the energy records are compared as
    statistics, sampled every 100 steps: same mean energy, same fluctuation,
    and a fitted drift over the run that is small against the fluctuation in
    both and equal within it."""
    spec = systems.water_c1(rotation_mode=mode)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision="mixed")
    eo, eg = [], []
    for _ in range(100):
        o.step(100)
        g.step(100)
        eo.append(o.kinetic_energies().sum() + o.potential_energy())
        eg.append(g.kinetic_energies().sum() + g.potential_energy())
    eo, eg = np.array(eo), np.array(eg)
    t = np.arange(100)
    drift_o = np.polyfit(t, eo, 1)[0] * 100
    drift_g = np.polyfit(t, eg, 1)[0] * 100
    so, sg = eo.std(), eg.std()
    assert so < 5e-5 * abs(eo.mean())                   # the reference conserves energy
    assert abs(eg.mean() - eo.mean()) < 0.5 * so
    assert sg == pytest.approx(so, rel=0.25)
    assert abs(drift_o) < 2 * so and abs(drift_g) < 2 * so
    assert abs(drift_g - drift_o) < 2 * so


def test_graph_and_launch_paths_bitwise_equal(hip):
    spec = systems.heterogeneous(500, 800, seed=12, rotation_mode=10)
    a = hip.Context.from_spec(spec)
    b = hip.Context.from_spec(spec)
    b.use_graph(False)
    a.step(7)
    b.step(7)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.kinetic_energies(), b.kinetic_energies())


def test_part_entry_points_equal_execute(hip):
    spec = systems.water_box(300, 0.3125, num_free=40, seed=13)
    a = hip.Context.from_spec(spec)
    b = hip.Context.from_spec(spec)
    a.step(1)
    b.part1()
    b.compute_forces()
    b.part2()
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())


def test_set_velocities_after_steps_and_step_size_change(hip, oracle):
    """Velocity-only rebuild on the live geometry, then a new step size."""
    spec = systems.heterogeneous(400, 300, seed=14, rotation_mode=10)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision="double")
    o.step(5)
    g.step(5)
    v = o.velocities() * 0.9
    o.set_velocities(v)
    g.set_velocities(v)
    o.set_step_size(0.0015)
    g.set_step_size(0.0015)
    o.step(5)
    g.step(5)
    compare_states(spec, o, g, "double")


def test_autotune_keeps_state(hip, oracle):
    spec = systems.heterogeneous(3000, 2000, seed=15, rotation_mode=10)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec)
    g.step(3)
    o.step(3)
    ke_before = g.kinetic_energies()
    g.autotune(reps=2)
    assert np.array_equal(g.kinetic_energies(), ke_before)
    assert "L" in g.describe()
    g.step(7)
    o.step(7)
    compare_states(spec, o, g, "mixed")


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lanes_per_body_tiles(hip, oracle, lanes):
    spec = systems.heterogeneous(600, 200, seed=16, rotation_mode=0)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision="double")
    for n in (2, 3, 6, 20):
        g.set_lanes(n, lanes)
    g.step(10)
    o.step(10)
    compare_states(spec, o, g, "double")


def test_errors_are_reported(hip):
    with pytest.raises(hip.RbhipError, match="single atom"):
        hip.Context([1.0, 1.0, 1.0], [1, 2, 2], 0.002)
    with pytest.raises(hip.RbhipError, match="massless"):
        hip.Context([1.0, 0.0], [1, 1], 0.002)
    with pytest.raises(hip.RbhipError, match="negative"):
        hip.Context([1.0, 1.0], [1, 1], 0.002, rotation_mode=-1)
    g = hip.Context([1.0, 2.0, 3.0], [1, 1, 0], 0.002)
    with pytest.raises(hip.RbhipError, match="positions must be set"):
        g.step(1)
    with pytest.raises(hip.RbhipError):
        g.set_launch(block=96)


@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("mode", [10, 0])
def test_fused_steps_equal_two_pass(hip, precision, mode):
    """step(k) with fused interior boundaries = k x [part1, forces, part2]:
    bit for bit in mixed precision (the state is rounded to fp32 at the same
    points); in FP64 the compiler fuses multiply-adds differently across the
    single kernel, so the two paths agree to a few ulp."""
    spec = systems.heterogeneous(700, 900, seed=17, rotation_mode=mode)
    a = hip.Context.from_spec(spec, precision=precision)
    b = hip.Context.from_spec(spec, precision=precision)
    b.use_fusion(False)
    if mode == 10:
        for c in (a, b):
            c.set_lanes(20, 4)
    for c in (a, b):
        c.step(9)
        c.step(1)
        c.step(5)
    if precision == "mixed":
        assert np.array_equal(a.positions(), b.positions())
        assert np.array_equal(a.velocities(), b.velocities())
        assert np.array_equal(a.kinetic_energies(), b.kinetic_energies())
        assert np.array_equal(a.body_states_array(), b.body_states_array())
    else:
        assert np.abs(a.positions() - b.positions()).max() < 1e-14
        assert np.abs(a.velocities() - b.velocities()).max() < 1e-13
        assert np.allclose(a.kinetic_energies(), b.kinetic_energies(), rtol=1e-14, atol=0)


@pytest.mark.parametrize("graph,fusion", [(True, True), (False, True), (True, False), (False, False)])
def test_potential_energy_is_the_last_evaluations(hip, oracle, graph, fusion):
    """A step asks for forces alone; the library takes the potential energy
    at the last force evaluation of each call only.  After calls of every
    length (1, 2, one fused chunk and more, with and without the
    refined-energy tail) the reported energy is that of the final positions:
    the oracle's within the mixed tolerance, and the same bits on every
    launch path."""
    spec = systems.heterogeneous(700, 900, seed=19, rotation_mode=10)
    # (anchors off the start positions, so the energy is not zero)
    spec.tether_r0 = spec.tether_r0 + 0.03 * np.random.default_rng(19).normal(size=spec.tether_r0.shape)
    o = oracle.OracleSim.from_spec(spec)
    g = hip.Context.from_spec(spec, precision="mixed")
    ref = hip.Context.from_spec(spec, precision="mixed")
    ref.use_graph(False)
    ref.use_fusion(False)
    g.use_graph(graph)
    g.use_fusion(fusion)
    assert g.potential_energy() == pytest.approx(o.potential_energy(), rel=1e-5)
    for k, refined in ((1, False), (2, False), (3, True), (17, False), (40, False), (2, True), (1, True)):
        for c in (g, ref):
            c.set_refined(refined)
        o.step(k)
        g.step(k)
        ref.step(k)
        assert g.potential_energy() == pytest.approx(o.potential_energy(), rel=1e-5), k
        assert g.potential_energy() == ref.potential_energy(), k
