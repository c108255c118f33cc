"""System-level checks of the oracle: bookkeeping against a brute-force
statement, the reference's own known-answer test (testSingleBond), conservation
laws, reversibility, and every documented deviation from the reference."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

from test_oracle_math import clean_spec


def atom_index_spec(body_index):
    """RigidBodySystem::initialize ordering (RigidBodySystem.cpp:84-105):
    free atoms ascending, then body 1's atoms ascending, body 2's, ..."""
    free = [i for i, b in enumerate(body_index) if b == 0]
    bodies = [i for b in range(1, max(body_index, default=0) + 1) for i, x in enumerate(body_index) if x == b]
    return np.array(free + bodies, dtype=np.int32)


def test_bookkeeping_matches_spec(oracle):
    rng = np.random.default_rng(1)
    for trial in range(10):
        n = int(rng.integers(10, 200))
        labels = rng.integers(-2, 15, size=n)
        # give every positive label at least two atoms
        for v in np.unique(labels[labels > 0]):
            if np.count_nonzero(labels == v) == 1:
                labels[rng.integers(0, n)] = v
                labels[np.nonzero(labels == v)[0][0]] = v
        for v in np.unique(labels[labels > 0]):
            if np.count_nonzero(labels == v) == 1:
                labels[labels == v] = 0
        s = oracle.OracleSim(rng.uniform(1, 20, n), labels, 0.002, 10)
        bi = clean_spec(labels)
        assert np.array_equal(s.body_index(), bi)
        assert np.array_equal(s.atom_index(), atom_index_spec(bi))
        c = s.counts()
        assert c["numFree"] == np.count_nonzero(bi == 0)
        assert c["numBodies"] == bi.max()


def test_single_bond_known_answer(oracle):
    """tests/TestRigidBodyIntegrator.h:49-83: two free atoms (m = 2) joined by a
    harmonic bond (r0 = 1.5, k = 1) oscillate as 1.5 + 0.5 cos t."""
    s = oracle.OracleSim([2.0, 2.0], [0, 0], 0.01, 0)
    s.add_bond(0, 1, 1.5, 1.0)
    s.set_positions(np.array([[-1.0, 0, 0], [1.0, 0, 0]]))
    s.set_velocities(np.zeros((2, 3)))
    e0 = s.kinetic_energies().sum() + s.potential_energy()
    for _ in range(1000):
        t = s.time()
        x, v = s.positions(), s.velocities()
        d = 1.5 + 0.5 * np.cos(t)
        assert np.allclose(x[0], [-0.5 * d, 0, 0], atol=0.02) and np.allclose(x[1], [0.5 * d, 0, 0], atol=0.02)
        sp = -0.5 * np.sin(t)
        assert np.allclose(v[0], [-0.5 * sp, 0, 0], atol=0.02) and np.allclose(v[1], [0.5 * sp, 0, 0], atol=0.02)
        assert s.kinetic_energies().sum() + s.potential_energy() == pytest.approx(e0, abs=0.01)
        s.step(1)
    assert s.time() == pytest.approx(10.0, abs=1e-5)


def test_single_bond_planted_mistake_fails(oracle):
    """The KAT has teeth: a factor-2 error in the step (dt doubled) fails it."""
    s = oracle.OracleSim([2.0, 2.0], [0, 0], 0.02, 0)
    s.add_bond(0, 1, 1.5, 1.0)
    s.set_positions(np.array([[-1.0, 0, 0], [1.0, 0, 0]]))
    s.set_velocities(np.zeros((2, 3)))
    s.step(250)
    x = s.positions()
    assert abs(x[1, 0] - 0.5 * (1.5 + 0.5 * np.cos(2.5))) > 0.02


@pytest.mark.parametrize("mode", [0, 10])
def test_water_energy_and_kinetic_identity(oracle, mode):
    spec = systems.water_box(64, 0.3125, num_free=8, seed=3, rotation_mode=mode)
    s = oracle.OracleSim.from_spec(spec)
    e = []
    for _ in range(40):
        s.step(25)
        ke = s.kinetic_energies()
        e.append(ke.sum() + s.potential_energy())
        # atom velocities are rigid-body velocities after a step
        v = s.velocities()
        assert 0.5 * np.sum(spec.masses[:, None] * v * v) == pytest.approx(ke.sum(), rel=1e-12)
    e = np.array(e)
    # bounded fluctuation, no drift
    assert np.ptp(e) < 2e-3 * abs(e.mean())
    assert abs(e[:20].mean() - e[20:].mean()) < 5e-4 * abs(e.mean())
    # rigidity
    x = s.positions()
    x0 = spec.positions
    for b in range(64):
        sl = slice(3 * b, 3 * b + 3)
        d0 = np.linalg.norm(x0[sl][:, None] - x0[sl][None], axis=-1)
        d1 = np.linalg.norm(x[sl][:, None] - x[sl][None], axis=-1)
        assert np.allclose(d0, d1, atol=1e-12)


def test_time_reversibility(oracle):
    spec = systems.water_box(27, 0.3125, num_free=4, seed=4, rotation_mode=10)
    s = oracle.OracleSim.from_spec(spec)
    x0 = s.positions()
    s.step(50)
    s.set_step_size(-spec.dt)
    s.step(50)
    assert np.abs(s.positions() - x0).max() < 1e-11


def test_hetero_bodies_conserve_energy(oracle):
    spec = systems.heterogeneous(num_bodies=300, num_free=200, seed=9, rotation_mode=0)
    s = oracle.OracleSim.from_spec(spec)
    e0 = s.kinetic_energies().sum() + s.potential_energy()
    s.step(200)
    e1 = s.kinetic_energies().sum() + s.potential_energy()
    assert abs(e1 - e0) < 2e-3 * abs(e0)
    assert np.all(np.isfinite(s.positions()))


# ---- documented deviations (flag off reproduces the reference defect) ----

def test_pcm_reset_deviation(oracle):
    spec = systems.water_box(8, 0.3125, seed=5)
    fixed = oracle.OracleSim.from_spec(spec)
    faithful = oracle.OracleSim.from_spec(spec, flags=oracle.FIX_ALL & ~oracle.FIX_PCM_RESET)
    for s in (fixed, faithful):
        s.set_velocities(spec.velocities)
    # the reference accumulates p_cm on every rebuild (RigidBody.cpp:129)
    assert fixed.body_state(0)["pcm"] == pytest.approx(faithful.body_state(0)["pcm"] / 2, rel=1e-12)


def test_linear_frame_deviation(oracle):
    spec = systems.heterogeneous(num_bodies=40, num_free=0, seed=6, rotation_mode=10, shuffle=False)
    fixed = oracle.OracleSim.from_spec(spec)
    faithful = oracle.OracleSim.from_spec(spec, flags=oracle.FIX_ALL & ~oracle.FIX_LINEAR_FRAME)
    d_fixed, d_faith = fixed.body_fixed_positions(), faithful.body_fixed_positions()
    lin = [b for b in fixed.body_states() if b["dof"] == 5]
    assert lin, "expected linear bodies in the mix"
    n_checked = 0
    for b in fixed.body_states():
        if b["dof"] != 5:
            continue
        sl = slice(int(b["loc"]), int(b["loc"] + b["N"]))
        assert np.allclose(d_fixed[sl][:, :2], 0, atol=1e-14)       # on the zero-moment axis
        assert not np.allclose(d_faith[sl][:, :2], 0, atol=1e-6)    # the reference's transposed frame
        n_checked += 1
    assert n_checked == len(lin)


def test_linear_exact_deviation(oracle):
    spec = systems.heterogeneous(num_bodies=40, num_free=0, seed=6, rotation_mode=0, shuffle=False)
    faithful = oracle.OracleSim.from_spec(spec, flags=oracle.FIX_ALL & ~oracle.FIX_LINEAR_EXACT)
    faithful.step(1)
    assert not np.all(np.isfinite(faithful.positions()))   # divides by I0 - I1 = 0
    fixed = oracle.OracleSim.from_spec(spec)
    fixed.step(5)
    assert np.all(np.isfinite(fixed.positions()))


def test_orthonormal_deviation(oracle):
    """Reference orthonormal (RigidBody.cpp:51-58) projects e_m for m = the
    smallest signed component of the axis u: for a diatomic whose axis points
    along -x that e_m is parallel to u and the frame is 0/0.  The fixed rule
    (smallest |u_k|) gives a proper frame with the same dynamics."""
    masses = [3.0, 5.0]
    x = np.array([[0.75, 2.0, 3.0], [1.15, 2.0, 3.0]])   # axis from the centre of mass to atom 0 is -x
    v = np.array([[0.3, -0.5, 0.1], [0.3, 0.3, 0.1]])
    for mode in (0, 4):
        faithful = oracle.OracleSim(masses, [1, 1], 0.002, mode, flags=oracle.FIX_ALL & ~oracle.FIX_ORTHONORMAL)
        faithful.set_positions(x)
        assert not np.all(np.isfinite(faithful.body_state(0)["q"]))
        fixed = oracle.OracleSim(masses, [1, 1], 0.002, mode)
        fixed.set_positions(x)
        fixed.set_velocities(v)
        q = fixed.body_state(0)["q"]
        assert np.all(np.isfinite(q)) and np.linalg.norm(q) == pytest.approx(1.0, abs=1e-15)
        d = fixed.body_fixed_positions()
        assert np.allclose(d[:, :2], 0.0, atol=1e-15)        # atoms on body axis 3
        fixed.step(100)
        r = fixed.positions()
        assert np.linalg.norm(r[1] - r[0]) == pytest.approx(0.4, abs=1e-12)
    # an axis whose smallest signed component is not parallel to it: both
    # rules give a valid frame (rotated about the axis) and, with the exact
    # rotation (invariant under that choice), the same motion
    x2 = np.array([[0.8, 1.9, 3.05], [1.1, 2.07, 2.95]])
    runs = []
    for flags in (oracle.FIX_ALL, oracle.FIX_ALL & ~oracle.FIX_ORTHONORMAL):
        s = oracle.OracleSim(masses, [1, 1], 0.002, 0, flags=flags)
        s.set_positions(x2)
        s.set_velocities(v)
        s.step(50)
        runs.append(s.positions())
    assert np.abs(runs[0] - runs[1]).max() < 1e-12


def test_stale_system_deviation(oracle):
    """Reference: setVelocities after stepping rebuilds from the integrator's
    stale copy of the body system, so the bodies jump back to the geometry of
    the last setPositions (RigidBodyIntegrator.cpp:70-72)."""
    spec = systems.water_box(8, 0.3125, seed=7)
    for flags, expect_jump in ((oracle.FIX_ALL, False), (oracle.FIX_ALL & ~oracle.FIX_STALE_SYSTEM, True)):
        s = oracle.OracleSim.from_spec(spec, flags=flags)
        s.step(50)
        x = s.positions()
        s.set_velocities(s.velocities())
        s.step(1)
        jump = np.abs(s.positions() - x).max()
        assert (jump > 0.01) == expect_jump, jump


def test_rejections(oracle):
    with pytest.raises(oracle.OracleError, match="at least two atoms"):
        oracle.OracleSim([1.0, 1.0, 1.0], [1, 2, 2], 0.002, 0)
    with pytest.raises(oracle.OracleError, match="massless"):
        oracle.OracleSim([1.0, 0.0, 1.0], [1, 1, 0], 0.002, 0)
    with pytest.raises(oracle.OracleError, match="negative"):
        oracle.OracleSim([1.0, 1.0], [1, 1], 0.002, -1)
