"""Edge shapes of the step on the GPU against the FP64 oracle (10 steps,
mixed and double tolerances as in test_gpu_parity): no bodies at all, large
bodies (8 to 200 atoms, several lanes per body), a single body, bonds that
couple bodies to each other and to free atoms, and a start from rest."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd.systems import SystemSpec, maxwell_boltzmann

pytestmark = pytest.mark.gpu

TOL = {"mixed": dict(x=1e-6, v=1e-5, ke=1e-5), "double": dict(x=1e-10, v=1e-9, ke=1e-11)}


def clusters(sizes, num_free, seed, mode, spacing=2.0, radius=0.4):
    """Rigid clusters of the given atom counts (atoms uniform in a sphere) and
    free atoms, on a lattice, with random masses, 300 K velocities and tethers."""
    rng = np.random.default_rng(seed)
    n_sites = len(sizes) + num_free
    side = int(np.ceil(n_sites ** (1 / 3)))
    sites = np.array([(i, j, k) for i in range(side) for j in range(side) for k in range(side)],
                     dtype=float)[:n_sites] * spacing
    rng.shuffle(sites)
    pos, bi = [], []
    for b, n in enumerate(sizes):
        d = rng.normal(size=(n, 3))
        d *= (radius * rng.uniform(0.2, 1.0, size=(n, 1)) ** (1 / 3)) / np.linalg.norm(d, axis=1, keepdims=True)
        pos.append(sites[b] + d)
        bi += [b + 1] * n
    if num_free:
        pos.append(sites[len(sizes):])
        bi += [0] * num_free
    x = np.concatenate(pos)
    m = rng.uniform(1.0, 20.0, size=x.shape[0])
    v = maxwell_boltzmann(rng, m, 300.0)
    return SystemSpec(masses=m, body_indices=np.array(bi, dtype=np.int32), positions=x, velocities=v,
                      tether_k=np.full(x.shape[0], 800.0), tether_r0=x.copy(), dt=0.002, rotation_mode=mode)


def compare(oracle, hip, spec, precision, bonds=None, start_at_rest=False):
    o = oracle.OracleSim(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode)
    g = hip.Context(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode, precision)
    for s in (o, g):
        s.set_tether(spec.tether_k, spec.tether_r0)
    if bonds:
        for a, b, L, k in bonds:
            o.add_bond(a, b, L, k)
        g.add_bonds([(a, b) for a, b, _, _ in bonds], [L for _, _, L, _ in bonds], [k for _, _, _, k in bonds])
    for s in (o, g):
        s.set_positions(spec.positions)
        if not start_at_rest:
            s.set_velocities(spec.velocities)
    o.step(10)
    g.step(10)
    tol = TOL[precision]
    xo, xg = o.positions(), g.positions()
    vo, vg = o.velocities(), g.velocities()
    assert np.all(np.isfinite(xg)) and np.all(np.isfinite(vg))
    assert np.abs(xo - xg).max() <= tol["x"], np.abs(xo - xg).max()
    vscale = max(np.sqrt((vo ** 2).sum(1).mean()), 1e-3)
    assert np.abs(vo - vg).max() <= tol["v"] * vscale, np.abs(vo - vg).max() / vscale
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=tol["ke"], atol=1e-9)
    np.testing.assert_array_equal(g.atom_index(), o.atom_index())
    assert g.counts()["numDOF"] == o.counts()["numDOF"]
    return g


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_free_atoms_only(hip, oracle, precision):
    spec = clusters([], 900, seed=1, mode=10)
    bonds = [(i, i + 1, 0.5, 300.0) for i in range(0, 898, 2)]
    compare(oracle, hip, spec, precision, bonds=bonds)


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_large_bodies(hip, oracle, mode, precision):
    sizes = [8] * 20 + [13] * 15 + [33] * 10 + [64] * 6 + [200] * 3 + [3] * 30
    spec = clusters(sizes, 50, seed=2, mode=mode)
    g = compare(oracle, hip, spec, precision)
    assert "L8" in g.describe() or "L4" in g.describe()     # several lanes per body for the big classes


@pytest.mark.parametrize("mode", [10, 0])
def test_single_body(hip, oracle, mode):
    compare(oracle, hip, clusters([5], 0, seed=3, mode=mode), "mixed")


@pytest.mark.parametrize("mode", [10, 0])
def test_bonds_couple_bodies_and_free_atoms(hip, oracle, mode):
    spec = clusters([3] * 40 + [6] * 10, 30, seed=4, mode=mode, spacing=0.9)
    n = spec.num_atoms
    rng = np.random.default_rng(5)
    bonds = []
    for _ in range(60):
        a, b = rng.choice(n, size=2, replace=False)
        if spec.body_indices[a] != spec.body_indices[b] or spec.body_indices[a] == 0:
            d = np.linalg.norm(spec.positions[a] - spec.positions[b])
            bonds.append((int(a), int(b), float(d * rng.uniform(0.9, 1.1)), 50.0))
    compare(oracle, hip, spec, "mixed", bonds=bonds)
    compare(oracle, hip, spec, "double", bonds=bonds)


@pytest.mark.parametrize("mode", [10, 0])
def test_start_at_rest(hip, oracle, mode):
    spec = clusters([3] * 30 + [2] * 20, 20, seed=6, mode=mode)
    spec.tether_r0 = spec.tether_r0 + 0.05     # displaced anchors: the bodies start to move
    compare(oracle, hip, spec, "mixed", start_at_rest=True)


def rotation_arg_step(I, L, dt):
    """|wp dt|: how far the exact free-rotor solution advances the Jacobi
    argument in one step (the exact rotation evaluates short advances, up to
    0.2, by series and quadrature and longer ones by the closed form)."""
    invI = 1.0 / I
    w0 = invI * L
    Lsq, twoKr = L @ L, L @ w0
    r1, r3 = Lsq - twoKr * I[2], twoKr * I[0] - Lsq
    l1, l3 = r1 * invI[1] / (I[1] - I[2]), r3 * invI[1] / (I[0] - I[1])
    c13 = 1.0 / (I[0] - I[2])
    a = np.sqrt([r1 * invI[0] * c13, min(l1, l3), r3 * invI[2] * c13])
    if abs(w0[1] / a[1]) >= 1:
        a[1] = abs(w0[1])
    return abs(invI[1] * a[0] * a[2] / (a[1] * c13)) * dt


@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_fast_rotors_take_both_exact_paths(hip, oracle, precision):
    spec = clusters([5] * 40 + [9] * 24, 10, seed=9, mode=0)
    rng = np.random.default_rng(9)
    # spin the bodies: rigid rotation at |w| up to 600 rad/ps on top of 300 K
    v = spec.velocities.copy()
    for b in range(1, 65):
        sel = spec.body_indices == b
        m = spec.masses[sel]
        com = (m[:, None] * spec.positions[sel]).sum(0) / m.sum()
        w = rng.normal(size=3)
        w *= rng.uniform(5.0, 600.0) / np.linalg.norm(w)
        v[sel] += np.cross(w, spec.positions[sel] - com)
    spec.velocities = v
    o = oracle.OracleSim.from_spec(spec)
    steps = []
    for st in o.body_states():
        a, b_, c, d = st["q"]
        Bq = np.array([[-b_, -c, -d], [a, -d, c], [d, a, -b_], [-c, b_, a]])
        steps.append(rotation_arg_step(st["I"], 0.5 * Bq.T @ st["pi"], spec.dt))
    steps = np.array(steps)
    assert (steps <= 0.2).sum() >= 8 and (steps > 0.2).sum() >= 8, steps
    compare(oracle, hip, spec, precision)


@pytest.mark.parametrize("mode", [0, 10])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_spinning_dimer_atom_velocities(hip, mode, precision):
    """Rigid-body kinematics alone, no oracle: a dimer spun about z at w keeps
    every atom at v_cm + w e_z x (r_j - c) (RigidBody.cpp:159-168), which a
    sign slip (delta x w) reverses."""
    m = np.array([3.0, 1.0])
    x = np.array([[1.00, 2.00, 3.00], [1.20, 2.00, 3.00]])
    c0 = (m[:, None] * x).sum(0) / m.sum()
    vcm, w, ez = np.array([0.1, -0.2, 0.3]), 5.0, np.array([0.0, 0.0, 1.0])
    g = hip.Context(m, np.array([1, 1], np.int32), 0.002, mode, precision)
    g.set_positions(x)
    g.set_velocities(vcm + w * np.cross(ez, x - c0))
    g.step(7)
    xg, vg = g.positions(), g.velocities()
    c = (m[:, None] * xg).sum(0) / m.sum()
    tol = 1e-5 if precision == "mixed" else 1e-9
    np.testing.assert_allclose(vg, vcm + w * np.cross(ez, xg - c), rtol=0, atol=tol)
    np.testing.assert_allclose(c, c0 + vcm * 0.014, rtol=0, atol=10 * tol)
