"""Distance constraints between free atoms on the GPU
(rbhip_create_constrained) against the FP64 oracle, whose SHAKE/RATTLE is
pinned in test_oracle_constraints.py.

At a tight constraint tolerance both converge to the constrained solution, so
the GPU matches the oracle to the precision tolerances (the iteration counts
may differ where a check sits on the tolerance).  Covered: chains, triangles,
branched clusters and pairs beside rigid bodies and unconstrained free atoms,
both rotation modes and precisions, fused and two-pass steps, graphs on and
off, the split part1/forces/part2 calls, the DOF count, and the refusals
(constraints on body atoms, clusters over 16 atoms, batches, refined
energies)."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu

TOL = {"mixed": dict(pos=1e-6, vel=1e-4), "double": dict(pos=1e-10, vel=1e-8)}


def constrained_spec(mode, seed=5):
    """Waters as rigid bodies plus 60 free atoms, 43 of them in clusters:
    chains of 4, triangles, branched 5-atom clusters, pairs and a
    tetrahedron; the constraint lengths are the starting distances."""
    spec = systems.water_box(27, 0.5, num_free=60, seed=seed, rotation_mode=mode)
    rng = np.random.default_rng(seed)
    spec.tether_r0 = spec.tether_r0 + 0.03 * rng.normal(size=spec.tether_r0.shape)
    free = np.flatnonzero(spec.body_indices == 0)
    x = spec.positions
    # pull each cluster's atoms close together so the bonds are short
    # (the last: a tetrahedron with one edge given twice, 7 constraints on 4
    # atoms, which takes the path for larger clusters)
    shapes = [[(0, 1), (1, 2), (2, 3)]] * 3 + [[(0, 1), (1, 2), (0, 2)]] * 3 + \
             [[(0, 1), (0, 2), (0, 3), (2, 3), (3, 4)]] * 2 + [[(0, 1)]] * 4 + \
             [[(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 1)]]
    cons, at = [], 0
    for sh in shapes:
        n = max(max(p) for p in sh) + 1
        atoms = free[at:at + n]
        at += n
        if len(sh) == 7:
            # a regular tetrahedron (a flat one makes SHAKE converge slowly)
            tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) * 0.05
            x[atoms] = x[atoms[0]] + tet - tet[0]
        else:
            x[atoms[1:]] = x[atoms[0]] + rng.normal(size=(n - 1, 3)) * 0.08
        for a, b in sh:
            cons.append((int(atoms[a]), int(atoms[b]), float(np.linalg.norm(x[atoms[a]] - x[atoms[b]]))))
    spec.positions = x
    return spec, cons


def pair(oracle, hip, spec, cons, precision, tol):
    o = oracle.OracleSim.from_spec(spec)
    for a, b, d in cons:
        o.add_constraint(a, b, d)
    o.set_constraint_tolerance(tol)
    o.set_positions(spec.positions)
    o.set_velocities(spec.velocities)
    g = hip.Context.from_spec(spec, precision=precision, constraints=cons)
    g.set_constraint_tolerance(tol)
    return o, g


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_constrained_steps_match_oracle(hip, oracle, mode, precision):
    spec, cons = constrained_spec(mode)
    o, g = pair(oracle, hip, spec, cons, precision, 1e-10)
    assert g.counts()["numDOF"] == o.counts()["numDOF"]
    for steps in (1, 9):
        o.step(steps)
        g.step(steps)
        t = TOL[precision]
        np.testing.assert_allclose(g.positions(), o.positions(), rtol=0, atol=t["pos"])
        vo = o.velocities()
        np.testing.assert_allclose(g.velocities(), vo, rtol=0, atol=t["vel"] * np.abs(vo).max())
        np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=t["vel"])
    x = g.positions()
    for a, b, d in cons:
        assert abs(np.linalg.norm(x[a] - x[b]) - d) <= 1e-8 * d
    assert g.constraint_failures() == 0 and o.constraint_failures() == 0


@pytest.mark.parametrize("graph,fusion", [(True, True), (False, True), (True, False), (False, False)])
def test_constrained_launch_paths_agree(hip, graph, fusion):
    spec, cons = constrained_spec(10, seed=6)
    ref = hip.Context.from_spec(spec, precision="mixed", constraints=cons)
    ref.use_graph(False)
    ref.use_fusion(False)
    g = hip.Context.from_spec(spec, precision="mixed", constraints=cons)
    g.use_graph(graph)
    g.use_fusion(fusion)
    for c in (ref, g):
        c.step(7)
    np.testing.assert_array_equal(g.positions(), ref.positions())
    np.testing.assert_array_equal(g.velocities(), ref.velocities())
    np.testing.assert_array_equal(g.kinetic_energies(), ref.kinetic_energies())


def test_constrained_split_calls(hip, oracle):
    spec, cons = constrained_spec(10, seed=7)
    o, g = pair(oracle, hip, spec, cons, "double", 1e-11)
    for _ in range(3):
        o.step(1)
        g.part1()
        g.compute_forces()
        g.part2()
    np.testing.assert_allclose(g.positions(), o.positions(), rtol=0, atol=1e-10)
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=1e-8)


def test_default_tolerance_holds_distances(hip):
    spec, cons = constrained_spec(10, seed=8)
    g = hip.Context.from_spec(spec, precision="mixed", constraints=cons)
    g.step(200)
    x = g.positions()
    for a, b, d in cons:
        assert abs(np.linalg.norm(x[a] - x[b]) - d) <= 2.5e-5 * d
    assert g.constraint_failures() == 0


def test_constraint_refusals(hip):
    spec, cons = constrained_spec(10, seed=9)
    body_atom = int(np.flatnonzero(spec.body_indices > 0)[0])
    free = np.flatnonzero(spec.body_indices == 0)
    with pytest.raises(hip.RbhipError, match="rigid-body atoms"):
        hip.Context.from_spec(spec, constraints=[(body_atom, int(free[0]), 0.1)])
    with pytest.raises(hip.RbhipError):
        hip.Context.from_spec(spec, constraints=[(int(free[0]), int(free[0]), 0.1)])
    chain = [(int(free[i]), int(free[i + 1]), 0.1) for i in range(16)]   # 17 atoms
    with pytest.raises(hip.RbhipError, match="at most 16"):
        hip.Context.from_spec(spec, constraints=chain)
    g = hip.Context.from_spec(spec, constraints=cons)
    with pytest.raises(hip.RbhipError, match="refined"):
        g.set_refined(True)
    with pytest.raises(hip.RbhipError, match="batched"):
        b = hip.Batch()
        b.set_members([g])
    with pytest.raises(hip.RbhipError):
        g.set_constraint_tolerance(0.0)
