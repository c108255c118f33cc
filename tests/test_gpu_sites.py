"""Virtual sites and massless particles on the GPU (rbhip_set_virtual_sites)
against the oracle, whose sites are pinned in test_oracle_sites.py: rigid
waters carrying a tethered massless site of each kind (two- and
three-particle averages, out of plane, local coordinates), free atoms with
sites, a massless particle that is no site, with and without constraints;
both rotation modes and precisions, fused and two-pass steps, graphs on and
off."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from test_oracle_sites import SITES as KINDS, site_position   # the four site kinds

pytestmark = pytest.mark.gpu

TOL = {"mixed": dict(pos=2e-6, vel=1e-4), "double": dict(pos=1e-10, vel=1e-8)}


def site_spec(mode, seed=3, free_sites=True):
    """27 waters and 12 argon atoms; one massless site per water (the four
    kinds in turn), sites on triples of free atoms, and one massless particle
    that is no site.  Returns the spec and the site list."""
    base = systems.water_box(27, 0.5, num_free=12, seed=seed, rotation_mode=mode)
    rng = np.random.default_rng(seed)
    x = list(base.positions)
    m = list(base.masses)
    lab = list(base.body_indices)
    sites = []
    owners = [(3 * b, [3 * b, 3 * b + 1, 3 * b + 2]) for b in range(27)]
    free = list(np.flatnonzero(base.body_indices == 0))
    if free_sites:
        owners += [(None, free[i:i + 3]) for i in range(0, 12, 3)]
    for i, (_, par) in enumerate(owners):
        # local-coordinates sites on the waters only (the argon triples are collinear)
        kind, w = KINDS[i % len(KINDS)] if i < 27 else KINDS[i % 3]
        par = par[:2] if kind == 1 else par
        site = len(x)
        x.append(site_position(kind, [base.positions[p] for p in par], w))
        m.append(0.0)
        lab.append(0)
        sites.append((kind, site, [int(p) for p in par], w))
    x.append(np.array([0.1, 0.2, 0.3]))
    m.append(0.0)
    lab.append(0)
    n = len(x)
    x = np.array(x)
    v = np.zeros_like(x)
    v[:base.num_atoms] = base.velocities
    spec = systems.SystemSpec(masses=np.array(m), body_indices=np.array(lab, np.int32), positions=x,
                              velocities=v, tether_k=np.full(n, 800.0),
                              tether_r0=x + 0.02 * rng.normal(size=x.shape), dt=base.dt, rotation_mode=mode)
    return spec, sites


def pair(oracle, hip, spec, sites, precision, constraints=None):
    o = oracle.OracleSim.from_spec(spec)
    for kind, site, par, w in sites:
        o.add_virtual_site(kind, site, par, w)
    for a, b, d in constraints or []:
        o.add_constraint(a, b, d)
    if constraints:
        o.set_constraint_tolerance(1e-10)
    o.set_positions(spec.positions)
    o.set_velocities(spec.velocities)
    g = hip.Context(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode, precision, constraints=constraints)
    g.set_tether(spec.tether_k, spec.tether_r0)
    g.set_virtual_sites(sites)
    if constraints:
        g.set_constraint_tolerance(1e-10)
    g.set_positions(spec.positions)
    g.set_velocities(spec.velocities)
    return o, g


def check(o, g, precision):
    t = TOL[precision]
    np.testing.assert_allclose(g.positions(), o.positions(), rtol=0, atol=t["pos"])
    vo = o.velocities()
    np.testing.assert_allclose(g.velocities(), vo, rtol=0, atol=t["vel"] * np.abs(vo).max())
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=t["vel"])


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_sites_match_oracle(hip, oracle, mode, precision):
    spec, sites = site_spec(mode)
    o, g = pair(oracle, hip, spec, sites, precision)
    assert g.counts() == o.counts()
    np.testing.assert_array_equal(g.atom_index(), o.atom_index())
    np.testing.assert_allclose(g.forces(), o.forces(), rtol=0, atol=1e-3 if precision == "mixed" else 1e-9)
    for steps in (1, 9):
        o.step(steps)
        g.step(steps)
        check(o, g, precision)
    x = g.positions()
    for kind, site, par, w in sites:
        np.testing.assert_allclose(x[site], site_position(kind, x[par], w), rtol=0, atol=TOL[precision]["pos"])
    # the massless particle that is no site stays where it was (mixed
    # precision keeps positions as fp32 hi + lo)
    np.testing.assert_allclose(x[-1], spec.positions[-1], rtol=0, atol=1e-12)


def test_sites_with_constraints(hip, oracle):
    spec, sites = site_spec(10, seed=4)
    free = np.flatnonzero((spec.body_indices == 0) & (spec.masses > 0))
    x = spec.positions
    cons = [(int(free[i]), int(free[i + 1]), float(np.linalg.norm(x[free[i]] - x[free[i + 1]]))) for i in (0, 3, 6)]
    o, g = pair(oracle, hip, spec, sites, "double", cons)
    o.step(10)
    g.step(10)
    check(o, g, "double")


@pytest.mark.parametrize("graph,fusion", [(True, True), (False, True), (True, False)])
def test_site_launch_paths_agree(hip, graph, fusion):
    spec, sites = site_spec(10, seed=5)

    def run(graph, fusion):
        g = hip.Context.from_spec(spec, precision="mixed", set_state=False)
        g.set_virtual_sites(sites)
        g.set_positions(spec.positions)
        g.set_velocities(spec.velocities)
        g.use_graph(graph)
        g.use_fusion(fusion)
        g.step(7)
        return g.positions(), g.velocities()

    ref = run(False, False)
    got = run(graph, fusion)
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


def test_site_refusals(hip):
    spec, sites = site_spec(10, seed=6)
    with pytest.raises(hip.RbhipError, match="rigid bodies"):
        lab = spec.body_indices.copy()
        lab[-1] = 1
        hip.Context(spec.masses, lab, spec.dt)
    g = hip.Context.from_spec(spec, set_state=False)
    with pytest.raises(hip.RbhipError, match="massless"):
        g.set_virtual_sites([(1, 0, [1, 2], [0.5, 0.5])])          # a massive particle as a site
    with pytest.raises(hip.RbhipError, match="massive"):
        g.set_virtual_sites([(1, sites[0][1], [sites[1][1], 0], [0.5, 0.5])])   # a massless parent
    g.set_virtual_sites(sites)
    g.set_positions(spec.positions)
    g.set_velocities(spec.velocities)
    with pytest.raises(hip.RbhipError, match="refined"):
        g.set_refined(True)
    with pytest.raises(hip.RbhipError, match="batched"):
        b = hip.Batch()
        b.set_members([g])
