"""Virtual sites and massless particles in the oracle.

The reference places OpenMM's virtual sites after part 1 and the constraints
(ReferenceRigidBodyKernels.cpp:100) and leaves massless particles and sites
out of the free atoms and the bodies (RigidBodySystem.cpp:65-76); OpenMM
moves a site's force to its parents.  Pinned by:
- the site definitions (two- and three-particle averages, out of plane,
  local coordinates) after every step;
- the redistributed forces against central differences of the potential
  energy with the site following its parents (a wrong sign or a transposed
  cross product fails);
- energy conservation of a rigid water whose massless fourth site carries a
  tether, with the O(dt^2) fluctuation of velocity Verlet;
- bookkeeping: numActualAtoms excludes sites, a massless particle that is no
  site stays where it is, labelled massless particles are refused.
"""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd.systems import water_template

O_MASS, H_MASS = 15.9994, 1.008


def site_position(kind, r, w):
    r = np.asarray(r)
    if kind == 1:
        return w[0] * r[0] + w[1] * r[1]
    if kind == 2:
        return w[0] * r[0] + w[1] * r[1] + w[2] * r[2]
    if kind == 3:
        r12, r13 = r[1] - r[0], r[2] - r[0]
        return r[0] + w[0] * r12 + w[1] * r13 + w[2] * np.cross(r12, r13)
    w = np.asarray(w)
    o, x, y = w[0:3] @ r, w[3:6] @ r, w[6:9] @ r
    ex = x / np.linalg.norm(x)
    z = np.cross(x, y)
    ez = z / np.linalg.norm(z)
    ey = np.cross(ez, ex)
    return o + w[9] * ex + w[10] * ey + w[11] * ez


# two- and three-particle averages, out of plane, local coordinates
SITES = [(1, [0.3, 0.7]), (2, [0.6, 0.25, 0.15]), (3, [0.2, 0.35, 2.5]),
         (4, [0.5, 0.25, 0.25, -1.0, 1.0, 0.0, -1.0, 0.0, 1.0, 0.02, 0.03, 0.015])]
NS = len(SITES)


def water_with_sites(oracle, rigid=True, tether_site=True, dt=0.001, seed=1):
    """Two waters (O, H, H) plus one site of each kind on the first water
    (particles 6 ..) and a massless particle that is not a site (the last)."""
    rng = np.random.default_rng(seed)
    t = water_template()
    x = np.concatenate([t, t + [0.4, 0.0, 0.0]])
    masses = [O_MASS, H_MASS, H_MASS] * 2
    labels = [1, 1, 1, 2, 2, 2] if rigid else [0] * 6
    for kind, w in SITES:
        x = np.vstack([x, site_position(kind, x[:3], w)])
        masses.append(0.0)
        labels.append(0)
    x = np.vstack([x, [[1.0, 1.0, 1.0]]])   # massless, no site
    masses.append(0.0)
    labels.append(0)
    n = len(masses)
    o = oracle.OracleSim(np.array(masses), np.array(labels, np.int32), dt, 10)
    for s, (kind, w) in enumerate(SITES):
        o.add_virtual_site(kind, 6 + s, [0, 1, 2][:2 if kind == 1 else 3], w)
    k = np.full(n, 200.0)
    if not tether_site:
        k[6:6 + NS] = 0.0
    o.set_tether(k, x + rng.normal(size=x.shape) * 0.02)
    o.set_positions(x)
    v = rng.normal(size=x.shape) * 0.3
    v[6:] = 0.0
    o.set_velocities(v)
    return o, x


def test_sites_follow_their_definitions(oracle):
    o, _ = water_with_sites(oracle)
    for _ in range(5):
        o.step(3)
        x = o.positions()
        for s, (kind, w) in enumerate(SITES):
            np.testing.assert_allclose(x[6 + s], site_position(kind, x[:3], w), rtol=0, atol=1e-14)
        np.testing.assert_array_equal(x[-1], [1.0, 1.0, 1.0])


@pytest.mark.parametrize("kind,w", SITES)
def test_site_forces_are_the_energy_gradient(oracle, kind, w):
    rng = np.random.default_rng(kind)
    x = rng.normal(size=(4, 3)) * 0.1
    x[3] = site_position(kind, x[:3], w)
    o = oracle.OracleSim(np.array([2.0, 3.0, 5.0, 0.0]), np.zeros(4, np.int32), 0.001, 10)
    o.add_virtual_site(kind, 3, [0, 1, 2][:2 if kind == 1 else 3], w)
    r0 = rng.normal(size=(4, 3)) * 0.2
    o.set_tether(np.array([0.0, 0.0, 0.0, 300.0]), r0)   # only the site is tethered

    def energy(y):
        y = y.copy()
        y[3] = site_position(kind, y[:3], w)
        o.set_positions(y)
        return o.potential_energy()

    o.set_positions(x)
    f = o.forces()
    np.testing.assert_array_equal(f[3], [0.0, 0.0, 0.0])
    h = 1e-6
    for p in range(3):
        for c in range(3):
            d = np.zeros((4, 3))
            d[p, c] = h
            num = -(energy(x + d) - energy(x - d)) / (2 * h)
            assert abs(f[p, c] - num) <= 1e-7 * max(1.0, abs(num)), (p, c, f[p, c], num)


def test_energy_conservation_with_a_tethered_site(oracle):
    spread = []
    for dt in (0.002, 0.001):
        o, _ = water_with_sites(oracle, dt=dt)
        e = []
        for _ in range(int(round(0.4 / dt))):
            o.step(1)
            e.append(o.kinetic_energies().sum() + o.potential_energy())
        spread.append(np.ptp(e))
    assert spread[1] < 0.01 * abs(np.mean(e))
    assert 3.0 < spread[0] / spread[1] < 5.0


def test_site_bookkeeping(oracle):
    o, _ = water_with_sites(oracle)
    c = o.counts()
    assert c["numActualAtoms"] == (6 + NS + 1) - NS     # every particle but the sites
    assert c["numFree"] == 0 and c["numBodies"] == 2
    free, _ = water_with_sites(oracle, rigid=False)
    assert free.counts()["numFree"] == 6
    with pytest.raises(RuntimeError):
        oracle.OracleSim(np.array([1.0, 0.0]), np.array([1, 1], np.int32), 0.001, 10)
    o2 = oracle.OracleSim(np.array([1.0, 1.0, 0.0]), np.zeros(3, np.int32), 0.001, 10)
    with pytest.raises(RuntimeError):
        o2.add_virtual_site(1, 0, [1, 2], [0.5, 0.5])     # a massive particle as a site
    with pytest.raises(RuntimeError):
        o2.add_constraint(0, 2, 0.1)                    # a massless particle in a constraint
