"""Full-size workloads on the GPU: the 1M-body water box (NO-SQUISH, mixed)
and the 200k-body heterogeneous mix with 500k free atoms and permuted body
indices (exact rotation, mixed).

Parity at full size: each body feels only its own atoms' tethers, so a random
sample of bodies stepped alone by the FP64 oracle must match the same bodies
inside the full GPU run, at the mixed tolerances of tests/tolerances.py
(positions <= 1e-6 nm; each body's momentum and angular momentum within 1e-5
of its own magnitude above the floor the force field's fp32 rounding sets,
eight times larger in the 21.5 nm box than in the 2.5 nm one).
Size-independent invariants over every body: rigidity, |q| = 1, and
1/2 sum m v^2 = KE_trans + KE_rot."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from subsample import sample_atoms, subsystem
from tolerances import check_state

pytestmark = pytest.mark.gpu


def run_and_check(hip, oracle, spec, steps, nsample_bodies, nsample_free):
    g = hip.Context.from_spec(spec, precision="mixed")
    g.step(steps)
    x, v = g.positions(), g.velocities()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    # kinetic-energy identity over all atoms (velocities are rigid after a step)
    ke = g.kinetic_energies()
    assert 0.5 * np.sum(spec.masses[:, None] * v * v) == pytest.approx(ke.sum(), rel=1e-6)
    # rigidity of every body: distances to the body's first atom are preserved
    bi = spec.body_indices
    order = np.argsort(bi, kind="stable")
    sb = bi[order]
    starts = np.r_[0, np.nonzero(np.diff(sb))[0] + 1]
    first = np.repeat(order[starts], np.diff(np.r_[starts, sb.size]))
    body = sb > 0
    d0 = np.linalg.norm(spec.positions[order] - spec.positions[first], axis=1)[body]
    d1 = np.linalg.norm(x[order] - x[first], axis=1)[body]
    assert np.abs(d0 - d1).max() < 2e-6
    # sampled parity against the FP64 oracle
    atoms = sample_atoms(spec, nsample_bodies, nsample_free, seed=3)
    sub = subsystem(spec, atoms)
    o = oracle.OracleSim.from_spec(sub)
    o.step(steps)
    # (the floors from the whole system's largest coordinate, at which the
    # GPU's fp32 values are rounded)
    check_state(sub, o.positions(), o.velocities(), x[atoms], v[atoms], "mixed", steps,
                coord_scale=float(np.abs(spec.positions).max()))
    return g


def test_million_water_bodies_nosquish(hip, oracle):
    spec = systems.water_c2()
    g = run_and_check(hip, oracle, spec, 10, 3000, 0)
    q = np.array([b["q"] for b in g.body_states()[:: 997]])
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)


def test_heterogeneous_200k_exact(hip, oracle):
    spec = systems.heterogeneous(200_000, 500_000, seed=3, rotation_mode=0)
    run_and_check(hip, oracle, spec, 10, 2000, 2000)
