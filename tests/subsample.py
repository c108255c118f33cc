"""Sub-systems of a synthetic spec for full-size parity checks.

Every force in these workloads is a per-atom tether, so a body (or free atom)
evolves exactly as it would alone: the FP64 oracle can step a sample of the
bodies of a 1M-body system and be compared atom by atom with the GPU's run of
the whole system."""
import numpy as np

from openmm_rigidbody_plugin_amd.systems import SystemSpec


def sample_atoms(spec, num_bodies, num_free, seed=0):
    """Original atom indices of `num_bodies` random bodies and `num_free` free atoms."""
    rng = np.random.default_rng(seed)
    bi = spec.body_indices
    labels = np.unique(bi[bi > 0])
    pick = rng.choice(labels, size=min(num_bodies, labels.size), replace=False)
    body_atoms = np.nonzero(np.isin(bi, pick))[0]
    free = np.nonzero(bi <= 0)[0]
    free_pick = rng.choice(free, size=min(num_free, free.size), replace=False) if free.size else free
    return np.sort(np.concatenate([body_atoms, free_pick]))


def subsystem(spec, atoms):
    return SystemSpec(masses=spec.masses[atoms], body_indices=spec.body_indices[atoms],
                      positions=spec.positions[atoms], velocities=spec.velocities[atoms],
                      tether_k=spec.tether_k[atoms], tether_r0=spec.tether_r0[atoms], dt=spec.dt,
                      rotation_mode=spec.rotation_mode, name=spec.name + "-sample")


def body_frames(masses, body_indices, x, v):
    """Per body: centre of mass, total momentum and space-frame angular momentum
    about the centre of mass (frame-invariant observables)."""
    out = {}
    for b in np.unique(body_indices[body_indices > 0]):
        idx = np.nonzero(body_indices == b)[0]
        m = masses[idx]
        M = m.sum()
        r = (m[:, None] * x[idx]).sum(0) / M
        p = (m[:, None] * v[idx]).sum(0)
        L = (m[:, None] * np.cross(x[idx] - r, v[idx] - p / M)).sum(0)
        out[int(b)] = (r, p, L)
    return out
