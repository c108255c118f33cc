"""Seeded synthetic rigid-body systems (the workloads the integrator is measured on).

Units are OpenMM's: nm, ps, amu, kJ/mol.  Every builder returns a ``SystemSpec``
holding per-atom masses, the reference-style ``bodyIndices`` list (0 = free
atom, equal positive labels = one rigid body), positions, Maxwell-Boltzmann
velocities and a harmonic tether (k, r0) per atom.  Nothing here depends on
the GPU; the same spec feeds the HIP path and the CPU checker.

Pinned generator choices (the prose workload descriptions leave them open):
  * water body: O at the origin, two H at 0.09572 nm, H-O-H 104.52 deg,
    masses 15.9994 / 1.008 / 1.008; atoms stored O, H, H;
  * body centres (centre of mass) on a cubic lattice, cell centres at
    (i + 1/2) * spacing, jittered uniformly by +-jitter per axis;
  * orientation: normalised 4-D Gaussian -> uniform random unit quaternion;
  * velocities: per-atom Maxwell-Boltzmann, sigma = sqrt(kT / m),
    kT = 0.0083144626 * T kJ/mol;
  * tether: F_i = -k (r_i - r0_i) with r0 = initial positions.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KB = 0.0083144626  # kJ/mol/K
WATER_MASSES = np.array([15.9994, 1.008, 1.008])
WATER_OH = 0.09572
WATER_ANGLE_DEG = 104.52
ARGON_MASS = 39.948


@dataclass
class SystemSpec:
    masses: np.ndarray            # (n,) float64
    body_indices: np.ndarray      # (n,) int32, 0 = free atom
    positions: np.ndarray         # (n, 3) float64
    velocities: np.ndarray        # (n, 3) float64
    tether_k: np.ndarray          # (n,) float64
    tether_r0: np.ndarray         # (n, 3) float64
    dt: float = 0.002
    rotation_mode: int = 10
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def num_atoms(self) -> int:
        return int(self.masses.shape[0])

    @property
    def num_bodies(self) -> int:
        return int(np.unique(self.body_indices[self.body_indices > 0]).size)

    @property
    def num_free(self) -> int:
        return int(np.count_nonzero(self.body_indices <= 0))


def water_template() -> np.ndarray:
    """Water atom coordinates (O, H, H) relative to the centre of mass."""
    half = np.deg2rad(WATER_ANGLE_DEG / 2.0)
    x = np.array([[0.0, 0.0, 0.0],
                  [WATER_OH * np.sin(half), 0.0, WATER_OH * np.cos(half)],
                  [-WATER_OH * np.sin(half), 0.0, WATER_OH * np.cos(half)]])
    com = (WATER_MASSES[:, None] * x).sum(0) / WATER_MASSES.sum()
    return x - com


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """(n, 3, 3) rotation matrices of uniformly random unit quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def lattice_points(rng: np.random.Generator, count: int, side: int, spacing: float,
                   jitter: float) -> np.ndarray:
    """First ``count`` cells of a side^3 cubic lattice, jittered by +-jitter."""
    idx = np.arange(count)
    ijk = np.stack([idx // (side * side), (idx // side) % side, idx % side], axis=1)
    pts = (ijk + 0.5) * spacing
    return pts + rng.uniform(-jitter, jitter, size=pts.shape)


def maxwell_boltzmann(rng: np.random.Generator, masses: np.ndarray, temperature: float) -> np.ndarray:
    sigma = np.sqrt(KB * temperature / masses)
    return rng.standard_normal((masses.size, 3)) * sigma[:, None]


def water_box(num_bodies: int, spacing: float, num_free: int = 0, seed: int = 2024,
              temperature: float = 300.0, k: float = 1000.0, jitter: float = 0.02,
              dt: float = 0.002, rotation_mode: int = 10, free_spacing: float | None = None,
              name: str = "water") -> SystemSpec:
    """Water-like rigid bodies on a jittered cubic lattice plus optional argon free atoms.

    Atom order: the bodies' atoms first (O, H, H per body, body labels 1..B),
    then the free atoms (label 0) on their own lattice, offset by a quarter
    cell from the water lattice.
    """
    rng = np.random.default_rng(seed)
    side = int(np.ceil(num_bodies ** (1.0 / 3.0) - 1e-9))
    centres = lattice_points(rng, num_bodies, side, spacing, jitter)
    rot = random_rotations(rng, num_bodies)
    tmpl = water_template()
    body_pos = centres[:, None, :] + np.einsum("bij,aj->bai", rot, tmpl)
    pos = [body_pos.reshape(-1, 3)]
    masses = [np.tile(WATER_MASSES, num_bodies)]
    labels = [np.repeat(np.arange(1, num_bodies + 1, dtype=np.int32), 3)]
    if num_free:
        fside = int(np.ceil(num_free ** (1.0 / 3.0) - 1e-9))
        fsp = free_spacing if free_spacing is not None else side * spacing / fside
        fpos = lattice_points(rng, num_free, fside, fsp, 0.0) + 0.25 * fsp
        pos.append(fpos)
        masses.append(np.full(num_free, ARGON_MASS))
        labels.append(np.zeros(num_free, dtype=np.int32))
    positions = np.concatenate(pos)
    m = np.concatenate(masses)
    vel = maxwell_boltzmann(rng, m, temperature)
    return SystemSpec(masses=m, body_indices=np.concatenate(labels), positions=positions,
                      velocities=vel, tether_k=np.full(m.size, k), tether_r0=positions.copy(),
                      dt=dt, rotation_mode=rotation_mode, name=name,
                      meta=dict(seed=seed, spacing=spacing, jitter=jitter, temperature=temperature,
                                box=side * spacing))


def water_c1(rotation_mode: int = 10, seed: int = 1) -> SystemSpec:
    """512 water bodies (8^3, 0.3125 nm, 2.5 nm box) + 64 argon atoms, k = 1000, dt = 2 fs."""
    return water_box(512, 0.3125, num_free=64, seed=seed, rotation_mode=rotation_mode,
                     name="water512+ar64")


def water_c2(num_bodies: int = 1_000_000, seed: int = 2, rotation_mode: int = 10) -> SystemSpec:
    """1M water bodies on a 100^3 lattice (0.215 nm, 21.5 nm box), no free atoms."""
    return water_box(num_bodies, 0.215, num_free=0, seed=seed, rotation_mode=rotation_mode,
                     name=f"water{num_bodies}")


# ---------------------------------------------------------------------------
# heterogeneous bodies
# ---------------------------------------------------------------------------
def _ring(n: int, radius: float) -> np.ndarray:
    ang = 2 * np.pi * np.arange(n) / n
    return np.stack([radius * np.cos(ang), radius * np.sin(ang), np.zeros(n)], axis=1)


def heterogeneous(num_bodies: int = 200_000, num_free: int = 500_000, seed: int = 3,
                  temperature: float = 300.0, k: float = 1000.0, dt: float = 0.001,
                  rotation_mode: int = 0, spacing: float = 0.3, mass_range=(1.0, 20.0),
                  shuffle: bool = True, name: str = "hetero") -> SystemSpec:
    """Mixed rigid bodies with per-atom random masses U[mass_range] plus free atoms.

    Body mix by count: 30% linear 2-atom rotors (0.11 nm bond), 40% 3-site bent
    bodies (water geometry), 20% planar 6-atom rings (radius 0.14 nm), 10%
    20-atom clusters (atoms uniform in a sphere of radius 0.15 nm).  Body and
    free-atom centres share one jittered cubic lattice.  Body labels are
    distinct random integers in [1, 8 B] (so labels have gaps), and with
    ``shuffle`` the atom order is a random permutation, so body atoms are
    non-contiguous.
    """
    rng = np.random.default_rng(seed)
    counts = [int(round(num_bodies * f)) for f in (0.3, 0.4, 0.2)]
    counts.append(num_bodies - sum(counts))
    sizes = np.concatenate([np.full(c, s) for c, s in zip(counts, (2, 3, 6, 20))])
    rng.shuffle(sizes)
    units = num_bodies + num_free
    side = int(np.ceil(units ** (1.0 / 3.0) - 1e-9))
    centres = lattice_points(rng, units, side, spacing, 0.02)
    cell = rng.permutation(units)
    body_centres = centres[cell[:num_bodies]]
    free_pos = centres[cell[num_bodies:]]

    tmpl_bent = water_template()
    pos_list, mass_list, label_list = [], [], []
    labels = rng.choice(np.arange(1, 8 * max(num_bodies, 1) + 1), size=num_bodies, replace=False)
    for size in (2, 3, 6, 20):
        sel = np.nonzero(sizes == size)[0]
        nb = sel.size
        if nb == 0:
            continue
        if size == 2:
            local = np.array([[0.0, 0.0, -0.055], [0.0, 0.0, 0.055]])
            local = np.broadcast_to(local, (nb, 2, 3))
        elif size == 3:
            local = np.broadcast_to(tmpl_bent, (nb, 3, 3))
        elif size == 6:
            local = np.broadcast_to(_ring(6, 0.14), (nb, 6, 3))
        else:
            v = rng.standard_normal((nb, 20, 3))
            v /= np.linalg.norm(v, axis=2, keepdims=True)
            r = 0.15 * rng.uniform(0, 1, (nb, 20, 1)) ** (1.0 / 3.0)
            local = v * r
        m = rng.uniform(*mass_range, size=(nb, size))
        rot = random_rotations(rng, nb)
        world = np.einsum("bij,baj->bai", rot, local)
        com = (m[:, :, None] * world).sum(1) / m.sum(1)[:, None]
        world = world - com[:, None, :] + body_centres[sel][:, None, :]
        pos_list.append(world.reshape(-1, 3))
        mass_list.append(m.reshape(-1))
        label_list.append(np.repeat(labels[sel], size))
    if num_free:
        pos_list.append(free_pos)
        mass_list.append(rng.uniform(*mass_range, size=num_free))
        label_list.append(np.zeros(num_free, dtype=np.int64))
    positions = np.concatenate(pos_list)
    masses = np.concatenate(mass_list)
    body_idx = np.concatenate(label_list).astype(np.int32)
    if shuffle:
        perm = rng.permutation(positions.shape[0])
        positions, masses, body_idx = positions[perm], masses[perm], body_idx[perm]
    vel = maxwell_boltzmann(rng, masses, temperature)
    return SystemSpec(masses=masses, body_indices=body_idx, positions=positions, velocities=vel,
                      tether_k=np.full(masses.size, k), tether_r0=positions.copy(), dt=dt,
                      rotation_mode=rotation_mode, name=name,
                      meta=dict(seed=seed, counts=counts, spacing=spacing, box=side * spacing))


def replica(units: int, seed: int, k: float, dt: float = 0.002, rotation_mode: int = 10,
            body_fraction: float = 0.8) -> SystemSpec:
    """One batched replica: ``units`` = bodies + free atoms, 80% water bodies, 20% argon."""
    nb = int(round(units * body_fraction))
    nf = units - nb
    spacing = 0.3125
    return water_box(nb, spacing, num_free=nf, seed=seed, k=k, dt=dt, rotation_mode=rotation_mode,
                     name=f"replica{units}")


def replica_set(count: int = 256, lo: float = 1e3, hi: float = 2e5, seed: int = 4,
                dt: float = 0.002, rotation_mode: int = 10, scale: float = 1.0):
    """Replicas with units ~ LogUniform(lo, hi) and k ~ U[500, 2000] (scaled by ``scale``)."""
    rng = np.random.default_rng(seed)
    units = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)) * scale
    ks = rng.uniform(500.0, 2000.0, size=count)
    seeds = rng.integers(0, 2**31 - 1, size=count)
    return [replica(max(int(u), 8), int(s), float(kk), dt=dt, rotation_mode=rotation_mode)
            for u, kk, s in zip(units, ks, seeds)]
