"""Per-result tolerances of the GPU-vs-oracle comparisons after a few steps.

Mixed precision (fp32 state, FP64 accumulation):
  - positions <= 1e-6 nm;
  - each body's momentum and space-frame angular momentum within 1e-5 of its
    own magnitude, plus a floor: the most the force field's fp32 rounding can
    move it in `steps` steps.  The device force field reads fp32 positions
    and holds its tether anchors at fp32, each off by up to half an ulp h of
    the largest coordinate, so an atom's force is off by up to
    k sqrt(3) 2h; over the steps a body's momentum by steps dt sum_j dF_j and
    its angular momentum by steps dt sum_j |r_j - c| dF_j.  In the 2.5 nm
    box h = 1.2e-7 nm and the momentum floor of a water is 2e-6 of its
    typical momentum; in the 21.5 nm box h = 9.5e-7 nm, eight times that;
  - free-atom velocities likewise, with floor steps dt dF / m;
  - kinetic energies <= 1e-5 relative.
Double precision: positions <= 1e-10 nm; momenta within 1e-9 relative above a
floor of 1e-10 of the RMS over the bodies; kinetic energies <= 1e-11.
"""
import numpy as np

from subsample import body_frames

TOL = {"mixed": dict(x=1e-6, p=1e-5, ke=1e-5), "double": dict(x=1e-10, p=1e-9, floor=1e-10, ke=1e-11)}


def half_ulp32(x):
    """Half an fp32 ulp at the largest magnitude in x."""
    m = float(np.abs(x).max())
    return 2.0 ** (np.floor(np.log2(m)) - 24) if m > 0 else 0.0


def force_field_floors(spec, steps, coord_scale=0.0):
    """Per body label (momentum floor, angular-momentum floor) and per free
    atom velocity floors, from the fp32 rounding of positions and anchors
    (coordinates up to coord_scale, or the spec's own largest)."""
    h = max(half_ulp32(spec.positions), half_ulp32(spec.tether_r0), half_ulp32(np.array([coord_scale])))
    df = np.asarray(spec.tether_k, dtype=np.float64) * np.sqrt(3.0) * 2.0 * h
    t = steps * spec.dt
    bodies = {}
    bi = spec.body_indices
    for b in np.unique(bi[bi > 0]):
        idx = np.nonzero(bi == b)[0]
        m = spec.masses[idx]
        c = (m[:, None] * spec.positions[idx]).sum(0) / m.sum()
        lever = np.linalg.norm(spec.positions[idx] - c, axis=1)
        bodies[int(b)] = (t * df[idx].sum(), t * (lever * df[idx]).sum())
    free = bi <= 0
    vfloor = t * df[free] / np.where(spec.masses[free] > 0, spec.masses[free], np.inf)
    return bodies, vfloor


def ratio_to_bound(ref, got, rel, floor):
    """max_i |got_i - ref_i| / (rel |ref_i| + floor_i)"""
    bound = rel * np.linalg.norm(ref, axis=1) + floor
    return float((np.linalg.norm(got - ref, axis=1) / bound).max())


def check_state(spec, xo, vo, xg, vg, precision, steps, coord_scale=0.0):
    """Positions, and per body momentum and angular momentum, and free-atom
    velocities of the GPU (xg, vg) against the oracle (xo, vo); returns the
    largest ratios of the errors to their bounds.  coord_scale: the largest
    coordinate of the system the GPU ran, when spec is a sample of it."""
    tol = TOL[precision]
    assert np.all(np.isfinite(xg)) and np.all(np.isfinite(vg))
    dx = np.abs(xo - xg).max()
    assert dx <= tol["x"], dx
    fo = body_frames(spec.masses, spec.body_indices, xo, vo)
    fg = body_frames(spec.masses, spec.body_indices, xg, vg)
    labels = list(fo)
    free = spec.body_indices <= 0
    ratios = {}
    if labels:
        P = np.array([fo[b][1] for b in labels]); Pg = np.array([fg[b][1] for b in labels])
        L = np.array([fo[b][2] for b in labels]); Lg = np.array([fg[b][2] for b in labels])
        if precision == "mixed":
            floors, _ = force_field_floors(spec, steps, coord_scale)
            fp = np.array([floors[b][0] for b in labels])
            fl = np.array([floors[b][1] for b in labels])
        else:
            fp = np.full(len(labels), tol["floor"] * np.sqrt((P ** 2).sum(1).mean()))
            fl = np.full(len(labels), tol["floor"] * np.sqrt((L ** 2).sum(1).mean()))
        ratios["p"] = ratio_to_bound(P, Pg, tol["p"], fp)
        ratios["L"] = ratio_to_bound(L, Lg, tol["p"], fl)
    if free.any():
        if precision == "mixed":
            _, fv = force_field_floors(spec, steps, coord_scale)
        else:
            fv = tol["floor"] * np.sqrt((vo[free] ** 2).sum(1).mean())
        ratios["v"] = ratio_to_bound(vo[free], vg[free], tol["p"], fv)
    print("bound ratios: " + ", ".join(f"{k} {v:.3g}" for k, v in ratios.items()))
    for k, v in ratios.items():
        assert v <= 1.0, f"{k} error {v:.3g} x its per-item bound"
    return ratios
