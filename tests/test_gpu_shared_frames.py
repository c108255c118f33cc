"""Bodies of one shape reading shared body-fixed coordinates
(rbhip_set_shared_frames, mixed precision).  Water bodies match a few shapes
(the frame's axis signs), the tiles of such a class read the shared rows, and
every body moves by at most four fp32 ulps of its size, so trajectories stay
within 1e-7 nm of the unshared ones and within the mixed tolerances of the
FP64 oracle.  Bodies of random masses have shapes of their own and keep their
rows.  A class with a minority of one-shape bodies keeps its rows but the
matched bodies still take the shared values, and the step passes read the
same values whichever rows they use."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from openmm_rigidbody_plugin_amd.systems import SystemSpec

pytestmark = pytest.mark.gpu


def shapes(ctx):
    d = ctx.describe()
    return d[d.index("sharedShapes=[") + 14:d.index("]", d.index("sharedShapes=["))] if "sharedShapes=[" in d else ""


@pytest.mark.parametrize("mode", [10, 0])
def test_water_bodies_share_frames_and_match_the_oracle(hip, oracle, mode):
    spec = systems.water_c1(rotation_mode=mode)
    a = hip.Context.from_spec(spec, precision="mixed")
    b = hip.Context.from_spec(spec, precision="mixed", shared_frames=False)
    assert shapes(a).startswith("N3:") and 1 <= int(shapes(a)[3:]) <= 8
    assert shapes(b) == ""
    # snapped coordinates: within four fp32 ulps of the body's size
    da, db = a.body_fixed_positions(), b.body_fixed_positions()
    assert 0 < np.abs(da - db).max() <= 4 * 2.0 ** -24 * np.abs(db).max()
    for c in (a, b):
        c.step(20)
    o = oracle.OracleSim.from_spec(spec)
    o.step(20)
    assert np.abs(a.positions() - b.positions()).max() <= 1e-7
    assert np.abs(a.positions() - o.positions()).max() <= 1e-6
    np.testing.assert_allclose(a.kinetic_energies(), o.kinetic_energies(), rtol=1e-5)


def test_fused_and_two_pass_agree_with_shared_frames(hip):
    spec = systems.water_box(3000, 0.3, num_free=500, seed=9, rotation_mode=10)
    a = hip.Context.from_spec(spec, precision="mixed")
    b = hip.Context.from_spec(spec, precision="mixed")
    b.use_fusion(False)
    b.use_graph(False)
    a.step(15)
    b.step(15)
    assert shapes(a).startswith("N3:")
    np.testing.assert_array_equal(a.positions(), b.positions())
    np.testing.assert_array_equal(a.velocities(), b.velocities())


def test_random_masses_keep_their_own_rows(hip):
    spec = systems.heterogeneous(3000, 1000, seed=10, rotation_mode=10)
    c = hip.Context.from_spec(spec, precision="mixed")
    assert shapes(c) == ""


def test_double_precision_never_shares(hip):
    c = hip.Context.from_spec(systems.water_c1(), precision="double")
    assert shapes(c) == ""


def test_minority_shape_keeps_own_rows_with_the_same_trajectory(hip):
    # 200 water bodies among 600 3-atom bodies of random masses: the class's
    # first bodies give the candidates (waters first in the body order), but
    # fewer than half of the class match, so its tiles read their own rows
    w = systems.water_box(200, 0.4, seed=11, rotation_mode=10)
    h = systems.heterogeneous(1500, 0, seed=12, rotation_mode=10)
    keep = np.isin(h.body_indices, np.unique(h.body_indices)[np.bincount(h.body_indices)[np.unique(h.body_indices)] == 3])
    hb = h.body_indices[keep]
    _, hb = np.unique(hb, return_inverse=True)
    hb = hb.astype(np.int32) + 201
    spec = SystemSpec(masses=np.concatenate([w.masses, h.masses[keep]]),
                      body_indices=np.concatenate([w.body_indices, hb]),
                      positions=np.concatenate([w.positions, h.positions[keep] + 10.0]),
                      velocities=np.concatenate([w.velocities, h.velocities[keep]]),
                      tether_k=np.full(w.masses.size + int(keep.sum()), 1000.0),
                      tether_r0=np.concatenate([w.positions, h.positions[keep] + 10.0]),
                      dt=0.002, rotation_mode=10)
    a = hip.Context.from_spec(spec, precision="mixed")
    b = hip.Context.from_spec(spec, precision="mixed", shared_frames=False)
    assert shapes(a) == ""
    a.step(10)
    b.step(10)
    assert np.abs(a.positions() - b.positions()).max() <= 1e-7


@pytest.mark.parametrize("heavy", ["alternate", "last"])
def test_equal_geometry_other_masses_keep_their_inertia(hip, oracle, heavy):
    """Bodies of one geometry whose masses are all scaled (their body-fixed
    coordinates are unchanged, their 1/I and 1/M are not) never take another
    body's inertia: the step passes of a sharing tile read 1/I and 1/M at the
    body they share rows with, so matching compares the inertia too.  With
    every second water twice as heavy, or the last 112 of 512, the
    trajectory stays with the unshared one and the FP64 oracle's; in the
    second case the light majority (whose bodies give the candidate shapes)
    shares, and the heavy bodies sit in sharing tiles."""
    spec = systems.water_c1(rotation_mode=10)
    bi = spec.body_indices
    sel = (bi % 2 == 0) if heavy == "alternate" else (bi > 400)
    spec.masses = np.where((bi > 0) & sel, 2.0 * spec.masses, spec.masses)
    a = hip.Context.from_spec(spec, precision="mixed")
    b = hip.Context.from_spec(spec, precision="mixed", shared_frames=False)
    if heavy == "last":
        assert shapes(a).startswith("N3:")
    for c in (a, b):
        c.step(20)
    o = oracle.OracleSim.from_spec(spec)
    o.step(20)
    assert np.abs(a.positions() - b.positions()).max() <= 1e-7
    assert np.abs(a.positions() - o.positions()).max() <= 1e-6
    np.testing.assert_allclose(a.kinetic_energies(), o.kinetic_energies(), rtol=1e-5)
