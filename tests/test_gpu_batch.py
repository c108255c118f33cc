"""Batched systems on the GPU (rbhip_batch_*): several independent systems of
uneven sizes stepped by one launch per phase over all their tiles, replayed
from a HIP graph.

A member's bodies go through the same kernels and tiles in the batch as alone,
so with the same launch shape its trajectory is bit-for-bit the one it has
stepped alone; after the batch autotuner changes tiles (lanes per body, grid
cap) the reductions run in another order and results agree to rounding."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu


def member_specs(mode):
    specs = [systems.replica(u, seed=s, k=k, rotation_mode=mode)
             for u, s, k in ((40, 1, 700.0), (1500, 2, 1200.0), (333, 3, 1900.0), (4000, 4, 500.0))]
    specs.append(systems.heterogeneous(600, 300, seed=8, rotation_mode=mode))
    return specs


def make(hip, specs, block):
    ctxs = [hip.Context.from_spec(s, precision="mixed") for s in specs]
    for c in ctxs:
        c.set_launch(block)
    return ctxs


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("graph,fusion", [(True, True), (False, True), (True, False)])
def test_batch_members_equal_alone(hip, mode, graph, fusion):
    block = 256 if mode else 64
    specs = member_specs(mode)
    alone = make(hip, specs, block)
    members = make(hip, specs, block)
    batch = hip.Batch()
    batch.set_launch(block)
    batch.set_members(members)
    batch.use_graph(graph)
    batch.use_fusion(fusion)
    for c in alone:
        c.use_graph(graph)
        c.use_fusion(fusion)
    for n in (9, 1, 5):
        batch.step(n)
        for c in alone:
            c.step(n)
    ke = batch.kinetic_energies()
    for i, (a, m) in enumerate(zip(alone, members)):
        np.testing.assert_array_equal(m.positions(), a.positions())
        np.testing.assert_array_equal(m.velocities(), a.velocities())
        np.testing.assert_array_equal(ke[i], a.kinetic_energies())
        np.testing.assert_array_equal(m.kinetic_energies(), a.kinetic_energies())
        assert m.potential_energy() == a.potential_energy()
        assert m.time() == a.time()


def test_batch_interleaves_with_member_calls(hip):
    """Batch steps, a member's own steps, a step-size change and queries stay
    ordered: the result equals the same sequence run alone."""
    specs = member_specs(10)[:3]
    alone = make(hip, specs, 256)
    members = make(hip, specs, 256)
    batch = hip.Batch()
    batch.set_members(members)
    batch.step(4)
    members[1].step(3)
    members[0].set_step_size(0.0015)
    batch.step(6)
    for i, c in enumerate(alone):
        c.step(4)
        if i == 1:
            c.step(3)
        if i == 0:
            c.set_step_size(0.0015)
        c.step(6)
    for a, m in zip(alone, members):
        np.testing.assert_array_equal(m.positions(), a.positions())
        np.testing.assert_array_equal(m.velocities(), a.velocities())


def test_batch_autotune_keeps_state_and_accuracy(hip, oracle):
    specs = member_specs(10)
    members = make(hip, specs, 256)
    alone = make(hip, specs, 256)
    batch = hip.Batch()
    batch.set_members(members)
    x0 = [m.positions() for m in members]
    batch.autotune(reps=2)
    desc = batch.describe()
    assert "block=" in desc and "maxBlocks=" in desc
    for m, x in zip(members, x0):
        np.testing.assert_array_equal(m.positions(), x)   # tuning restored the state
    batch.step(10)
    for c in alone:
        c.step(10)
    for a, m in zip(alone, members):
        assert np.abs(m.positions() - a.positions()).max() <= 1e-6
        np.testing.assert_allclose(m.kinetic_energies(), a.kinetic_energies(), rtol=1e-6)
    # and against the FP64 oracle for one member
    o = oracle.OracleSim.from_spec(specs[2])
    o.step(10)
    assert np.abs(members[2].positions() - o.positions()).max() <= 1e-6


def test_batch_rejects_mixed_kinds(hip):
    a = hip.Context.from_spec(systems.replica(50, seed=1, k=800.0, rotation_mode=10))
    b = hip.Context.from_spec(systems.replica(50, seed=2, k=800.0, rotation_mode=0))
    batch = hip.Batch()
    with pytest.raises(hip.RbhipError):
        batch.set_members([a, b])
    c = hip.Context.from_spec(systems.replica(50, seed=3, k=800.0, rotation_mode=10), precision="double")
    with pytest.raises(hip.RbhipError):
        batch.set_members([a, c])
