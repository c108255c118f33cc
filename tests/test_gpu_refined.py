"""Refined energies on the GPU (rbhip_set_refined_energies and the
refined-energy getters) against the FP64 oracle, whose refined terms are
pinned in test_oracle_refined.py.

Covered: free atoms and bodies of many sizes (linear 2-atom bodies, several
lanes per body), both rotation modes, both precisions, the part1 path
(step(1)) and the fused path (step(k)), graphs on and off, the split
part1/forces/part2 calls, state changes, a batch with refined and plain
members, and the plugin layer.  Tolerances: mixed precision 2e-5 relative
(fp32 state; the refined terms difference fp32 quaternions), double 1e-9."""
import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems
from test_gpu_edges import clusters

pytestmark = pytest.mark.gpu

RTOL = {"mixed": 2e-5, "double": 1e-9}


def spec_mix(mode, seed=11):
    # linear 2-atom bodies, waters, 5..40-atom bodies, free atoms
    spec = clusters([2] * 25 + [3] * 40 + [5] * 10 + [12] * 6 + [40] * 3, 60, seed=seed, mode=mode)
    rng = np.random.default_rng(seed)
    spec.tether_r0 = spec.tether_r0 + 0.03 * rng.normal(size=spec.tether_r0.shape)
    return spec


def pair(oracle, hip, spec, precision):
    o = oracle.OracleSim(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode)
    g = hip.Context(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode, precision)
    for s in (o, g):
        s.set_tether(spec.tether_k, spec.tether_r0)
        s.set_refined(True)
        s.set_positions(spec.positions)
        s.set_velocities(spec.velocities)
    return o, g


def check(o, g, precision):
    rt = RTOL[precision]
    np.testing.assert_allclose(g.refined_kinetic_energies(), o.refined_kinetic_energies(), rtol=rt)
    np.testing.assert_allclose(g.potential_energy_refinement(), o.potential_energy_refinement(), rtol=rt)
    np.testing.assert_allclose(g.kinetic_energies(), o.kinetic_energies(), rtol=rt)


@pytest.mark.parametrize("mode", [10, 0])
@pytest.mark.parametrize("precision", ["mixed", "double"])
@pytest.mark.parametrize("steps", [1, 6])
def test_refined_energies_match_oracle(hip, oracle, mode, precision, steps):
    o, g = pair(oracle, hip, spec_mix(mode), precision)
    # before the first step: empty accumulators, refinement of the initial forces
    np.testing.assert_array_equal(g.refined_kinetic_energies(), [0.0, 0.0])
    np.testing.assert_allclose(g.potential_energy_refinement(), o.potential_energy_refinement(), rtol=RTOL[precision])
    for _ in range(2):
        o.step(steps)
        g.step(steps)
        check(o, g, precision)


@pytest.mark.parametrize("graph,fusion", [(False, True), (True, False), (False, False)])
def test_refined_energies_launch_paths(hip, oracle, graph, fusion):
    o, g = pair(oracle, hip, spec_mix(10, seed=12), "mixed")
    g.use_graph(graph)
    g.use_fusion(fusion)
    o.step(5)
    g.step(5)
    check(o, g, "mixed")


def test_refined_energies_split_calls(hip, oracle):
    o, g = pair(oracle, hip, spec_mix(10, seed=13), "double")
    for _ in range(3):
        o.step(1)
        g.part1()
        g.compute_forces()
        g.part2()
    check(o, g, "double")


def test_refined_off_state_change_and_step_size(hip, oracle):
    spec = spec_mix(10, seed=14)
    g = hip.Context.from_spec(spec, precision="double")
    g.step(3)
    # off: the plain kinetic energies and no potential refinement
    np.testing.assert_array_equal(g.refined_kinetic_energies(), g.kinetic_energies())
    assert g.potential_energy_refinement() == 0.0
    o, g = pair(oracle, hip, spec, "double")
    o.step(3)
    g.step(3)
    check(o, g, "double")
    # a state change empties the accumulators
    g.set_velocities(spec.velocities)
    np.testing.assert_array_equal(g.refined_kinetic_energies(), [0.0, 0.0])
    o.set_velocities(spec.velocities)
    # a new step size: the next steps and the 1/(6 dt) factor use it
    for s in (o, g):
        s.set_step_size(0.0015)
        s.step(4)
    check(o, g, "double")
    # off again, then on: empty until the next step
    g.set_refined(False)
    np.testing.assert_array_equal(g.refined_kinetic_energies(), g.kinetic_energies())
    g.set_refined(True)
    np.testing.assert_array_equal(g.refined_kinetic_energies(), [0.0, 0.0])


@pytest.mark.parametrize("mode", [10, 0])
def test_batch_with_refined_and_plain_members(hip, mode):
    block = 256 if mode else 64
    specs = [systems.replica(u, seed=s, k=900.0, rotation_mode=mode) for u, s in ((300, 1), (1200, 2), (50, 3))]
    specs.append(systems.heterogeneous(400, 200, seed=9, rotation_mode=mode))
    refined = [True, False, True, True]

    def make():
        ctxs = [hip.Context.from_spec(s, precision="mixed") for s in specs]
        for c, r in zip(ctxs, refined):
            c.set_launch(block)
            if r:
                c.set_refined(True)
        return ctxs

    alone, members = make(), make()
    batch = hip.Batch()
    batch.set_launch(block)
    batch.set_members(members)
    for c in alone:
        c.step(7)
    batch.step(7)
    for a, m, r in zip(alone, members, refined):
        np.testing.assert_array_equal(m.positions(), a.positions())
        np.testing.assert_array_equal(m.refined_kinetic_energies(), a.refined_kinetic_energies())
        np.testing.assert_array_equal(m.potential_energy_refinement(), a.potential_energy_refinement())
        if not r:
            np.testing.assert_array_equal(m.refined_kinetic_energies(), m.kinetic_energies())
    batch.close()


def test_plugin_layer_refined_energies(oracle):
    from openmm_rigidbody_plugin_amd import plugin
    spec = spec_mix(0, seed=15)
    sim = plugin.Simulation.from_spec(spec, precision="double", refined_energies=True)
    o = oracle.OracleSim.from_spec(spec)
    o.set_refined(True)
    o.set_positions(spec.positions)
    o.set_velocities(spec.velocities)
    sim.step(4)
    o.step(4)
    e = sim.energies()
    np.testing.assert_allclose(e["refined_kinetic"], o.refined_kinetic_energies(), rtol=1e-9)
    np.testing.assert_allclose(e["potential_refinement"], o.potential_energy_refinement(), rtol=1e-9)
    plain = plugin.Simulation.from_spec(spec, precision="double")
    plain.step(4)
    e = plain.energies()
    np.testing.assert_array_equal(e["refined_kinetic"], e["kinetic"])
    assert e["potential_refinement"] == 0.0
    sim.close()
    plain.close()
