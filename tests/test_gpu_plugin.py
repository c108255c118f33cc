"""The C++ host plugin layer on the GPU: System + RigidBodyIntegrator + Context
on the HIP platform (the reference's API, openmmapi/include/RigidBodyIntegrator.h),
driving the gfx950 library through the rbhip C ABI.

- tests/cpp/TestHipRigidBodyIntegrator: the reference's known-answer test
  (testSingleBond) and API/exception behaviour, in C++;
- the plugin path against the FP64 oracle on the 512-water + 64-argon box
  and a heterogeneous mix (mixed tolerances: positions <= 1e-6 nm);
- the plugin path is the rbhip path: identical results to the C ABI driven
  directly (bit for bit), including a step-size change mid-run;
- host bookkeeping (cleanBodyIndices / atomIndex / DOF) equals the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "_bin", "TestHipRigidBodyIntegrator")


@pytest.fixture(scope="module")
def plugin(hip):
    from openmm_rigidbody_plugin_amd import plugin
    plugin.lib()
    return plugin


def test_cpp_host_layer(plugin):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("Done")


@pytest.mark.parametrize("mode", [10, 0])
def test_plugin_matches_oracle_c1(plugin, oracle, mode):
    spec = systems.water_c1(rotation_mode=mode)
    sim = plugin.Simulation.from_spec(spec, precision="mixed")
    ref = oracle.OracleSim.from_spec(spec)
    sim.step(10)
    ref.step(10)
    st = sim.state(forces=True)
    assert np.abs(st["positions"] - ref.positions()).max() <= 1e-6
    assert np.abs(st["velocities"] - ref.velocities()).max() <= 1e-5 * np.abs(ref.velocities()).max()
    assert st["time"] == pytest.approx(10 * spec.dt, rel=1e-12)
    ke = sim.energies()
    np.testing.assert_allclose(ke["kinetic"], ref.kinetic_energies(), rtol=1e-5)
    np.testing.assert_array_equal(ke["refined_kinetic"], ke["kinetic"])
    assert ke["potential_refinement"] == 0.0
    assert st["kinetic"] == pytest.approx(ke["kinetic"].sum(), rel=1e-12)
    assert st["potential"] == pytest.approx(ref.potential_energy(), rel=1e-5)


def test_plugin_is_the_rbhip_path(plugin, hip):
    spec = systems.heterogeneous(300, 400, seed=21, rotation_mode=6)
    sim = plugin.Simulation.from_spec(spec, precision="mixed")
    ctx = hip.Context.from_spec(spec, precision="mixed")
    for s in (sim, ctx):
        s.step(7)
    sim.set_step_size(0.0015)
    ctx.set_step_size(0.0015)
    sim.step(3)
    ctx.step(3)
    st = sim.state()
    np.testing.assert_array_equal(st["positions"], ctx.positions())
    np.testing.assert_array_equal(st["velocities"], ctx.velocities())
    np.testing.assert_array_equal(sim.energies()["kinetic"], ctx.kinetic_energies())
    b = sim.rigid_body(5)
    ref_states = ctx.body_states()
    for key in ("q", "pi", "rcm", "pcm"):
        np.testing.assert_array_equal(b[key], ref_states[5][key])


def test_plugin_bookkeeping(plugin, oracle):
    spec = systems.heterogeneous(200, 150, seed=5, rotation_mode=0)
    sim = plugin.Simulation.from_spec(spec, precision="double")
    ref = oracle.OracleSim.from_spec(spec)
    c, rc = sim.counts(), ref.counts()
    assert (c["numFree"], c["numBodies"], c["numActualAtoms"], c["numBodyAtoms"], c["numDOF"]) == \
        (rc["numFree"], rc["numBodies"], rc["numActualAtoms"], rc["numBodyAtoms"], rc["numDOF"])
    np.testing.assert_array_equal(sim.atom_index(), ref.atom_index())
    sim.step(5)
    ref.step(5)
    assert np.abs(sim.state()["positions"] - ref.positions()).max() <= 1e-9
