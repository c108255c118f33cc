"""The job server on the GPU (rbhip_server_*): admission by HBM budget,
resident state, continuous batching.

Each job runs in rounds shared with whatever else is resident; with the
server's workgroup size its bodies go through the same kernels and tiles as
when stepped alone, so every result equals the job stepped alone, bit for
bit, whatever the mix.  A short job is also checked against the FP64 oracle
(mixed tolerance)."""
import time

import numpy as np
import pytest

from openmm_rigidbody_plugin_amd import systems

pytestmark = pytest.mark.gpu


def jobs(mode=0):
    sizes = [(120, 40), (700, 250), (35, 0), (1500, 300), (260, 90)]
    steps = [30, 7, 55, 12, 40]
    specs = [systems.heterogeneous(nb, nf, seed=100 + i, rotation_mode=mode, name=f"job{i}")
             for i, (nb, nf) in enumerate(sizes)]
    return specs, steps


def alone(hip, spec, steps, block):
    c = hip.Context.from_spec(spec, precision="mixed")
    c.set_launch(block)
    c.step(steps)
    return c


def rounds_until(seconds=60.0):
    """Yields until a deadline: once every job has finished stepping, a round
    returns at once while the collection thread still downloads the last
    jobs (the first download of a process takes longest), so loops over
    rounds are bounded in time, not in count."""
    deadline = time.monotonic() + seconds
    while time.monotonic() < deadline:
        yield


def drain(server):
    results = {}
    for _ in rounds_until():
        server.run_round()
        for jid in server.completed():
            results[jid] = server.take(jid)
        st = server.status()
        if st["waiting"] == 0 and st["resident"] == 0:
            # collection runs on a thread of its own: a job can complete
            # after the last completed() call
            for jid in server.completed():
                results[jid] = server.take(jid)
            return results
    raise AssertionError(f"server did not drain: {server.status()} {server.profile()} taken {sorted(results)}")


@pytest.mark.parametrize("mode,block", [(0, 64), (10, 256)])
def test_server_results_equal_alone(hip, mode, block):
    specs, steps = jobs(mode)
    server = hip.Server(hbm_budget=8 << 30, steps_per_round=10, block=block)
    ids = [server.submit(s, n) for s, n in zip(specs[:3], steps[:3])]
    server.run_round()
    server.run_round()
    # arrivals while others are resident join in later rounds
    ids += [server.submit(s, n) for s, n in zip(specs[3:], steps[3:])]
    results = {}
    for jid in server.completed():
        results[jid] = server.take(jid)
    results.update(drain(server))
    assert sorted(results) == sorted(ids)
    for jid, spec, n in zip(ids, specs, steps):
        ref = alone(hip, spec, n, block)
        r = results[jid]
        np.testing.assert_array_equal(r["positions"], ref.positions())
        np.testing.assert_array_equal(r["velocities"], ref.velocities())
        np.testing.assert_array_equal(r["kinetic"], ref.kinetic_energies())
        assert r["latency"] >= r["resident"] >= 0.0


def test_submitted_arrays_outlive_the_caller(hip):
    # the server reads a job's arrays when it sets the job up; the wrapper
    # keeps them alive, including the converted copies of float32 and
    # non-contiguous inputs, after the caller has dropped every reference
    import copy
    import gc
    specs, steps = jobs(0)
    server = hip.Server(hbm_budget=2 << 20, steps_per_round=5, block=64)   # most jobs wait for admission
    ids = []
    for s, n in zip(specs, steps):
        t = copy.deepcopy(s)
        t.positions = np.asfortranarray(t.positions)
        t.velocities = t.velocities.astype(np.float32).astype(np.float64)[:, :]
        t.masses = t.masses.astype(np.float32)
        ids.append(server.submit(t, n))
        del t
    gc.collect()
    results = drain(server)
    for jid, spec, n in zip(ids, specs, steps):
        ref = copy.deepcopy(spec)
        ref.velocities = ref.velocities.astype(np.float32).astype(np.float64)
        ref.masses = ref.masses.astype(np.float32).astype(np.float64)
        np.testing.assert_array_equal(results[jid]["positions"], alone(hip, ref, n, 64).positions())
    assert server._inputs == {}


def test_longest_first_admission(hip):
    # equal-size jobs and a budget for one of them: jobs are admitted by step
    # count, most steps first, so they complete in that order; results equal
    # alone
    specs = [systems.heterogeneous(700, 250, seed=300 + i, rotation_mode=0, name=f"same{i}") for i in range(4)]
    steps = [12, 40, 25, 33]
    budget = max(hip.estimate_bytes(s.body_indices) for s in specs)
    server = hip.Server(hbm_budget=budget, steps_per_round=8, block=64)
    server.set_admission_order("longest_first")
    ids = [server.submit(s, n) for s, n in zip(specs, steps)]
    order = []
    for _ in rounds_until():
        server.run_round()
        st = server.status()
        assert st["resident"] <= 1
        for jid in server.completed():
            order.append(jid)
            r = server.take(jid)
            i = ids.index(jid)
            np.testing.assert_array_equal(r["positions"], alone(hip, specs[i], steps[i], 64).positions())
        if st["waiting"] == 0 and st["resident"] == 0 and len(order) == len(ids):
            break
    assert order == [ids[1], ids[3], ids[2], ids[0]]
    assert hip.lib().rbhip_server_set_admission_order(server._h, 7) != 0
    with pytest.raises(ValueError):
        server.set_admission_order("shortest_first")


def test_server_respects_budget(hip):
    specs, steps = jobs(0)
    est = [hip.estimate_bytes(s.body_indices) for s in specs]
    budget = int(max(est) * 1.3)
    server = hip.Server(hbm_budget=budget, steps_per_round=8, block=64)
    ids = [server.submit(s, n) for s, n in zip(specs, steps)]
    results = {}
    saw_waiting = False
    for _ in rounds_until():
        server.run_round()
        st = server.status()
        assert st["bytes_in_use"] <= budget
        saw_waiting |= st["waiting"] > 0 and st["resident"] > 0
        for jid in server.completed():
            results[jid] = server.take(jid)
        if st["waiting"] == 0 and st["resident"] == 0:
            for jid in server.completed():
                results[jid] = server.take(jid)
            break
    assert saw_waiting, "the budget should have held jobs back"
    assert sorted(results) == sorted(ids)
    for jid, spec, n in zip(ids, specs, steps):
        np.testing.assert_array_equal(results[jid]["positions"], alone(hip, spec, n, 64).positions())


def test_estimate_bounds_device_bytes(hip):
    for spec in jobs(0)[0] + [systems.water_c1(rotation_mode=10)]:
        c = hip.Context.from_spec(spec, precision="mixed")
        used = c.device_bytes()
        est = hip.estimate_bytes(spec.body_indices)
        assert used <= est <= used + 64 * 1024 + used // 50


def test_server_job_matches_oracle(hip, oracle):
    specs, _ = jobs(0)
    server = hip.Server(hbm_budget=4 << 30, steps_per_round=4, block=64)
    big = server.submit(specs[3], 40)
    probe = server.submit(specs[0], 10)
    results = drain(server)
    o = oracle.OracleSim.from_spec(specs[0])
    o.step(10)
    assert np.abs(results[probe]["positions"] - o.positions()).max() <= 1e-6
    np.testing.assert_allclose(results[probe]["kinetic"], o.kinetic_energies(), rtol=1e-5)
    assert big in results


def test_server_rejections(hip):
    specs, _ = jobs(0)
    server = hip.Server(hbm_budget=1 << 20, steps_per_round=4, block=64)
    with pytest.raises(hip.ServerError):
        server.submit(specs[3], 10)                  # larger than the whole budget
    server = hip.Server(hbm_budget=1 << 30, steps_per_round=4, block=64)
    server.submit(specs[0], 5)
    with pytest.raises(hip.ServerError):
        server.submit(jobs(10)[0][1], 5)              # NO-SQUISH job on an exact-rotation server
    with pytest.raises(hip.ServerError):
        server.submit(specs[1], 0)
