"""Benchmark of the gfx950 RigidBodyIntegrator step.

    python bench.py [--steps K] [--warmup W] [--config water1M] [--gpus 1]

Workload (default): 1,000,000 water-like 3-site rigid bodies (3M atoms) on a
jittered 100^3 lattice in a 21.5 nm box, 300 K velocities, harmonic tethers
k = 1000 kJ/mol/nm^2, dt = 2 fs, NO-SQUISH n = 10, mixed precision (fp32
state, FP64 accumulation); synthetic seeded data.  The timed region is one
RigidBodyIntegrator::step(K) call: part1, forces, (K-1) x [fused part2+part1,
forces], part2 -- one step per body per fused pass plus one tether force
evaluation per step.

Prints one JSON line: whole-job body-steps/s over K timed steps (inputs
resident in HBM), the HBM roofline of the dominant integrator kernel measured
with HIP events on its stream, and the FP64 CPU restatement timed on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, n_atoms_per_body, num_bodies, num_free, indexed=False):
    """Algorithmic bytes per launch, counted as the reference's canonical
    drop-in step moves them: fp32 4-byte I/O, the quaternion momentum pi as
    4 values, the next step's first half-kick folded into part 2.  Per body
    of N atoms:
      part1  reads r, p, q, pi (56), 1/M, 1/I (16), d (12N); writes r, q, pi
             (44), positions (12N): 116 + 24N
      part2  reads p, q, pi (44), 1/M, 1/I (16), d (12N), forces (12N);
             writes p, pi (28), velocities (12N): 88 + 36N
      (part1 + part2 = 204 + 60N = 384 B for a 3-site body)
      fused  (part2 of step n + part1 of step n+1) reads r, p, q, pi, 1/M,
             1/I (72), d (12N), forces (12N); writes r, p, q, pi (56),
             positions (12N): 128 + 36N = 236 B for a 3-site body
    Per free atom: part1 reads r, v, F, 1/m (40), writes r, v (24): 64;
    part2 reads v, F, 1/m (28), writes v (12): 40; fused 64.
    forces (the benchmark tether) read r and r0, write F: 36 B per atom.
    Atoms not stored body by body (`indexed`) add an int32 index read per
    atom and pass (+4 B).  What the implementation moves beyond these bytes
    (the FP64 centre of mass, the free atoms' low position parts) is
    precision overhead and is never counted here; what it moves below them
    (bodies of one shape reading shared body-fixed coordinates) is not
    subtracted either: these are the bytes of the canonical step, and the
    measured traffic is reported beside them."""
    N = n_atoms_per_body
    idx = 4 if indexed else 0
    atoms = num_bodies * N + num_free
    if kernel == "part1":
        b = num_bodies * (116 + 24 * N) + num_free * 64
    elif kernel == "part2":
        b = num_bodies * (88 + 36 * N) + num_free * 40
    elif kernel == "fused":
        b = num_bodies * (128 + 36 * N) + num_free * 64
    elif kernel == "forces":
        return atoms * 36
    else:
        raise ValueError(kernel)
    return b + idx * atoms


def implementation_bytes(kernel, n_atoms_per_body, num_bodies, num_free, shared_frames=False):
    """Bytes the mixed-precision kernels move by design: body state as q (16),
    body-frame L and p (12 each), 1/I and 1/M (16), r_cm in FP64 (24); free
    atoms' positions as fp32 hi + lo (24); the tether's anchor in fp32.
      part1: 132 + 24N per body, 88 per free atom
      part2: 80 + 36N per body, 40 per free atom
      fused: 144 + 36N per body, 88 per free atom
      forces: 36 per atom, + 12 per free atom (low position part)
    With shared frames (bodies of one shape read a few cached rows of
    body-fixed coordinates and one cached 1/I, 1/M) each pass reads a
    4-byte body index per body instead of its 12N bytes of coordinates and
    16 bytes of inertia."""
    N = n_atoms_per_body
    d = (4 - 12 * N - 16) if shared_frames else 0
    if kernel == "part1":
        return num_bodies * (132 + 24 * N + d) + num_free * 88
    if kernel == "part2":
        return num_bodies * (80 + 36 * N + d) + num_free * 40
    if kernel == "fused":
        return num_bodies * (144 + 36 * N + d) + num_free * 88
    if kernel == "forces":
        return (num_bodies * N + num_free) * 36 + num_free * 12
    raise ValueError(kernel)


def make_workload(name):
    from openmm_rigidbody_plugin_amd import systems
    if name == "water1M":
        return systems.water_c2(), dict(workload="water 3-site rigid bodies, NO-SQUISH n=10", bodies=1_000_000,
                                        atoms=3_000_000, box_nm=21.5, dt_ps=0.002, rotation_mode=10,
                                        precision="mixed")
    if name == "water4M":
        spec = systems.water_box(4_000_000, 0.215, seed=2, rotation_mode=10, name="water4M")
        return spec, dict(workload="water 3-site rigid bodies, NO-SQUISH n=10", bodies=4_000_000,
                          atoms=12_000_000, box_nm=34.2, dt_ps=0.002, rotation_mode=10, precision="mixed")
    if name == "hetero200k":
        spec = systems.heterogeneous(200_000, 500_000, seed=3, rotation_mode=0)
        return spec, dict(workload="heterogeneous bodies (2/3/6/20 atoms) + free atoms, exact rotation",
                          bodies=200_000, free_atoms=500_000, atoms=int(spec.num_atoms), dt_ps=0.001,
                          rotation_mode=0, precision="mixed")
    raise SystemExit(f"unknown config {name}")


def native_oracle():
    """Path of the oracle built for this host's CPU (-O3 -march=native,
    oracle/Makefile `native`), or the portable build if that fails."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True, text=True)
    path = os.path.join(ROOT, "oracle", "_native", "librbref_native.so")
    if r.returncode == 0 and os.path.exists(path):
        return path, "g++ -O3 -march=native -ffp-contract=off"
    log(f"native oracle build failed, timing the portable build: {r.stderr[-300:]}")
    return os.path.join(ROOT, "oracle", "librbref.so"), "g++ -O3 -ffp-contract=off"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(spec, budget_s=15.0, sample_bodies=100_000):
    """The FP64 CPU restatement of the reference platform (single thread, as
    the reference loops over bodies serially), built for this host: the
    first `sample_bodies` bodies of the headline workload for ~budget_s
    (value, body-steps/s), and the reference's own small case (512 waters +
    64 argon atoms) for all 1000 steps in both rotation modes."""
    lib_path, compiler = native_oracle()
    os.environ["RBREF_LIB"] = lib_path
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from openmm_rigidbody_plugin_amd import systems
    from openmm_rigidbody_plugin_amd.systems import SystemSpec
    bi = spec.body_indices
    keep = np.nonzero((bi > 0) & (bi <= sample_bodies))[0]
    sub = SystemSpec(masses=spec.masses[keep], body_indices=bi[keep], positions=spec.positions[keep],
                     velocities=spec.velocities[keep], tether_k=spec.tether_k[keep],
                     tether_r0=spec.tether_r0[keep], dt=spec.dt, rotation_mode=spec.rotation_mode)
    nb = int(np.unique(bi[keep]).size)
    o = pyoracle.OracleSim.from_spec(sub)
    t = time.perf_counter()
    o.step(2)
    per_step = (time.perf_counter() - t) / 2
    steps = int(max(3, min(500, budget_s / max(per_step, 1e-9))))
    t = time.perf_counter()
    o.step(steps)
    el = time.perf_counter() - t
    small = {}
    for mode in (10, 0):
        c1 = systems.water_c1(rotation_mode=mode)
        oc = pyoracle.OracleSim.from_spec(c1)
        oc.step(10)
        t1 = time.perf_counter()
        oc.step(1000)
        s1 = time.perf_counter() - t1
        small["nosquish10" if mode else "exact"] = {
            "steps": 1000, "seconds": round(s1, 3), "body_steps_per_s": 512 * 1000 / s1,
            "atom_steps_per_s_free": 64 * 1000 / s1}
    cpu = cpu_model()
    return {"value": nb * steps / el, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"first {nb} bodies of the same workload, {steps} steps, FP64 oracle (single thread, "
                      f"{compiler}) on {cpu}", "seconds": round(el, 2), "cpu_model": cpu,
            "nproc": os.cpu_count(), "compiler": compiler,
            "reference_small_case": {"workload": "512 water bodies + 64 argon atoms, tether k=1000, dt 2 fs, FP64",
                                     **small}}


def valu_from_profiles(config, kernel):
    """VALU issue fraction recorded in profiles/ (rocprofv3 counters of an
    earlier run of this kernel; not measured by this run)."""
    path = os.path.join(ROOT, "profiles", f"valu_{kernel}_{config}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("valu_issue_frac")


def traffic_from_profiles(config, kernel):
    """HBM bytes per launch recorded in profiles/ (two rocprofv3 --pmc runs,
    FETCH_SIZE and WRITE_SIZE scaled by a copy kernel's; not measured by this
    run)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{config}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def plain_run(ctxs, steps, per_step):
    """Steps each context in turn through its own calls.  per_step: the
    reference's execute pattern (part1, forces, part2 launched for every
    step, as an OpenMM integration does with OpenMM computing the forces);
    otherwise one step(K) call per context (fused boundaries, HIP-graph
    replay).  Returns the seconds taken."""
    from openmm_rigidbody_plugin_amd import rbhip
    for c in ctxs:
        c.use_fusion(not per_step)
        c.use_graph(not per_step)
    rbhip.synchronize()
    t = time.perf_counter()
    for c in ctxs:
        c.step(steps)
    rbhip.synchronize()
    dt = time.perf_counter() - t
    for c in ctxs:
        c.use_fusion(True)
        c.use_graph(True)
    return dt


def batched(steps, warmup, count=256, log_every=64, plain=True):
    """256 independent replicas (units ~ LogUniform(1e3, 2e5), 80% water bodies
    / 20% argon, k ~ U[500, 2000]) stepped `steps` times: one system at a time
    (each context on its own stream) with per-step launches and with step(K)
    (unless `plain` is false), then all of them as one batch (one launch per
    phase over every replica's tiles, HIP-graph replay, launch shape from the
    batch autotuner), and the batch's kernels timed on their own dispatches."""
    from openmm_rigidbody_plugin_amd import rbhip, systems
    t = time.perf_counter()
    specs = systems.replica_set(count)
    ctxs = []
    for i, s in enumerate(specs):
        ctxs.append(rbhip.Context.from_spec(s, precision="mixed"))
        if (i + 1) % log_every == 0:
            log(f"  replicas built: {i + 1}/{count}")
    rbhip.synchronize()
    nb = sum(s.num_bodies for s in specs)
    nf = sum(s.num_free for s in specs)
    log(f"batched: {count} replicas, {nb} bodies + {nf} free atoms, built in {time.perf_counter() - t:.1f}s")
    for c in ctxs:
        c.step(warmup)
    rbhip.synchronize()
    per_step_s = plain_s = None
    if plain:
        per_step_s = plain_run(ctxs, steps, per_step=True)
        log(f"  plain, per-step launches: {per_step_s:.3f}s")
        plain_s = plain_run(ctxs, steps, per_step=False)
        log(f"  plain, one step(K) call per system: {plain_s:.3f}s")
    batch = rbhip.Batch()
    batch.set_members(ctxs)
    t = time.perf_counter()
    batch.autotune(reps=5)
    log(f"  batch autotune {time.perf_counter() - t:.1f}s: {batch.describe()}")
    batch.step(warmup)
    rbhip.synchronize()
    t = time.perf_counter()
    batch.step(steps)
    rbhip.synchronize()
    batch_s = time.perf_counter() - t
    ke = batch.kinetic_energies()
    assert np.all(np.isfinite(ke)), "non-finite kinetic energy in the batch"
    log(f"  batched: {batch_s:.3f}s")
    kms, _ = batch.time_kernels(50)
    alg = {k: algorithmic_bytes(k, 3, nb, nf) for k in ("fused", "forces")}
    log("  batch kernel ms: " + ", ".join(f"{k}={v:.4f}" for k, v in kms.items()))
    units = nb + nf
    out = {
        "workload": f"{count} replicas, units LogUniform(1e3, 2e5), 80% 3-site bodies / 20% free atoms, "
                    "k ~ U[500, 2000], NO-SQUISH n=10, dt 2 fs, mixed",
        "replicas": count, "bodies": nb, "free_atoms": nf, "steps": steps,
        "value": nb * steps / batch_s, "unit": "body-steps/s",
        "units_per_s": units * steps / batch_s,
        "ms_per_step": batch_s / steps * 1e3,
        "launch": batch.describe(),
        "kernel_ms": kms,
        "kernel_algorithmic_bytes": alg,
        "kernel_frac_hbm": {k: alg[k] / (kms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS for k in alg},
    }
    if plain:
        out.update({
            "plain_per_step_value": nb * steps / per_step_s,
            "plain_step_k_value": nb * steps / plain_s,
            "speedup_vs_plain_per_step": per_step_s / batch_s,
            "speedup_vs_plain_step_k": plain_s / batch_s,
            "plain_per_step_ms_per_step": per_step_s / steps * 1e3,
            "plain_step_k_ms_per_step": plain_s / steps * 1e3,
        })
    return out


def served(num_jobs=96, rate=25.0, budget=1 << 30, steps_per_round=250, seed=11):
    """Serving: a Poisson stream of jobs (bodies ~ LogUniform(500, 1e5) with the
    heterogeneous body mix and 2.5 free atoms per body, steps ~
    LogUniform(100, 1e4), exact rotation, dt 1 fs, mixed) submitted to the job
    server in real time; admission against a fixed HBM budget, continuous
    batching.  Reports completion-latency percentiles and aggregate
    body-steps/s; then the same jobs all at once (saturated server) and one
    after another through the plain per-system path."""
    from openmm_rigidbody_plugin_amd import rbhip, systems
    rng = np.random.default_rng(seed)
    nbs = np.exp(rng.uniform(np.log(500), np.log(1e5), num_jobs)).astype(int)
    steps = np.exp(rng.uniform(np.log(100), np.log(1e4), num_jobs)).astype(int)
    t = time.perf_counter()
    specs = [systems.heterogeneous(int(nb), int(2.5 * nb), seed=int(seed * 1000 + i), rotation_mode=0,
                                   name=f"job{i}") for i, nb in enumerate(nbs)]
    log(f"served: {num_jobs} jobs generated in {time.perf_counter() - t:.1f}s "
        f"({int(nbs.sum())} bodies, {int((nbs * steps).sum())} body-steps)")
    work = int((nbs * steps).sum())
    block = 64

    def run_server(arrivals):
        srv = rbhip.Server(hbm_budget=budget, steps_per_round=steps_per_round, block=block)
        ids, lat, peak = {}, {}, 0
        nxt = 0
        t0 = time.perf_counter()
        while len(lat) < num_jobs:
            now = time.perf_counter() - t0
            while nxt < num_jobs and arrivals[nxt] <= now:
                ids[srv.submit(specs[nxt], int(steps[nxt]))] = nxt
                nxt += 1
            # finished jobs are collected on a server thread: take them whenever listed
            for jid in srv.completed():
                lat[ids[jid]] = srv.take(jid)
            st = srv.status()
            peak = max(peak, st["resident"])
            if st["waiting"] == 0 and st["resident"] == 0:
                if nxt < num_jobs:
                    time.sleep(max(0.0, min(arrivals[nxt] - now, 0.001)))
                continue
            srv.run_round()
        makespan = time.perf_counter() - t0
        log(f"    server phases: {srv.profile()}")
        return srv, lat, makespan, peak

    gaps = rng.exponential(1.0 / rate, num_jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    _, lat, makespan, peak = run_server(arrivals)
    L = np.array([lat[i]["latency"] for i in range(num_jobs)])
    log(f"  poisson: makespan {makespan:.2f}s, p50 {np.percentile(L, 50) * 1e3:.1f} ms, "
        f"p99 {np.percentile(L, 99) * 1e3:.1f} ms, peak resident {peak}")
    # saturated: every job waiting at t = 0
    _, lat_s, makespan_s, peak_s = run_server(np.zeros(num_jobs))
    log(f"  saturated: makespan {makespan_s:.2f}s, peak resident {peak_s}")
    # serial plain: one job at a time through its own context, with per-step
    # launches (the reference's execute pattern) and with one step(K) call
    sample = {int(np.argmin(nbs * steps)), int(np.argsort(nbs * steps)[num_jobs // 2])}
    alone = {}

    def serial_plain(per_step):
        t = time.perf_counter()
        for i, spec in enumerate(specs):
            c = rbhip.Context.from_spec(spec, precision="mixed")
            c.set_launch(block)
            c.use_fusion(not per_step)
            c.use_graph(not per_step)
            c.step(int(steps[i]))
            x = c.positions()
            if i in sample:
                alone[(i, per_step)] = x
            del c
        return time.perf_counter() - t

    serial_ps = serial_plain(True)
    log(f"  serial plain, per-step launches: {serial_ps:.2f}s")
    serial = serial_plain(False)
    log(f"  serial plain, step(K): {serial:.2f}s")
    same = all(np.array_equal(x, lat_s[i]["positions"]) and np.array_equal(x, lat[i]["positions"])
               for (i, _), x in alone.items())
    # the jobs with the fewest steps, kept for the FP64 check of the CPU leg
    check = [(specs[int(i)], int(steps[int(i)]), lat_s[int(i)]["positions"], int(i)) for i in np.argsort(steps)[:3]]
    return {
        "workload": f"{num_jobs} jobs, bodies LogUniform(500, 1e5) (30% 2-atom, 40% 3-site, 20% 6-ring, "
                    "10% 20-atom) + 2.5 free atoms per body, steps LogUniform(100, 1e4), exact rotation, "
                    "dt 1 fs, mixed",
        "arrival_rate_per_s": rate, "hbm_budget_bytes": budget, "steps_per_round": steps_per_round,
        "block": block, "jobs": num_jobs, "body_steps": work,
        "p50_latency_ms": float(np.percentile(L, 50) * 1e3), "p99_latency_ms": float(np.percentile(L, 99) * 1e3),
        "p50_wait_ms": float(np.percentile([lat[i]["wait"] for i in range(num_jobs)], 50) * 1e3),
        "aggregate_body_steps_per_s": work / makespan, "makespan_s": makespan, "peak_resident": peak,
        "saturated_body_steps_per_s": work / makespan_s,
        "serial_plain_per_step_body_steps_per_s": work / serial_ps,
        "serial_plain_step_k_body_steps_per_s": work / serial,
        "saturated_speedup_vs_serial_plain_per_step": serial_ps / makespan_s,
        "saturated_speedup_vs_serial_plain_step_k": serial / makespan_s,
        "sampled_jobs_equal_alone": bool(same), "sampled_jobs": sorted({i for i, _ in alone}),
        "_fp64_check_jobs": check,
    }


def served_fp64_check(check, bodies=64, free=64, tol_nm=1e-5):
    """Per-job agreement of served results with the FP64 oracle on a sample
    (part of the CPU leg): for each (spec, steps, served positions, job) in
    `check`, `bodies` bodies and `free` free atoms stepped alone by the
    oracle (every force is a per-atom tether, so a sample evolves as it
    does in the whole job), compared atom by atom with the served result.
    The checked jobs run 100-120 steps, and mixed-precision rounding grows
    with the step count, so the bound is 1e-5 nm rather than the 1e-6 nm
    the parity tests apply after 10 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle
    from subsample import sample_atoms, subsystem
    out = []
    for spec, steps, positions, job in check:
        atoms = sample_atoms(spec, bodies, free, seed=job)
        o = pyoracle.OracleSim.from_spec(subsystem(spec, atoms))
        o.step(steps)
        dx = float(np.abs(o.positions() - positions[atoms]).max())
        out.append({"job": job, "steps": steps, "atoms": int(atoms.size), "max_dx_nm": dx})
        log(f"  fp64 check: job {job}, {steps} steps, {atoms.size} atoms, max |dx| {dx:.2e} nm")
    return {"tolerance_nm": tol_nm, "jobs": out, "ok": all(j["max_dx_nm"] <= tol_nm for j in out)}


def kernel_roofline(kms, natoms_body, nb, nf, config, shared_frames=False):
    """The HBM roofline of the dominant kernel of step(K) (the fused pass)
    from its duration (HIP events on its own dispatch) and its algorithmic
    bytes; the two-pass step
    (part1 + part2, what one execute() with forces computed outside the
    library runs) and the full step with the tether beside it."""
    alg = {k: algorithmic_bytes(k, natoms_body, nb, nf) for k in ("part1", "part2", "fused", "forces")}
    impl = {k: implementation_bytes(k, natoms_body, nb, nf, shared_frames) for k in ("part1", "part2", "fused", "forces")}
    achieved = {k: alg[k] / (kms[k] * 1e-3) / 1e9 for k in alg}
    dominant = "fused"
    two_ms = kms["part1"] + kms["part2"]
    two_alg = alg["part1"] + alg["part2"]
    return {
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved[dominant],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved[dominant] / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(config, dominant),
            "traffic_source": "recorded: profiles/traffic_fused_" + config + ".json (rocprofv3 --pmc, not this run)",
            "algorithmic_bytes_per_launch": alg[dominant],
            "implementation_bytes_per_launch": impl[dominant],
            "implementation_over_algorithmic": impl[dominant] / alg[dominant],
            "shared_frames": shared_frames,
            "avg_launch_ms": kms[dominant],
            "valu_issue_frac_recorded": valu_from_profiles(config, dominant),
        },
        "integrator": {
            "note": "step(K) runs one fused pass and one force evaluation per step; the two-pass step "
                    "(part1 + part2, one execute() with the forces computed outside the library) beside it",
            "fused_ms_per_step": kms["fused"],
            "fused_body_steps_per_s": nb / (kms["fused"] * 1e-3),
            "two_pass_ms_per_step": two_ms,
            "two_pass_algorithmic_bytes": two_alg,
            "two_pass_frac_hbm": two_alg / (two_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "two_pass_body_steps_per_s": nb / (two_ms * 1e-3),
            "forces_ms": kms["forces"],
            "forces_frac_hbm": achieved["forces"] / HBM_PEAK_GBS,
            "kernel_achieved_GBs": achieved,
            "kernel_algorithmic_bytes": alg,
            "kernel_implementation_bytes": impl,
        },
    }


def adopted_leg(spec, precision, kernel_steps):
    """The drop-in path on the same system: the atom arrays in OpenMM's HIP
    layout, owned by the caller (posq + posqCorrection float4, velm double4
    with 1/m in w, the fixed-point force array; slot order = atom order) and
    handed to rbhip_adopt_arrays, the library's tether writing the
    fixed-point forces.  part1 and part2 are the kernels of rbhip_part1 /
    rbhip_part2, the per-step calls of an OpenMM binding's execute()."""
    from openmm_rigidbody_plugin_amd import rbhip
    from openmm_rigidbody_plugin_amd.hipmem import DeviceArray
    n = spec.num_atoms
    padded = (n + 31) // 32 * 32
    m = spec.masses
    rt = np.float32 if precision == "mixed" else np.float64
    velm = np.zeros((padded, 4))
    velm[:n, 3] = np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 0.0)
    posq = DeviceArray(np.zeros((padded, 4), rt))
    corr = DeviceArray(np.zeros((padded, 4), np.float32)) if precision == "mixed" else None
    velm = DeviceArray(velm)
    force = DeviceArray(np.zeros((3, padded), np.int64))
    ctx = rbhip.Context(spec.masses, spec.body_indices, spec.dt, spec.rotation_mode, precision)
    ctx.set_tether(spec.tether_k, spec.tether_r0)
    ctx.adopt_arrays(posq.ptr.value, corr.ptr.value if corr else None, velm.ptr.value, force.ptr.value, padded,
                     np.arange(n, dtype=np.int32))
    ctx.set_positions(spec.positions)
    ctx.set_velocities(spec.velocities)
    ctx.autotune(reps=3)
    kms, _ = ctx.time_kernels(kernel_steps)
    counts = ctx.counts()
    nb, nf = counts["numBodies"], counts["numFree"]
    N = counts["numBodyAtoms"] / max(nb, 1)
    alg = {k: algorithmic_bytes(k, N, nb, nf, indexed=True) for k in ("part1", "part2", "forces")}
    two_ms = kms["part1"] + kms["part2"]
    ke = ctx.kinetic_energies()
    assert np.all(np.isfinite(ke)), "non-finite kinetic energy"
    launch = ctx.describe()
    ctx.close()
    for a in (posq, corr, velm, force):
        if a:
            a.free()
    log(f"  adopted arrays: part1 {kms['part1']:.4f} ms, part2 {kms['part2']:.4f} ms, forces {kms['forces']:.4f} ms")
    return {
        "note": "the same system on OpenMM-layout arrays owned by the caller (rbhip_adopt_arrays); "
                "algorithmic bytes add the int32 slot read per atom and pass",
        "kernel_ms": {k: kms[k] for k in ("part1", "forces", "part2")},
        "two_pass_ms_per_step": two_ms,
        "two_pass_body_steps_per_s": nb / (two_ms * 1e-3),
        "two_pass_frac_hbm": (alg["part1"] + alg["part2"]) / (two_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "with_forces_ms_per_step": two_ms + kms["forces"],
        "kernel_algorithmic_bytes": alg,
        "launch": launch,
    }


def hetero_leg(steps, warmup):
    """The heterogeneous configuration timed by the same clock: 200k bodies
    (30% 2-atom rotors, 40% 3-site, 20% 6-rings, 10% 20-atom clusters, random
    masses) + 500k free atoms, atom order permuted, exact rotation, dt 1 fs,
    mixed precision."""
    from openmm_rigidbody_plugin_amd import rbhip
    spec, cfg = make_workload("hetero200k")
    ctx = rbhip.Context.from_spec(spec, precision="mixed")
    counts = ctx.counts()
    nb, nf = counts["numBodies"], counts["numFree"]
    ctx.autotune(reps=3)
    ctx.step(warmup)
    rbhip.synchronize()
    t = time.perf_counter()
    ctx.step(steps)
    rbhip.synchronize()
    el = time.perf_counter() - t
    kms, _ = ctx.time_kernels(100)
    roof = kernel_roofline(kms, counts["numBodyAtoms"] / max(nb, 1), nb, nf, "hetero200k",
                           "sharedShapes=" in ctx.describe())
    ke = ctx.kinetic_energies()
    assert np.all(np.isfinite(ke)), "non-finite kinetic energy"
    log(f"  hetero200k: {el / steps * 1e3:.4f} ms/step, fused {kms['fused']:.4f} ms, "
        f"frac {roof['roofline']['frac']:.3f}")
    return {"workload": cfg["workload"], "bodies": nb, "free_atoms": nf, "steps": steps,
            "ms_per_step": el / steps * 1e3, "value": nb * steps / el, "unit": "body-steps/s",
            "free_atom_steps_per_s": nf * steps / el, "launch": ctx.describe(), "kernel_ms": kms,
            "roofline": roof["roofline"], "integrator": roof["integrator"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="water1M")
    ap.add_argument("--kernel-steps", type=int, default=100)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default=None, metavar="BLOCK[,MAXBLOCKS]",
                    help="a fixed launch shape for the main workload instead of the autotuner's")
    ap.add_argument("--serve-jobs", type=int, default=96, help="jobs in the served stream")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-replica batched measurement")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-serve", action="store_true", help="skip the served job-stream measurement")
    ap.add_argument("--no-hetero", action="store_true",
                    help="skip the heterogeneous 200k-body exact-rotation measurement")
    ap.add_argument("--hetero-steps", type=int, default=1000)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched replicas, without the plain baselines (for profiling the batch)")
    ap.add_argument("--no-adopted", action="store_true",
                    help="skip the timing of the step kernels on OpenMM-layout (adopted) arrays")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only stream MIB MiB through the plain copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the integrator has no cross-GPU exchange to measure)", file=sys.stderr)
        sys.exit(2)

    from openmm_rigidbody_plugin_amd import rbhip
    if rbhip.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        ms = rbhip.copy_calibration(nbytes, reps=20)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "avg_ms": ms,
                          "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched": batched(args.batch_steps, args.warmup, plain=False)}), flush=True)
        return
    t0 = time.perf_counter()
    spec, cfg = make_workload(args.config)
    log(f"workload {args.config}: {spec.num_atoms} atoms built in {time.perf_counter() - t0:.1f}s")
    ctx = rbhip.Context.from_spec(spec, precision=cfg["precision"])
    counts = ctx.counts()
    nb, nf = counts["numBodies"], counts["numFree"]
    if args.launch:
        shape = [int(v) for v in args.launch.split(",")]
        ctx.set_launch(shape[0], shape[1] if len(shape) > 1 else 0)
        log(f"launch {ctx.describe()}")
    elif not args.no_autotune:
        t = time.perf_counter()
        ctx.autotune(reps=3)
        log(f"autotune {time.perf_counter() - t:.1f}s: {ctx.describe()}")

    ctx.step(args.warmup)
    rbhip.synchronize()
    t = time.perf_counter()
    ctx.step(args.steps)
    rbhip.synchronize()
    elapsed = time.perf_counter() - t
    ms_per_step = elapsed / args.steps * 1e3
    value = nb * args.steps / elapsed
    log(f"{args.steps} steps in {elapsed:.4f}s: {ms_per_step:.4f} ms/step, {value:.4e} body-steps/s")

    # per-kernel timing: HIP events attached to each step kernel's dispatch
    # (hipExtLaunchKernelGGL) on the context's stream, so each mean is the
    # kernel's own duration, as rocprofv3's kernel trace reports it
    kms, span = ctx.time_kernels(args.kernel_steps)
    log("kernel ms: " + ", ".join(f"{k}={v:.4f}" for k, v in kms.items()))
    natoms_body = counts["numBodyAtoms"] / max(nb, 1)
    # (one body class in these workloads: shared frames cover all bodies or none)
    roof = kernel_roofline(kms, natoms_body, nb, nf, args.config, "sharedShapes=" in ctx.describe())
    ke = ctx.kinetic_energies()
    assert np.all(np.isfinite(ke)), "non-finite kinetic energy"

    result = {
        "metric": "rigid body-steps/sec at 1M 3-site bodies, % HBM peak; batched p50/p99 latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 accumulation (mixed)",
        "data": "synthetic (seeded lattice, Maxwell-Boltzmann 300 K, harmonic tethers)",
        "config": {**cfg, "parallelism": "single GPU", "launch": ctx.describe()},
        "roofline": roof["roofline"],
        "kernel_ms": kms,
        "integrator": roof["integrator"],
    }
    if args.config == "water1M" and not args.no_adopted:
        del ctx
        ctx = None
        log("the same system on adopted OpenMM-layout arrays ...")
        result["integrator"]["adopted"] = adopted_leg(spec, cfg["precision"], args.kernel_steps)
    if args.config == "water1M" and not args.no_hetero:
        del ctx
        log("heterogeneous 200k bodies + 500k free atoms, exact rotation ...")
        result["hetero200k"] = hetero_leg(args.hetero_steps, args.warmup)
    if not args.no_batch and args.config == "water1M":
        ctx = None
        log("batched replicas ...")
        result["batched"] = batched(args.batch_steps, args.warmup)
    if not args.no_serve and args.config == "water1M":
        log("served job stream ...")
        result["served"] = served(num_jobs=args.serve_jobs)
    check = result.get("served", {}).pop("_fp64_check_jobs", None)
    if not args.no_cpu_baseline:
        # the CPU leg: the oracle timed on this host, and the served jobs'
        # FP64 agreement on a sample
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(spec)
        log(f"cpu baseline: {result['cpu_baseline']}")
        if check:
            result["served"]["fp64_check"] = served_fp64_check(check)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
