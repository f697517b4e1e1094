#!/usr/bin/env python3
"""Benchmark of the per-step particle update on one MI355X.

Headline workload (the plain-method GPU configuration): ellipsoid with semi-axes (2, 1, 1) from
an icosphere of frequency 32 (V = 10,242, F = 20,480), 2,000 particles placed uniformly by area
(seed 1), sigma = 0.05, v0 = 1, mu = 1, k = 10, tau = 1, dt = 0.02, angular noise N(0, 0.1^2)
pre-generated (seed 1). A step is one update of all 2,000 particles. The noise is resident in
HBM before the timed region; the timed region is K steps bracketed by device synchronisation.
The library's default execution (one launch of the step kernel per step for a system this size)
is `value`; the persistent LDS-resident kernel on the same system is reported beside it.

Also reported:
  batched: 256 replicas on the same mesh, sizes uniform in [100, 4000] (seed 3), tau log-uniform
    in [0.1, 10], noise std uniform in [0, 0.5], v0 uniform in [0.5, 2]; replica r uses init seed
    2000 + r and noise seed 1000 + r; 500 steps in one hipGraph-captured run, after the autotuner
    has picked the execution mode (tiled step kernel or persistent kernel) and launch shape.
    Compared with launching the plain per-step kernel once per replica (timed on 16 replicas
    spread over the size range).
  adjoint: forward + adjoint of 1,000 particles over 200 steps (checkpoints every 20), loss
    (phi_T - 0.8)^2 + 0.1 mean |rdot - u|^2, particle-steps/s of forward and reverse together.
  large: perturbed sphere (F = 200,000), 100,000 particles with sigma 0.01 for 500 steps, and
    64 x 2,000 particles forward + adjoint together over 200 steps.
  cpu_baseline: the FP64 oracle's dense reference algorithm on a bounded sample, single thread;
    cpu_baseline_all_cores: the same on every core this process may use (affinity mask, capped
    by the cgroup quota), with the host's `lscpu` model name.
A secondary measurement that fails is reported as {"error": ...} in its own field; the headline
line is printed regardless.
Roofline: algorithmic bytes per particle-step (64 B: read x, y, theta, vertex id, eta; write x,
y, theta, vx, vy, r3d, vertex id, face, neighbour count) over the kernel's HIP-event time, against
8 TB/s.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_STEP = 64
FP32_PEAK_GFLOPS = 157300.0  # MI355X FP32 vector (packed); 78,600 unpacked
# algorithmic flops of one particle-step (transcendentals count 1): per neighbour pair within
# 2 sigma, per particle (speed, Euler step, heading), per candidate face of the remap
FLOP_PER_PAIR, FLOP_PER_PARTICLE, FLOP_PER_FACE = 20, 40, 40

WORKLOAD = dict(freq=32, sigma=0.05, n=2000, seed=1, noise_seed=1, noise_std=0.1, v0=1.0, k=10.0, mu=1.0, tau=1.0,
                dt=0.02)


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


COUNTERS = os.path.join(ROOT, "profiles", "step_kernel_counters.json")


def measured(workload: str | None, launch: dict | None):
    """rocprofv3 counters recorded for this workload's step kernel (profiles/step_kernel_counters.json:
    FETCH_SIZE / WRITE_SIZE scaled by the copy calibration, VALU instruction issue), used only when
    the launch shape is the profiled one."""
    if workload is None or launch is None:
        return None
    try:
        with open(COUNTERS) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    m = d.get("workloads", {}).get(workload)
    if not m or launch.get("lanes") != m["lanes"] or launch.get("block") != m["block"]:
        return None
    return m, d.get("source_commit")


def step_kernels(nmax: int, uv_vertices: int, what: str = "one step") -> str:
    """The kernels of one tiled step: replicas of up to 8,192 particles at most one per two vertices
    keep their vertex buckets raw (the library's rule, engine.cpp), so no finalize pass runs."""
    if nmax <= 8192 and 2 * nmax <= uv_vertices:
        return f"k_step_tiled ({what}; raw vertex buckets, no finalize pass)"
    return f"k_step_tiled + k_bucket_finalize ({what})"


def roofline(particles: int, avg_launch_s: float, kernel: str, steps_per_launch: int = 1,
             workload: str | None = None, launch: dict | None = None) -> dict:
    """HBM roofline of one step: algorithmic bytes (64 B per particle-step: the state read and
    written) over the step's average duration from HIP events on the stream the kernels run on."""
    bytes_per_launch = BYTES_PER_PARTICLE_STEP * particles * steps_per_launch
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
           "avg_launch_us": avg_launch_s * 1e6, "bytes_per_launch": bytes_per_launch,
           "steps_per_launch": steps_per_launch}
    m = measured(workload, launch)
    if m:
        c, commit = m
        out["traffic"] = c["traffic_bytes"]
        out["traffic_kernel"] = c["kernel"]
        out["valu_issue_frac"] = c["valu_issue_frac"]
        out["counters_source"] = f"profiles/step_kernel_counters.json (commit {commit})"
    return out


def valu_axis(ps: float, pairs_per_particle: float, faces_per_locate: float, issue_frac) -> dict:
    """Secondary axis: FP32 VALU throughput from the algorithmic flop count of a particle-step
    (FLOP_PER_PAIR x neighbour pairs + FLOP_PER_PARTICLE + FLOP_PER_FACE x candidate faces)."""
    flops = FLOP_PER_PAIR * pairs_per_particle + FLOP_PER_PARTICLE + FLOP_PER_FACE * faces_per_locate
    achieved = ps * flops / 1e9
    return {"bound": "fp32_valu", "achieved": achieved, "peak": FP32_PEAK_GFLOPS, "unit": "GFLOP/s",
            "frac": achieved / FP32_PEAK_GFLOPS, "flop_per_particle_step": flops,
            "pairs_per_particle": pairs_per_particle, "faces_per_locate": faces_per_locate,
            "valu_issue_frac": issue_frac}


def cpu_model() -> str:
    """The host CPU's model name as `lscpu` reports it, with its family / model / stepping numbers
    (a virtual machine may report only a generic name); falls back to /proc/cpuinfo."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f.setdefault(k.strip().lower(), v.strip())
        if "model name" in f:
            ids = [f"{k} {f[k]}" for k in ("cpu family", "model", "stepping") if k in f]
            return f["model name"] + (f" ({', '.join(ids)})" if ids else "")
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> dict:
    """Cores this process may run on: `nproc` (the affinity mask), os.cpu_count() (every core of
    the machine) and the cgroup CPU quota; `usable` is the affinity count capped by the quota."""
    nproc = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    usable = nproc if quota is None else max(1, min(nproc, int(quota + 0.5)))
    return dict(nproc=nproc, cpu_count=os.cpu_count(), cgroup_quota_cpus=quota, usable=usable)


def warm_graph(run, steps: int, row: int) -> None:
    """Captures and instantiates, outside the timed region, the step graph that a timed
    `run(steps, row)` will launch. The engine keys its graphs by step count, noise rows and the
    phase of its rotating buffers (state: 2, vertex buckets: 3), so after one untimed call it runs
    (-steps) % 6 more steps to bring the phase back."""
    run(steps, row)
    extra = (-steps) % 6
    if extra:
        run(extra, row)


def cpu_baseline(T, tables, s0, steps, threads=1, min_seconds=0.0):
    """The oracle's FP64 dense reference algorithm (N x N matrices, brute-force remap over all
    faces, restart tiling loop), timed on this host over a bounded sample: `steps` steps from the
    plain workload's initial state with its own noise rows (the workload's noise seed), continued
    up to 4 x `steps` while less than `min_seconds` have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(threads)
    om = O.Mesh(tables)
    w = WORKLOAD
    p = dict(sigma=w["sigma"], v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    noise = T.gen_noise(w["noise_seed"], w["noise_std"], 4 * steps, w["n"]).astype(np.float64)
    x, y, th, vid = (s0.x.astype(np.float64), s0.y.astype(np.float64), s0.theta.astype(np.float64), s0.vid)
    t0 = time.perf_counter()
    done = 0
    while done < len(noise) and (done < steps or time.perf_counter() - t0 < min_seconds):
        st, s, _ = O.step(om, p, x, y, th, vid, noise[done], remap_mode=0)
        x, y, th, vid = s["x"], s["y"], s["theta"], s["vid"]
        done += 1
    dt = time.perf_counter() - t0
    return dict(value=w["n"] * done / dt, unit="particle-steps/s", cores=threads, kind="port",
                cpu_model=cpu_model(), host=host_cores(),
                sample=f"{done} steps of the {w['n']}-particle workload (FP64, dense N^2 forces/alignment, "
                       f"brute-force remap over all {tables['nf']} faces) on {threads} thread(s), {dt:.1f} s wall")


def plain(T, mesh, mode, lanes, steps, warmup):
    w = WORKLOAD
    p = T.params(sigma=w["sigma"], v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    sim = T.Sim(mesh, p, w["n"])
    sim.set_mode(mode)
    sim.set_launch(lanes, 0)
    sim.init_random(w["seed"])
    s0 = sim.get_state()
    noise_h = T.gen_noise(w["noise_seed"], w["noise_std"], warmup + steps, w["n"])
    sim.stage_noise(noise_h)
    if warmup:
        sim.step_staged(warmup, 0)
    warm_graph(sim.step_staged, steps, warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    code = sim.step_staged(steps, warmup)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    pairs = float(np.mean([len(a) for a in sim.neighbor_lists()]))  # |A_i| of the final state
    res = dict(code=code, elapsed=elapsed, kernel_ms=sim.last_step_ms(), flags=sim.flags()[0],
               phi_polar_last=float(sim.order()[1][-1]), launch=sim.launch(), pairs=pairs)
    sim.close()
    return res, s0


def batched(T, mesh, steps, n_replicas=256, plain_sample=16):
    w = WORKLOAD
    rng = np.random.default_rng(3)
    sizes = rng.integers(100, 4001, n_replicas)
    tau = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n_replicas))
    std = rng.uniform(0.0, 0.5, n_replicas)
    v0 = rng.uniform(0.5, 2.0, n_replicas)
    ps = [T.params(sigma=w["sigma"], v0=float(v0[r]), k=w["k"], mu=w["mu"], tau=float(tau[r]), dt=w["dt"])
          for r in range(n_replicas)]
    init_seeds = np.arange(n_replicas, dtype=np.uint64) + 2000
    noise_seeds = np.arange(n_replicas, dtype=np.uint64) + 1000
    b = T.Batch(mesh, sizes, ps)
    b.init_random(init_seeds)
    log(f"batched: generating {steps} noise rows for {b.ntot} particles")
    b.gen_noise(noise_seeds, std, steps)
    log("batched: autotuning the execution mode and launch shape")
    b.autotune(20)
    shape = b.launch()
    log(f"batched: {shape}; timing {steps} steps")
    warm_graph(b.run, steps, 0)
    b.init_random(init_seeds)
    t0 = time.perf_counter()
    code = b.run(steps, 0)
    elapsed = time.perf_counter() - t0
    run_ms = b.last_run_ms()
    flags, _ = b.flags()
    ps_batched = b.ntot * steps / elapsed
    out = dict(workload=f"{n_replicas} replicas, sizes U[100,4000] (total {b.ntot}), {steps} steps",
               value=ps_batched, unit="particle-steps/s", total_particles=int(b.ntot), steps=steps,
               elapsed_s=elapsed, graph_ms=run_ms, status=code, replicas_flagged=int((flags != 0).sum()),
               launch=shape,
               roofline=roofline(b.ntot, run_ms / 1e3 / steps,
                                 step_kernels(int(np.max(sizes)), mesh.info()["nv"], "one step, all replicas")
                                 if shape["mode"] == 1 else "k_resident (one step, all replicas)",
                                 workload="batched" if shape["mode"] == 1 else None, launch=shape))
    b.close()
    # the same replicas, each launched alone through the plain per-step kernel
    pick = np.argsort(sizes)[np.linspace(0, n_replicas - 1, plain_sample).astype(int)]
    log(f"batched: timing {plain_sample} replicas alone with the per-step kernel")
    tot_ps, tot_t = 0, 0.0
    for r in pick:
        s = T.Sim(mesh, ps[r], int(sizes[r]))
        s.set_mode(T.MODE_STEP_KERNEL)
        s.init_random(int(init_seeds[r]))
        s.stage_noise(T.gen_noise(int(noise_seeds[r]), float(std[r]), steps, int(sizes[r])))
        warm_graph(s.step_staged, steps, 0)
        s.init_random(int(init_seeds[r]))
        s.synchronize()
        t0 = time.perf_counter()
        s.step_staged(steps, 0)
        s.synchronize()
        tot_t += time.perf_counter() - t0
        tot_ps += int(sizes[r]) * steps
        s.close()
    out["plain_per_replica"] = dict(value=tot_ps / tot_t, unit="particle-steps/s",
                                    sample=f"{plain_sample} replicas spread over the size range, {steps} steps each")
    out["speedup_vs_plain_per_replica"] = ps_batched / (tot_ps / tot_t)
    return out


def batched_sweep(T, mesh, steps, n_replicas=256, n=200):
    """A parameter sweep: 256 replicas of 200 particles each with the parameter spread of the
    batched workload (parameter seed 3, init seed 2000 + r, noise seed 1000 + r), `steps` steps.
    The autotuner's choice is timed, and beside it the persistent LDS-resident kernel and the tiled
    step kernel, each at the best of three launch shapes. One equal replica per CU is where the
    persistent kernel (a workgroup per replica, all steps of a launch in LDS) wins; with uneven
    sizes the largest replica's workgroup sets its pace and the tiled kernel wins."""
    w = WORKLOAD
    sizes = np.full(n_replicas, n)
    rng = np.random.default_rng(3)
    tau = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n_replicas))
    std = rng.uniform(0.0, 0.5, n_replicas)
    v0 = rng.uniform(0.5, 2.0, n_replicas)
    ps = [T.params(sigma=w["sigma"], v0=float(v0[r]), k=w["k"], mu=w["mu"], tau=float(tau[r]), dt=w["dt"])
          for r in range(n_replicas)]
    init_seeds = np.arange(n_replicas, dtype=np.uint64) + 2000
    noise_seeds = np.arange(n_replicas, dtype=np.uint64) + 1000
    b = T.Batch(mesh, sizes, ps)
    b.init_random(init_seeds)
    b.gen_noise(noise_seeds, std, steps)

    def timed():
        warm_graph(b.run, steps, 0)
        b.init_random(init_seeds)
        t0 = time.perf_counter()
        code = b.run(steps, 0)
        return code, b.ntot * steps / (time.perf_counter() - t0), b.last_run_ms() * 1e3 / steps, b.launch()

    log("batched, parameter sweep of equal replicas: autotuning")
    b.autotune(20)
    code, value, us, shape = timed()
    per_mode = {}
    for mode, name in ((T.MODE_RESIDENT, "resident"), (T.MODE_STEP_KERNEL, "tiled")):
        best = None
        for lanes, block in ((1, 512), (2, 512), (4, 512)) if mode == T.MODE_RESIDENT else ((2, 128), (4, 128), (8, 128)):
            b.set_mode(mode, -1)
            b.set_launch(lanes, block, 0)
            c, v, u, s = timed()
            if c == T.OK and (best is None or v > best["value"]):
                best = dict(value=v, us_per_step=u, launch=s)
        per_mode[name] = best
    b.close()
    return dict(workload=f"{n_replicas} replicas of {n} particles (total {int(sizes.sum())}), {steps} steps",
                value=value, unit="particle-steps/s", status=code, us_per_step=us, launch=shape,
                best_resident=per_mode["resident"], best_tiled=per_mode["tiled"],
                resident_over_tiled=per_mode["resident"]["value"] / per_mode["tiled"]["value"])


def adjoint(T, mesh, n=1000, steps=200, ckpt=20):
    """Forward + adjoint: 1,000 particles, 200 steps, noise seed 2, checkpoints every 20 steps,
    loss (phi_T - 0.8)^2 + 0.1 mean |rdot - u|^2 with the seeded field u (seed 4); gradients with
    respect to tau, mu, v0, k and the initial headings and positions."""
    w = WORKLOAD
    p = T.params(sigma=w["sigma"], v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    L = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    noise = T.gen_noise(2, 0.1, steps, n)
    sim = T.Sim(mesh, p, n)
    sim.stage_noise(noise)
    sim.init_random(2)
    sim.adjoint(steps, L, 0, ckpt)  # warm: graphs captured, buffers allocated
    sim.init_random(2)
    sim.synchronize()
    t0 = time.perf_counter()
    code, g, dth, _, _ = sim.adjoint(steps, L, 0, ckpt)
    elapsed = time.perf_counter() - t0
    sim.close()
    gpu_ms = g["forward_ms"] + g["reverse_ms"]
    r = roofline(n, gpu_ms / 1e3 / steps, "checkpointed forward + recompute + reverse (one step)")
    r["achieved"] *= 2  # forward + adjoint touch the state twice: 128 B per particle-step
    r["frac"] *= 2
    r["bytes_per_launch"] *= 2
    return dict(workload=f"{n} particles, {steps} steps, checkpoints every {ckpt}", status=code,
                value=n * steps / elapsed, unit="particle-steps/s (forward + adjoint)", elapsed_s=elapsed,
                forward_ms=g["forward_ms"], reverse_ms=g["reverse_ms"], loss=g["loss"],
                d_tau=g["d_tau"], d_mu=g["d_mu"], d_v0=g["d_v0"], d_k=g["d_k"],
                max_abs_d_theta0=float(np.abs(dth).max()), roofline=r)


def large(T, steps=500, adj_steps=200, cpu_threads=0):
    """The large surface: perturbed sphere r = 1 + 0.2 Y_3^2 (icosphere frequency 100, F = 200,000),
    sigma = 0.01. 100,000 particles (seed 5, noise seed 5) for `steps` forward steps; then 64
    replicas of 2,000 particles (tau log-uniform in [0.1, 10], seed 6) forward + adjoint together
    over `adj_steps` steps with per-replica gradients."""
    w = WORKLOAD
    log("large surface: building the frequency-100 perturbed sphere")
    t0 = time.perf_counter()
    mesh = T.Mesh.perturbed_sphere(0.2, 100).build_tables(0.01)
    build_s = time.perf_counter() - t0
    info = mesh.info()
    n = 100000
    p = T.params(sigma=0.01, v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    sim = T.Sim(mesh, p, n)
    sim.init_random(5)
    sim.stage_noise(T.gen_noise(5, 0.1, steps + 20, n))
    sim.step_staged(20, 0)
    warm_graph(sim.step_staged, steps, 20)
    sim.synchronize()
    s_cpu = sim.get_state() if cpu_threads else None
    log(f"large surface: {n} particles, {steps} steps")
    t0 = time.perf_counter()
    code = sim.step_staged(steps, 20)
    sim.synchronize()
    el = time.perf_counter() - t0
    shape = sim.launch()
    fwd = dict(value=n * steps / el, unit="particle-steps/s", status=code, elapsed_s=el, launch=shape,
               roofline=roofline(n, sim.last_step_ms() / 1e3 / steps, step_kernels(n, info["nv"]),
                                 workload="large", launch=shape))
    sim.close()
    cpu = None
    if cpu_threads:
        log(f"large surface: CPU, one step of the sparse oracle on {cpu_threads} threads")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.set_threads(cpu_threads)
        om = O.Mesh(mesh.export())
        eta = T.gen_noise(5, 0.1, 200, n).astype(np.float64)
        po = dict(sigma=0.01, v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
        x, y, th, vid = (s_cpu.x.astype(np.float64), s_cpu.y.astype(np.float64), s_cpu.theta.astype(np.float64),
                         s_cpu.vid)
        t0 = time.perf_counter()
        done, status = 0, 0
        while done < len(eta) and time.perf_counter() - t0 < 10.0:  # bounded sample: about 10 s
            status, st_, _ = O.step_sparse(om, po, x, y, th, vid, eta[done])
            x, y, th, vid = st_["x"], st_["y"], st_["theta"], st_["vid"]
            done += 1
            if status != O.OK:
                break
        dt = time.perf_counter() - t0
        cpu = dict(value=n * done / dt, unit="particle-steps/s", cores=cpu_threads, kind="port", cpu_model=cpu_model(),
                   status=int(status), sample=f"{done} steps of the {n}-particle state, {dt:.1f} s wall",
                   algorithm="scalable CPU algorithm: the oracle's sparse O(N) restatement (geodesic-table rows "
                             "instead of N x N matrices, bitwise equal to the dense one); the reference's dense "
                             "algorithm is infeasible here (N x N matrices of 80 GB)")
    R, nr = 64, 2000
    rng = np.random.default_rng(6)
    taus = np.exp(rng.uniform(np.log(0.1), np.log(10.0), R))
    ps = [T.params(sigma=0.01, v0=w["v0"], k=w["k"], mu=w["mu"], tau=float(t), dt=w["dt"]) for t in taus]
    seeds = np.arange(R, dtype=np.uint64) + 300
    b = T.Batch(mesh, [nr] * R, ps)
    b.gen_noise(seeds, np.full(R, 0.1), adj_steps)
    L = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    b.init_random(seeds)
    b.adjoint(adj_steps, L, 0, 20)  # warm
    b.init_random(seeds)
    log(f"large surface: {R} x {nr} forward + adjoint, {adj_steps} steps")
    t0 = time.perf_counter()
    code, gs, _, _, _ = b.adjoint(adj_steps, L, 0, 20)
    el = time.perf_counter() - t0
    b.close()
    adj = dict(value=R * nr * adj_steps / el, unit="particle-steps/s (forward + adjoint)", status=code, elapsed_s=el,
               forward_ms=gs[0]["forward_ms"], reverse_ms=gs[0]["reverse_ms"],
               d_tau_first=[g["d_tau"] for g in gs[:4]])
    return dict(mesh=dict(faces=info["nf"], uv_vertices=info["nv"], table_entries=info["nnz"], build_s=build_s),
                cpu_baseline_scalable=cpu,
                forward=dict(workload=f"{n} particles, sigma 0.01, {steps} steps", **fwd),
                batched_adjoint=dict(workload=f"{R} x {nr} particles, {adj_steps} steps, checkpoints every 20", **adj))


def copy_calibration(mib: int, reps: int = 5) -> dict:
    """A plain device-to-device copy of known size (torch's copy kernel), for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profile to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return dict(bytes_per_copy=4 * n, copies=reps)


def leg(out: dict, key: str, fn, *args, **kw) -> None:
    """Runs one secondary measurement into out[key]; a failure is recorded there as
    {"error": ...} and the other legs, and the JSON line, still go ahead."""
    try:
        out[key] = fn(*args, **kw)
    except Exception as e:  # noqa: BLE001 -- every failure of a leg is reported, none is fatal
        log(f"{key}: FAILED: {type(e).__name__}: {e}")
        out[key] = {"error": f"{type(e).__name__}: {e}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-resident", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a device copy of MIB MiB (counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (replicas would shard "
              "across GPUs without per-step communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    calib = copy_calibration(args.copy_calibration) if args.copy_calibration > 0 else None

    T = importlib.import_module("2dtissue_amd")
    w = WORKLOAD
    log("building mesh tables")
    mesh = T.Mesh.ellipsoid(2.0, 1.0, 1.0, w["freq"]).build_tables(w["sigma"])
    tables = mesh.export()
    info = mesh.info()

    log(f"plain, default mode: warmup {args.warmup}, timing {args.steps} steps")
    main_res, s0 = plain(T, mesh, T.MODE_AUTO, args.lanes, args.steps, args.warmup)

    n = w["n"]
    ps = n * args.steps / main_res["elapsed"]
    out = {
        "metric": "particle-steps/s on 1 GPU (plain method, 2000-particle system)",
        "value": ps,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["elapsed"] * 1e3 / args.steps,
        "gpu_ms_per_step": main_res["kernel_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain step: ellipsoid (2,1,1) icosphere freq 32 (V=10242, F=20480), 2000 particles, "
                        "sigma=0.05, v0=1, mu=1, k=10, tau=1, dt=0.02, noise N(0,0.1^2); one launch per step",
            "particles": n, "faces": tables["nf"], "uv_vertices": tables["nv"], "status": main_res["code"],
            "flags": main_res["flags"],
            "timing": "value and ms_per_step: host wall clock around one call of `steps` steps between device "
                      "synchronisations; gpu_ms_per_step: HIP events on the simulation's stream around the "
                      "same steps",
        },
        "roofline": roofline(n, main_res["kernel_ms"] / 1e3 / args.steps, step_kernels(n, tables["nv"]),
                             workload="plain", launch=main_res["launch"]),
        "valu": valu_axis(ps, main_res["pairs"], info["n_cell_faces"] / info["grid"] ** 2,
                          (measured("plain", main_res["launch"]) or ({},))[0].get("valu_issue_frac")),
        "launch": main_res["launch"],
    }

    def resident():
        r, _ = plain(T, mesh, T.MODE_RESIDENT, args.lanes, args.steps, args.warmup)
        return {"value": n * args.steps / r["elapsed"], "unit": "particle-steps/s",
                "ms_per_step": r["elapsed"] * 1e3 / args.steps, "status": r["code"],
                "roofline": roofline(n, r["kernel_ms"] / 1e3 / args.steps, "k_resident (one step)")}

    if not args.no_resident:
        log(f"plain, persistent kernel: timing {args.steps} steps")
        leg(out, "plain_resident_kernel", resident)
    if not args.no_batched:
        leg(out, "batched", batched, T, mesh, args.batched_steps)
        leg(out, "batched_sweep", batched_sweep, T, mesh, args.batched_steps)
    if not args.no_adjoint:
        log("forward + adjoint: 1000 particles, 200 steps")
        leg(out, "adjoint", adjoint, T, mesh)
    cores = host_cores()
    threads = cores["usable"]
    if not args.no_large:
        leg(out, "large", large, T, cpu_threads=0 if args.no_cpu else threads)
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_steps} oracle steps on 1 thread")
        leg(out, "cpu_baseline", cpu_baseline, T, tables, s0, args.cpu_steps)
        log(f"CPU baseline on all {threads} usable cores ({cores})")
        leg(out, "cpu_baseline_all_cores", cpu_baseline, T, tables, s0, args.cpu_steps, threads=threads,
            min_seconds=10.0)
    if calib:
        out["copy_calibration"] = calib
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
