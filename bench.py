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
    has picked the execution mode (tiled step kernel or persistent kernel) and launch shape. Compared with launching the plain
    per-step kernel once per replica (timed on 16 replicas spread over the size range).
  cpu_baseline: the FP64 oracle's dense reference algorithm on a bounded sample, single thread.
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

WORKLOAD = dict(freq=32, sigma=0.05, n=2000, seed=1, noise_seed=1, noise_std=0.1, v0=1.0, k=10.0, mu=1.0, tau=1.0,
                dt=0.02)


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def roofline(particles: int, avg_launch_s: float, kernel: str, steps_per_launch: int = 1) -> dict:
    bytes_per_launch = BYTES_PER_PARTICLE_STEP * particles * steps_per_launch
    achieved = bytes_per_launch / avg_launch_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
            "avg_launch_us": avg_launch_s * 1e6, "bytes_per_launch": bytes_per_launch,
            "steps_per_launch": steps_per_launch}


def cpu_baseline(tables, x, y, th, vid, noise, steps, threads=1):
    """The oracle's FP64 dense reference algorithm (N x N matrices, brute-force remap over all
    faces, restart tiling loop), timed on this host for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.set_threads(threads)
    om = O.Mesh(tables)
    w = WORKLOAD
    p = dict(sigma=w["sigma"], v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    t0 = time.perf_counter()
    for t in range(steps):
        st, s, _ = O.step(om, p, x, y, th, vid, noise[t].astype(np.float64), remap_mode=0)
        x, y, th, vid = s["x"], s["y"], s["theta"], s["vid"]
    dt = time.perf_counter() - t0
    return dict(value=w["n"] * steps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"{steps} steps of the {w['n']}-particle workload (FP64, dense N^2 forces/alignment, "
                       f"brute-force remap over all {tables['nf']} faces), {dt:.1f} s wall")


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
    sim.synchronize()
    t0 = time.perf_counter()
    code = sim.step_staged(steps, warmup)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    res = dict(code=code, elapsed=elapsed, kernel_ms=sim.last_step_ms(), flags=sim.flags()[0],
               phi_polar_last=float(sim.order()[1][-1]))
    sim.close()
    return res, s0, noise_h


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
    lanes, block = shape["lanes"], shape["block"]
    log(f"batched: {shape}; timing {steps} steps")
    b.run(steps, 0)  # warm the graph (captured and instantiated on first use)
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
                                 "k_step_tiled (one step, all replicas)" if shape["mode"] == 1
                                 else "k_resident (one step, all replicas)"))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (replicas would shard "
              "across GPUs without per-step communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    T = importlib.import_module("2dtissue_amd")
    w = WORKLOAD
    log("building mesh tables")
    mesh = T.Mesh.ellipsoid(2.0, 1.0, 1.0, w["freq"]).build_tables(w["sigma"])
    tables = mesh.export()

    log(f"plain, default mode: warmup {args.warmup}, timing {args.steps} steps")
    main_res, s0, noise_h = plain(T, mesh, T.MODE_AUTO, args.lanes, args.steps, args.warmup)
    log(f"plain, persistent kernel: timing {args.steps} steps")
    res_res, _, _ = plain(T, mesh, T.MODE_RESIDENT, args.lanes, args.steps, args.warmup)

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
        },
        "roofline": roofline(n, main_res["kernel_ms"] / 1e3 / args.steps, "k_step"),
        "plain_resident_kernel": {
            "value": n * args.steps / res_res["elapsed"], "unit": "particle-steps/s",
            "ms_per_step": res_res["elapsed"] * 1e3 / args.steps, "status": res_res["code"],
            "roofline": roofline(n, res_res["kernel_ms"] / 1e3 / args.steps, "k_resident (one step)"),
        },
    }
    if not args.no_batched:
        out["batched"] = batched(T, mesh, args.batched_steps)
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_steps} oracle steps")
        out["cpu_baseline"] = cpu_baseline(tables, s0.x.astype(np.float64), s0.y.astype(np.float64),
                                           s0.theta.astype(np.float64), s0.vid, noise_h, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
