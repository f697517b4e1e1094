#!/usr/bin/env python3
"""Benchmark of the per-step particle update on one MI355X.

Workload (the plain-method GPU configuration): ellipsoid with semi-axes (2, 1, 1) from an
icosphere of frequency 32 (V = 10,242, F = 20,480), 2,000 particles placed uniformly by area
(seed 1), sigma = 0.05, v0 = 1, mu = 1, k = 10, tau = 1, dt = 0.02, angular noise N(0, 0.1^2)
pre-generated (seed 1). A step is one update of all 2,000 particles. The noise array is resident
in HBM before the timed region; the timed region is K steps bracketed by device synchronisation.

Prints one JSON line: particle-steps/s, the roofline of the per-step kernel (algorithmic bytes
per particle-step / HIP-event kernel time vs 8 TB/s), and the FP64 CPU baseline (the oracle's
dense reference algorithm on a bounded sample, single-threaded).
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
BYTES_PER_PARTICLE_STEP = 64  # read x,y,theta,vid,eta (20 B); write x,y,theta,vx,vy,r3d,vid,face,nbr (44 B)

WORKLOAD = dict(freq=32, sigma=0.05, n=2000, seed=1, noise_seed=1, noise_std=0.1, v0=1.0, k=10.0, mu=1.0, tau=1.0,
                dt=0.02)


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
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
    p = T.params(sigma=w["sigma"], v0=w["v0"], k=w["k"], mu=w["mu"], tau=w["tau"], dt=w["dt"])
    log("creating the simulation on the GPU")
    sim = T.Sim(mesh, p, w["n"])
    sim.set_launch(args.lanes, 0)
    sim.init_random(w["seed"])
    s0 = sim.get_state()
    total = args.warmup + args.steps
    noise_h = T.gen_noise(w["noise_seed"], w["noise_std"], total, w["n"])
    sim.stage_noise(noise_h)  # eta resident in HBM before the timed region

    log(f"warmup {args.warmup} steps")
    if args.warmup:
        sim.step_staged(args.warmup, 0)
    sim.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    code = sim.step_staged(args.steps, args.warmup)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = sim.last_step_ms()
    flags, _ = sim.flags()

    ps = w["n"] * args.steps / elapsed
    avg_launch_s = kernel_ms / 1e3 / args.steps
    achieved = BYTES_PER_PARTICLE_STEP * w["n"] / avg_launch_s / 1e9
    out = {
        "metric": "particle-steps/s on 1 GPU (plain per-step kernel)",
        "value": ps,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain step: ellipsoid (2,1,1) icosphere freq 32 (V=10242, F=20480), 2000 particles, "
                        "sigma=0.05, v0=1, mu=1, k=10, tau=1, dt=0.02, noise N(0,0.1^2)",
            "particles": w["n"], "faces": tables["nf"], "uv_vertices": tables["nv"], "status": code,
            "flags": flags, "lanes": args.lanes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_step",
            "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_launch": BYTES_PER_PARTICLE_STEP * w["n"],
        },
    }
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_steps} oracle steps")
        out["cpu_baseline"] = cpu_baseline(tables, s0.x.astype(np.float64), s0.y.astype(np.float64),
                                           s0.theta.astype(np.float64), s0.vid, noise_h, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
