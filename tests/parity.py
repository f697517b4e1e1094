"""Shared helpers of the parity tests: the workloads and the GPU-vs-oracle step comparison.

Tolerances (FP32 GPU vs FP64 oracle on identical FP32 inputs):
  positions, velocities, r3d: |gpu - ref| <= 1e-5 * max(1, |ref|)
  headings: compared modulo 2 pi, same bound
  neighbour sets, neighbour counts: exact
  point location (face, vertex id): exact, against the oracle's remap of the GPU's stored position
"""
from __future__ import annotations

import functools
import importlib

import numpy as np

TOL = 1e-5

# workloads: (ellipsoid frequency, sigma, particles, init seed, noise seed)
C1 = dict(freq=22, sigma=0.1, n=100, seed=0, noise_seed=0, steps=50)
C2 = dict(freq=32, sigma=0.05, n=2000, seed=1, noise_seed=1, steps=1000)
# large surface: unit sphere with radius 1 + 0.2 Y_3^2, icosphere frequency 100 (F = 200,000)
C5 = dict(freq=100, amp=0.2, sigma=0.01, n=100000, seed=5, noise_seed=5, steps=500)
PHYS = dict(v0=1.0, k=10.0, mu=1.0, tau=1.0, dt=0.02)


def T():
    return importlib.import_module("2dtissue_amd")


@functools.lru_cache(maxsize=None)
def mesh_and_tables(freq: int, sigma_max: float):
    t = T()
    m = t.Mesh.ellipsoid(2.0, 1.0, 1.0, freq).build_tables(sigma_max)
    return m, m.export()


@functools.lru_cache(maxsize=None)
def sphere_and_tables(freq: int, amp: float, sigma_max: float):
    t = T()
    m = t.Mesh.perturbed_sphere(amp, freq).build_tables(sigma_max)
    return m, m.export()


def angle_diff(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return np.abs((d + np.pi) % (2 * np.pi) - np.pi)


def close(gpu, ref, tol=TOL):
    gpu = np.asarray(gpu, np.float64); ref = np.asarray(ref, np.float64)
    return np.abs(gpu - ref) <= tol * np.maximum(1.0, np.abs(ref))


def compare_step(O, om, p, s_in, s_out, out, nbrs_gpu, eta, sparse=False):
    """Compare one GPU step (s_in -> s_out, outputs) with the oracle's step from the same inputs.
    sparse selects the oracle's O(N) step, bitwise equal to the dense one. Returns a dict of
    failure counts (all zero when the step matches)."""
    if sparse:
        st, ref_s, ref_o = O.step_sparse(om, p, s_in.x, s_in.y, s_in.theta, s_in.vid, eta)
    else:
        st, ref_s, ref_o = O.step(om, p, s_in.x, s_in.y, s_in.theta, s_in.vid, eta, remap_mode=1)
    fails = {}
    fails["status"] = int(st != 0)
    fails["x"] = int((~close(s_out.x, ref_s["x"])).sum())
    fails["y"] = int((~close(s_out.y, ref_s["y"])).sum())
    fails["vx"] = int((~close(out.vx, ref_o["vx"])).sum())
    fails["vy"] = int((~close(out.vy, ref_o["vy"])).sum())
    fails["theta"] = int((angle_diff(s_out.theta, ref_s["theta"]) > TOL * np.pi).sum())
    fails["nbr"] = int((out.nbr != ref_o["nbr"]).sum())
    ref_sets = (O.neighbor_sets_sparse if sparse else O.neighbor_sets)(om, s_in.vid, p["sigma"])
    fails["sets"] = int(sum(not np.array_equal(a, b) for a, b in zip(nbrs_gpu, ref_sets)))
    r3d, vid, face = O.remap(om, s_out.x, s_out.y, 1)
    fails["vid"] = int((vid != s_out.vid).sum())
    fails["face"] = int((face != out.face).sum())
    fails["r3d"] = int((~close(out.r3d, r3d)).any(axis=1).sum())
    return fails, ref_o
