"""Shared helpers of the parity tests: the workloads and the GPU-vs-oracle step comparison.

Tolerances (FP32 GPU vs FP64 oracle on identical FP32 inputs):
  positions, velocities, r3d: |gpu - ref| <= 1e-5 * max(1, |ref|)
  headings: compared modulo 2 pi, same bound
  neighbour sets, neighbour counts: exact
  point location (face, vertex id): exact, against the oracle's remap of the GPU's stored position
  order parameters (LinearAlgebra.cpp:90-105 as called at 2DTissue.cpp:242, z := 0):
    phi_polar = |sum (cos theta, sin theta)| / N: |gpu - ref| <= 1e-5 * ref
    phi_ref = |sum sign(x vy - y vx)| / N: exact. The GPU's value must equal the host's FP64
      evaluation of the same formula on the GPU's own FP32 outputs bit for bit, and each particle's
      sign must equal the oracle's, except for particles whose |x vy - y vx| lies within the
      rounding bound of the compared quantities (the TOL bounds on x, y, vx, vy propagated
      through the cross product); those are counted and reported, and N |phi_gpu - phi_ref| may
      differ by at most twice their number.
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


def last_phi(sim):
    """(phi_ref, phi_polar) of the last step a Sim ran (its order-parameter history's last entry)."""
    a, b = sim.order()
    return float(a[-1]), float(b[-1])


def cross_sign_bound(x, y, vx, vy, tol=TOL):
    """Per particle: how far x vy - y vx can move when each of x, y, vx, vy moves by the parity
    tolerance tol * max(1, |value|) (first-order propagation, doubled for the second-order term)."""
    ax, ay, avx, avy = (np.abs(np.asarray(a, np.float64)) for a in (x, y, vx, vy))
    ex, ey, evx, evy = (tol * np.maximum(1.0, a) for a in (ax, ay, avx, avy))
    return 2.0 * (ex * avy + ax * evy + ey * avx + ay * evx + ex * evy + ey * evx)


def compare_phi(s_out, out, ref_s, ref_o, phi_gpu):
    """phi_ref and phi_polar of one GPU step against the oracle (tolerances in the module doc).
    Returns (failure counts, number of particles exempt from the sign comparison)."""
    n = len(s_out.x)
    x, y = s_out.x.astype(np.float64), s_out.y.astype(np.float64)
    vx, vy = out.vx.astype(np.float64), out.vy.astype(np.float64)
    cz_gpu = x * vy - y * vx  # the kernel's FP64 evaluation on its FP32 outputs (no contraction)
    sg_gpu = np.sign(cz_gpu).astype(np.int64)
    cz_ref = ref_s["x"] * ref_o["vy"] - ref_s["y"] * ref_o["vx"]
    sg_ref = np.sign(cz_ref).astype(np.int64)
    exempt = np.abs(cz_ref) <= cross_sign_bound(ref_s["x"], ref_s["y"], ref_o["vx"], ref_o["vy"])
    fails = {}
    fails["phi_ref_sign"] = int(((sg_gpu != sg_ref) & ~exempt).sum())
    fails["phi_ref_self"] = int(phi_gpu[0] != (1.0 / n) * abs(float(sg_gpu.sum())))
    fails["phi_ref"] = int(abs(n * phi_gpu[0] - n * ref_o["phi_ref"]) > 2 * int(exempt.sum()) + 1e-9)
    fails["phi_polar"] = int(abs(phi_gpu[1] - ref_o["phi_polar"]) > TOL * ref_o["phi_polar"])
    return fails, int(exempt.sum())


def visit_mix(j):
    """The MurmurHash3 32-bit finaliser, elementwise on uint32 (the step kernel's visit_mix)."""
    h = np.asarray(j, dtype=np.uint32).copy()
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x85EBCA6B)
    h ^= h >> np.uint32(13)
    h *= np.uint32(0xC2B2AE35)
    h ^= h >> np.uint32(16)
    return h


def visit_records(sets, off=0):
    """(na, hash) the visit check must report for neighbour sets given as ids within a replica
    starting at global index off: the set sizes and the sums mod 2^32 of visit_mix(off + j)."""
    na = np.array([len(s) for s in sets], np.int32)
    h = np.array([int(visit_mix(np.asarray(s, np.int64) + off).sum(dtype=np.uint64) & 0xFFFFFFFF) for s in sets],
                 np.uint32)
    return na, h


def compare_step(O, om, p, s_in, s_out, out, nbrs_gpu, eta, sparse=False, phi_gpu=None, visits=None):
    """Compare one GPU step (s_in -> s_out, outputs) with the oracle's step from the same inputs.
    sparse selects the oracle's O(N) step, bitwise equal to the dense one. phi_gpu: the step's
    (phi_ref, phi_polar) from the GPU (compared when given). visits: the step's own (na, hash)
    records (Sim.visits), compared exactly with the oracle's neighbour sets of the input state when
    given. Returns a dict of failure counts (all
    zero when the step matches); the count of particles exempt from the phi_ref sign comparison is
    returned in the second value's "phi_exempt"."""
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
    if visits is not None:
        na, h = visit_records(ref_sets)
        fails["visits"] = int(((visits[0] != na) | (visits[1] != h)).sum())
    r3d, vid, face = O.remap(om, s_out.x, s_out.y, 1)
    fails["vid"] = int((vid != s_out.vid).sum())
    fails["face"] = int((face != out.face).sum())
    fails["r3d"] = int((~close(out.r3d, r3d)).any(axis=1).sum())
    if phi_gpu is not None:
        f_phi, exempt = compare_phi(s_out, out, ref_s, ref_o, phi_gpu)
        fails.update(f_phi)
        ref_o = dict(ref_o, phi_exempt=exempt)
    return fails, ref_o


def teacher_forced(O, om, p, sim, noise, visits=False, sparse=False):
    """Teacher-forced run of a Sim: len(noise) single steps, each recomputed by the oracle from
    the GPU's own input state and compared with compare_step (with the step's own neighbour walk
    when visits is set: the Sim must have its visit check on). Returns (failure counts summed over
    the steps, particle-steps exempt from the phi_ref sign comparison)."""
    total, exempt = {}, 0
    for t in range(len(noise)):
        s_in = sim.get_state()
        nb = sim.neighbor_lists()
        assert sim.step(1, noise[t]) == T().OK, t
        fails, ref = compare_step(O, om, p, s_in, sim.get_state(), sim.get_outputs(), nb,
                                  np.asarray(noise[t], np.float64), sparse, phi_gpu=last_phi(sim),
                                  visits=sim.visits() if visits else None)
        exempt += ref["phi_exempt"]
        for k, v in fails.items():
            total[k] = total.get(k, 0) + v
        if any(fails.values()):
            print(f"step {t}: {fails}", flush=True)
    return total, exempt
