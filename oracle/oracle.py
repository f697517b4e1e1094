"""ctypes wrapper of liboracle.so, the FP64 CPU oracle (oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OK, E_INVALID, E_LOST_PARTICLE, E_NONFINITE, E_TILING = 0, 1, 2, 3, 4


class OrMesh(C.Structure):
    _fields_ = [("nv", C.c_int32), ("nf", C.c_int32), ("uv", C.c_void_p), ("xyz", C.c_void_p),
                ("faces", C.c_void_p), ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("dist", C.c_void_p),
                ("grid", C.c_int32), ("cell_start", C.c_void_p), ("cell_faces", C.c_void_p)]


class OrParams(C.Structure):
    _fields_ = [("sigma", C.c_double), ("v0", C.c_double), ("k", C.c_double), ("mu", C.c_double),
                ("tau", C.c_double), ("dt", C.c_double), ("max_crossings", C.c_int32)]


_lib = None
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_point_triangle_distance.restype = C.c_double
        L.or_order_parameter.restype = C.c_double
        L.or_dist_lookup.restype = C.c_double
        L.or_field_eval.restype = C.c_double
        L.or_repulsive_adhesion_motion.argtypes = [C.c_double] * 7 + [_D]
        L.or_repulsive_adhesion_motion.restype = None
        L.or_count_particle_neighbors.argtypes = [_D, C.c_int32, C.c_double, _I]
        L.or_average_n_within_distance_deg.argtypes = [_D, C.c_int32, _I, C.c_double]
        L.or_calculate_forces_between_particles.argtypes = [_D, _D, _D, C.c_int32, C.c_double, C.c_double,
                                                            C.c_double, C.c_double, _D]
        L.or_neighbor_sets.argtypes = [C.c_void_p, _I, C.c_int32, C.c_double, _I, _I, C.c_int32]
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


class Mesh:
    """The tables of one tissue mesh (dict from the product's tissue_mesh_export) as an or_mesh."""

    def __init__(self, tables: dict):
        self.t = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in tables.items()}
        t = self.t
        self.s = OrMesh(t["nv"], t["nf"], t["uv"].ctypes.data, t["xyz"].ctypes.data, t["faces"].ctypes.data,
                        t["row_ptr"].ctypes.data, t["col"].ctypes.data, t["dist"].ctypes.data, t["grid"],
                        t["cell_start"].ctypes.data, t["cell_faces"].ctypes.data)

    @property
    def ref(self):
        return C.byref(self.s)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def step(mesh: Mesh, p: dict, x, y, theta, vid, noise=None, remap_mode=1):
    """One continuous-mode FP64 step. Inputs are copied; returns (status, state, outputs)."""
    n = len(x)
    x = np.array(x, np.float64); y = np.array(y, np.float64); th = np.array(theta, np.float64)
    v = np.array(vid, np.int32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
    vx = np.zeros(n); vy = np.zeros(n); r3d = np.zeros((n, 3)); face = np.zeros(n, np.int32)
    nbr = np.zeros(n, np.int32); phi = np.zeros(2)
    prm = OrParams(p["sigma"], p["v0"], p["k"], p["mu"], p["tau"], p["dt"], p.get("max_crossings", 0) or 16384)
    st = lib().or_step(mesh.ref, C.byref(prm), n, dp(x), dp(y), dp(th), ip(v), None if nz is None else dp(nz),
                       remap_mode, dp(vx), dp(vy), dp(r3d), ip(face), ip(nbr), dp(phi))
    return st, dict(x=x, y=y, theta=th, vid=v), dict(vx=vx, vy=vy, r3d=r3d, face=face, nbr=nbr, phi_ref=phi[0],
                                                    phi_polar=phi[1])


def step_sparse(mesh: Mesh, p: dict, x, y, theta, vid, noise=None):
    """or_step_sparse: the same step as step(..., remap_mode=1), O(N) per step."""
    n = len(x)
    x = np.array(x, np.float64); y = np.array(y, np.float64); th = np.array(theta, np.float64)
    v = np.array(vid, np.int32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
    vx = np.zeros(n); vy = np.zeros(n); r3d = np.zeros((n, 3)); face = np.zeros(n, np.int32)
    nbr = np.zeros(n, np.int32); phi = np.zeros(2)
    prm = OrParams(p["sigma"], p["v0"], p["k"], p["mu"], p["tau"], p["dt"], p.get("max_crossings", 0) or 16384)
    st = lib().or_step_sparse(mesh.ref, C.byref(prm), n, dp(x), dp(y), dp(th), ip(v), None if nz is None else dp(nz),
                              dp(vx), dp(vy), dp(r3d), ip(face), ip(nbr), dp(phi))
    return st, dict(x=x, y=y, theta=th, vid=v), dict(vx=vx, vy=vy, r3d=r3d, face=face, nbr=nbr, phi_ref=phi[0],
                                                    phi_polar=phi[1])


def neighbor_sets_sparse(mesh: Mesh, vid, sigma):
    n = len(vid)
    v = np.ascontiguousarray(vid, np.int32)
    cap = max(64 * n, 1024)
    while True:
        rp = np.zeros(n + 1, np.int32); cols = np.zeros(cap, np.int32)
        st = lib().or_neighbor_sets_sparse(mesh.ref, ip(v), n, C.c_double(sigma), ip(rp), ip(cols), cap)
        if st == OK:
            return [cols[rp[i]:rp[i + 1]].copy() for i in range(n)]
        cap *= 4


def step_int_degrees(mesh: Mesh, p: dict, x, y, n_deg, vid, remap_mode=1):
    n = len(x)
    x = np.array(x, np.float64); y = np.array(y, np.float64)
    nd = np.array(n_deg, np.int32); v = np.array(vid, np.int32)
    vx = np.zeros(n); vy = np.zeros(n); r3d = np.zeros((n, 3)); face = np.zeros(n, np.int32)
    nbr = np.zeros(n, np.int32); phi = np.zeros(2)
    prm = OrParams(p["sigma"], p["v0"], p["k"], p.get("mu", 1.0), p.get("tau", 1.0), p["dt"], 64)
    st = lib().or_step_int_degrees(mesh.ref, C.byref(prm), n, dp(x), dp(y), ip(nd), ip(v), remap_mode, dp(vx),
                                   dp(vy), dp(r3d), ip(face), ip(nbr), dp(phi))
    return st, dict(x=x, y=y, n=nd, vid=v), dict(vx=vx, vy=vy, r3d=r3d, face=face, nbr=nbr, phi_ref=phi[0],
                                                phi_polar=phi[1])


def remap(mesh: Mesh, x, y, mode=1):
    n = len(x)
    r = np.ascontiguousarray(np.stack([np.asarray(x, np.float64), np.asarray(y, np.float64)], 1))
    r3d = np.zeros((n, 3)); vid = np.zeros(n, np.int32); face = np.zeros(n, np.int32)
    lib().or_remap(mesh.ref, dp(r), n, mode, dp(r3d), ip(vid), ip(face))
    return r3d, vid, face


def seam_wrap(r, r_old, heading, unit_deg, max_crossings=64):
    r = np.array(r, np.float64).reshape(-1, 2).copy(); ro = np.array(r_old, np.float64).reshape(-1, 2).copy()
    h = np.array(heading, np.float64).copy(); n = len(h); cr = np.zeros(n, np.int32)
    st = lib().or_seam_wrap(dp(r), dp(ro), dp(h), n, int(unit_deg), int(max_crossings), ip(cr))
    return st, r, ro, h, cr


def neighbor_sets(mesh: Mesh, vid, sigma):
    n = len(vid)
    v = np.ascontiguousarray(vid, np.int32)
    cap = max(64 * n, 1024)
    while True:
        rp = np.zeros(n + 1, np.int32); cols = np.zeros(cap, np.int32)
        st = lib().or_neighbor_sets(C.addressof(mesh.s), ip(v), n, sigma, ip(rp), ip(cols), cap)
        if st == OK:
            return [cols[rp[i]:rp[i + 1]].copy() for i in range(n)]
        cap *= 4


# ---- reverse pass (oracle/adjoint.cpp) ----

class OrLoss(C.Structure):
    _fields_ = [("target_order", C.c_double), ("w_order", C.c_double), ("w_velocity", C.c_double),
                ("n_modes", C.c_int32), ("component", C.c_void_p), ("amp", C.c_void_p), ("kx", C.c_void_p),
                ("ky", C.c_void_p), ("phase", C.c_void_p)]


class Loss:
    """Trajectory loss w_order (phi_T - target)^2 + w_velocity mean |rdot - u(r)|^2; u is a sum of
    modes u_c(x, y) += amp sin(kx x + ky y + phase), c = component (0: x, 1: y)."""

    def __init__(self, target_order=0.8, w_order=1.0, w_velocity=0.1, component=(), amp=(), kx=(), ky=(),
                 phase=()):
        self.target_order, self.w_order, self.w_velocity = target_order, w_order, w_velocity
        self.component = np.ascontiguousarray(component, np.int32)
        self.amp = np.ascontiguousarray(amp, np.float64)
        self.kx = np.ascontiguousarray(kx, np.float64)
        self.ky = np.ascontiguousarray(ky, np.float64)
        self.phase = np.ascontiguousarray(phase, np.float64)
        self.s = OrLoss(target_order, w_order, w_velocity, len(self.amp), self.component.ctypes.data,
                        self.amp.ctypes.data, self.kx.ctypes.data, self.ky.ctypes.data, self.phase.ctypes.data)

    def field(self, x, y):
        """u(x, y) for arrays x, y (numpy restatement of the same sum)."""
        x = np.asarray(x, np.float64); y = np.asarray(y, np.float64)
        u = [np.zeros_like(x), np.zeros_like(x)]
        for c, a, kx, ky, ph in zip(self.component, self.amp, self.kx, self.ky, self.phase):
            u[c] = u[c] + a * np.sin(kx * x + ky * y + ph)
        return u[0], u[1]


def field_eval(loss: Loss, x: float, y: float):
    u = np.zeros(2); jac = np.zeros(4)
    lib().or_field_eval(C.byref(loss.s), C.c_double(x), C.c_double(y), dp(u), dp(jac))
    return u, jac.reshape(2, 2)


def trajectory(mesh: Mesh, p: dict, x, y, theta, vid, noise):
    """FP64 forward of len(noise) steps; returns (status, states dict of [T+1][n] arrays, velocities
    vx, vy [T][n] of each step)."""
    T, n = np.shape(noise)
    xs = np.zeros((T + 1, n)); ys = np.zeros((T + 1, n)); ths = np.zeros((T + 1, n))
    vids = np.zeros((T + 1, n), np.int32)
    vx = np.zeros((T, n)); vy = np.zeros((T, n))
    s = dict(x=np.array(x, np.float64), y=np.array(y, np.float64), theta=np.array(theta, np.float64),
             vid=np.array(vid, np.int32))
    status = OK
    for t in range(T):
        xs[t], ys[t], ths[t], vids[t] = s["x"], s["y"], s["theta"], s["vid"]
        st, s, out = step(mesh, p, s["x"], s["y"], s["theta"], s["vid"], noise[t])
        vx[t], vy[t] = out["vx"], out["vy"]
        if st != OK and status == OK:
            status = st
    xs[T], ys[T], ths[T], vids[T] = s["x"], s["y"], s["theta"], s["vid"]
    return status, dict(x=xs, y=ys, theta=ths, vid=vids), (vx, vy)


def trajectory_loss(states: dict, vel, loss: Loss) -> dict:
    """The loss of a trajectory from its states and the velocities of its steps (numpy)."""
    th = states["theta"][-1]
    n = th.shape[0]
    phi = np.hypot(np.cos(th).sum(), np.sin(th).sum()) / n
    lo = loss.w_order * (phi - loss.target_order) ** 2
    vx, vy = vel
    T = vx.shape[0]
    ux, uy = loss.field(states["x"][:T], states["y"][:T])
    lv = loss.w_velocity * np.mean((vx - ux) ** 2 + (vy - uy) ** 2)
    return dict(loss=lo + lv, loss_order=lo, loss_velocity=lv, phi_final=phi)


def adjoint_reverse(mesh: Mesh, p: dict, states: dict, loss: Loss) -> dict:
    """FP64 gradients along the given trajectory (decisions of step t from state t)."""
    xs = np.ascontiguousarray(states["x"], np.float64); ys = np.ascontiguousarray(states["y"], np.float64)
    ths = np.ascontiguousarray(states["theta"], np.float64); vids = np.ascontiguousarray(states["vid"], np.int32)
    T = xs.shape[0] - 1
    n = xs.shape[1]
    prm = OrParams(p["sigma"], p["v0"], p["k"], p["mu"], p["tau"], p["dt"], p.get("max_crossings", 0) or 16384)
    out = np.zeros(8); dth = np.zeros(n); dx = np.zeros(n); dy = np.zeros(n)
    st = lib().or_adjoint_reverse(mesh.ref, C.byref(prm), n, T, dp(xs), dp(ys), dp(ths), ip(vids), C.byref(loss.s),
                                  dp(out), dp(dth), dp(dx), dp(dy))
    if st != OK:
        raise RuntimeError(f"or_adjoint_reverse: status {st}")
    keys = ("loss", "loss_order", "loss_velocity", "phi_final", "d_tau", "d_mu", "d_v0", "d_k")
    r = {k: float(v) for k, v in zip(keys, out)}
    r.update(d_theta0=dth, d_x0=dx, d_y0=dy)
    return r
