"""Python host binding of libtissue (include/tissue.h) over ctypes.

The package directory starts with a digit, so import it with
``importlib.import_module("2dtissue_amd")``. The library is built in-tree by
``make -C 2dtissue_amd`` (``__graft_entry__.build()``); loading fails loudly when it is missing:
there is no CPU fallback.

Classes: ``Mesh`` (surface, UV chart and tables), ``Sim`` (one system: the per-step update of
KAIZN-io/2DTissue src/simulation/2DTissue.cpp:136-206 and its outputs), ``Batch`` (replicas with
their own sizes and parameters in one launch sequence), plus ``loss`` / ``Sim.adjoint`` for
gradients. ``SIGNATURES`` lists every C-ABI entry point with its ctypes argument types. The C++
drop-in of the reference's ``_2DTissue`` class is the header include/2DTissue.h.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtissue.so")

OK = 0
BUILD_AUTO, BUILD_HOST, BUILD_GPU = -1, 0, 1
E_INVALID_ARG = 1
E_LOST_PARTICLE = 2
E_NONFINITE = 3
E_TILING_NO_CONVERGE = 4
E_HIP = 5
E_IO = 6
E_MESH = 7
FLAG_LOST_PARTICLE = 1
FLAG_NONFINITE = 2
FLAG_TILING = 4
FLAG_BAD_INPUT = 8

_STATUS_NAMES = {1: "E_INVALID_ARG", 2: "E_LOST_PARTICLE", 3: "E_NONFINITE", 4: "E_TILING_NO_CONVERGE",
                 5: "E_HIP", 6: "E_IO", 7: "E_MESH"}


class TissueError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("sigma", C.c_double), ("v0", C.c_double), ("k", C.c_double), ("mu", C.c_double),
                ("tau", C.c_double), ("dt", C.c_double), ("max_crossings", C.c_int32), ("reserved", C.c_int32)]


def params(sigma=0.05, v0=1.0, k=10.0, mu=1.0, tau=1.0, dt=0.02, max_crossings=0) -> Params:
    return Params(sigma, v0, k, mu, tau, dt, max_crossings, 0)


MAX_MODES = 16


class Loss(C.Structure):
    """tissue_loss: w_order (phi_T - target_order)^2 + w_velocity mean |rdot - u(r)|^2, with
    u_c(x, y) = sum of amp sin(kx x + ky y + phase) over the modes of component c."""
    _fields_ = [("target_order", C.c_double), ("w_order", C.c_double), ("w_velocity", C.c_double),
                ("n_modes", C.c_int32), ("reserved", C.c_int32), ("component", C.c_int32 * MAX_MODES),
                ("amp", C.c_double * MAX_MODES), ("kx", C.c_double * MAX_MODES), ("ky", C.c_double * MAX_MODES),
                ("phase", C.c_double * MAX_MODES)]


class Grads(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("loss", "loss_order", "loss_velocity", "phi_final", "d_tau", "d_mu", "d_v0",
                                          "d_k", "forward_ms", "reverse_ms")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def loss(target_order=0.8, w_order=1.0, w_velocity=0.1, modes=()) -> Loss:
    """modes: (component, amp, kx, ky, phase) tuples of the target field u."""
    if len(modes) > MAX_MODES:
        raise ValueError(f"at most {MAX_MODES} modes")
    L = Loss()
    L.target_order, L.w_order, L.w_velocity, L.n_modes = target_order, w_order, w_velocity, len(modes)
    for m, (c, a, kx, ky, ph) in enumerate(modes):
        L.component[m], L.amp[m], L.kx[m], L.ky[m], L.phase[m] = int(c), a, kx, ky, ph
    return L


def target_field(seed: int = 4, per_component: int = 4) -> list:
    """A seeded smooth velocity field on the UV square: per component, modes with integer
    wave numbers in {-2..2} (times 2 pi, not both zero), amplitudes U[-1, 1], phases U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    modes = []
    for c in (0, 1):
        for _ in range(per_component):
            while True:
                kx, ky = rng.integers(-2, 3, 2)
                if kx or ky:
                    break
            modes.append((c, float(rng.uniform(-1, 1)), 2 * np.pi * float(kx), 2 * np.pi * float(ky),
                          float(rng.uniform(0, 2 * np.pi))))
    return modes


_lib = None

# (name, restype, argtypes) of every entry point declared in include/tissue.h
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
SIGNATURES = [
    ("tissue_error_string", C.c_char_p, []),
    ("tissue_mesh_ellipsoid", C.c_int32, [C.c_double, C.c_double, C.c_double, C.c_int32, C.POINTER(_P)]),
    ("tissue_mesh_perturbed_sphere", C.c_int32, [C.c_double, C.c_int32, C.POINTER(_P)]),
    ("tissue_mesh_load_off", C.c_int32, [C.c_char_p, C.POINTER(_P)]),
    ("tissue_mesh_build_tables", C.c_int32, [_P, C.c_double, C.c_int32]),
    ("tissue_mesh_build_tables_on", C.c_int32, [_P, C.c_double, C.c_int32, C.c_int32]),
    ("tissue_set_uv_solve", C.c_int32, [C.c_int32]),
    ("tissue_get_uv_solve", C.c_int32, [_I]),
    ("tissue_mesh_destroy", None, [_P]),
    ("tissue_mesh_info", C.c_int32, [_P, _I, _I, _I, C.POINTER(C.c_int64), _I, C.POINTER(C.c_int64)]),
    ("tissue_mesh_export", C.c_int32, [_P, _D, _D, _I, _I, _I, _F, _I, _I]),
    ("tissue_mesh_save_off", C.c_int32, [_P, C.c_char_p, C.c_char_p]),
    ("tissue_create", C.c_int32, [_P, C.POINTER(Params), C.c_int32, C.POINTER(_P)]),
    ("tissue_destroy", None, [_P]),
    ("tissue_init_random", C.c_int32, [_P, C.c_uint64]),
    ("tissue_set_state", C.c_int32, [_P, _F, _F, _F, _I]),
    ("tissue_get_state", C.c_int32, [_P, _F, _F, _F, _I]),
    ("tissue_get_outputs", C.c_int32, [_P, _F, _F, _F, _I, _I]),
    ("tissue_step", C.c_int32, [_P, C.c_int32, _F]),
    ("tissue_step_device", C.c_int32, [_P, C.c_int32, C.c_void_p]),
    ("tissue_stage_noise", C.c_int32, [_P, _F, C.c_int32]),
    ("tissue_step_staged", C.c_int32, [_P, C.c_int32, C.c_int32]),
    ("tissue_get_order", C.c_int32, [_P, _D, _D, C.c_int32, _I]),
    ("tissue_reset_order", C.c_int32, [_P]),
    ("tissue_get_flags", C.c_int32, [_P, C.POINTER(C.c_uint32), _I]),
    ("tissue_neighbor_lists", C.c_int32, [_P, _I, _I, C.c_int32]),
    ("tissue_set_visit_check", C.c_int32, [_P, C.c_int32]),
    ("tissue_get_visits", C.c_int32, [_P, _I, _I]),
    ("tissue_set_launch", C.c_int32, [_P, C.c_int32, C.c_int32]),
    ("tissue_synchronize", C.c_int32, [_P]),
    ("tissue_last_step_ms", C.c_int32, [_P, _D]),
    ("tissue_set_mode", C.c_int32, [_P, C.c_int32]),
    ("tissue_resident_capacity", C.c_int32, [_P, _I]),
    ("tissue_get_launch", C.c_int32, [_P, _I]),
    ("tissue_autotune", C.c_int32, [_P, C.c_int32, _I]),
    ("tissue_batch_create", C.c_int32, [_P, C.c_int32, _I, C.POINTER(Params), C.POINTER(_P)]),
    ("tissue_batch_destroy", None, [_P]),
    ("tissue_batch_init_random", C.c_int32, [_P, C.POINTER(C.c_uint64)]),
    ("tissue_batch_set_state", C.c_int32, [_P, _F, _F, _F, _I]),
    ("tissue_batch_get_state", C.c_int32, [_P, _F, _F, _F, _I]),
    ("tissue_batch_get_outputs", C.c_int32, [_P, _F, _F, _F, _I, _I]),
    ("tissue_batch_stage_noise", C.c_int32, [_P, _F, C.c_int32]),
    ("tissue_batch_gen_noise", C.c_int32, [_P, C.POINTER(C.c_uint64), _D, C.c_int32]),
    ("tissue_batch_run", C.c_int32, [_P, C.c_int32, C.c_int32]),
    ("tissue_batch_get_order", C.c_int32, [_P, C.c_int32, _D, _D, C.c_int32, _I]),
    ("tissue_batch_get_flags", C.c_int32, [_P, C.POINTER(C.c_uint32), _I]),
    ("tissue_batch_set_launch", C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32]),
    ("tissue_batch_set_mode", C.c_int32, [_P, C.c_int32, C.c_int32]),
    ("tissue_batch_get_launch", C.c_int32, [_P, _I]),
    ("tissue_batch_autotune", C.c_int32, [_P, C.c_int32, _I, _I]),
    ("tissue_batch_last_run_ms", C.c_int32, [_P, _D]),
    ("tissue_batch_set_visit_check", C.c_int32, [_P, C.c_int32]),
    ("tissue_batch_get_visits", C.c_int32, [_P, _I, _I]),
    ("tissue_adjoint", C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Loss), C.POINTER(Grads), _D, _D,
                                   _D]),
    ("tissue_batch_adjoint", C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Loss), C.POINTER(Grads), _D,
                                         _D, _D]),
    ("tissue_gen_noise", C.c_int32, [C.c_uint64, C.c_double, C.c_int32, C.c_int32, _F]),
    ("tissue_philox4x32", None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("tissue_seam_wrap", C.c_int32, [C.c_int32, _D, _D, _D, _I, C.c_int32, C.c_int32, C.c_int32]),
    ("tissue_random_init_host", C.c_int32, [_P, C.c_uint64, C.c_int32, _F, _F, _F]),
    ("tissue_record_start", C.c_int32, [_P, C.c_char_p, C.c_int32]),
    ("tissue_record_stop", C.c_int32, [_P]),
    ("tissue_step_noalloc", C.c_int32, [_P, C.POINTER(Params), C.c_int32] + [C.c_void_p] * 19),
]


def lib() -> C.CDLL:
    """Load libtissue.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C 2dtissue_amd`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code: int) -> None:
    if code != OK:
        raise TissueError(code, lib().tissue_error_string().decode(errors="replace"))


def _ptr(a: np.ndarray | None, ct):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def set_uv_solve(where: int) -> None:
    """Where the mesh constructors solve the UV map (tissue_set_uv_solve): BUILD_AUTO, BUILD_HOST or
    BUILD_GPU; the result is the same bit for bit."""
    _check(lib().tissue_set_uv_solve(where))


def get_uv_solve() -> int:
    w = C.c_int32()
    _check(lib().tissue_get_uv_solve(C.byref(w)))
    return w.value


class Mesh:
    """UV-parameterised closed surface with its step tables (tissue_mesh)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def ellipsoid(cls, a=2.0, b=1.0, c=1.0, freq=22) -> "Mesh":
        h = C.c_void_p()
        _check(lib().tissue_mesh_ellipsoid(a, b, c, freq, C.byref(h)))
        return cls(h)

    @classmethod
    def perturbed_sphere(cls, amp=0.2, freq=100) -> "Mesh":
        h = C.c_void_p()
        _check(lib().tissue_mesh_perturbed_sphere(amp, freq, C.byref(h)))
        return cls(h)

    @classmethod
    def load_off(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        _check(lib().tissue_mesh_load_off(path.encode(), C.byref(h)))
        return cls(h)

    def build_tables(self, sigma_max: float, grid: int = 0, where: int | None = None) -> "Mesh":
        """Face buckets and the exact geodesic table (tissue_mesh_build_tables); where = BUILD_HOST
        or BUILD_AUTO through tissue_mesh_build_tables_on (BUILD_GPU is rejected: the table is built
        on the host's threads), None: automatic."""
        if where is None:
            _check(lib().tissue_mesh_build_tables(self.h, sigma_max, grid))
        else:
            _check(lib().tissue_mesh_build_tables_on(self.h, sigma_max, grid, where))
        return self

    def info(self) -> dict:
        nv, nf, nv3d, grid = (C.c_int32() for _ in range(4))
        nnz, ncf = C.c_int64(), C.c_int64()
        _check(lib().tissue_mesh_info(self.h, C.byref(nv), C.byref(nf), C.byref(nv3d), C.byref(nnz), C.byref(grid),
                                      C.byref(ncf)))
        return dict(nv=nv.value, nf=nf.value, nv3d=nv3d.value, nnz=nnz.value, grid=grid.value,
                    n_cell_faces=ncf.value)

    def random_init_host(self, seed: int, n: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The initial (x, y, theta) that Sim.init_random(seed) draws on the GPU, computed on the host."""
        x = np.zeros(n, np.float32); y = np.zeros(n, np.float32); t = np.zeros(n, np.float32)
        _check(lib().tissue_random_init_host(self.h, seed, n, _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(t, C.c_float)))
        return x, y, t

    def export(self) -> dict:
        i = self.info()
        uv = np.zeros((i["nv"], 2)); xyz = np.zeros((i["nv"], 3))
        faces = np.zeros((i["nf"], 3), np.int32)
        row_ptr = np.zeros(i["nv"] + 1, np.int32)
        col = np.zeros(max(i["nnz"], 1), np.int32); dist = np.zeros(max(i["nnz"], 1), np.float32)
        g = i["grid"]
        cell_start = np.zeros(g * g + 1 if g else 1, np.int32)
        cell_faces = np.zeros(max(i["n_cell_faces"], 1), np.int32)
        _check(lib().tissue_mesh_export(self.h, _ptr(uv, C.c_double), _ptr(xyz, C.c_double), _ptr(faces, C.c_int32),
                                        _ptr(row_ptr, C.c_int32), _ptr(col, C.c_int32), _ptr(dist, C.c_float),
                                        _ptr(cell_start, C.c_int32), _ptr(cell_faces, C.c_int32)))
        out = dict(i)
        out.update(uv=uv, xyz=xyz, faces=faces, row_ptr=row_ptr, col=col[: i["nnz"]], dist=dist[: i["nnz"]],
                   cell_start=cell_start, cell_faces=cell_faces[: i["n_cell_faces"]])
        return out

    def save_off(self, path_3d: str | None, path_uv: str | None) -> None:
        _check(lib().tissue_mesh_save_off(self.h, path_3d.encode() if path_3d else None,
                                          path_uv.encode() if path_uv else None))

    def close(self):
        if self.h:
            lib().tissue_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class State:
    x: np.ndarray
    y: np.ndarray
    theta: np.ndarray
    vid: np.ndarray


@dataclass
class Outputs:
    vx: np.ndarray
    vy: np.ndarray
    r3d: np.ndarray
    face: np.ndarray
    nbr: np.ndarray


class Sim:
    """One simulation on the current GPU (tissue_sim)."""

    def __init__(self, mesh: Mesh, p: Params, n: int):
        self.mesh = mesh  # keeps the mesh alive
        self.n = n
        self.p = p
        h = C.c_void_p()
        _check(lib().tissue_create(mesh.h, C.byref(p), n, C.byref(h)))
        self.h = h

    def init_random(self, seed: int) -> None:
        _check(lib().tissue_init_random(self.h, seed))

    def set_state(self, x, y, theta, vid=None) -> None:
        x = np.ascontiguousarray(x, np.float32); y = np.ascontiguousarray(y, np.float32)
        t = np.ascontiguousarray(theta, np.float32)
        v = None if vid is None else np.ascontiguousarray(vid, np.int32)
        _check(lib().tissue_set_state(self.h, _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(t, C.c_float),
                                      _ptr(v, C.c_int32)))

    def get_state(self) -> State:
        n = self.n
        s = State(np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32))
        _check(lib().tissue_get_state(self.h, _ptr(s.x, C.c_float), _ptr(s.y, C.c_float), _ptr(s.theta, C.c_float),
                                      _ptr(s.vid, C.c_int32)))
        return s

    def get_outputs(self) -> Outputs:
        n = self.n
        o = Outputs(np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32),
                    np.zeros(n, np.int32), np.zeros(n, np.int32))
        _check(lib().tissue_get_outputs(self.h, _ptr(o.vx, C.c_float), _ptr(o.vy, C.c_float), _ptr(o.r3d, C.c_float),
                                        _ptr(o.face, C.c_int32), _ptr(o.nbr, C.c_int32)))
        return o

    def step(self, n_steps: int, noise: np.ndarray | None = None) -> int:
        """Run n_steps; returns the status code (validation errors are returned, not raised)."""
        if noise is not None:
            noise = np.ascontiguousarray(noise, np.float32)
            assert noise.size >= n_steps * self.n
        code = lib().tissue_step(self.h, n_steps, _ptr(noise, C.c_float))
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code

    def step_device(self, n_steps: int, noise_dev_ptr: int | None) -> int:
        code = lib().tissue_step_device(self.h, n_steps, C.c_void_p(noise_dev_ptr) if noise_dev_ptr else None)
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code

    def stage_noise(self, noise: np.ndarray) -> None:
        noise = np.ascontiguousarray(noise, np.float32).reshape(-1, self.n)
        _check(lib().tissue_stage_noise(self.h, _ptr(noise, C.c_float), noise.shape[0]))

    def step_staged(self, n_steps: int, first_row: int) -> int:
        code = lib().tissue_step_staged(self.h, n_steps, first_row)
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code

    def order(self) -> tuple[np.ndarray, np.ndarray]:
        cnt = C.c_int32()
        _check(lib().tissue_get_order(self.h, None, None, 0, C.byref(cnt)))
        a = np.zeros(cnt.value); b = np.zeros(cnt.value)
        _check(lib().tissue_get_order(self.h, _ptr(a, C.c_double), _ptr(b, C.c_double), cnt.value, C.byref(cnt)))
        return a, b

    def reset_order(self) -> None:
        _check(lib().tissue_reset_order(self.h))

    def flags(self) -> tuple[int, int]:
        f = C.c_uint32(); s = C.c_int32()
        _check(lib().tissue_get_flags(self.h, C.byref(f), C.byref(s)))
        return f.value, s.value

    def neighbor_lists(self, cap: int = 64) -> list[np.ndarray]:
        counts = np.zeros(self.n, np.int32); cols = np.zeros((self.n, cap), np.int32)
        _check(lib().tissue_neighbor_lists(self.h, _ptr(counts, C.c_int32), _ptr(cols, C.c_int32), cap))
        if counts.max(initial=0) > cap:
            return self.neighbor_lists(int(counts.max()))
        return [cols[i, : counts[i]].copy() for i in range(self.n)]

    def set_visit_check(self, enable: bool = True) -> None:
        """While on, each step of the step kernel records what its own neighbour walk visited
        (tissue_set_visit_check)."""
        _check(lib().tissue_set_visit_check(self.h, 1 if enable else 0))

    def visits(self) -> tuple[np.ndarray, np.ndarray]:
        """(na, hash) of the last step: the number of particles each particle's walk visited within
        2 sigma (self included) and the sum mod 2^32 of visit_mix over their indices."""
        na = np.zeros(self.n, np.int32); h = np.zeros(self.n, np.uint32)
        _check(lib().tissue_get_visits(self.h, _ptr(na, C.c_int32), _ptr(h, C.c_int32)))
        return na, h

    def set_launch(self, lanes: int = 0, block: int = 0) -> None:
        _check(lib().tissue_set_launch(self.h, lanes, block))

    def record_start(self, directory: str, every: int = 1) -> None:
        """Per-step CSV output stream (tissue_record_start) into an existing directory."""
        _check(lib().tissue_record_start(self.h, os.fsencode(directory), every))

    def record_stop(self) -> None:
        _check(lib().tissue_record_stop(self.h))

    def set_mode(self, mode: int) -> None:
        _check(lib().tissue_set_mode(self.h, mode))

    def launch(self) -> dict:
        """Launch shape the next step uses (tissue_get_launch)."""
        s = np.zeros(5, np.int32)
        _check(lib().tissue_get_launch(self.h, _ptr(s, C.c_int32)))
        return _shape_dict(s)

    def autotune(self, n_steps: int) -> dict:
        s = np.zeros(5, np.int32)
        _check(lib().tissue_autotune(self.h, n_steps, _ptr(s, C.c_int32)))
        return _shape_dict(s)

    def adjoint(self, n_steps: int, loss: Loss, first_row: int = -1, ckpt_every: int = 20):
        """Forward n_steps (state advances) and reverse; returns (status, gradients dict,
        d_theta0, d_x0, d_y0)."""
        g = Grads()
        dth = np.zeros(self.n); dx = np.zeros(self.n); dy = np.zeros(self.n)
        code = lib().tissue_adjoint(self.h, n_steps, first_row, ckpt_every, C.byref(loss), C.byref(g),
                                    _ptr(dth, C.c_double), _ptr(dx, C.c_double), _ptr(dy, C.c_double))
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code, g.as_dict(), dth, dx, dy

    def synchronize(self) -> None:
        _check(lib().tissue_synchronize(self.h))

    def last_step_ms(self) -> float:
        ms = C.c_double()
        _check(lib().tissue_last_step_ms(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().tissue_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MODE_AUTO, MODE_STEP_KERNEL, MODE_RESIDENT = 0, 1, 2


def _shape_dict(s: np.ndarray) -> dict:
    return dict(mode=int(s[0]), lanes=int(s[1]), block=int(s[2]), steps_per_launch=int(s[3]), xcd_order=int(s[4]))


def resident_capacity(mesh: Mesh) -> int:
    c = C.c_int32()
    _check(lib().tissue_resident_capacity(mesh.h, C.byref(c)))
    return c.value


class Batch:
    """Independent replicas run together (tissue_batch): tiled step kernel or persistent kernel."""

    def __init__(self, mesh: Mesh, sizes, params_list):
        self.mesh = mesh
        self.sizes = np.ascontiguousarray(sizes, np.int32)
        self.R = len(self.sizes)
        self.off = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        self.ntot = int(self.sizes.sum())
        ps = (Params * self.R)(*params_list)
        h = C.c_void_p()
        _check(lib().tissue_batch_create(mesh.h, self.R, _ptr(self.sizes, C.c_int32), ps, C.byref(h)))
        self.h = h

    def init_random(self, seeds) -> None:
        s = np.ascontiguousarray(seeds, np.uint64)
        _check(lib().tissue_batch_init_random(self.h, s.ctypes.data_as(C.POINTER(C.c_uint64))))

    def set_state(self, x, y, theta, vid=None) -> None:
        x = np.ascontiguousarray(x, np.float32); y = np.ascontiguousarray(y, np.float32)
        t = np.ascontiguousarray(theta, np.float32)
        v = None if vid is None else np.ascontiguousarray(vid, np.int32)
        _check(lib().tissue_batch_set_state(self.h, _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(t, C.c_float),
                                            _ptr(v, C.c_int32)))

    def get_state(self) -> State:
        n = self.ntot
        s = State(np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.int32))
        _check(lib().tissue_batch_get_state(self.h, _ptr(s.x, C.c_float), _ptr(s.y, C.c_float),
                                            _ptr(s.theta, C.c_float), _ptr(s.vid, C.c_int32)))
        return s

    def get_outputs(self) -> Outputs:
        n = self.ntot
        o = Outputs(np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32),
                    np.zeros(n, np.int32), np.zeros(n, np.int32))
        _check(lib().tissue_batch_get_outputs(self.h, _ptr(o.vx, C.c_float), _ptr(o.vy, C.c_float),
                                              _ptr(o.r3d, C.c_float), _ptr(o.face, C.c_int32),
                                              _ptr(o.nbr, C.c_int32)))
        return o

    def stage_noise(self, noise: np.ndarray) -> None:
        noise = np.ascontiguousarray(noise, np.float32).reshape(-1, self.ntot)
        _check(lib().tissue_batch_stage_noise(self.h, _ptr(noise, C.c_float), noise.shape[0]))

    def gen_noise(self, seeds, stds, n_rows: int) -> None:
        s = np.ascontiguousarray(seeds, np.uint64); d = np.ascontiguousarray(stds, np.float64)
        _check(lib().tissue_batch_gen_noise(self.h, s.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(d, C.c_double),
                                            n_rows))

    def run(self, n_steps: int, first_row: int = -1) -> int:
        code = lib().tissue_batch_run(self.h, n_steps, first_row)
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code

    def order(self, r: int) -> tuple[np.ndarray, np.ndarray]:
        cnt = C.c_int32()
        _check(lib().tissue_batch_get_order(self.h, r, None, None, 0, C.byref(cnt)))
        a = np.zeros(cnt.value); b = np.zeros(cnt.value)
        _check(lib().tissue_batch_get_order(self.h, r, _ptr(a, C.c_double), _ptr(b, C.c_double), cnt.value,
                                            C.byref(cnt)))
        return a, b

    def flags(self) -> tuple[np.ndarray, np.ndarray]:
        f = np.zeros(self.R, np.uint32); s = np.zeros(self.R, np.int32)
        _check(lib().tissue_batch_get_flags(self.h, f.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(s, C.c_int32)))
        return f, s

    def set_launch(self, lanes: int = 0, block: int = 0, steps_per_launch: int = 0) -> None:
        _check(lib().tissue_batch_set_launch(self.h, lanes, block, steps_per_launch))

    def set_mode(self, mode: int, xcd_order: int = -1) -> None:
        _check(lib().tissue_batch_set_mode(self.h, mode, xcd_order))

    def launch(self) -> dict:
        s = np.zeros(5, np.int32)
        _check(lib().tissue_batch_get_launch(self.h, _ptr(s, C.c_int32)))
        return _shape_dict(s)

    def adjoint(self, n_steps: int, loss: Loss, first_row: int = -1, ckpt_every: int = 20):
        """Every replica forward and reverse together; returns (status, [gradients dict per
        replica], d_theta0, d_x0, d_y0) with the d_* arrays over all particles."""
        g = (Grads * self.R)()
        dth = np.zeros(self.ntot); dx = np.zeros(self.ntot); dy = np.zeros(self.ntot)
        code = lib().tissue_batch_adjoint(self.h, n_steps, first_row, ckpt_every, C.byref(loss), g,
                                          _ptr(dth, C.c_double), _ptr(dx, C.c_double), _ptr(dy, C.c_double))
        if code not in (OK, E_LOST_PARTICLE, E_NONFINITE, E_TILING_NO_CONVERGE):
            _check(code)
        return code, [x.as_dict() for x in g], dth, dx, dy

    def autotune(self, n_steps: int) -> tuple[int, int]:
        l = C.c_int32(); b = C.c_int32()
        _check(lib().tissue_batch_autotune(self.h, n_steps, C.byref(l), C.byref(b)))
        return l.value, b.value

    def last_run_ms(self) -> float:
        ms = C.c_double()
        _check(lib().tissue_batch_last_run_ms(self.h, C.byref(ms)))
        return ms.value

    def set_visit_check(self, enable: bool = True) -> None:
        """Sim.set_visit_check over every replica (tissue_batch_set_visit_check)."""
        _check(lib().tissue_batch_set_visit_check(self.h, 1 if enable else 0))

    def visits(self) -> tuple[np.ndarray, np.ndarray]:
        """(na, hash) over all particles for the last step (indices hashed globally)."""
        na = np.zeros(self.ntot, np.int32); h = np.zeros(self.ntot, np.uint32)
        _check(lib().tissue_batch_get_visits(self.h, _ptr(na, C.c_int32), _ptr(h, C.c_int32)))
        return na, h

    def close(self):
        if getattr(self, "h", None):
            lib().tissue_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_noise(seed: int, std: float, n_steps: int, n: int) -> np.ndarray:
    out = np.zeros((n_steps, n), np.float32)
    _check(lib().tissue_gen_noise(seed, std, n_steps, n, _ptr(out, C.c_float)))
    return out


def seam_wrap(start, end, lanes: int = 1, max_crossings: int = 0, wave: bool = False):
    """The step's seam wrap of moves start[n, 2] -> end[n, 2] on the GPU
    (EuclideanTiling::diagonal_seam_edges_square_border, EuclideanTiling.cpp:31-69); wave selects
    the loop of crowded systems (same results).
    Returns (status, wrapped end [n, 2], heading rotation [n] in radians, quarter turns [n])."""
    a = np.ascontiguousarray(start, np.float64).reshape(-1, 2)
    e = np.ascontiguousarray(end, np.float64).reshape(-1, 2).copy()
    n = len(a)
    if len(e) != n:
        raise ValueError("start and end must have the same number of moves")
    d = np.zeros(n); q = np.zeros(n, np.int32)
    code = lib().tissue_seam_wrap(n, _ptr(a, C.c_double), _ptr(e, C.c_double), _ptr(d, C.c_double), _ptr(q, C.c_int32),
                                  max_crossings, lanes, 1 if wave else 0)
    if code not in (OK, E_TILING_NO_CONVERGE):
        _check(code)
    return code, e, d, q


def philox4x32(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr); k = (C.c_uint32 * 2)(*key); o = (C.c_uint32 * 4)()
    lib().tissue_philox4x32(c, k, o)
    return list(o)
