"""tissue_step_noalloc (the step entry point that allocates nothing) against tissue_step, through
the C-ABI: two calls in a row on one GPU, the second with fewer particles (the scratch array then
holds the first call's contents), each equal bit for bit to one step of a simulation with the same
state; plus its size limits and the mesh-residency rule."""
import ctypes as C

import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers for the test, from the HIP runtime the library itself links (ctypes)."""

    H2D, D2H = 1, 2

    def __init__(self, T):
        T.lib()  # the library first: its HIP runtime is the one opened below (by soname)
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def ok(self, code):
        assert code == 0, f"HIP error {code}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), max(a.nbytes, 4)))
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D))
        self.bufs.append(p)
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.rt.hipDeviceSynchronize())
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H))
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def device_step(T, hip, mesh, p, st, noise):
    """One tissue_step_noalloc call on device copies of state st (default stream); returns host
    copies of its outputs."""
    n = len(st.x)
    x, y, th = (hip.upload(np.asarray(a, np.float32)) for a in (st.x, st.y, st.theta))
    vid = hip.upload(np.asarray(st.vid, np.int32))
    eta = None if noise is None else hip.upload(np.asarray(noise, np.float32))
    f_like = np.full(n, np.nan, np.float32)
    i_like = np.full(n, -7, np.int32)
    out_f = {k: hip.upload(f_like) for k in ("x", "y", "theta", "vx", "vy", "r3x", "r3y", "r3z")}
    out_i = {k: hip.upload(i_like) for k in ("vid", "face", "nbr")}
    phi = hip.upload(np.full(2, -1.0))
    flags = hip.upload(np.full(1, 0x5A5A, np.uint32))
    code = T.lib().tissue_step_noalloc(
        mesh.h, C.byref(p), n, x, y, th, vid, eta, out_f["x"], out_f["y"], out_f["theta"], out_i["vid"], out_f["vx"],
        out_f["vy"], out_f["r3x"], out_f["r3y"], out_f["r3z"], out_i["face"], out_i["nbr"], phi, flags, None)
    assert code == T.OK, T.lib().tissue_error_string()
    res = {k: hip.download(v, f_like) for k, v in out_f.items()}
    res.update({k: hip.download(v, i_like) for k, v in out_i.items()})
    res["phi"] = hip.download(phi, np.zeros(2))
    res["flags"] = int(hip.download(flags, np.zeros(1, np.uint32))[0])
    hip.free()
    return res


def sim_step(T, mesh, p, st, noise):
    n = len(st.x)
    s = T.Sim(mesh, p, n)
    s.set_mode(T.MODE_STEP_KERNEL)
    s.set_state(st.x, st.y, st.theta, st.vid)
    code = s.step(1, None if noise is None else noise[None, :])
    out = dict(code=code, state=s.get_state(), outputs=s.get_outputs(), phi=P.last_phi(s), flags=s.flags()[0])
    s.close()
    return out


def assert_equal(na, ref):
    st, o = ref["state"], ref["outputs"]
    for k, a in (("x", st.x), ("y", st.y), ("theta", st.theta), ("vid", st.vid), ("vx", o.vx), ("vy", o.vy),
                 ("face", o.face), ("nbr", o.nbr)):
        assert np.array_equal(na[k], a), k
    for c, k in enumerate(("r3x", "r3y", "r3z")):
        assert np.array_equal(na[k], o.r3d[:, c]), k
    assert na["phi"][0] == ref["phi"][0] and na["phi"][1] == ref["phi"][1], (na["phi"], ref["phi"])
    assert na["flags"] == ref["flags"]


def evolved_state(T, mesh, p, n, seed, steps):
    s = T.Sim(mesh, p, n)
    s.init_random(seed)
    assert s.step(steps, T.gen_noise(seed, 0.1, steps, n)) == T.OK
    st = s.get_state()
    s.close()
    return st


def test_noalloc_step_equals_tissue_step_twice(T):
    hip = Hip(T)
    cfg = P.C2
    mesh, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = T.params(**dict(P.PHYS, sigma=cfg["sigma"]))
    # first call: the 2,000-particle system after 30 steps; second: 700 particles, zero noise
    for n, seed, noisy in ((2000, 1, True), (700, 9, False)):
        st = evolved_state(T, mesh, p, n, seed, 30)
        noise = T.gen_noise(seed + 100, 0.1, 1, n)[0] if noisy else None
        ref = sim_step(T, mesh, p, st, noise)
        assert ref["code"] == T.OK
        na = device_step(T, hip, mesh, p, st, noise)
        print(f"n={n}: phi={na['phi']}, flags={na['flags']}", flush=True)
        assert_equal(na, ref)


def test_noalloc_sorted_buckets_case(T):
    """A system denser than one particle per two vertices: the vertex buckets are sorted by the
    finalize pass (the engine's rule), still equal to tissue_step."""
    hip = Hip(T)
    cfg = P.C1
    mesh, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = T.params(**dict(P.PHYS, sigma=0.05))
    n = 2600
    assert 2 * n > mesh.info()["nv"]
    st = evolved_state(T, mesh, p, n, 4, 10)
    noise = T.gen_noise(44, 0.1, 1, n)[0]
    assert_equal(device_step(T, hip, mesh, p, st, noise), sim_step(T, mesh, p, st, noise))


def test_noalloc_rejects_sizes_and_absent_tables(T):
    hip = Hip(T)
    cfg = P.C2
    mesh, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = T.params(**dict(P.PHYS, sigma=cfg["sigma"]))
    big = hip.upload(np.zeros(20000, np.float32))
    ptrs = [big] * 18
    code = T.lib().tissue_step_noalloc(mesh.h, C.byref(p), 16385, *ptrs[:4], None, *ptrs[5:18], C.c_void_p(0))
    assert code == T.E_INVALID_ARG
    assert b"particle count" in T.lib().tissue_error_string()
    # tables built but never uploaded to this GPU: rejected, not uploaded
    fresh = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 8).build_tables(0.2)
    q = T.params(**dict(P.PHYS, sigma=0.1))
    code = T.lib().tissue_step_noalloc(fresh.h, C.byref(q), 10, *ptrs[:4], None, *ptrs[5:18], C.c_void_p(0))
    assert code == T.E_INVALID_ARG
    assert b"not on this GPU" in T.lib().tissue_error_string()
    fresh.close()
    hip.free()
