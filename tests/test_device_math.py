"""Exactness of the device geometry (2dtissue_amd/csrc/geom.h) against the FP64 oracle. The same
source is compiled for the host (tests/hostcheck) and must take the oracle's decisions on every
input: seam-wrap crossings (division-free border tests), and point location / nearest vertex
(FP32 ranking with FP64 fallback on close calls)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HC_DIR = os.path.join(HERE, "hostcheck")


@pytest.fixture(scope="module")
def H():
    subprocess.run(["make", "-s", "-C", HC_DIR], check=True)
    L = C.CDLL(os.path.join(HC_DIR, "libhostcheck.so"))
    L.hc_quotient_in_unit.argtypes = [C.c_double, C.c_double]
    return L


def p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def test_quotient_in_unit_matches_division(H):
    rng = np.random.default_rng(1)
    d = rng.normal(size=20000) * 10.0 ** rng.integers(-12, 3, 20000)
    n = d * rng.choice([1.0, -1.0, 0.5, 1.5], 20000)
    # adversarial: numerators one ulp around the denominator, and zeros
    n[:4000] = np.nextafter(d[:4000], np.where(rng.random(4000) < 0.5, np.inf, -np.inf))
    n[4000:4100] = 0.0
    n[4100:4200] = -0.0
    for nn, dd in zip(n, d):
        if dd == 0:
            continue
        q = nn / dd
        assert H.hc_quotient_in_unit(nn, dd) == int(0.0 <= q <= 1.0), (nn, dd)


def wrap_both(H, O, a, b, maxc=16384):
    n = len(a)
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    out = np.zeros((n, 2)); dth = np.zeros(n); ok = np.zeros(n, np.int32); cr = np.zeros(n, np.int32)
    H.hc_seam_wrap_many(n, p(a, C.c_double), p(b, C.c_double), maxc, p(out, C.c_double), p(dth, C.c_double),
                        p(ok, C.c_int32), p(cr, C.c_int32))
    st, r, _, h, ocr = O.seam_wrap(b, a, np.zeros(n), unit_deg=0, max_crossings=maxc)
    return out, dth, ok, cr, st, r, h, ocr


def test_seam_wrap_matches_oracle_random_flights(H, O):
    rng = np.random.default_rng(2)
    n = 20000
    a = rng.random((n, 2))
    L = np.exp(rng.uniform(np.log(1e-4), np.log(60.0), n))
    ang = rng.uniform(-np.pi, np.pi, n)
    b = a + L[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    out, dth, ok, cr, st, r, h, ocr = wrap_both(H, O, a, b)
    assert st == O.OK and ok.all()
    assert np.array_equal(out, r)        # same crossings, same FP64 images
    assert np.array_equal(cr, ocr)
    assert np.array_equal(dth, h)        # same rotations, accumulated in the same order
    assert cr.max() > 40                 # long flights are exercised


def test_seam_wrap_matches_oracle_adversarial(H, O):
    rng = np.random.default_rng(3)
    pts = []
    for _ in range(3000):
        # starts on a border or a corner, moves through or along the square's edges and corners
        e = rng.integers(0, 6)
        s = rng.random()
        start = [(0.0, s), (1.0, s), (s, 0.0), (s, 1.0), (0.0, 0.0), (1.0, 1.0)][e]
        tgt = rng.choice([-1.0, 0.0, 1.0, 2.0], 2) + rng.choice([0.0, 1e-3, -1e-3, 5e-4, 1e-12], 2)
        pts.append((start, tuple(tgt + rng.normal(0, 1e-9, 2) * (rng.random() < 0.5))))
        # lines passing within ~1e-3 of a corner (exit snapping)
        c = rng.choice([0.0, 1.0], 2)
        off = rng.choice([1e-3, 9.99e-4, 1.001e-3, 1e-6]) * rng.choice([-1, 1], 2)
        st = rng.random(2)
        d = (c + off) - st
        pts.append((tuple(st), tuple(st + d * rng.uniform(1.0, 30.0))))
    a = np.array([q[0] for q in pts]); b = np.array([q[1] for q in pts])
    keep = ~(((b >= 0) & (b <= 1)).all(1))  # moves that end outside
    out, dth, ok, cr, st, r, h, ocr = wrap_both(H, O, a[keep], b[keep], maxc=512)
    both_ok = ok.astype(bool)
    assert np.array_equal(out[both_ok], r[both_ok])
    assert np.array_equal(cr, ocr)
    assert np.array_equal(dth[both_ok], h[both_ok])


def wrap_flip_vs_direct(H, a, b, maxc):
    n = len(a)
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    res = []
    for fn in ("direct", "flip"):
        out = np.zeros((n, 2)); dth = np.zeros(n); ok = np.zeros(n, np.int32); cr = np.zeros(n, np.int32)
        q = np.zeros(n, np.int32)
        if fn == "direct":
            H.hc_seam_wrap_many(n, p(a, C.c_double), p(b, C.c_double), maxc, p(out, C.c_double), p(dth, C.c_double),
                                p(ok, C.c_int32), p(cr, C.c_int32))
            H.hc_seam_wrap_quarter_many(n, p(a, C.c_double), p(b, C.c_double), maxc, p(q, C.c_int32))
        else:
            H.hc_seam_wrap_flip_many(n, p(a, C.c_double), p(b, C.c_double), maxc, p(out, C.c_double),
                                     p(dth, C.c_double), p(ok, C.c_int32), p(cr, C.c_int32), p(q, C.c_int32))
        res.append((out, dth, ok, cr, q))
    return res


def test_seam_wrap_flip_form_bitwise(H):
    """The lane-parallel loop's border tests and crossing (sign flips instead of selects) against
    the direct form, bit for bit (zero signs included), on flights, edge and corner starts, exact
    2.0 / 1.0 / -0.0 coordinates and near-corner passes."""
    rng = np.random.default_rng(5)
    n = 20000
    a = rng.random((n, 2))
    k = np.arange(n) % 10
    a[k == 1, 0] = 0.0
    a[k == 2, 1] = 0.0
    a[k == 3, 0] = 1.0
    a[k == 4, 1] = 1.0
    a[k == 5] = 0.0
    a[k == 6] = 1.0
    a[k == 7] = (0.0, 1.0)
    a[k == 8, 0] = -0.0
    L = np.exp(rng.uniform(np.log(1e-4), np.log(80.0), n))
    ang = rng.uniform(-np.pi, np.pi, n)
    b = a + L[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    # end points on exact values: 2.0, -1.0, 1.0, -0.0 in one coordinate
    m = rng.random(n) < 0.15
    b[m, rng.integers(0, 2)] = rng.choice([2.0, -1.0, 1.0, -0.0, 3.0], m.sum())
    # near-corner passes (exit snapping) and moves along the edges
    c = rng.choice([0.0, 1.0], (4000, 2))
    off = rng.choice([1e-3, 9.99e-4, 1.001e-3, 1e-6, 0.0], (4000, 2)) * rng.choice([-1.0, 1.0], (4000, 2))
    st = rng.random((4000, 2))
    st[:1000, 0] = rng.choice([0.0, 1.0], 1000)
    b2 = st + ((c + off) - st) * rng.uniform(1.0, 30.0, (4000, 1))
    b2[:500, 0] = st[:500, 0]             # along an edge
    A = np.concatenate([a, st]); B = np.concatenate([b, b2])
    keep = ~(((B >= 0) & (B <= 1)).all(1))
    (o1, t1, k1, c1, q1), (o2, t2, k2, c2, q2) = wrap_flip_vs_direct(H, A[keep], B[keep], 4096)
    assert np.array_equal(k1, k2) and np.array_equal(c1, c2) and np.array_equal(q1, q2)
    assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
    assert np.array_equal(t1.view(np.uint64), t2.view(np.uint64))
    assert c1.max() > 40 and (c1 > 0).sum() > 10000


def wrap_passes(H, a, b, maxc, before):
    n = len(a)
    a = np.ascontiguousarray(a, np.float64); b = np.ascontiguousarray(b, np.float64)
    out = np.zeros((n, 2)); dth = np.zeros(n); ok = np.zeros(n, np.int32); cr = np.zeros(n, np.int32)
    q = np.zeros(n, np.int32); nones = np.zeros(n, np.int32)
    H.hc_seam_wrap_passes_many(n, p(a, C.c_double), p(b, C.c_double), maxc, before, p(out, C.c_double),
                               p(dth, C.c_double), p(ok, C.c_int32), p(cr, C.c_int32), p(q, C.c_int32),
                               p(nones, C.c_int32))
    return out, dth, ok, cr, q, nones


@pytest.mark.parametrize("maxc", [1, 2, 3, 4, 5, 7, 4096])
def test_seam_wrap_passes_bitwise(H, maxc):
    """The last flight of a wave, three crossings per pass by 64 speculating lanes (host emulation
    of the GPU's loop, entered after 0, 1, 2 or 5 ordinary crossings), against the sequential
    loop, bit for bit: same crossing counts, end points, rotations and quarter turns, with the
    crossing bound met at every level of a pass; includes flights whose tests all fail at some
    crossing (the fall-through to the down mapping)."""
    rng = np.random.default_rng(17)
    n = 6000
    a = rng.random((n, 2))
    k = np.arange(n) % 8
    a[k == 1, 0] = 0.0
    a[k == 2, 1] = 1.0
    a[k == 3] = 0.0
    a[k == 4] = (1.0, 0.0)
    L = np.exp(rng.uniform(np.log(1e-3), np.log(120.0), n))
    ang = rng.uniform(-np.pi, np.pi, n)
    b = a + L[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    m = rng.random(n) < 0.1
    b[m, rng.integers(0, 2)] = rng.choice([2.0, -1.0, 1.0, -0.0, 3.0], m.sum())
    c = rng.choice([0.0, 1.0], (2000, 2))
    off = rng.choice([1e-3, 9.99e-4, 1.001e-3, 1e-6, 0.0], (2000, 2)) * rng.choice([-1.0, 1.0], (2000, 2))
    st = rng.random((2000, 2))
    b2 = st + ((c + off) - st) * rng.uniform(1.0, 30.0, (2000, 1))
    b2[:300, 0] = st[:300, 0]
    A = np.concatenate([a, st]); B = np.concatenate([b, b2])
    keep = ~(((B >= 0) & (B <= 1)).all(1))
    A, B = A[keep], B[keep]
    (o1, t1, k1, c1, q1), _ = wrap_flip_vs_direct(H, A, B, maxc)
    for before in (0, 1, 2, 5):
        o2, t2, k2, c2, q2, _ = wrap_passes(H, A, B, maxc, before)
        assert np.array_equal(k1, k2) and np.array_equal(c1, c2) and np.array_equal(q1, q2), before
        assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64)), before
        assert np.array_equal(t1.view(np.uint64), t2.view(np.uint64)), before
    if maxc == 4096:
        assert c1.max() > 100
        # crossings whose four tests all fail (counted by the sequential loop) occur mid-flight
        nones = wrap_passes(H, A, B, maxc, 1 << 30)[5]
        print("moves with a fall-through crossing:", int((nones > 0).sum()), flush=True)
        assert (nones > 0).sum() >= 10
    else:
        assert (k1 == 0).sum() > 100  # the bound is met


def test_seam_wrap_reference_vector(H, O):
    import json
    g = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))["seam_wrap"]
    out, dth, ok, cr, *_ = wrap_both(H, O, np.array(g["r_old"]), np.array(g["r"]))
    assert np.abs(out - np.array(g["expected_r"])).max() <= g["tol"]
    assert np.allclose(np.degrees(dth) + np.array(g["n"]), [-280, -60, -48], atol=1e-9)


@pytest.fixture(scope="module")
def mesh32(T):
    import parity as P
    return P.mesh_and_tables(32, 0.05)[1]


def locate_both(H, O, t, x, y):
    n = len(x)
    pts = np.ascontiguousarray(np.stack([x, y], 1), np.float64)
    uv = np.ascontiguousarray(t["uv"], np.float32); xyz = np.ascontiguousarray(t["xyz"], np.float32)
    faces = np.ascontiguousarray(t["faces"], np.int32)
    face = np.zeros(n, np.int32); vid = np.zeros(n, np.int32); r3d = np.zeros((n, 3)); fp = np.zeros(n, np.int32)
    H.hc_locate_many(n, p(pts, C.c_double), t["nv"], t["nf"], p(uv, C.c_float), p(xyz, C.c_float),
                     p(faces, C.c_int32), t["grid"], p(t["cell_start"], C.c_int32), p(t["cell_faces"], C.c_int32),
                     p(face, C.c_int32), p(vid, C.c_int32), p(r3d, C.c_double), p(fp, C.c_int32))
    om = O.Mesh(t)
    r3, ov, of = O.remap(om, x, y, 0)
    return face, vid, r3d, fp, of, ov, r3


def test_locate_matches_oracle_random(H, O, mesh32):
    rng = np.random.default_rng(4)
    x = rng.random(100000).astype(np.float32).astype(np.float64)
    y = rng.random(100000).astype(np.float32).astype(np.float64)
    face, vid, r3d, fp, of, ov, r3 = locate_both(H, O, mesh32, x, y)
    assert np.array_equal(face, of) and np.array_equal(vid, ov)
    assert np.abs(r3d - r3).max() < 1e-5
    assert (fp == 0).mean() > 0.95  # almost every point is decided in FP32


def test_locate_matches_oracle_near_edges_and_vertices(H, O, mesh32):
    """Points on and within 1e-7..1e-4 of edges and vertices: the close calls."""
    t = mesh32
    rng = np.random.default_rng(5)
    uv, f = t["uv"], t["faces"]
    k = rng.integers(0, len(f), 20000)
    a, b = uv[f[k, 0]], uv[f[k, 1]]
    s = rng.random(20000)[:, None]
    e = a + s * (b - a)
    nrm = np.stack([-(b - a)[:, 1], (b - a)[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    h = rng.choice([0.0, 1e-7, 3e-6, 1e-5, 1e-4], 20000)[:, None] * rng.choice([-1, 1], 20000)[:, None]
    q = np.clip(e + h * nrm, 0, 1)
    q = np.vstack([q, uv[rng.integers(0, len(uv), 2000)]]).astype(np.float32).astype(np.float64)
    face, vid, r3d, fp, of, ov, r3 = locate_both(H, O, t, q[:, 0], q[:, 1])
    assert np.array_equal(face, of) and np.array_equal(vid, ov)
    assert np.abs(r3d - r3).max() < 1e-5
    assert (fp & 1).any() and (fp & 2).any()  # the FP64 fallbacks are exercised


def test_triangle_distance_branch_free(H):
    """The FP32 ranking's branch-free squared distance against the region-by-region FP64 distance
    (the reference's case analysis), over points in every Voronoi region of random triangles."""
    rng = np.random.default_rng(9)
    n = 200000
    c = rng.random((n, 2))
    tri = (c[:, None, :] + rng.normal(0, 0.01, (n, 3, 2))).reshape(n, 6)
    pts = c + rng.normal(0, 0.02, (n, 2))
    tri32 = np.ascontiguousarray(tri, np.float32)
    pts32 = np.ascontiguousarray(pts, np.float32)
    d2 = np.zeros(n, np.float32)
    d64 = np.zeros(n)
    H.hc_tri_dist_many(n, p(pts32, C.c_float), p(tri32, C.c_float), p(d2, C.c_float), p(d64, C.c_double))
    err = np.abs(np.sqrt(d2.astype(np.float64)) - d64)
    assert err.max() < 1e-6, err.max()  # far below the ranking margin (1e-5)
    assert (d64 == 0).mean() > 0.02 and (d64 > 0).mean() > 0.5  # interior and exterior points


def test_sort8_network(H):
    """The bucket sort network sorts every 0-1 input of length 8 (0-1 principle: then it sorts
    every input), and random and padded (INT_MAX) inputs as the buckets use it."""
    rows = np.array([[(m >> k) & 1 for k in range(8)] for m in range(256)], np.int32)
    rng = np.random.default_rng(5)
    rand = rng.integers(0, 1 << 30, (2000, 8)).astype(np.int32)
    pad = rand.copy()
    c = rng.integers(0, 9, 2000)
    for i, k in enumerate(c):
        pad[i, k:] = np.iinfo(np.int32).max
    for v in (rows, rand, pad):
        a = np.ascontiguousarray(v.copy())
        H.hc_sort8_many(len(a), p(a, C.c_int32))
        assert np.array_equal(a, np.sort(v, axis=1))
