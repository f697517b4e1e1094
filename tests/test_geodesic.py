"""The geodesic table (tissue_mesh_build_tables; 2dtissue_amd/csrc/geodesic.cpp) against what the
mathematics fixes. The reference reads its table from MeshCartographyLib's
CachedGeodesicDistanceHelper on the cut-open mesh (2DTissue.cpp:102-105), whose source is absent;
the table here is the exact polyhedral geodesic distance (the shortest path on the
piecewise-flat surface), so it is pinned by closed forms:

- on a box, two vertices on one face are at their Euclidean distance, and two vertices on
  adjacent faces at the length of the straight line of the two faces unfolded into one plane;
- two vertices joined by a mesh edge are at the edge's length (nothing is shorter than the chord);
- on the unit sphere, the polyhedral distance tends to the great-circle distance at second order
  in the edge length;
- on a flat sheet cut by a slit, a path between the slit's two sides goes around its tip (the cut
  is a boundary: the tip is where the shortest path bends).

Table values are FP32 roundings of FP64 distances: comparisons allow one FP32 rounding."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
F32 = 2.0 ** -24  # relative rounding of one FP32 conversion


def box_off(path, a, b, c, h):
    """Closed triangulated box [0,a] x [0,b] x [0,c], square grid of spacing h on each face, the
    diagonals alternating, outward orientation; written as an OFF file."""
    index, verts, faces = {}, [], []

    def vid(p):
        key = tuple(np.round(np.asarray(p) * 1e9).astype(np.int64))
        if key not in index:
            index[key] = len(verts)
            verts.append(p)
        return index[key]

    def face(o, du, dv, flip):
        o, du, dv = (np.array(x, float) for x in (o, du, dv))
        nu, nv = int(round(np.linalg.norm(du) / h)), int(round(np.linalg.norm(dv) / h))
        for i in range(nu):
            for j in range(nv):
                q = [[vid(o + du * (i + s) / nu + dv * (j + t) / nv) for t in (0, 1)] for s in (0, 1)]
                tris = ([(q[0][0], q[1][0], q[1][1]), (q[0][0], q[1][1], q[0][1])] if (i + j) % 2 == 0
                        else [(q[0][0], q[1][0], q[0][1]), (q[1][0], q[1][1], q[0][1])])
                faces.extend(t[::-1] if flip else t for t in tris)

    face((0, 0, 0), (a, 0, 0), (0, b, 0), True)
    face((0, 0, c), (a, 0, 0), (0, b, 0), False)
    face((0, 0, 0), (a, 0, 0), (0, 0, c), False)
    face((0, b, 0), (a, 0, 0), (0, 0, c), True)
    face((0, 0, 0), (0, b, 0), (0, 0, c), True)
    face((a, 0, 0), (0, b, 0), (0, 0, c), False)
    with open(path, "w") as f:
        f.write(f"OFF\n{len(verts)} {len(faces)} 0\n")
        for p in verts:
            f.write("%.17g %.17g %.17g\n" % tuple(p))
        for t in faces:
            f.write("3 %d %d %d\n" % t)


def box_distance(P, Q, dims, eps=1e-9):
    """Exact distance on the box surface for points on one face or on two adjacent faces (the
    straight line of the unfolded faces), valid when both are farther from the box's corners than
    the distance itself."""
    def planes(p):
        return {(ax, v) for ax in range(3) for v in (0.0, dims[ax]) if abs(p[ax] - v) < eps}
    fp, fq = planes(P), planes(Q)
    if fp & fq:
        return float(np.linalg.norm(P - Q)), "face"
    best = np.inf
    for i1, c1 in fp:
        for i2, c2 in fq:
            if i1 != i2:
                i3 = 3 - i1 - i2
                best = min(best, np.hypot(abs(P[i2] - c2) + abs(Q[i1] - c1), P[i3] - Q[i3]))
    return float(best), "edge"


def seam_vertices(xyz):
    """UV vertices whose 3-D point appears more than once (the two sides of the cut)."""
    key = np.round(xyz * 1e9).astype(np.int64)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    return cnt[inv.ravel()] > 1


def rows_of(t):
    rp, col, dist = t["row_ptr"], t["col"], t["dist"].astype(np.float64)
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return src, col, dist


def test_box_closed_forms(T, tmp_path):
    a, b, c, h, sigma = 2.0, 1.0, 1.0, 0.05, 0.1
    box_off(str(tmp_path / "box.off"), a, b, c, h)
    m = T.Mesh.load_off(str(tmp_path / "box.off")).build_tables(sigma)
    t = m.export()
    xyz = t["xyz"]
    lim = 2.4 * sigma
    corners = np.array([[x, y, z] for x in (0, a) for y in (0, b) for z in (0, c)])
    seam = xyz[seam_vertices(xyz)]
    away = np.array([np.linalg.norm(corners - p, axis=1).min() > 2 * lim and np.linalg.norm(seam - p, axis=1).min() > 2 * lim
                     for p in xyz])
    src, col, dist = rows_of(t)
    keep = away[src] & away[col] & (src != col)
    kinds, errs = [], []
    for s, v, d in zip(src[keep], col[keep], dist[keep]):
        ex, kind = box_distance(xyz[s], xyz[v], (a, b, c))
        kinds.append(kind)
        errs.append(abs(d - ex) / ex)
    errs, kinds = np.array(errs), np.array(kinds)
    print(f"{len(errs)} pairs ({(kinds == 'edge').sum()} across a box edge): max relative error {errs.max():.2e}")
    assert len(errs) > 50000 and (kinds == "edge").sum() > 2000
    assert errs.max() <= F32


def test_mesh_edges_are_exact(T):
    """Vertices joined by an edge are at the edge's length; every entry is at least the chord."""
    import parity as P
    _, t = P.mesh_and_tables(32, 0.05)
    xyz, faces = t["xyz"], t["faces"]
    src, col, dist = rows_of(t)
    chord = np.linalg.norm(xyz[src] - xyz[col], axis=1)
    assert np.all(dist >= chord * (1 - F32))
    seam = seam_vertices(xyz)
    table = {(s, v): d for s, v, d in zip(src, col, dist)}
    checked = 0
    for f in faces:
        for u, v in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            if (u, v) in table and not (seam[u] and seam[v]):
                L = np.linalg.norm(xyz[u] - xyz[v])
                assert abs(table[(u, v)] - L) <= F32 * L, (u, v, table[(u, v)], L)
                checked += 1
    assert checked > 50000


def test_sphere_second_order_convergence(T):
    """Unit icosphere, table cutoff 0.24: the largest relative deviation from the great-circle
    distance (pairs away from the cut) falls about fourfold per halving of the edge length."""
    worst = []
    for freq in (8, 16, 32):
        t = T.Mesh.perturbed_sphere(0.0, freq).build_tables(0.1).export()
        xyz = t["xyz"]
        seam = xyz[seam_vertices(xyz)]
        away = np.array([np.linalg.norm(seam - p, axis=1).min() > 0.5 for p in xyz])
        src, col, dist = rows_of(t)
        keep = away[src] & away[col] & (src != col)
        s, v, d = src[keep], col[keep], dist[keep]
        cosang = np.einsum("ij,ij->i", xyz[s], xyz[v]) / (np.linalg.norm(xyz[s], axis=1) * np.linalg.norm(xyz[v], axis=1))
        gc = np.arccos(np.clip(cosang, -1, 1))
        worst.append(np.max(np.abs(d - gc) / gc))
    print("max relative deviation from the great circle at freq 8, 16, 32:", worst)
    assert worst[1] < worst[0] / 2.5 and worst[2] < worst[1] / 2.5
    assert worst[2] < 3e-4


@pytest.fixture(scope="module")
def H():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "hostcheck")], check=True)
    L = C.CDLL(os.path.join(HERE, "hostcheck", "libhostcheck.so"))
    L.hc_geodesic_row.restype = C.c_int
    L.hc_geodesic_row.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p]
    return L


def slit_sheet(n, slit_to):
    """Unit square, n x n grid cells, cut along x = 0.5 from y = 0 up to y = slit_to: the grid
    vertices on the cut below its tip are doubled (one copy per side). Returns xyz, faces, the
    tip's index and each vertex's side (-1 left of the cut, +1 right, 0 elsewhere)."""
    h = 1.0 / n
    xyz, ids, side = [], {}, []
    half = n // 2
    for j in range(n + 1):
        for i in range(n + 1):
            on_cut = i == half and j * h < slit_to - 1e-12
            for s in ((-1, 1) if on_cut else (0,)):
                ids[(i, j, s)] = len(xyz)
                xyz.append((i * h, j * h, 0.0))
                side.append(s)
    faces = []
    for j in range(n):
        for i in range(n):
            def vid(ii, jj):
                # a cell left of the cut takes the left copy of a doubled vertex, right of it the right one
                s = -1 if i < half else 1
                return ids.get((ii, jj, s), ids.get((ii, jj, 0)))
            q00, q10, q01, q11 = vid(i, j), vid(i + 1, j), vid(i, j + 1), vid(i + 1, j + 1)
            faces += [(q00, q10, q11), (q00, q11, q01)]
    tip = ids[(half, int(round(slit_to * n)), 0)]
    return np.array(xyz, float), np.array(faces, np.int32), tip, np.array(side)


def test_slit_paths_bend_at_the_tip(H):
    """Across a slit the shortest path runs around its tip: |s - tip| + |tip - t|; elsewhere it is
    the straight segment. Every vertex of a row is checked against that closed form."""
    xyz, faces, tip, side = slit_sheet(40, 0.5)
    T = xyz[tip]
    lim = 0.3
    cap = len(xyz)
    ov = np.zeros(cap, np.int32)
    od = np.zeros(cap)
    sources = [i for i in range(len(xyz)) if abs(xyz[i, 0] - 0.5) < 0.2 and 0.2 < xyz[i, 1] < 0.7]
    checked = bent = 0
    for s in sources:
        k = H.hc_geodesic_row(len(xyz), len(faces), xyz.ctypes.data, faces.ctypes.data, lim, s, cap, ov.ctypes.data,
                              od.ctypes.data)
        for v, d in zip(ov[:k], od[:k]):
            P, Q = xyz[s], xyz[v]
            # the straight segment crosses the slit when the two lie on its two sides below the tip
            ps = side[s] if side[s] else np.sign(P[0] - 0.5)
            qs = side[v] if side[v] else np.sign(Q[0] - 0.5)
            cross = False
            if ps * qs < 0:  # (the two copies of a cut vertex are both at x = 0.5)
                y_at = P[1] if Q[0] == P[0] else P[1] + (Q[1] - P[1]) * (0.5 - P[0]) / (Q[0] - P[0])
                cross = y_at < T[1] - 1e-12
            expect = np.linalg.norm(P - T) + np.linalg.norm(T - Q) if cross else np.linalg.norm(P - Q)
            assert abs(d - expect) <= 1e-12 * max(expect, 1e-3), (s, v, d, expect, cross)
            checked += 1
            bent += cross
    print(f"{checked} distances checked, {bent} around the tip")
    assert bent > 100
