"""CPU tests of the product library: it loads without a GPU and exports every entry point of
include/tissue.h; the host RNG; the mesh toolkit (UV chart with the reference's gluing, face
buckets, geodesic table). No GPU compute is called here."""
import math
import os
import re
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "tissue.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tissue_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(T):
    names = header_functions()
    assert len(names) >= 25
    lib = T.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in T.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_philox_known_answers(T):
    """Philox4x32-10 known-answer vectors of Random123 (Salmon et al., SC'11)."""
    assert T.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert T.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert T.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_noise_statistics_and_determinism(T):
    a = T.gen_noise(7, 0.1, 200, 1000)
    assert np.array_equal(a, T.gen_noise(7, 0.1, 200, 1000))
    assert not np.array_equal(a, T.gen_noise(8, 0.1, 200, 1000))
    assert abs(a.mean()) < 0.1 * 4 / math.sqrt(a.size)
    assert abs(a.std() - 0.1) < 0.1 * 0.01
    # counter-based: a prefix of particles/steps does not depend on the array shape
    assert np.array_equal(T.gen_noise(7, 0.1, 3, 10), a[:3, :10])


@pytest.fixture(scope="module")
def ell(T):
    m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 12).build_tables(0.1)
    return m, m.export()


def test_mesh_counts_and_uv_chart(ell):
    m, t = ell
    f = 12
    assert t["nv3d"] == 10 * f * f + 2 and t["nf"] == 20 * f * f
    uv, faces = t["uv"], t["faces"]
    assert uv.min() >= 0 and uv.max() <= 1
    assert np.array_equal(uv, uv.astype(np.float32).astype(np.float64))  # FP32-representable
    a, b, c = uv[faces[:, 0]], uv[faces[:, 1]], uv[faces[:, 2]]
    area = 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    assert (area > 0).all()                     # counter-clockwise, no folds
    assert abs(area.sum() - 1.0) < 1e-9         # the triangles tile the unit square


def test_mesh_seam_gluing(ell):
    """Seam duplicates sit where the seam wrap sends their twin: (t,0)~(0,t) and (1,t)~(t,1)."""
    m, t = ell
    uv, xyz = t["uv"], t["xyz"]
    nv3d, nv = t["nv3d"], t["nv"]
    key = {tuple(xyz[v]): v for v in range(nv3d)}
    n_dup = 0
    for d in range(nv3d, nv):
        v = key[tuple(xyz[d])]
        (x0, y0), (x1, y1) = uv[v], uv[d]
        if y0 == 0.0:
            assert (x1, y1) == (0.0, x0)
        else:
            assert x0 == 1.0 and (x1, y1) == (y0, 1.0)
        n_dup += 1
    assert n_dup == nv - nv3d > 0
    # the boundary of the chart is exactly the unit square's edges
    edges = {}
    for f in t["faces"]:
        for q in range(3):
            e = tuple(sorted((f[q], f[(q + 1) % 3])))
            edges[e] = edges.get(e, 0) + 1
    bverts = {v for e, c in edges.items() if c == 1 for v in e}
    for v in bverts:
        x, y = uv[v]
        assert x in (0.0, 1.0) or y in (0.0, 1.0)


def test_face_buckets_cover_faces(ell):
    m, t = ell
    g, cs, cf = t["grid"], t["cell_start"], t["cell_faces"]
    uv, faces = t["uv"], t["faces"]
    cen = uv[faces].mean(1)
    cells = np.minimum((cen * g).astype(int), g - 1)
    for fi, (cx, cy) in enumerate(cells):
        c = cy * g + cx
        assert fi in cf[cs[c]:cs[c + 1]]


def test_distance_table_structure(ell):
    m, t = ell
    rp, col, dist = t["row_ptr"], t["col"], t["dist"]
    nv = t["nv"]
    lookup = {}
    for v in range(nv):
        row = list(zip(dist[rp[v]:rp[v + 1]], col[rp[v]:rp[v + 1]]))
        assert row == sorted(row)               # ascending (distance, vertex)
        assert (0.0, v) in row                  # diagonal present, zero
        for d, w in row:
            lookup[(v, w)] = d
            assert d <= 2.4 * 0.1 * (1 + 1e-6)
    for (v, w), d in lookup.items():
        assert lookup[(w, v)] == d              # symmetric (min of both directions)


def test_mesh_build_independent_of_thread_count(tmp_path):
    """The UV solve, the face buckets and the geodesic table are split over host threads; their
    results must not depend on how many: the same mesh built with 1 and with 5 threads, byte for
    byte."""
    import subprocess
    import sys
    code = (
        "import importlib, sys, numpy as np\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "T = importlib.import_module('2dtissue_amd')\n"
        "m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 28).build_tables(0.1)\n"
        "e = m.export()\n"
        "np.savez(sys.argv[1], **{k: e[k] for k in ('uv', 'xyz', 'faces', 'row_ptr', 'col', 'dist', 'cell_start', 'cell_faces')})\n")
    out = {}
    for n in (1, 5):
        f = str(tmp_path / f"mesh{n}.npz")
        env = dict(os.environ, TISSUE_HOST_THREADS=str(n))
        subprocess.run([sys.executable, "-c", code, f], check=True, env=env, timeout=300)
        out[n] = np.load(f)
    for k in out[1].files:
        assert out[1][k].tobytes() == out[5][k].tobytes(), k


def test_off_roundtrip(T, ell):
    m, t = ell
    with tempfile.TemporaryDirectory() as d:
        p3, puv = os.path.join(d, "e.off"), os.path.join(d, "e_uv.off")
        m.save_off(p3, puv)
        assert open(puv).readline().strip() == "OFF"
        with pytest.raises(T.TissueError):
            T.Mesh.load_off(p3)  # the UV mesh's duplicated seam vertices make it open, not genus 0
        src = os.path.join(d, "closed.off")  # a closed icosahedron written by hand
        phi = (1 + 5 ** 0.5) / 2
        V = [(0, s1, s2 * phi) for s1 in (-1, 1) for s2 in (-1, 1)] + \
            [(s1, s2 * phi, 0) for s1 in (-1, 1) for s2 in (-1, 1)] + \
            [(s2 * phi, 0, s1) for s1 in (-1, 1) for s2 in (-1, 1)]
        V = np.array(V, float)
        F = []
        for i in range(12):
            for j in range(i + 1, 12):
                for k in range(j + 1, 12):
                    if all(abs(np.linalg.norm(V[a] - V[b]) - 2) < 1e-9 for a, b in ((i, j), (j, k), (i, k))):
                        n = np.cross(V[j] - V[i], V[k] - V[i])
                        F.append((i, j, k) if n @ (V[i] + V[j] + V[k]) > 0 else (i, k, j))
        with open(src, "w") as fh:
            fh.write(f"OFF\n12 {len(F)} 0\n")
            for v in V:
                fh.write("%.17g %.17g %.17g\n" % tuple(v))
            for f in F:
                fh.write("3 %d %d %d\n" % f)
        mm = T.Mesh.load_off(src)
        assert mm.info()["nv3d"] == 12 and mm.info()["nf"] == 20


def test_mesh_errors(T):
    with pytest.raises(T.TissueError):
        T.Mesh.ellipsoid(-1.0, 1.0, 1.0, 4)
    with pytest.raises(T.TissueError):
        T.Mesh.load_off("/nonexistent/mesh.off")
    m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 6)
    with pytest.raises(T.TissueError):
        m.build_tables(0.1, where=7)  # no such place
    with pytest.raises(T.TissueError):
        m.build_tables(-0.1, where=T.BUILD_HOST)
    m.build_tables(0.1, where=T.BUILD_HOST)
    assert m.info()["nnz"] > 0
    with pytest.raises(T.TissueError) as e:  # the table has no GPU build
        T.Mesh.ellipsoid(2.0, 1.0, 1.0, 6).build_tables(0.1, where=T.BUILD_GPU)
    assert "host" in str(e.value)


def test_empty_and_invalid_systems_are_rejected(T):
    """Systems without particles, batches without replicas or with an empty replica, and
    parameters out of range are refused with TISSUE_E_INVALID_ARG before any device call (so
    this holds on a host without a GPU too)."""
    m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 6)
    m.build_tables(0.1, where=T.BUILD_HOST)
    for n in (0, -3):
        with pytest.raises(T.TissueError) as e:
            T.Sim(m, T.params(sigma=0.1), n)
        assert e.value.code == T.E_INVALID_ARG and "positive" in str(e.value)
    with pytest.raises(T.TissueError) as e:
        T.Batch(m, [], [])
    assert e.value.code == T.E_INVALID_ARG
    with pytest.raises(T.TissueError) as e:
        T.Batch(m, [10, 0, 5], [T.params(sigma=0.1)] * 3)
    assert e.value.code == T.E_INVALID_ARG and "positive" in str(e.value)
    with pytest.raises(T.TissueError) as e:  # sigma above the one the table was built for
        T.Batch(m, [10, 5], [T.params(sigma=0.1), T.params(sigma=0.2)])
    assert e.value.code == T.E_INVALID_ARG


def restated_init(T, tb, seed, n):
    """The area-uniform initialisation restated in numpy from its definition: cumulative FP64 3-D
    face areas, upper-bound face choice, folded barycentrics, heading 2 pi u wrapped to [-pi, pi),
    all from the Philox4x32-10 block (pinned by its known answers above)."""
    xyz, faces, uv = np.asarray(tb["xyz"]), np.asarray(tb["faces"]), np.asarray(tb["uv"])
    A, B, Cc = xyz[faces[:, 0]], xyz[faces[:, 1]], xyz[faces[:, 2]]
    e1, e2 = B - A, Cc - A
    cx = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    cy = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    cz = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    area = 0.5 * np.sqrt(cx * cx + cy * cy + cz * cz)
    cdf = np.cumsum(area)
    u53 = lambda a, b: ((a >> 5) * 67108864.0 + (b >> 6)) / 9007199254740992.0
    key = [seed & 0xFFFFFFFF, seed >> 32]
    x = np.zeros(n, np.float32); y = np.zeros(n, np.float32); th = np.zeros(n, np.float32)
    for i in range(n):
        d = T.philox4x32([i, 0, 0, 0], key)
        e = T.philox4x32([i, 0, 2, 0], key)
        f = min(int(np.searchsorted(cdf, u53(d[0], d[1]) * cdf[-1], side="right")), len(faces) - 1)
        r1, r2 = u53(d[2], d[3]), u53(e[0], e[1])
        if r1 + r2 > 1.0:
            r1, r2 = 1.0 - r1, 1.0 - r2
        a, b, c = uv[faces[f, 0]], uv[faces[f, 1]], uv[faces[f, 2]]
        px = a[0] + r1 * (b[0] - a[0]) + r2 * (c[0] - a[0])
        py = a[1] + r1 * (b[1] - a[1]) + r2 * (c[1] - a[1])
        x[i] = np.float32(min(max(px, 0.0), 1.0)); y[i] = np.float32(min(max(py, 0.0), 1.0))
        t = 2.0 * math.pi * u53(e[2], e[3])
        th[i] = np.float32(t - 2.0 * math.pi * math.floor((t + math.pi) / (2.0 * math.pi)))
    return x, y, th


def test_random_init_host_matches_definition(T, ell):
    m, tb = ell
    x, y, th = m.random_init_host(7, 400)
    rx, ry, rth = restated_init(T, tb, 7, 400)
    assert np.array_equal(x, rx) and np.array_equal(y, ry) and np.array_equal(th, rth)
    x2, _, _ = m.random_init_host(8, 400)
    assert not np.array_equal(x, x2)


def test_random_init_host_is_area_uniform(T, O, ell):
    """Faces are chosen with probability proportional to their 3-D area: the share of particles
    whose remapped point lies on faces with centroid z > 0 matches those faces' share of the area."""
    m, tb = ell
    n = 20000
    x, y, th = m.random_init_host(3, n)
    assert ((x >= 0) & (x <= 1) & (y >= 0) & (y <= 1)).all()
    assert ((th >= -np.pi) & (th < np.pi)).all()
    assert abs(float(th.mean())) < 0.05
    _, _, face = O.remap(O.Mesh(tb), x.astype(np.float64), y.astype(np.float64))
    xyz, faces = np.asarray(tb["xyz"]), np.asarray(tb["faces"])
    A, B, Cc = xyz[faces[:, 0]], xyz[faces[:, 1]], xyz[faces[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(B - A, Cc - A), axis=1)
    upper = (A + B + Cc)[:, 2] > 0
    share = area[upper].sum() / area.sum()
    got = upper[face].mean()
    assert abs(got - share) < 4 * np.sqrt(share * (1 - share) / n), (got, share)


def test_uv_solve_place_is_validated(T):
    """tissue_set_uv_solve takes AUTO, HOST or GPU; anything else is rejected and leaves it as it
    was; HOST builds on the host in any case."""
    assert T.get_uv_solve() == T.BUILD_AUTO
    with pytest.raises(T.TissueError):
        T.set_uv_solve(7)
    assert T.get_uv_solve() == T.BUILD_AUTO
    T.set_uv_solve(T.BUILD_HOST)
    try:
        m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, 8)
        assert m.info()["nf"] == 20 * 8 * 8
    finally:
        T.set_uv_solve(T.BUILD_AUTO)
