"""The UV map's harmonic solve on the GPU (tissue_set_uv_solve) against the host's: the GPU runs the
host's Jacobi-preconditioned CG operation for operation (row-wise products and updates one thread
per row, every dot product as the host forms it: per-block sums in row order, then the block sums
in block order), so the UV coordinates, and with them every table built from them, are equal bit
for bit. Checked on the two parity meshes and the large surface."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build(T, kind, freq, where):
    T.set_uv_solve(where)
    try:
        t0 = time.perf_counter()
        m = T.Mesh.ellipsoid(2.0, 1.0, 1.0, freq) if kind == "ellipsoid" else T.Mesh.perturbed_sphere(0.2, freq)
        return m, time.perf_counter() - t0
    finally:
        T.set_uv_solve(T.BUILD_AUTO)


@pytest.mark.parametrize("kind,freq,sigma", [("ellipsoid", 22, 0.1), ("ellipsoid", 32, 0.05), ("sphere", 100, 0.01)])
def test_gpu_uv_solve_equals_host(T, kind, freq, sigma):
    mh, th = build(T, kind, freq, T.BUILD_HOST)
    mg, tg = build(T, kind, freq, T.BUILD_GPU)
    print(f"{kind} freq {freq}: mesh with the UV solve on the host {th:.3f} s, on the GPU {tg:.3f} s", flush=True)
    eh = mh.build_tables(sigma).export()
    eg = mg.build_tables(sigma).export()
    for k in ("uv", "xyz", "faces", "cell_start", "cell_faces", "row_ptr", "col", "dist"):
        assert np.array_equal(eh[k], eg[k]), k

