"""The geodesic table's local marches on the GPU (tissue_mesh_build_tables_on, mesh_gpu.hip)
against the host marches: the same tables byte for byte, on the ellipsoids of the plain and
batched workloads, the large perturbed sphere, and a cutoff so wide that many marches outgrow the
kernel's tables and return to the host."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("row_ptr", "col", "dist", "cell_start", "cell_faces", "uv")


def build_both(T, make, sigma_max):
    out, secs = {}, {}
    for where in (T.BUILD_HOST, T.BUILD_GPU):
        m = make()
        t0 = time.perf_counter()
        m.build_tables(sigma_max, where=where)
        secs[where] = time.perf_counter() - t0
        out[where] = m.export()
        m.close()
    return out[T.BUILD_HOST], out[T.BUILD_GPU], secs


@pytest.mark.parametrize("case", ["ellipsoid22_s0.1", "ellipsoid32_s0.05", "ellipsoid22_s0.5"])
def test_gpu_table_equals_host(T, case):
    freq = int(case.split("_")[0][len("ellipsoid"):])
    sigma = float(case.split("_s")[1])
    h, g, secs = build_both(T, lambda: T.Mesh.ellipsoid(2.0, 1.0, 1.0, freq), sigma)
    print(case, "host %.3f s, GPU %.3f s" % (secs[T.BUILD_HOST], secs[T.BUILD_GPU]), "entries", h["nnz"], flush=True)
    for k in FIELDS:
        assert h[k].tobytes() == g[k].tobytes(), (case, k)
    if sigma >= 0.5:  # rows longer than the kernel keeps: those marches ran on the host
        assert np.diff(h["row_ptr"]).max() > 128


def test_gpu_table_equals_host_large_surface(T):
    h, g, secs = build_both(T, lambda: T.Mesh.perturbed_sphere(0.2, 100), 0.01)
    print("perturbed sphere freq 100: host %.3f s, GPU %.3f s" % (secs[T.BUILD_HOST], secs[T.BUILD_GPU]),
          "entries", h["nnz"], flush=True)
    for k in FIELDS:
        assert h[k].tobytes() == g[k].tobytes(), k
