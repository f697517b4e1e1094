"""The drop-in class include/2DTissue.h (the reference's _2DTissue API over the C-ABI) and the
command-line driver built on it (2dtissue_amd/tissue_main, the reference's main.cpp flags)."""
import os
import subprocess

import numpy as np
import pytest

import parity as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "2dtissue_amd", "tissue_main")


def test_compat_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "2DTissue.h"\n'
                   'int main() {\n'
                   '  _2DTissue t(false, false, false, "ellipsoid:2,1,1,8", 10, 3, 1.0);\n'
                   '  t.start();\n'
                   '  while (!t.is_finished()) { System s = t.update(); (void)s.particles.size(); }\n'
                   '  return (int)t.get_order_parameter().size() - 3;\n'
                   '}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


def test_cli_usage():
    r = subprocess.run([CLI, "--no-such-flag"], capture_output=True, text=True)
    assert r.returncode == 1 and "--step-count" in r.stderr


@pytest.mark.gpu
def test_cli_matches_c_abi(T, tmp_path):
    """tissue_main (compat class) and the Python binding of the C-ABI step the same system: the
    order parameters and the per-step CSV files agree."""
    n, steps = 100, 10
    data = tmp_path / "data"
    data.mkdir()
    r = subprocess.run([CLI, "--mesh-path", "ellipsoid:2,1,1,22", "--particle-count", str(n), "--step-count",
                        str(steps), "--sigma", "0.1", "--step-time", "1.0", "--seed", "0", "--save-data",
                        "--data-dir", str(data)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    phi_cli = np.array([float(x) for x in lines[:steps]])
    assert lines[steps].startswith("Time taken:")
    # the same run through the Python binding: v0 = --step-time (main.cpp:76-84), k = 1, mu = 1,
    # dt = tau = the reference's default step_size 0.001
    m, _ = P.mesh_and_tables(22, 0.1)
    s = T.Sim(m, T.params(sigma=0.1, v0=1.0, k=1.0, mu=1.0, tau=0.001, dt=0.001), n)
    s.init_random(0)
    for _ in range(steps):
        assert s.step(1) == T.OK
    phi = s.order()[0]
    assert np.allclose(phi_cli, phi, rtol=1e-14, atol=0)
    st = s.get_state()
    xy = np.loadtxt(data / f"r_data_{steps}.csv", delimiter=",")
    assert np.allclose(xy[:, 0], st.x, rtol=1e-14) and np.allclose(xy[:, 1], st.y, rtol=1e-14)
    nb = np.loadtxt(data / f"particles_color_{steps}.csv")
    assert np.array_equal(nb.astype(int), s.get_outputs().nbr)
    r3 = np.loadtxt(data / f"r_data_3D_{steps}.csv", delimiter=",")
    assert np.allclose(r3, s.get_outputs().r3d, rtol=1e-14)
