"""The per-step output stream (tissue_record_start): the CSV files it writes in the background hold,
for every recorded step, the same values as reading the state back after that step, in the
reference's format (2DTissue.cpp:270-280, IO.h:40-69), and runs longer than its ring are split
without losing a step."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


def test_record_stream_matches_state(T, tmp_path):
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    n = 500
    s = T.Sim(m, T.params(**dict(P.PHYS, sigma=cfg["sigma"])), n)
    s.init_random(4)
    noise = T.gen_noise(4, 0.1, 12, n)
    s.record_start(str(tmp_path), every=2)
    expect = {}
    for t in range(12):
        assert s.step(1, noise[t]) == T.OK
        if (t + 1) % 2 == 0:
            st, out = s.get_state(), s.get_outputs()
            expect[t + 1] = (st.x.copy(), st.y.copy(), out.r3d.copy(), out.nbr.copy())
    s.record_stop()
    files = sorted(p.name for p in tmp_path.iterdir())
    assert files == sorted(f"{k}_{t}.csv" for t in expect for k in ("r_data", "r_data_3D", "particles_color"))
    for t, (x, y, r3, nb) in expect.items():
        xy = np.loadtxt(tmp_path / f"r_data_{t}.csv", delimiter=",")
        assert np.array_equal(xy[:, 0].astype(np.float32), x) and np.array_equal(xy[:, 1].astype(np.float32), y)
        r = np.loadtxt(tmp_path / f"r_data_3D_{t}.csv", delimiter=",")
        assert np.array_equal(r.astype(np.float32), r3.astype(np.float32))
        assert np.array_equal(np.loadtxt(tmp_path / f"particles_color_{t}.csv").astype(int), nb)
    line = (tmp_path / "r_data_2.csv").read_text().splitlines()[0]
    assert line == "%.15g,%.15g" % (float(expect[2][0][0]), float(expect[2][1][0]))
    s.close()


def test_record_stream_long_run_is_split(T, tmp_path):
    """One call of many more steps than the ring holds: every step is written, in order."""
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    n = 100
    s = T.Sim(m, T.params(**dict(P.PHYS, sigma=cfg["sigma"])), n)
    s.init_random(5)
    s.record_start(str(tmp_path), every=1)
    assert s.step(600) == T.OK
    last = s.get_state()
    s.record_stop()
    import re
    steps = sorted(int(m.group(1)) for m in (re.fullmatch(r"r_data_(\d+)\.csv", p.name) for p in tmp_path.iterdir()) if m)
    assert steps == list(range(1, 601))
    xy = np.loadtxt(tmp_path / "r_data_600.csv", delimiter=",")
    assert np.array_equal(xy[:, 0].astype(np.float32), last.x)
    s.close()
