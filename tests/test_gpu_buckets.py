"""Raw vertex buckets (no finalize pass; each reader orders a bucket's particles itself: the mode of
replicas of up to 8,192 particles at most one per two vertices) against sorted buckets (the
finalize pass): both visit a
bucket's particles in ascending id order, so trajectories, outputs, neighbour lists and gradients
agree bit for bit, also when buckets are crowded (more particles on a vertex than its 8 slots,
read through the overflow list). TISSUE_SORTED_BUCKETS chooses the mode of a new simulation."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


def same(a, b, what):
    assert a.tobytes() == b.tobytes(), what


def snapshot(s):
    st, o = s.get_state(), s.get_outputs()
    return [st.x, st.y, st.theta, st.vid, o.vx, o.vy, o.r3d, o.face, o.nbr], s.neighbor_lists()


def compare(a, b):
    for k, (u, v) in enumerate(zip(a[0], b[0])):
        same(u, v, k)
    assert len(a[1]) == len(b[1])
    for u, v in zip(a[1], b[1]):
        same(u, v, "neighbours")


def test_raw_buckets_equal_sorted_plain(T, monkeypatch):
    m, _ = P.mesh_and_tables(32, 0.05)
    p = dict(P.PHYS, sigma=0.05)
    n, steps = 2000, 60
    noise = T.gen_noise(1, 0.1, steps, n)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TISSUE_SORTED_BUCKETS", mode)
        s = T.Sim(m, T.params(**p), n)
        s.init_random(1)
        s.stage_noise(noise)
        assert s.step_staged(steps, 0) == T.OK
        out[mode] = snapshot(s)
        s.close()
    compare(out["1"], out["0"])


def test_raw_buckets_equal_sorted_crowded(T, monkeypatch):
    m, tb = P.mesh_and_tables(22, 0.1)
    p = dict(P.PHYS, sigma=0.1)
    uv = tb["uv"]
    rng = np.random.default_rng(7)
    inner = np.flatnonzero(((uv > 0.2) & (uv < 0.8)).all(1))
    piles = rng.choice(inner, 20, replace=False)
    n = 400
    x = rng.random(n); y = rng.random(n)
    # 300 particles on 20 vertices, 15 each (each within 1e-4 of its vertex, so it stays the nearest)
    k = np.repeat(piles, 15)
    x[:300] = uv[k, 0] + rng.uniform(-1e-4, 1e-4, 300)
    y[:300] = uv[k, 1] + rng.uniform(-1e-4, 1e-4, 300)
    th = rng.uniform(-np.pi, np.pi, n)
    noise = T.gen_noise(3, 0.1, 5, n)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TISSUE_SORTED_BUCKETS", mode)
        s = T.Sim(m, T.params(**p), n)
        s.set_state(x, y, th)
        v0 = s.get_state().vid
        assert np.bincount(v0).max() > 8  # crowded buckets at the start
        first = snapshot(s)[1]
        s.stage_noise(noise)
        code = s.step_staged(5, 0)
        out[mode] = (first, snapshot(s), code)
        s.close()
    assert out["1"][2] == out["0"][2]
    for u, v in zip(out["1"][0], out["0"][0]):
        same(u, v, "neighbours of the crowded start")
    compare(out["1"][1], out["0"][1])


def test_raw_buckets_equal_sorted_adjoint(T, monkeypatch):
    cfg = P.C1
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = dict(P.PHYS, sigma=cfg["sigma"])
    Lc = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    noise = T.gen_noise(2, 0.1, 15, cfg["n"])
    res = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TISSUE_SORTED_BUCKETS", mode)
        s = T.Sim(m, T.params(**p), cfg["n"])
        s.init_random(0)
        s.stage_noise(noise)
        code, g, dth, dx, dy = s.adjoint(15, Lc, 0, 4)
        assert code == T.OK
        res[mode] = (g, dth, dx, dy)
        s.close()
    g1, g0 = res["1"][0], res["0"][0]
    for key in ("d_tau", "d_mu", "d_v0", "d_k", "loss"):
        assert g1[key] == g0[key], key
    for u, v in zip(res["1"][1:], res["0"][1:]):
        same(u, v, "initial-state gradients")


def test_raw_buckets_equal_sorted_batch(T, monkeypatch):
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    sizes = [300, 1000, 650, 2400]
    ps = [T.params(sigma=sigma, v0=v, tau=t) for v, t in zip((0.7, 1.0, 1.5, 2.0), (0.3, 1.0, 3.0, 0.1))]
    seeds = np.array([31, 32, 33, 34], np.uint64)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("TISSUE_SORTED_BUCKETS", mode)
        b = T.Batch(m, sizes, ps)
        b.set_mode(1, 0)
        b.init_random(seeds)
        b.gen_noise(seeds, np.array([0.1, 0.2, 0.3, 0.4]), 40)
        code = b.run(40, 0)
        st = b.get_state()
        out[mode] = (code, [st.x, st.y, st.theta, st.vid])
        b.close()
    assert out["1"][0] == out["0"][0]
    for u, v in zip(out["1"][1], out["0"][1]):
        same(u, v, "batch state")
