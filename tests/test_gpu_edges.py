"""GPU edge cases of the step, each against the FP64 oracle or against the same system run alone:

- the smallest systems (1, 2 and 3 particles), where a tile holds fewer particles than it has
  slots, at every lane count and in both execution modes;
- particles piled onto one mesh vertex (the reference's d == 0 rule, ForceHelper.cpp:60-63, with
  its |F| ~ k |Delta| / 0.001) and two particles at the very same position (Delta = 0), whose long
  flights cross the seam many times in one step;
- a batch whose replica sizes straddle the tile sizes of the launch shapes (1, 2, 3, 31 ... 257
  particles), which must equal each replica run alone, and whose neighbour walk must visit
  exactly the oracle's sets."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu

TILED, RESIDENT = 1, 2


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
@pytest.mark.parametrize("lanes", [1, 2, 4, 32, 64])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_smallest_systems_teacher_forced(T, O, n, lanes, mode):
    cfg = dict(P.C1, n=n)
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"])
    sim = T.Sim(m, T.params(**p), n)
    sim.set_mode(mode)
    sim.set_launch(lanes, 0)
    if mode == TILED:
        sim.set_visit_check(True)
    sim.init_random(7 + n)
    noise = T.gen_noise(3, 0.1, 30, n)
    total, _ = P.teacher_forced(O, om, p, sim, noise, visits=mode == TILED)
    sim.close()
    assert all(v == 0 for v in total.values()), total


def pile_on_vertex(O, om, tb, n_piled):
    """n_piled positions that share one nearest vertex (the remap's vertex id, so every pair among
    them has d = 0), spread as far apart as that vertex's region allows, the first two at the very
    same point: points of a disc around an interior UV vertex are located with the oracle's
    remap, the most frequent vertex id is kept and its points are taken in order of angle.
    Returns (vertex id, x, y)."""
    uv = tb["uv"]
    u = int(np.argmin(np.hypot(uv[:, 0] - 0.37, uv[:, 1] - 0.61)))
    faces = tb["faces"]
    ring = np.unique(faces[(faces == u).any(axis=1)])
    ring = ring[ring != u]
    e = float(np.hypot(uv[ring, 0] - uv[u, 0], uv[ring, 1] - uv[u, 1]).min())
    rng = np.random.default_rng(21)
    rad = e * np.sqrt(rng.uniform(0.0, 0.36, 400))
    ang = rng.uniform(-np.pi, np.pi, 400)
    x = (uv[u, 0] + rad * np.cos(ang)).astype(np.float32)
    y = (uv[u, 1] + rad * np.sin(ang)).astype(np.float32)
    _, vid, _ = O.remap(om, x, y)
    v = int(np.bincount(vid).argmax())
    sel = np.flatnonzero(vid == v)
    cx, cy = x[sel].mean(), y[sel].mean()
    sel = sel[np.argsort(np.arctan2(y[sel] - cy, x[sel] - cx))]
    pick = sel[np.linspace(0, len(sel) - 1, n_piled - 1).astype(int)]
    pick = np.concatenate([pick[:1], pick])
    return v, x[pick], y[pick]


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
@pytest.mark.parametrize("lanes", [1, 2, 32])
def test_particles_piled_on_one_vertex(T, O, lanes, mode):
    """Six particles with one nearest vertex on the C1 mesh (two at one point) among 40 others,
    with a stiff repulsion (k = 300): the d = 0 pairs fling the piled particles across the seam
    many times in the first step; 12 teacher-forced steps."""
    cfg = P.C1
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"], k=300.0)
    u, xp, yp = pile_on_vertex(O, om, tb, 6)
    rng = np.random.default_rng(5)
    n_free = 40
    x = np.concatenate([xp, rng.uniform(0.05, 0.95, n_free).astype(np.float32)])
    y = np.concatenate([yp, rng.uniform(0.05, 0.95, n_free).astype(np.float32)])
    th = rng.uniform(-np.pi, np.pi, len(x)).astype(np.float32)
    sim = T.Sim(m, T.params(**p), len(x))
    sim.set_mode(mode)
    sim.set_launch(lanes, 0)
    if mode == TILED:
        sim.set_visit_check(True)
    sim.set_state(x, y, th)
    assert (sim.get_state().vid[:6] == u).all()
    # the pile's first step is a long flight: count its crossings with the oracle's own loop
    _, _, out = O.step(om, p, x, y, th, sim.get_state().vid, np.zeros(len(x)), remap_mode=1)
    pre = np.stack([x + out["vx"] * p["dt"], y + out["vy"] * p["dt"]], 1)
    _, _, _, _, cr = O.seam_wrap(pre, np.stack([x, y], 1).astype(np.float64), np.zeros(len(x)), 0, 100000)
    print("crossings of the piled particles in the first step:", cr[:6].tolist(), flush=True)
    assert cr[:6].max() >= 4
    noise = T.gen_noise(11, 0.1, 12, len(x))
    total, _ = P.teacher_forced(O, om, p, sim, noise, visits=mode == TILED)
    sim.close()
    assert all(v == 0 for v in total.values()), total


RAGGED = [1, 2, 3, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 257]


def ragged_batch(T, m, sigma, steps, mode, lanes, block, xcd):
    R = len(RAGGED)
    rng = np.random.default_rng(9)
    ps = [T.params(sigma=sigma, v0=float(v), k=10.0, mu=1.0, tau=float(t), dt=0.02)
          for v, t in zip(rng.uniform(0.5, 2.0, R), np.exp(rng.uniform(np.log(0.1), np.log(10.0), R)))]
    std = rng.uniform(0.05, 0.5, R)
    seeds = np.arange(R, dtype=np.uint64) + 300
    b = T.Batch(m, RAGGED, ps)
    b.set_mode(mode, xcd)
    b.set_launch(lanes, block, 0)
    b.init_random(seeds)
    b.gen_noise(seeds, std, steps)
    return b, ps, std, seeds


@pytest.mark.parametrize("mode,lanes,block,xcd", [(TILED, 1, 64, 0), (TILED, 2, 128, 1), (TILED, 4, 256, 0),
                                                  (TILED, 32, 64, 1), (TILED, 64, 128, 0), (RESIDENT, 2, 512, 0)])
def test_ragged_batch_equals_solo_runs(T, mode, lanes, block, xcd):
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    steps = 12
    b, ps, std, seeds = ragged_batch(T, m, sigma, steps, mode, lanes, block, xcd)
    assert b.run(steps, 0) == T.OK
    st, out = b.get_state(), b.get_outputs()
    for r, n in enumerate(RAGGED):
        o = int(b.off[r])
        s = T.Sim(m, ps[r], n)
        s.set_mode(mode)
        s.set_launch(lanes, block)
        s.init_random(int(seeds[r]))
        assert s.step(steps, T.gen_noise(int(seeds[r]), float(std[r]), steps, n)) == T.OK
        ss, so = s.get_state(), s.get_outputs()
        for k in ("x", "y", "theta", "vid"):
            assert np.array_equal(getattr(ss, k), getattr(st, k)[o:o + n]), (n, k)
        for k in ("vx", "vy", "r3d", "face", "nbr"):
            assert np.array_equal(getattr(so, k), getattr(out, k)[o:o + n]), (n, k)
        (ra, pa), (rb, pb) = s.order(), b.order(r)
        assert np.array_equal(ra, rb) and np.allclose(pa, pb, rtol=0, atol=1e-12), n
        s.close()
    b.close()


@pytest.mark.parametrize("lanes,block", [(2, 128), (32, 64)])
def test_ragged_batch_walk_visits_the_oracle_sets(T, O, lanes, block):
    sigma = 0.05
    m, tb = P.mesh_and_tables(32, sigma)
    om = O.Mesh(tb)
    steps = 4
    b, _, _, _ = ragged_batch(T, m, sigma, steps, TILED, lanes, block, 1)
    b.set_visit_check(True)
    for t in range(steps):
        st = b.get_state()
        assert b.run(1, t) == T.OK
        na, h = b.visits()
        for r, n in enumerate(RAGGED):
            o = int(b.off[r])
            ena, eh = P.visit_records(O.neighbor_sets_sparse(om, st.vid[o:o + n], sigma), off=o)
            assert np.array_equal(na[o:o + n], ena) and np.array_equal(h[o:o + n], eh), (t, n)
    b.close()
