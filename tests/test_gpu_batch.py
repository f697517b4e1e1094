"""GPU tests of the two execution modes and of batches of replicas.

Mode 1 is the tiled step kernel (one launch per step over every replica), mode 2 the persistent
LDS-resident kernel. Both call the same per-particle update and add neighbours in the same
(ascending id) order, so with the same lanes per particle their states are bitwise equal; a
batch must reproduce each replica run alone, in either mode and either workgroup order."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu

TILED, RESIDENT = 1, 2


def make_sim(T, m, cfg, mode, lanes, block=0):
    p = dict(P.PHYS, sigma=cfg["sigma"])
    s = T.Sim(m, T.params(**p), cfg["n"])
    s.set_mode(mode)
    s.set_launch(lanes, block)
    s.init_random(cfg["seed"])
    return s


def same_state(a, b):
    return all(np.array_equal(getattr(a, k), getattr(b, k)) for k in ("x", "y", "theta", "vid"))


@pytest.mark.parametrize("lanes", [1, 2, 8])
def test_resident_equals_step_kernel(T, lanes):
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    noise = T.gen_noise(cfg["noise_seed"], 0.1, 30, cfg["n"])
    a = make_sim(T, m, cfg, TILED, lanes)
    b = make_sim(T, m, cfg, RESIDENT, lanes)
    assert a.step(30, noise) == T.OK
    assert b.step(30, noise) == T.OK
    assert same_state(a.get_state(), b.get_state())
    oa, ob = a.get_outputs(), b.get_outputs()
    for k in ("vx", "vy", "r3d", "face", "nbr"):
        assert np.array_equal(getattr(oa, k), getattr(ob, k)), k
    (ra, pa), (rb, pb) = a.order(), b.order()
    assert np.array_equal(ra, rb)
    assert np.allclose(pa, pb, rtol=0, atol=1e-12)


def test_resident_teacher_forced_c1(T, O):
    cfg = P.C1
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"])
    sim = make_sim(T, m, cfg, RESIDENT, 0)
    noise = T.gen_noise(cfg["noise_seed"], 0.1, 50, cfg["n"])
    total = {}
    for t in range(50):
        s_in = sim.get_state()
        nb = sim.neighbor_lists()
        assert sim.step(1, noise[t]) == T.OK
        fails, _ = P.compare_step(O, om, p, s_in, sim.get_state(), sim.get_outputs(), nb,
                                  noise[t].astype(np.float64), phi_gpu=P.last_phi(sim))
        for k, v in fails.items():
            total[k] = total.get(k, 0) + v
    assert all(v == 0 for v in total.values()), total


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
def test_split_runs_equal_one_run(T, mode):
    """Runs are replayed hipGraphs: 3 runs of 7 steps (the first graph replayed) and one run
    of 21 steps from the same state give bitwise the same state and order history."""
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    noise = T.gen_noise(cfg["noise_seed"], 0.1, 21, cfg["n"])
    a = make_sim(T, m, cfg, mode, 2)
    b = make_sim(T, m, cfg, mode, 2)
    a.stage_noise(noise)
    b.stage_noise(noise)
    for k in range(3):
        assert a.step_staged(7, 7 * k) == T.OK
    assert b.step_staged(21, 0) == T.OK
    assert same_state(a.get_state(), b.get_state())
    assert np.array_equal(a.order()[0], b.order()[0])
    assert np.allclose(a.order()[1], b.order()[1], rtol=0, atol=1e-12)


BATCH_SIZES = [100, 517, 1000, 2000, 3100, 4000, 250, 1500, 777]


def batch_params(T, sigma, R):
    rng = np.random.default_rng(3)
    tau = np.exp(rng.uniform(np.log(0.1), np.log(10.0), R))
    std = rng.uniform(0.0, 0.5, R)
    v0 = rng.uniform(0.5, 2.0, R)
    ps = [T.params(sigma=sigma, v0=float(v0[r]), k=10.0, mu=1.0, tau=float(tau[r]), dt=0.02) for r in range(R)]
    return ps, std


@pytest.mark.parametrize("mode,xcd", [(TILED, 0), (TILED, 1), (RESIDENT, 0)])
def test_batch_equals_solo_runs(T, mode, xcd):
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    ps, std = batch_params(T, sigma, len(BATCH_SIZES))
    seeds = np.array([1000 + r for r in range(len(BATCH_SIZES))], np.uint64)
    steps = 40
    lanes = 1 if mode == RESIDENT else 2
    b = T.Batch(m, BATCH_SIZES, ps)
    b.set_mode(mode, xcd)
    b.set_launch(lanes, 512, 16)  # several launches / reductions per run
    b.init_random(seeds)
    b.gen_noise(seeds, std, steps)
    assert b.run(steps, 0) == T.OK
    st = b.get_state()
    out = b.get_outputs()
    for r, n in enumerate(BATCH_SIZES):
        o = int(b.off[r])
        s = T.Sim(m, ps[r], n)
        s.set_mode(mode)
        s.set_launch(lanes, 512)
        s.init_random(int(seeds[r]))
        assert s.step(steps, T.gen_noise(int(seeds[r]), float(std[r]), steps, n)) == T.OK
        ss = s.get_state()
        for k in ("x", "y", "theta", "vid"):
            assert np.array_equal(getattr(ss, k), getattr(st, k)[o:o + n]), (r, k)
        so = s.get_outputs()
        for k in ("vx", "vy", "face", "nbr"):
            assert np.array_equal(getattr(so, k), getattr(out, k)[o:o + n]), (r, k)
        (ra, pa), (rb, pb) = s.order(), b.order(r)
        assert np.array_equal(ra, rb) and np.allclose(pa, pb, rtol=0, atol=1e-12), r
        s.close()


def test_batch_autotune_restores_state(T):
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    ps, std = batch_params(T, sigma, len(BATCH_SIZES))
    seeds = np.arange(len(BATCH_SIZES), dtype=np.uint64) + 7
    b = T.Batch(m, BATCH_SIZES, ps)
    b.init_random(seeds)
    b.gen_noise(seeds, std, 10)
    before = b.get_state()
    lanes, block = b.autotune(10)
    assert lanes in (1, 2, 4, 8, 16, 32, 64) and block in (64, 128, 256, 512)
    assert b.launch()["mode"] in (TILED, RESIDENT)
    assert same_state(before, b.get_state())
    assert len(b.order(0)[0]) == 0
    assert b.run(10, 0) == T.OK
    assert len(b.order(0)[0]) == 10


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
def test_batch_failure_is_per_replica(T, mode):
    """A replica whose seam wrap exceeds its crossing bound (forced from the host with
    max_crossings = 1) stops and reports; the others run on exactly as alone. The stopped
    replica stays stopped in later runs."""
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    sizes = [800, 800, 800]
    ps = [T.params(sigma=sigma, v0=1.0, k=10.0, mu=1.0, tau=1.0, dt=0.02) for _ in sizes]
    ps[1] = T.params(sigma=sigma, v0=1.0, k=10.0, mu=1.0, tau=1.0, dt=0.02, max_crossings=1)
    seeds = np.array([11, 12, 13], np.uint64)
    b = T.Batch(m, sizes, ps)
    b.set_mode(mode, 0)
    b.set_launch(2, 512, 0)
    b.init_random(seeds)
    b.gen_noise(seeds, np.array([0.1, 0.1, 0.1]), 40)
    code = b.run(30, 0)
    flags, first = b.flags()
    assert code == T.E_TILING_NO_CONVERGE
    assert flags[1] & T.FLAG_TILING and flags[0] == 0 and flags[2] == 0
    assert 0 <= first[1] < 30
    s = T.Sim(m, ps[2], 800)
    s.set_mode(mode)
    s.set_launch(2, 512)
    s.init_random(13)
    assert s.step(30, T.gen_noise(13, 0.1, 30, 800)) == T.OK
    st = b.get_state()
    assert np.array_equal(s.get_state().x, st.x[1600:2400])
    # a later run leaves the stopped replica as it was and raises nothing new
    assert b.run(10, 30) == T.OK
    st2 = b.get_state()
    assert np.array_equal(st.x[800:1600], st2.x[800:1600])
    assert b.flags()[0][1] == 0
    assert not np.array_equal(st.x[1600:2400], st2.x[1600:2400])


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
def test_sim_nonfinite_noise_reported_and_halts(T, mode):
    cfg = P.C1
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    s = make_sim(T, m, cfg, mode, 0)
    noise = T.gen_noise(0, 0.1, 5, cfg["n"])
    noise[2, 17] = np.nan
    assert s.step(5, noise) == T.E_NONFINITE
    f, first = s.flags()
    assert f & T.FLAG_NONFINITE and first == 2
    # steps after the failure pass the state through: order history of the halted steps is zero
    r, p = s.order()
    assert len(r) == 5 and r[3] == 0.0 and r[4] == 0.0


@pytest.mark.parametrize("mode", [TILED, RESIDENT])
def test_sim_crossing_bound_reported(T, mode):
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = dict(P.PHYS, sigma=cfg["sigma"], max_crossings=1)
    s = T.Sim(m, T.params(**p), cfg["n"])
    s.set_mode(mode)
    s.init_random(cfg["seed"])
    assert s.step(20, T.gen_noise(1, 0.1, 20, cfg["n"])) == T.E_TILING_NO_CONVERGE
    f, first = s.flags()
    assert f & T.FLAG_TILING and first >= 0


def test_set_state_rejects_bad_input(T):
    m, _ = P.mesh_and_tables(22, 0.1)
    s = T.Sim(m, T.params(sigma=0.1), 3)
    with pytest.raises(T.TissueError):
        s.set_state([0.5, 1.5, 0.2], [0.5, 0.5, 0.5], [0, 0, 0])
    with pytest.raises(T.TissueError):
        s.set_state([0.5, 0.5, 0.2], [0.5, float("nan"), 0.5], [0, 0, 0])
    with pytest.raises(T.TissueError):
        s.step(1)  # no state yet


def test_ensemble_workload_equals_solo_runs(T):
    """The ensemble workload at full size: 256 replicas (sizes U[100, 4000], tau log-uniform in
    [0.1, 10], noise std U[0, 0.5], v0 U[0.5, 2]; parameter seed 3, init seed 2000 + r, noise seed
    1000 + r), 500 steps in one run at the autotuned mode and launch shape. Eight replicas spread
    over the size range, each run alone with the same lanes per particle, give its final state and
    both order-parameter histories. (Another lane count splits each particle's neighbour sums into
    other partial sums: its trajectories agree with these within the parity tolerance per step, not
    bit for bit.)"""
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    R, steps = 256, 500
    rng = np.random.default_rng(3)
    sizes = rng.integers(100, 4001, R)
    tau = np.exp(rng.uniform(np.log(0.1), np.log(10.0), R))
    std = rng.uniform(0.0, 0.5, R)
    v0 = rng.uniform(0.5, 2.0, R)
    ps = [T.params(sigma=sigma, v0=float(v0[r]), k=10.0, mu=1.0, tau=float(tau[r]), dt=0.02) for r in range(R)]
    init_seeds = np.arange(R, dtype=np.uint64) + 2000
    noise_seeds = np.arange(R, dtype=np.uint64) + 1000
    b = T.Batch(m, sizes, ps)
    b.init_random(init_seeds)
    b.gen_noise(noise_seeds, std, steps)
    b.autotune(20)
    b.init_random(init_seeds)
    assert b.run(steps, 0) == T.OK
    flags, _ = b.flags()
    assert not flags.any()
    st = b.get_state()
    shape = b.launch()
    print("ensemble launch", shape, flush=True)
    for r in np.argsort(sizes)[np.linspace(0, R - 1, 8).astype(int)]:
        n, o = int(sizes[r]), int(b.off[r])
        s = T.Sim(m, ps[r], n)
        s.set_mode(shape["mode"])
        s.set_launch(shape["lanes"], shape["block"])  # same lanes per particle: same summation order
        s.init_random(int(init_seeds[r]))
        s.stage_noise(T.gen_noise(int(noise_seeds[r]), float(std[r]), steps, n))
        assert s.step_staged(steps, 0) == T.OK
        ss = s.get_state()
        for k in ("x", "y", "theta", "vid"):
            assert np.array_equal(getattr(ss, k), getattr(st, k)[o:o + n]), (r, k)
        (ra, pa), (rb, pb) = s.order(), b.order(int(r))
        assert len(ra) == steps and np.array_equal(ra, rb) and np.allclose(pa, pb, rtol=0, atol=1e-12), r
        s.close()
    b.close()


@pytest.mark.parametrize("xcd", [0, 1])
def test_batch_walk_visits_the_oracle_sets(T, O, xcd):
    """The batch's own neighbour walk (visit check) at the batched launch shape, 2 lanes in
    128-thread workgroups with the vertex-occupancy mask staged in LDS, in both workgroup orders:
    every particle of every replica visits exactly its replica's oracle neighbour set (size and
    id hash) at each of 8 steps."""
    sigma = 0.05
    m, tb = P.mesh_and_tables(32, sigma)
    om = O.Mesh(tb)
    ps, std = batch_params(T, sigma, len(BATCH_SIZES))
    seeds = np.arange(len(BATCH_SIZES), dtype=np.uint64) + 40
    b = T.Batch(m, BATCH_SIZES, ps)
    b.set_mode(TILED, xcd)
    b.set_launch(2, 128, 1)
    b.set_visit_check(True)
    b.init_random(seeds)
    b.gen_noise(seeds, std, 8)
    for t in range(8):
        st = b.get_state()
        assert b.run(1, t) == T.OK
        na, h = b.visits()
        for r, n in enumerate(BATCH_SIZES):
            o = int(b.off[r])
            ena, eh = P.visit_records(O.neighbor_sets_sparse(om, st.vid[o:o + n], sigma), off=o)
            assert np.array_equal(na[o:o + n], ena) and np.array_equal(h[o:o + n], eh), (t, r)
    b.close()
