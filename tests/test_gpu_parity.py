"""GPU parity: the HIP step (through the C-ABI) against the FP64 oracle, teacher-forced: every
step of the GPU trajectory is recomputed by the oracle from the GPU's own input state. Both
execution modes are covered: one kernel launch per step, and the persistent LDS-resident kernel."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


def run_teacher_forced(T, O, cfg, steps, lanes=0, mode=0, sparse=False, block=0):
    """Teacher-forced run; with the step kernel (mode 1) the step's own neighbour walk is checked
    too: each particle's visited set (size and order-free id hash) against the oracle's set."""
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"])
    sim = T.Sim(m, T.params(**p), cfg["n"])
    sim.set_mode(mode)
    sim.set_launch(lanes, block)
    check_visits = mode == 1
    if check_visits:
        sim.set_visit_check(True)
    sim.init_random(cfg["seed"])
    noise = T.gen_noise(cfg["noise_seed"], 0.1, steps, cfg["n"])
    total, exempt = P.teacher_forced(O, om, p, sim, noise, visits=check_visits, sparse=sparse)
    print(f"{cfg['n']} particles x {steps} steps: {exempt} particle-steps exempt from the phi_ref sign "
          f"comparison (|x vy - y vx| within its rounding bound)", flush=True)
    sim.close()
    return total


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("lanes", [1, 8, 32, 64])
def test_c1_teacher_forced_50_steps(T, O, lanes, mode):
    total = run_teacher_forced(T, O, P.C1, 50, lanes, mode)
    assert all(v == 0 for v in total.values()), total


@pytest.mark.parametrize("mode", [1, 2])
def test_c2_teacher_forced_20_steps(T, O, mode):
    total = run_teacher_forced(T, O, P.C2, 20, 0, mode)
    assert all(v == 0 for v in total.values()), total


@pytest.mark.parametrize("cfg,lanes,block", [("C1", 2, 128), ("C2", 2, 128), ("C2", 4, 256), ("C2", 8, 512),
                                              ("C1", 16, 512)])
def test_occupancy_mask_walk_teacher_forced(T, O, cfg, lanes, block):
    """Launch shapes whose workgroups stage the vertex-occupancy mask in LDS (the batched shape,
    2 lanes in 128 threads; 4 lanes in 256; and at 8 and 16 lanes the walk that loads bucket
    heads in its scan): the walk that skips unoccupied entries from LDS visits exactly the
    oracle's neighbour sets (visit check) and the step matches the oracle."""
    total = run_teacher_forced(T, O, getattr(P, cfg), 20, lanes, 1, block=block)
    assert total.get("visits", -1) == 0
    assert all(v == 0 for v in total.values()), total


def test_c2_teacher_forced_1000_steps(T, O):
    """The whole 1,000-step run of the 2,000-particle system, every step checked against the
    oracle's O(N) step (bitwise the dense one): exact neighbour sets, counts, vertex and face ids,
    positions, velocities and headings within 1e-5."""
    total = run_teacher_forced(T, O, P.C2, 1000, 0, 1, sparse=True)
    assert all(v == 0 for v in total.values()), total


@pytest.mark.parametrize("mode", [1, 2])
def test_staged_noise_rows_match_host_noise(T, mode):
    """Staged noise read from a moving first row (the captured step graphs take the row from device
    memory) gives, bit for bit, the trajectory of the same rows passed from the host."""
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = T.params(**dict(P.PHYS, sigma=cfg["sigma"]))
    noise = T.gen_noise(cfg["noise_seed"], 0.1, 40, cfg["n"])
    a, b = T.Sim(m, p, cfg["n"]), T.Sim(m, p, cfg["n"])
    for s in (a, b):
        s.set_mode(mode)
        s.init_random(cfg["seed"])
    a.stage_noise(noise)
    for row in (3, 11, 29, 11):  # the same step count from different rows, one row twice
        assert a.step_staged(6, row) == T.OK
        assert b.step(6, noise[row:row + 6]) == T.OK
        sa, sb = a.get_state(), b.get_state()
        for k in ("x", "y", "theta", "vid"):
            assert np.array_equal(getattr(sa, k), getattr(sb, k)), (row, k)
    a.close(); b.close()


def test_gpu_initialisation_matches_host(T):
    """init_random draws the state on the GPU; it equals the host initialisation bit for bit, for a
    single system and for every replica of a batch (replica r from its own seed)."""
    cfg = P.C2
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = T.params(**dict(P.PHYS, sigma=cfg["sigma"]))
    s = T.Sim(m, p, cfg["n"])
    s.init_random(123)
    st = s.get_state()
    x, y, th = m.random_init_host(123, cfg["n"])
    assert np.array_equal(st.x, x) and np.array_equal(st.y, y) and np.array_equal(st.theta, th)
    s.close()
    sizes = [100, 2500, 731]
    b = T.Batch(m, sizes, [p] * len(sizes))
    seeds = np.array([5, 6, 2**40 + 3], np.uint64)
    b.init_random(seeds)
    bs = b.get_state()
    o = 0
    for r, nr in enumerate(sizes):
        x, y, th = m.random_init_host(int(seeds[r]), nr)
        assert np.array_equal(bs.x[o:o + nr], x) and np.array_equal(bs.y[o:o + nr], y), r
        assert np.array_equal(bs.theta[o:o + nr], th), r
        o += nr
    b.close()


def test_restaged_fewer_rows_are_the_valid_range(T):
    """Staging 40 noise rows and then 10 leaves only the 10 in range: a run that asks for rows the
    second staging did not write is rejected instead of reading the first staging's leftovers."""
    cfg = P.C1
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    s = T.Sim(m, T.params(**dict(P.PHYS, sigma=cfg["sigma"])), cfg["n"])
    s.init_random(cfg["seed"])
    s.stage_noise(T.gen_noise(1, 0.1, 40, cfg["n"]))
    assert s.step_staged(5, 20) == T.OK
    s.stage_noise(T.gen_noise(2, 0.1, 10, cfg["n"]))
    with pytest.raises(T.TissueError) as e:
        s.step_staged(5, 20)
    assert e.value.code == T.E_INVALID_ARG
    assert s.step_staged(5, 5) == T.OK
    s.close()
