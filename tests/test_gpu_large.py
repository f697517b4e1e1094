"""The large surface: perturbed sphere r = 1 + 0.2 Y_3^2 at icosphere frequency 100 (F = 200,000),
sigma = 0.01. 100,000 particles teacher-forced against the oracle's O(N) step (bitwise equal to
the reference-literal dense step, see test_oracle.py), and 64 replicas of 2,000 particles run
forward and reverse together, each replica's gradients against the FP64 reverse pass along its
own trajectory."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("first", [2, 498])
def test_large_surface_teacher_forced(T, O, first):
    """Two steps teacher-forced, early (after leaving the initial placement) and at the end of the
    500-step run (after hundreds of long flights have piled particles onto the chart's corners)."""
    cfg = P.C5
    m, tb = P.sphere_and_tables(cfg["freq"], cfg["amp"], cfg["sigma"])
    om = O.Mesh(tb)
    O.set_threads(16)
    p = dict(P.PHYS, sigma=cfg["sigma"])
    sim = T.Sim(m, T.params(**p), cfg["n"])
    sim.set_mode(T.MODE_STEP_KERNEL)
    sim.set_visit_check(True)
    sim.init_random(cfg["seed"])
    noise = T.gen_noise(cfg["noise_seed"], 0.1, first + 2, cfg["n"])
    sim.stage_noise(noise)
    assert sim.step_staged(first, 0) == T.OK
    total = {}
    for t in range(first, first + 2):
        s_in = sim.get_state()
        nb = sim.neighbor_lists()
        assert sim.step(1, noise[t]) == T.OK
        fails, ref = P.compare_step(O, om, p, s_in, sim.get_state(), sim.get_outputs(), nb,
                                    noise[t].astype(np.float64), sparse=True, phi_gpu=P.last_phi(sim),
                                    visits=sim.visits())
        print(t, fails, f"phi_ref sign exempt: {ref['phi_exempt']}", flush=True)
        for k, v in fails.items():
            total[k] = total.get(k, 0) + v
    assert all(v == 0 for v in total.values()), total


@pytest.mark.parametrize("steps,ckpt", [(10, 5), (200, 20)])
def test_large_surface_batched_adjoint(T, O, steps, ckpt):
    """64 replicas of 2,000 particles forward and reverse together; three replicas' gradients
    against the FP64 reverse pass along each one's own trajectory, over 10 steps and over the 200
    steps (checkpoints every 20) the benchmark times, where the gradients reach ~1e130."""
    cfg = P.C5
    m, tb = P.sphere_and_tables(cfg["freq"], cfg["amp"], cfg["sigma"])
    om = O.Mesh(tb)
    R, n = 64, 2000
    rng = np.random.default_rng(6)
    taus = np.exp(rng.uniform(np.log(0.1), np.log(10.0), R))
    ps = [T.params(sigma=cfg["sigma"], v0=1.0, k=10.0, mu=1.0, tau=float(t), dt=0.02) for t in taus]
    seeds = np.arange(R, dtype=np.uint64) + 300
    b = T.Batch(m, [n] * R, ps)
    b.init_random(seeds)
    b.gen_noise(seeds, np.full(R, 0.1), steps)
    L = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    code, gs, dth, _, _ = b.adjoint(steps, L, 0, ckpt)
    assert code == T.OK
    shape = b.launch()
    fin = b.get_state()
    modes = T.target_field(4)
    comp, amp, kx, ky, ph = (np.array(v) for v in zip(*modes))
    Lo = O.Loss(0.8, 1.0, 0.1, component=comp, amp=amp, kx=kx, ky=ky, phase=ph)
    for r in (0, 17, 63):
        s = T.Sim(m, ps[r], n)
        s.set_mode(T.MODE_STEP_KERNEL)
        s.set_launch(shape["lanes"], shape["block"])  # the batch's lanes per particle: its trajectory
        s.init_random(int(seeds[r]))
        eta = T.gen_noise(int(seeds[r]), 0.1, steps, n)
        xs, ys, ths, vids = [], [], [], []
        for t in range(steps + 1):
            st = s.get_state()
            xs.append(st.x); ys.append(st.y); ths.append(st.theta); vids.append(st.vid)
            if t < steps:
                assert s.step(1, eta[t]) == T.OK
        states = dict(x=np.array(xs, np.float64), y=np.array(ys, np.float64), theta=np.array(ths, np.float64),
                      vid=np.array(vids, np.int32))
        o = r * n
        assert np.array_equal(xs[-1], fin.x[o:o + n]) and np.array_equal(vids[-1], fin.vid[o:o + n]), r
        pr = dict(sigma=cfg["sigma"], v0=1.0, k=10.0, mu=1.0, tau=float(taus[r]), dt=0.02)
        ref = O.adjoint_reverse(om, pr, states, Lo)
        for key in ("d_tau", "d_mu", "d_v0"):
            assert abs(gs[r][key] - ref[key]) <= 1e-4 * abs(ref[key]), (r, key, gs[r][key], ref[key])
        d = dth[o:o + n]
        sc = np.abs(ref["d_theta0"]).max()  # scaled first: squares of ~1e130 overflow FP64
        assert np.linalg.norm(d / sc - ref["d_theta0"] / sc) <= 1e-4 * np.linalg.norm(ref["d_theta0"] / sc), r
        print(steps, r, {k: ref[k] for k in ("d_tau", "d_mu", "d_v0")}, flush=True)
        s.close()
