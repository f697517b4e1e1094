"""GPU adjoint (tissue_adjoint / tissue_batch_adjoint) against the FP64 oracle's reverse pass,
evaluated along the GPU's own trajectory (the states are recorded by stepping an identical
simulation one step at a time, which the checkpointed forward reproduces bitwise). The oracle
reverse pass itself is pinned by central finite differences in test_adjoint_oracle.py."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu

TOL = 1e-4  # relative, FP32 adjoint vs FP64 reference


def c4_loss(T, O, w_order=1.0, w_velocity=0.1):
    modes = T.target_field(4)
    Lc = T.loss(0.8, w_order, w_velocity, modes)
    comp, amp, kx, ky, ph = (np.array(v) for v in zip(*modes))
    Lo = O.Loss(0.8, w_order, w_velocity, component=comp, amp=amp, kx=kx, ky=ky, phase=ph)
    return Lc, Lo


def make(T, m, p, n, seed, mode=1):
    s = T.Sim(m, T.params(**p), n)
    s.set_mode(mode)
    s.init_random(seed)
    return s


def record(T, sim, noise):
    """States 0..steps of a simulation stepped one step at a time."""
    steps = noise.shape[0]
    xs, ys, ths, vids = [], [], [], []
    for t in range(steps + 1):
        st = sim.get_state()
        xs.append(st.x); ys.append(st.y); ths.append(st.theta); vids.append(st.vid)
        if t < steps:
            assert sim.step(1, noise[t]) == T.OK
    return dict(x=np.array(xs, np.float64), y=np.array(ys, np.float64), theta=np.array(ths, np.float64),
                vid=np.array(vids, np.int32))


def rel(a, b):
    """||a - b|| / ||b||, scaled first: the gradients reach 1e248 at 200 steps, whose squares
    overflow FP64."""
    a = np.atleast_1d(np.asarray(a, np.float64)); b = np.atleast_1d(np.asarray(b, np.float64))
    s = np.abs(b).max()
    if s == 0:
        return float(np.abs(a).max())
    return float(np.linalg.norm(a / s - b / s) / np.linalg.norm(b / s))


def compare(g, dth, dx, dy, ref, keys=("d_tau", "d_mu", "d_v0", "d_k")):
    out = {k: rel(g[k], ref[k]) for k in keys}
    out["d_theta0"] = rel(dth, ref["d_theta0"])
    out["d_r0"] = rel(np.concatenate([dx, dy]), np.concatenate([ref["d_x0"], ref["d_y0"]]))
    out["loss"] = rel(g["loss"], ref["loss"])
    return out


@pytest.mark.parametrize("steps,ckpt", [(6, 4), (12, 5)])
def test_adjoint_matches_oracle_on_gpu_trajectory(T, O, steps, ckpt):
    cfg = P.C1
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"], tau=0.5, mu=0.8)
    Lc, Lo = c4_loss(T, O)
    noise = T.gen_noise(2, 0.1, steps, cfg["n"])
    a = make(T, m, p, cfg["n"], 0)
    a.stage_noise(noise)
    code, g, dth, dx, dy = a.adjoint(steps, Lc, 0, ckpt)
    assert code == T.OK
    states = record(T, make(T, m, p, cfg["n"], 0), noise)
    fin = a.get_state()
    assert np.array_equal(fin.x, states["x"][-1].astype(np.float32)) and np.array_equal(fin.vid, states["vid"][-1])
    ref = O.adjoint_reverse(om, p, states, Lo)
    errs = compare(g, dth, dx, dy, ref)
    print(errs, {k: ref[k] for k in ("loss", "d_tau", "d_mu", "d_v0", "d_k")}, flush=True)
    assert all(v <= TOL for v in errs.values()), errs


def test_adjoint_checkpoint_interval_is_exact(T):
    """The recomputed segments reproduce the forward bitwise, so every checkpoint interval gives
    bitwise the same gradients."""
    cfg = P.C1
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = dict(P.PHYS, sigma=cfg["sigma"])
    Lc = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    noise = T.gen_noise(2, 0.1, 15, cfg["n"])
    res = []
    for k in (1, 4, 15):
        s = make(T, m, p, cfg["n"], 0)
        s.stage_noise(noise)
        code, g, dth, dx, dy = s.adjoint(15, Lc, 0, k)
        assert code == T.OK
        res.append((g, dth, dx, dy))
    for g, dth, dx, dy in res[1:]:
        for key in ("d_tau", "d_mu", "d_v0", "d_k", "loss"):
            assert g[key] == res[0][0][key], key
        assert np.array_equal(dth, res[0][1]) and np.array_equal(dx, res[0][2]) and np.array_equal(dy, res[0][3])


def test_adjoint_tape_forms_are_exact(T, monkeypatch):
    """The reverse pass in its three forms gives bitwise the same gradients: several tapes whose
    recomputes overlap the reverse passes (the default), one tape with its states' bucket sets,
    and one tape that re-buckets each state (memory budgets set by TISSUE_ADJOINT_BUCKET_KB)."""
    cfg = P.C1
    m, _ = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    p = dict(P.PHYS, sigma=cfg["sigma"])
    Lc = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    noise = T.gen_noise(2, 0.1, 15, cfg["n"])
    K = 4
    set_bytes = m.info()["nv"] * 36 + cfg["n"] * 16 + 12
    one_tape_kb = int(1.5 * (K + 2) * set_bytes) // 1024
    res = []
    for budget in (None, one_tape_kb, 0):
        if budget is None:
            monkeypatch.delenv("TISSUE_ADJOINT_BUCKET_KB", raising=False)
        else:
            monkeypatch.setenv("TISSUE_ADJOINT_BUCKET_KB", str(budget))
        s = make(T, m, p, cfg["n"], 0)
        s.stage_noise(noise)
        code, g, dth, dx, dy = s.adjoint(15, Lc, 0, K)
        assert code == T.OK
        res.append((g, dth, dx, dy))
        s.close()
    for g, dth, dx, dy in res[1:]:
        for key in ("d_tau", "d_mu", "d_v0", "d_k", "loss"):
            assert g[key] == res[0][0][key], key
        assert np.array_equal(dth, res[0][1]) and np.array_equal(dx, res[0][2]) and np.array_equal(dy, res[0][3])


def test_batch_adjoint_equals_solo(T):
    sigma = 0.05
    m, _ = P.mesh_and_tables(32, sigma)
    sizes = [300, 1000, 650]
    taus = [0.3, 1.0, 3.0]
    ps = [T.params(sigma=sigma, v0=1.0, k=10.0, mu=1.0, tau=t, dt=0.02) for t in taus]
    seeds = np.array([21, 22, 23], np.uint64)
    steps = 10
    Lc = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    b = T.Batch(m, sizes, ps)
    b.set_mode(1, 0)
    b.init_random(seeds)
    b.gen_noise(seeds, np.array([0.1, 0.2, 0.3]), steps)
    code, gs, dth, dx, dy = b.adjoint(steps, Lc, 0, 4)
    assert code == T.OK
    shape = b.launch()
    for r, n in enumerate(sizes):
        o = int(b.off[r])
        s = T.Sim(m, ps[r], n)
        s.set_mode(1)
        s.set_launch(shape["lanes"], shape["block"])  # same lanes per particle: same sums
        s.init_random(int(seeds[r]))
        s.stage_noise(T.gen_noise(int(seeds[r]), [0.1, 0.2, 0.3][r], steps, n))
        c2, g, sth, sx, sy = s.adjoint(steps, Lc, 0, 4)
        assert c2 == T.OK
        for key in ("d_tau", "d_mu", "d_v0", "d_k", "loss", "phi_final"):
            assert g[key] == gs[r][key], (r, key)
        assert np.array_equal(sth, dth[o:o + n]) and np.array_equal(sx, dx[o:o + n]), r
        s.close()


def test_adjoint_differentiation_workload(T, O):
    """The differentiation workload: 1,000 particles on the frequency-32 ellipsoid, sigma 0.05,
    200 steps, noise seed 2, checkpoints every 20 steps, loss (phi_T - 0.8)^2 + 0.1 mean |rdot - u|^2
    with the seeded field u (seed 4). The gradient grows geometrically with the horizon (the d == 0
    repulsion rule), so it is compared with the FP64 reverse pass along the same trajectory; the
    heading-only loss (w_velocity = 0) is compared as well."""
    sigma = 0.05
    m, tb = P.mesh_and_tables(32, sigma)
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=sigma)
    n, steps = 1000, 200
    noise = T.gen_noise(2, 0.1, steps, n)
    states = record(T, make(T, m, p, n, 2), noise)
    for w_v in (0.1, 0.0):
        Lc, Lo = c4_loss(T, O, 1.0, w_v)
        a = make(T, m, p, n, 2)
        a.stage_noise(noise)
        code, g, dth, dx, dy = a.adjoint(steps, Lc, 0, 20)
        assert code == T.OK
        ref = O.adjoint_reverse(om, p, states, Lo)
        keys = ("d_tau", "d_mu", "d_v0", "d_k") if w_v > 0 else ("d_tau",)
        errs = compare(g, dth, dx, dy, ref, keys)
        if w_v == 0:
            errs.pop("d_r0")
            assert g["d_mu"] == 0.0 and g["d_v0"] == 0.0 and not dx.any()
        print(w_v, errs, {k: ref[k] for k in ("loss", "d_tau", "d_mu", "d_v0", "d_k")},
              "fwd ms", g["forward_ms"], "rev ms", g["reverse_ms"], flush=True)
        assert all(v <= TOL for v in errs.values()), errs


def test_torch_autograd_wrapper(T):
    """torch.autograd through the C-ABI adjoint: backward returns tissue_adjoint's gradients, and
    a small descent step along them lowers the (heading-only) loss."""
    import torch
    from torch_adjoint import TrajectoryLoss

    cfg0 = P.C1
    m, _ = P.mesh_and_tables(cfg0["freq"], cfg0["sigma"])
    n, steps = cfg0["n"], 5
    sim = make(T, m, dict(P.PHYS, sigma=cfg0["sigma"]), n, 0)
    st = sim.get_state()
    cfg = dict(lib=T, mesh=m, params=dict(P.PHYS, sigma=cfg0["sigma"]), x0=st.x, y0=st.y,
               noise=T.gen_noise(2, 0.1, steps, n), steps=steps, loss=T.loss(0.8, 1.0, 0.0, []))
    theta0 = torch.tensor(st.theta.astype(np.float64), requires_grad=True)
    tau = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    L = TrajectoryLoss.apply(theta0, tau, mu, v0, cfg)
    L.backward()
    # the same call straight through the C-ABI
    s = T.Sim(m, T.params(**dict(cfg["params"], tau=0.5)), n)
    s.set_mode(T.MODE_STEP_KERNEL)
    s.set_state(st.x, st.y, st.theta)
    s.stage_noise(cfg["noise"])
    code, g, dth, _, _ = s.adjoint(steps, cfg["loss"], 0, 20)
    assert code == T.OK
    assert np.array_equal(theta0.grad.numpy(), dth) and tau.grad.item() == g["d_tau"]
    assert mu.grad.item() == 0.0 and v0.grad.item() == 0.0  # heading-only loss
    # descent: a step of 1e-4 in the largest gradient entry
    lr = 1e-4 / float(theta0.grad.abs().max())
    with torch.no_grad():
        L2 = TrajectoryLoss.apply(theta0 - lr * theta0.grad, tau, mu, v0, cfg)
    assert L2.item() < L.item()


def test_gpu_adjoint_matches_fd_differentiation_workload(T, O):
    """The GPU adjoint on the differentiation workload (1,000 particles, frequency-32 ellipsoid,
    init and noise seed 2, target field of seed 4) against central differences of the FP64
    oracle's forward from the same initial state: within 1e-4 for d tau, d mu, d v0 and a random
    direction of the initial headings. The horizon is 2 steps, the longest at which the GPU's FP32
    trajectory makes the FP64 trajectory's decisions (a vertex id differs at step 3), so that both
    differentiate the same map; beyond it the GPU gradient is checked against the FP64 reverse
    pass along its own trajectory (above), and that reverse pass against central differences up
    to 4 steps (test_adjoint_oracle.py)."""
    from test_adjoint_oracle import FD_STEPS, c4_setup, central_difference, sparse_run
    om, p, Lo, s0, noise = c4_setup(T, O)
    H, n = 2, len(s0[0])
    m, _ = P.mesh_and_tables(32, 0.05)
    Lc = T.loss(0.8, 1.0, 0.1, T.target_field(4))
    sim = make(T, m, p, n, 2)
    st0 = sim.get_state()
    assert np.array_equal(st0.x.astype(np.float64), s0[0]) and np.array_equal(st0.vid, s0[3])
    sim.stage_noise(noise[:H].astype(np.float32))
    code, g, dth, _, _ = sim.adjoint(H, Lc, 0, 20)
    assert code == T.OK
    gpu_states = record(T, make(T, m, p, n, 2), noise[:H].astype(np.float32))
    base, states = sparse_run(O, om, p, s0, noise[:H], Lo)
    assert np.array_equal(gpu_states["vid"], states["vid"])  # the same decisions at every step
    errs = {}
    for name, key in (("tau", "d_tau"), ("mu", "d_mu"), ("v0", "d_v0")):
        def f(h, name=name):
            return sparse_run(O, om, dict(p, **{name: p[name] + h}), s0, noise[:H], Lo)

        fd = central_difference(f, states["vid"], FD_STEPS)
        errs[key] = abs(fd - g[key]) / abs(fd)
    v = np.random.default_rng(11).standard_normal(n)
    fd = central_difference(lambda h: sparse_run(O, om, p, s0, noise[:H], Lo, dth=h * v), states["vid"],
                            tuple(h * 0.1 for h in FD_STEPS))
    errs["d_theta0"] = abs(fd - dth @ v) / abs(fd)
    print(errs, {k: g[k] for k in ("loss", "d_tau", "d_mu", "d_v0")}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs
