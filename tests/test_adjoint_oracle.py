"""The FP64 oracle's reverse pass (oracle/adjoint.cpp) pinned by central finite differences of
the oracle's own FP64 forward, over horizons on which the perturbed runs make the same discrete
decisions (vertex ids, hence neighbour sets) as the base run. The reference has no derivative, so
finite differences of the (reference-restating) forward are the independent check; the loss
formula is checked against a numpy restatement."""
import numpy as np
import pytest

import parity as P


@pytest.fixture(scope="module")
def c1(O):
    m, tb = P.mesh_and_tables(22, 0.1)
    return O.Mesh(tb)


def modes(seed=4):
    rng = np.random.default_rng(seed)
    comp = [0, 0, 1, 1]
    return dict(component=comp, amp=rng.uniform(-1, 1, 4), kx=2 * np.pi * rng.integers(-2, 3, 4),
                ky=2 * np.pi * rng.integers(-2, 3, 4), phase=rng.uniform(0, 2 * np.pi, 4))


def initial(O, om, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, n)
    y = rng.uniform(0.05, 0.95, n)
    th = rng.uniform(-np.pi, np.pi, n)
    _, vid, _ = O.remap(om, x, y)
    return x, y, th, vid


def test_field_eval_and_jacobian(O):
    L = O.Loss(**modes())
    rng = np.random.default_rng(0)
    for x, y in rng.uniform(0, 1, (20, 2)):
        u, J = O.field_eval(L, x, y)
        ux, uy = L.field(np.array([x]), np.array([y]))
        assert np.allclose(u, [ux[0], uy[0]], rtol=0, atol=1e-14)
        h = 1e-6
        for k, (dx, dy) in enumerate(((h, 0), (0, h))):
            up, _ = O.field_eval(L, x + dx, y + dy)
            um, _ = O.field_eval(L, x - dx, y - dy)
            assert np.allclose((up - um) / (2 * h), J[:, k], rtol=1e-6, atol=1e-6)


def run_loss(O, om, p, x, y, th, vid, noise, L):
    st, s, vel = O.trajectory(om, p, x, y, th, vid, noise)
    return O.trajectory_loss(s, vel, L)["loss"], s, st


def central_difference(f, base_vid, hs):
    """(f(+h) - f(-h)) / 2h for the largest h in hs whose two runs keep every vertex id of the
    base run (the same discrete decisions, so the frozen-decision derivative applies)."""
    for h in hs:
        lp, sp = f(h)
        lm, sm = f(-h)
        if np.array_equal(sp["vid"], base_vid) and np.array_equal(sm["vid"], base_vid):
            return (lp - lm) / (2 * h)
    raise AssertionError("every step size changed a decision")


CASES = [
    # (particles, steps, parameter overrides, loss weights (w_order, w_velocity))
    (100, 5, {}, (1.0, 0.1)),
    (60, 8, dict(mu=0.7, tau=0.3, v0=1.3), (1.0, 0.1)),
    (100, 6, dict(tau=0.5), (1.0, 0.0)),
]


@pytest.mark.parametrize("n,steps,over,w", CASES)
def test_oracle_adjoint_matches_central_differences(O, c1, n, steps, over, w):
    p = dict(P.PHYS, sigma=0.1, **over)
    L = O.Loss(0.8, w[0], w[1], **modes())
    x, y, th, vid = initial(O, c1, n, 7)
    noise = np.random.default_rng(2).normal(0, 0.1, (steps, n))
    base, states, st = run_loss(O, c1, p, x, y, th, vid, noise, L)
    assert st == O.OK
    g = O.adjoint_reverse(c1, p, states, L)
    assert abs(g["loss"] - base) <= 1e-12 * abs(base)
    # parameters
    for name, key in (("tau", "d_tau"), ("mu", "d_mu"), ("v0", "d_v0"), ("k", "d_k")):
        scale = max(abs(p[name]), 1.0)

        def f(h, name=name, scale=scale):
            l, s, _ = run_loss(O, c1, dict(p, **{name: p[name] + h * scale}), x, y, th, vid, noise, L)
            return l, s

        fd = central_difference(f, states["vid"], (1e-6, 1e-7, 1e-8)) / scale
        assert abs(fd - g[key]) <= 1e-5 * max(abs(fd), 1e-3 * abs(base)), (name, fd, g[key])
    # initial headings and positions, along random directions
    rng = np.random.default_rng(11)
    v = rng.standard_normal(n)

    def f_th(h):
        l, s, _ = run_loss(O, c1, p, x, y, th + h * v, vid, noise, L)
        return l, s

    fd = central_difference(f_th, states["vid"], (1e-7, 1e-8, 1e-9))
    ad = g["d_theta0"] @ v
    assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)
    vx, vy = rng.standard_normal(n), rng.standard_normal(n)

    def f_r(h):
        l, s, _ = run_loss(O, c1, p, x + h * vx, y + h * vy, th, vid, noise, L)
        return l, s

    fd = central_difference(f_r, states["vid"], (1e-8, 1e-9))
    ad = g["d_x0"] @ vx + g["d_y0"] @ vy
    assert abs(fd - ad) <= 1e-4 * abs(fd), (fd, ad)


def test_order_only_loss_has_no_position_gradient(O, c1):
    """With w_velocity = 0 the loss depends on headings only, and in the frozen map headings do
    not depend on positions: d mu, d v0, d k and d r0 vanish exactly; d tau does not."""
    p = dict(P.PHYS, sigma=0.1, tau=0.5)
    L = O.Loss(0.8, 1.0, 0.0, **modes())
    x, y, th, vid = initial(O, c1, 80, 3)
    noise = np.random.default_rng(2).normal(0, 0.1, (10, 80))
    _, states, _ = run_loss(O, c1, p, x, y, th, vid, noise, L)
    g = O.adjoint_reverse(c1, p, states, L)
    assert g["d_mu"] == 0.0 and g["d_v0"] == 0.0 and g["d_k"] == 0.0
    assert not g["d_x0"].any() and not g["d_y0"].any()
    assert g["d_tau"] != 0.0 and np.abs(g["d_theta0"]).max() > 0


def test_isolated_particle_closed_form(O, c1):
    """One particle, no neighbours: theta_T = theta_0 + sum eta (the mean is its own heading) and
    the order loss (1 - 0.8)^2 has zero gradient; the velocity loss of a constant field u = (a, 0)
    is mean |v0 e(theta_t) - u|^2, whose theta_0 derivative is 2 w v0 a sin(theta_t) averaged."""
    p = dict(P.PHYS, sigma=0.1)
    a = 0.3
    L = O.Loss(0.8, 1.0, 0.5, component=[0], amp=[a], kx=[0.0], ky=[0.0], phase=[0.5 * np.pi])
    x, y, th, vid = np.array([0.4]), np.array([0.6]), np.array([0.25]), None
    _, vid, _ = O.remap(c1, x, y)
    noise = np.random.default_rng(5).normal(0, 0.1, (6, 1))
    _, states, _ = run_loss(O, c1, p, x, y, th, vid, noise, L)
    g = O.adjoint_reverse(c1, p, states, L)
    thetas = states["theta"][:-1, 0]
    expect = np.mean(2 * 0.5 * p["v0"] * a * np.sin(thetas))
    assert abs(g["d_theta0"][0] - expect) <= 1e-12
    assert abs(g["d_v0"] - np.mean(2 * 0.5 * (p["v0"] - a * np.cos(thetas)))) <= 1e-12


def c4_setup(T, O):
    """The differentiation workload: 1,000 particles on the frequency-32 ellipsoid, sigma 0.05,
    initial state of seed 2 (tissue_init_random's draw, computed on the host), noise seed 2
    (std 0.1), loss (phi_T - 0.8)^2 + 0.1 mean |rdot - u|^2 with the target field of seed 4."""
    m, tb = P.mesh_and_tables(32, 0.05)
    om = O.Mesh(tb)
    n = 1000
    x, y, th = (a.astype(np.float64) for a in m.random_init_host(2, n))
    _, vid, _ = O.remap(om, x, y, 1)
    fm = T.target_field(4)
    comp, amp, kx, ky, ph = (np.array(v) for v in zip(*fm))
    L = O.Loss(0.8, 1.0, 0.1, component=comp, amp=amp, kx=kx, ky=ky, phase=ph)
    p = dict(P.PHYS, sigma=0.05)
    noise = T.gen_noise(2, 0.1, 8, n).astype(np.float64)
    return om, p, L, (x, y, th, vid), noise


def sparse_run(O, om, p, s0, noise, L, dth=None):
    """FP64 forward (the oracle's O(N) step, bitwise its dense one) and the loss of its trajectory."""
    x, y, th, vid = s0
    s = dict(x=x, y=y, theta=th if dth is None else th + dth, vid=vid)
    keys = ("x", "y", "theta", "vid")
    hist = {k: [] for k in keys}
    vx, vy = [], []
    for t in range(len(noise)):
        for k in keys:
            hist[k].append(s[k])
        st, s, out = O.step_sparse(om, p, s["x"], s["y"], s["theta"], s["vid"], noise[t])
        assert st == O.OK
        vx.append(out["vx"]); vy.append(out["vy"])
    for k in keys:
        hist[k].append(s[k])
    states = {k: np.array(v) for k, v in hist.items()}
    return O.trajectory_loss(states, (np.array(vx), np.array(vy)), L)["loss"], states


FD_STEPS = (1e-5, 3e-6, 1e-6, 3e-7, 1e-7, 3e-8, 1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11)


def test_oracle_adjoint_matches_fd_differentiation_workload(T, O):
    """On the differentiation workload itself, the reverse pass equals central differences of the
    FP64 forward within 1e-4 for d tau, d mu, d v0 and a random direction of the initial headings,
    over 4 steps: the longest horizon at which, for all four, a step size from FD_STEPS keeps every
    vertex id of every step (at 5 steps d v0 agrees only to 1.5e-4, and from 6 steps on no step
    size keeps the decisions, as the d == 0 rule multiplies sensitivities about 10x per step)."""
    om, p, L, s0, noise = c4_setup(T, O)
    H = 4
    base, states = sparse_run(O, om, p, s0, noise[:H], L)
    g = O.adjoint_reverse(om, p, states, L)
    assert abs(g["loss"] - base) <= 1e-12 * abs(base)
    errs = {}
    for name, key in (("tau", "d_tau"), ("mu", "d_mu"), ("v0", "d_v0")):
        def f(h, name=name):
            return sparse_run(O, om, dict(p, **{name: p[name] + h}), s0, noise[:H], L)

        fd = central_difference(f, states["vid"], FD_STEPS)
        errs[key] = abs(fd - g[key]) / abs(fd)
    v = np.random.default_rng(11).standard_normal(len(s0[0]))
    fd = central_difference(lambda h: sparse_run(O, om, p, s0, noise[:H], L, dth=h * v), states["vid"],
                            tuple(h * 0.1 for h in FD_STEPS))
    errs["d_theta0"] = abs(fd - g["d_theta0"] @ v) / abs(fd)
    print(errs, {k: g[k] for k in ("loss", "d_tau", "d_mu", "d_v0")}, flush=True)
    assert all(e <= 1e-4 for e in errs.values()), errs
