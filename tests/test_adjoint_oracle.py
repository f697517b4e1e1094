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
