"""The seam wrap on the GPU (tissue_seam_wrap, the loop the step kernel runs) against the oracle's
restatement of EuclideanTiling::diagonal_seam_edges_square_border (EuclideanTiling.cpp:31-69), bit
for bit, for every lane-group width the step kernel uses: long random flights (hundreds of
crossings), moves that start on edges and corners, and lines passing within about 1e-3 of a corner
(where the exit point snaps). Both loops are covered: the plain one, and the one of crowded
systems, where every wave's last flight is finished by the wave's 64 lanes three crossings per
pass, entered after however many crossings the other flights of the wave took."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LANES = (1, 2, 4, 8, 16, 32, 64)


def flights(seed=11, n=8000):
    rng = np.random.default_rng(seed)
    a = rng.random((n, 2))
    k = np.arange(n) % 8
    a[k == 1, 0] = 0.0
    a[k == 2, 1] = 0.0
    a[k == 3, 0] = 1.0
    a[k == 4, 1] = 1.0
    a[k == 5] = 0.0
    a[k == 6] = 1.0
    a[k == 7] = (0.0, 1.0)
    L = np.where(np.arange(n) % 3 == 0, 400.0 * rng.random(n), np.exp(rng.uniform(np.log(1e-4), np.log(30.0), n)))
    ang = rng.uniform(-np.pi, np.pi, n)
    b = a + L[:, None] * np.stack([np.cos(ang), np.sin(ang)], 1)
    return a, b


def corner_passes(seed=12, n=3000):
    rng = np.random.default_rng(seed)
    st = rng.random((n, 2))
    c = rng.choice([0.0, 1.0], (n, 2))
    off = rng.choice([1e-3, 9.99e-4, 1.001e-3, 1e-6, 0.0], (n, 2)) * rng.choice([-1.0, 1.0], (n, 2))
    b = st + ((c + off) - st) * rng.uniform(1.0, 30.0, (n, 1))
    return st, b


@pytest.fixture(scope="module")
def cases(O):
    out = []
    for name, (a, b) in (("flights", flights()), ("corner passes", corner_passes())):
        st, r, _, h, cr = O.seam_wrap(b, a, np.zeros(len(a)), unit_deg=0, max_crossings=16384)
        assert st == O.OK, name
        out.append((name, a, b, r, h, cr))
    return out


@pytest.mark.parametrize("wave", [False, True])
@pytest.mark.parametrize("lanes", LANES)
def test_seam_wrap_gpu_matches_oracle(T, cases, lanes, wave):
    for name, a, b, r, h, cr in cases:
        code, out, dth, q = T.seam_wrap(a, b, lanes=lanes, wave=wave)
        assert code == T.OK, name
        assert np.array_equal(out, r), name    # same crossings, same FP64 images
        assert np.array_equal(dth, h), name    # same rotations, accumulated in the same order
        if name == "flights":
            assert cr.max() > 200                 # long flights are exercised


def test_seam_wrap_gpu_bound(T):
    # a flight that needs more crossings than allowed reports it
    for wave in (False, True):
        code, out, dth, q = T.seam_wrap([[0.5, 0.5]], [[300.5, 120.5]], lanes=4, max_crossings=10, wave=wave)
        assert code == T.E_TILING_NO_CONVERGE


@pytest.mark.parametrize("wave", [False, True])
@pytest.mark.parametrize("lanes", [1, 16, 64])
def test_seam_wrap_gpu_bound_at_every_level(T, O, cases, lanes, wave):
    """Crossing bounds of 1 to 7 end flights at every level of the three-crossing passes: a move
    fails exactly when the oracle needs more crossings than allowed, and the others equal the
    oracle's bit for bit."""
    name, a, b, r, h, cr = cases[0]
    for maxc in range(1, 8):
        code, out, dth, q = T.seam_wrap(a, b, lanes=lanes, max_crossings=maxc, wave=wave)
        bad = cr > maxc
        assert (code == T.E_TILING_NO_CONVERGE) == bool(bad.any()), maxc
        good = ~bad
        assert np.array_equal(out[good], r[good]), maxc
        assert np.array_equal(dth[good], h[good]), maxc
