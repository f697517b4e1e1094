"""Long runs are compared through order-parameter statistics. With the reference's d == 0 rule a particle
sharing a vertex with another is thrown across the chart, so two trajectories from the same inputs
part within a few steps: storing the FP64 oracle's own state in FP32 moves particles by more than
1e-5 after two steps and decorrelates the whole system after about six (test_oracle.py pins this
horizon). Per-step parity is therefore tested teacher-forced (test_gpu_parity.py); here the GPU's
free-running 1,000-step runs of the 2,000-particle system are compared with the oracle's
free-running runs from the same initial states and noise, through the statistics of the two order
parameters."""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu

SEEDS = (1, 2, 3)
STEPS = 1000
SKIP, BLOCK = 100, 50  # transient, and the block length of the block averages (several correlation times)
Z_MAX = 4.0


def block_means(series):
    """Block averages of each run after the transient, pooled over the runs."""
    return np.concatenate([np.asarray(a[SKIP:]).reshape(-1, BLOCK).mean(1) for a in series])


def zscore(a, b):
    """Difference of two pooled block-mean samples in units of its standard error."""
    se = np.hypot(a.std(ddof=1) / np.sqrt(len(a)), b.std(ddof=1) / np.sqrt(len(b)))
    return (a.mean() - b.mean()) / se


def test_c2_long_run_order_statistics(T, O):
    cfg = P.C2
    m, tb = P.mesh_and_tables(cfg["freq"], cfg["sigma"])
    om = O.Mesh(tb)
    p = dict(P.PHYS, sigma=cfg["sigma"])
    n = cfg["n"]
    gpu = {"phi_ref": [], "phi_polar": []}
    ref = {"phi_ref": [], "phi_polar": []}
    for seed in SEEDS:
        noise = T.gen_noise(seed, 0.1, STEPS, n)
        s = T.Sim(m, T.params(**p), n)
        s.init_random(seed)
        st = s.get_state()
        s.stage_noise(noise)
        assert s.step_staged(STEPS, 0) == T.OK
        phi_ref, phi_polar = s.order()
        assert len(phi_polar) == STEPS
        fin = s.get_state()
        assert ((fin.x >= 0) & (fin.x <= 1) & (fin.y >= 0) & (fin.y <= 1)).all()
        s.close()
        gpu["phi_ref"].append(phi_ref)
        gpu["phi_polar"].append(phi_polar)
        # the oracle from the same FP32 initial state and noise
        _, vid, _ = O.remap(om, st.x, st.y, 1)
        assert np.array_equal(vid, st.vid)
        o = dict(x=st.x.astype(np.float64), y=st.y.astype(np.float64), theta=st.theta.astype(np.float64), vid=vid)
        pr, pp = [], []
        nz = noise.astype(np.float64)
        for t in range(STEPS):
            code, o, out = O.step_sparse(om, p, o["x"], o["y"], o["theta"], o["vid"], nz[t])
            assert code == 0
            pr.append(out["phi_ref"])
            pp.append(out["phi_polar"])
        ref["phi_ref"].append(np.array(pr))
        ref["phi_polar"].append(np.array(pp))
        # the first step is the same computation on the same inputs
        # (phi_ref counts signs of cross products: a near-zero one may round the other way)
        assert abs(phi_polar[0] - pp[0]) <= 1e-4 and abs(phi_ref[0] - pr[0]) <= 4.0 / n + 1e-12
    report = {}
    for key in ("phi_ref", "phi_polar"):
        a, b = block_means(gpu[key]), block_means(ref[key])
        sd_a = np.concatenate([g[SKIP:] for g in gpu[key]]).std()
        sd_b = np.concatenate([r[SKIP:] for r in ref[key]]).std()
        report[key] = dict(gpu=a.mean(), oracle=b.mean(), z=zscore(a, b), sd_ratio=sd_a / sd_b)
    print(report, flush=True)
    for key, r in report.items():
        assert abs(r["z"]) <= Z_MAX, (key, r)
        assert 0.8 <= r["sd_ratio"] <= 1.25, (key, r)
