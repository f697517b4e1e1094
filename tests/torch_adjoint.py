"""A thin torch.autograd.Function over the C-ABI adjoint (test wrapper).

TrajectoryLoss.apply(theta0, tau, mu, v0, ctx) runs the forward and reverse of ctx["steps"]
steps through tissue_adjoint and returns the loss; backward hands back the adjoint's gradients
with respect to the initial headings and the three parameters."""
import numpy as np
import torch


class TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta0, tau, mu, v0, cfg):
        T = cfg["lib"]
        p = T.params(**dict(cfg["params"], tau=float(tau), mu=float(mu), v0=float(v0)))
        sim = T.Sim(cfg["mesh"], p, len(cfg["x0"]))
        try:
            sim.set_mode(T.MODE_STEP_KERNEL)
            sim.set_state(cfg["x0"], cfg["y0"], theta0.detach().cpu().numpy().astype(np.float32))
            sim.stage_noise(cfg["noise"])
            code, g, dth, _, _ = sim.adjoint(cfg["steps"], cfg["loss"], 0, cfg.get("ckpt", 20))
        finally:
            sim.close()
        if code != T.OK:
            raise RuntimeError(f"tissue_adjoint returned status {code}")
        fctx.save_for_backward(torch.from_numpy(dth), torch.tensor([g["d_tau"], g["d_mu"], g["d_v0"]],
                                                                   dtype=torch.float64))
        return torch.tensor(g["loss"], dtype=torch.float64)

    @staticmethod
    def backward(fctx, grad_out):
        dth, dp = fctx.saved_tensors
        return grad_out * dth, grad_out * dp[0], grad_out * dp[1], grad_out * dp[2], None
