"""PyTorch harness for the differentiable MC-Glauber batch (no counterpart
in the reference, which has no derivative code).

`BatchEccentricity` is a torch.autograd.Function over theta = (w, R, a,
alpha, k) whose forward runs the frozen-topology batch on the device
(smc_diff_forward) and whose backward runs the hand-written adjoint
(smc_diff_backward: cell cotangents, per-source gather, chain rule).  It
returns the batch means (<|eps_2|>, <|eps_3|>, <dS/dy>).

    with Context(params) as ctx:
        theta0 = torch.tensor(ctx.diff_setup(seed=1, event0=0, nev=4096), requires_grad=True)
        obs = BatchEccentricity.apply(theta0, ctx)
        obs[0].backward()            # theta0.grad = d<|eps_2|>/d(w, R, a, alpha, k)

`calibrate` fits one parameter to a target observable by Newton steps on
the adjoint gradient.
"""
import numpy as np
import torch

PARAM_NAMES = ("w", "R", "a", "alpha", "k")


class BatchEccentricity(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, smc):
        th = theta.detach().to("cpu", torch.float64).numpy().copy()
        out = smc.diff_forward(th)
        fctx.smc = smc
        fctx.theta = th
        fctx.device = theta.device
        means = out[:, :3].mean(0)
        return torch.tensor(means, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(fctx, grad):
        smc = fctx.smc
        if not np.array_equal(getattr(smc, "_diff_theta", None), fctx.theta):
            smc.diff_forward(fctx.theta)  # the adjoint needs this theta's forward state
        g = grad.detach().to("cpu", torch.float64).numpy()
        cot = np.tile(g / smc.diff_nev, (smc.diff_nev, 1))
        dtheta = smc.diff_backward(cot)
        return torch.tensor(dtheta, dtype=torch.float64, device=fctx.device), None


def calibrate(smc, theta0, param, target, observable=0, steps=8, tol=1e-7):
    """Newton iteration on one parameter (index or name) so that the batch
    observable (0 <|eps_2|>, 1 <|eps_3|>, 2 <dS/dy>) reaches target."""
    q = PARAM_NAMES.index(param) if isinstance(param, str) else int(param)
    th = torch.tensor(np.asarray(theta0, np.float64), requires_grad=True)
    hist = []
    for _ in range(steps):
        th.grad = None
        obs = BatchEccentricity.apply(th, smc)
        val = obs[observable]
        val.backward()
        resid = float(val) - target
        hist.append((float(th[q]), float(val)))
        if abs(resid) < tol:
            break
        slope = float(th.grad[q])
        with torch.no_grad():
            th[q] -= resid / slope
    return th.detach().numpy(), hist
