"""Differentiable Pb+Pb batch (frozen topology) on the HIP path against the
FP64 restatement (oracle/diff_oracle.py, pinned in test_diff_oracle_pins).

Bars (written here, checked below):
  forward |eps_2|, |eps_3|: absolute <= 1e-6; dS/dy: relative <= 1e-5
  adjoint gradient d(<|eps_2|> + <|eps_3|> + 0.01 <dS/dy>)/d(w, R, a, alpha, k):
    relative <= 1e-4 per component against the FP64 adjoint and against
    central finite differences (h = 1e-4) of the FP64 forward
  torch.autograd.Function: the same gradient through backward()
"""
import numpy as np
import pytest
import torch

import diff_oracle as D
import parity_helpers as H

pytestmark = pytest.mark.gpu

NEV = 4
COT = np.array([1.0, 1.0, 0.01])


@pytest.fixture(scope="module")
def batch():
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=8)
    ctx = Context(p)
    th0 = ctx.diff_setup(seed=77, event0=0, nev=NEV)
    evs = [D.event_from_context(ctx, e) for e in range(NEV)]
    yield ctx, th0, evs
    ctx.close()


def test_theta0_is_the_context_parameters(batch):
    ctx, th0, _ = batch
    np.testing.assert_allclose(th0, [0.5, 6.62, 0.546, 0.14, 0.8], rtol=1e-12)


@pytest.mark.parametrize("theta", [None, (0.6, 6.5, 0.5, 0.2, 1.1)])
def test_forward_matches_fp64(batch, theta):
    ctx, th0, evs = batch
    th = th0 if theta is None else np.array(theta)
    out = ctx.diff_forward(th)
    assert (out[:, 7] == 0).all()
    for e, ev in enumerate(evs):
        ref = ev.forward(th)
        assert abs(out[e, 0] - ref[0]) <= 1e-6, (e, out[e, 0], ref[0])
        assert abs(out[e, 1] - ref[1]) <= 1e-6, (e, out[e, 1], ref[1])
        assert out[e, 2] == pytest.approx(ref[2], rel=1e-5)


def test_positions_at_theta0_reproduce_the_topology(batch):
    # the frozen uniforms mapped back at theta0 give the positions the
    # collision step saw (to the 2^-24 fm quantum)
    ctx, th0, evs = batch
    for e, ev in enumerate(evs):
        nuc = ctx.nucleons(e)
        pos = ev.positions(th0[1], th0[2])
        assert np.abs(pos[0][0] - nuc["xy_proj"]).max() < 1e-7
        assert np.abs(pos[1][0] - nuc["xy_targ"]).max() < 1e-7


@pytest.mark.parametrize("theta", [None, (0.6, 6.5, 0.5, 0.2, 1.1)])
def test_adjoint_matches_fp64_and_finite_differences(batch, theta):
    ctx, th0, evs = batch
    th = th0 if theta is None else np.array(theta)
    ctx.diff_forward(th)
    g = ctx.diff_backward(np.tile(COT, (NEV, 1)))
    ref = sum(ev.adjoint(th, COT) for ev in evs)
    fd = sum(ev.fd_gradient(th, COT, h=1e-4) for ev in evs)
    scale = np.abs(ref).max()
    for q in range(5):
        assert abs(g[q] - ref[q]) <= 1e-4 * max(abs(ref[q]), 1e-3 * scale), (q, g[q], ref[q])
        assert abs(g[q] - fd[q]) <= 1e-4 * max(abs(fd[q]), 1e-3 * scale), (q, g[q], fd[q])


def test_autograd_function(batch):
    from supermc_amd.adjoint import BatchEccentricity
    ctx, th0, evs = batch
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    obs = BatchEccentricity.apply(th, ctx)
    (obs * torch.tensor(COT)).sum().backward()
    ref = sum(ev.adjoint(th0, COT) for ev in evs) / NEV
    np.testing.assert_allclose(th.grad.numpy(), ref, rtol=1e-4, atol=1e-7 * np.abs(ref).max())
    out = ctx.diff_forward(th0)
    np.testing.assert_allclose(obs.detach().numpy(), out[:, :3].mean(0), rtol=1e-15)
