"""FP64 density grids (smc_density_f64) against the oracle's double
grids: MCnucl::setDensity for which_mc_model 5 (MCnucl.cpp:689-760) and
7 (:762-811), each (source, cell) term in FP64 with the reference's
expressions and summation order, so the only difference left is the
device exp against the host's (an ulp per term).

Bars:
  FP64 grid vs oracle       1e-12 relative L-inf (measured near 1e-16)
  FP32 grid vs FP64 grid    1e-5 relative L-inf (the FP32 path's bar)
  any event range / chunk   bit-identical to the same event alone
"""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _check_batch(ctx, nev, model=5):
    s32, _, _ = ctx.density()
    s64 = ctx.density_f64()
    assert s64.shape == (nev, ctx.nx, ctx.ny)
    worst = 0.0
    for e in range(nev):
        o = H.oracle_event(ctx, e)
        err = H.rel_linf(s64[e], o["s"])
        worst = max(worst, err)
        assert err <= TOL64, (e, err)
        assert H.rel_linf(s32[e], s64[e]) <= 1e-5
    print(f"which_mc_model {model}: max rel L-inf FP64 grid vs oracle {worst:.3g} over {nev} events", flush=True)


def test_fp64_grids_auau_minbias_match_oracle():
    """Au+Au, the plain MC-Glauber set (R 6.38, a 0.535, d 0.4, 42 mb,
    alpha 0.14, no fluctuations), sampled minimum-bias events."""
    p = H.glauber_params("AuAu", max_events=24)
    from supermc_amd import Context
    with Context(p) as ctx:
        ctx.run(seed=2024, event0=0, nev=24, fetch=False)
        _check_batch(ctx, 24)


@pytest.mark.parametrize("grid", [{}, dict(maxy=9.0, dy=0.08)], ids=["261", "261x226"])
@pytest.mark.parametrize("model", [5, 7])
def test_fp64_grids_pbpb_fluct_fed_positions(model, grid):
    """Pb+Pb with Gamma factors on fed positions at b = 0 .. 14.5 fm, the
    two-component model and sqrt(T_A T_B), on 0.1 fm cells and on
    rectangular 0.1 x 0.08 fm cells."""
    p = H.glauber_params("PbPb", max_events=8, which_mc_model=model, **grid)
    bvals = np.array([0.0, 2.0, 4.5, 7.0, 9.0, 11.0, 13.0, 14.5])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=31 + model, bvals=bvals)
    from supermc_amd import Context
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=7)
        ctx.collide(allow=(4,))
        _check_batch(ctx, len(bvals), model)


def test_fp64_grid_ranges_and_chunks_are_event_local():
    """Ranges across a launch chunk (2048 events on small grids, about 1 GiB
    of grids on large ones) give each event the grid it has alone;
    out-of-batch ranges and other models are refused."""
    from supermc_amd import Context, SmcError
    from supermc_amd._native import SMC_ERR_ARG, SMC_ERR_UNSUPPORTED
    n = 2100
    p = H.glauber_params("AuAu", max_events=n, maxx=5.0, maxy=5.0, dx=0.2, dy=0.2)
    with Context(p) as ctx:
        ctx.run(seed=11, event0=0, nev=n, fetch=False)
        full = ctx.density_f64()
        for e in (0, 1, 2046, 2047, 2048, 2049, n - 1):
            one = ctx.density_f64(e, 1)
            assert np.array_equal(one[0], full[e]), e
        part = ctx.density_f64(2040, 20)
        assert np.array_equal(part, full[2040:2060])
        for ev0, m in ((n - 1, 2), (-1, 1), (0, n + 1)):
            with pytest.raises(SmcError) as ei:
                ctx.density_f64(ev0, m, out=np.zeros((max(m, 1), ctx.nx, ctx.ny)))
            assert ei.value.code == SMC_ERR_ARG
    # 521^2 grids: about 1 GiB of FP64 grids per launch, 494 events
    n = 500
    p = H.glauber_params("AuAu", max_events=n, dx=0.05, dy=0.05)
    with Context(p) as ctx:
        ctx.run(seed=12, event0=0, nev=n, fetch=False)
        full = ctx.density_f64()
        for e in (0, 493, 494, n - 1):
            assert np.array_equal(ctx.density_f64(e, 1)[0], full[e]), e
        o = H.oracle_event(ctx, 494)
        assert H.rel_linf(full[494], o["s"]) <= TOL64
    q = H.glauber_params("AuAu", max_events=2, which_mc_model=1, sub_model=7, ecm=2760.0, lambda_=0.138, tmax=71,
                         tmax_subdivision=3, kln_n_pt=8, kln_n_kt=8, kln_n_phi=8)
    xa, xb, bs = H.cpu_configs(q, 2, seed=3, bvals=np.array([0.0, 5.0]))
    with Context(q) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide()
        with pytest.raises(SmcError) as ei:
            ctx.density_f64()
        assert ei.value.code == SMC_ERR_UNSUPPORTED
