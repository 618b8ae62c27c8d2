"""smc_run_noalloc through the C-ABI: records equal smc_run's for the same
events (bit for bit: the same kernels, the moment scratch in the library's
static device array instead of context buffers), on two calls in a row,
the second with fewer events; nev above the scratch's limit and a context
without its batch buffers are refused."""
import ctypes as C

import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu


def _noalloc(ctx, seed, event0, nev, dev):
    return ctx.lib.smc_run_noalloc(ctx.h, seed, event0, nev, C.c_void_p(dev.data_ptr()))


def test_run_noalloc_matches_run():
    import torch
    from supermc_amd import Context
    from supermc_amd._native import SMC_ERR_CAPACITY, SMC_ERR_STATE, SMC_OK
    p = H.glauber_params("PbPb", max_events=3000)
    with Context(p) as ctx:
        dev = torch.full((3000, 64), float("nan"), dtype=torch.float64, device="cuda")
        assert _noalloc(ctx, 5, 0, 64, dev) == SMC_ERR_STATE  # no batch buffers yet
        ctx.run(seed=5, event0=0, nev=16)  # creates them
        for seed, event0, nev in ((77, 1000, 1500), (78, 40, 333)):
            assert _noalloc(ctx, seed, event0, nev, dev) == SMC_OK, ctx.lib.smc_last_error(ctx.h)
            ctx.synchronize()
            got = dev[:nev].cpu().numpy()
            want = ctx.run(seed=seed, event0=event0, nev=nev)
            assert np.array_equal(got[:, :58], want[:, :58]), (seed, event0, nev)
            assert (got[:, 57] == 0).all()
        assert _noalloc(ctx, 1, 0, 2049, dev) == SMC_ERR_CAPACITY


def test_run_noalloc_refuses_per_cell_kln():
    """The per-cell KLN quadrature keeps an active-cell list in context
    memory, so the allocation-free entry point refuses such a context."""
    import torch
    from supermc_amd import Context
    from supermc_amd._native import SMC_ERR_UNSUPPORTED
    p = H.glauber_params("PbPb", max_events=8, which_mc_model=1, sub_model=7, ecm=2760.0, lambda_=0.138,
                         kln_n_pt=8, kln_n_kt=8, kln_n_phi=8, kln_direct=1)
    with Context(p) as ctx:
        ctx.run(seed=5, event0=0, nev=8)
        dev = torch.zeros((8, 64), dtype=torch.float64, device="cuda")
        assert _noalloc(ctx, 5, 0, 8, dev) == SMC_ERR_UNSUPPORTED
