"""smc_run_graph -- smc_run replayed as a hipGraph -- against eager smc_run:
bit-exact event records and grids on every replay, with new seeds and
event0 per replay, eager work and synchronizations of the same context
between replays, and a batch-size change (eager call, then a new capture)."""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu


def _ctx(p):
    from supermc_amd import Context
    return Context(p)


@pytest.mark.parametrize("model", [5, 1, "kln_direct"])
def test_graph_replays_equal_eager_runs(model):
    # "kln_direct": MC-KLN by the per-cell quadrature, whose captured
    # sequence resets the active-cell counter with a kernel
    extra = dict(which_mc_model=1, sub_model=7, ecm=2760.0, lambda_=0.138, kln_n_pt=16, kln_n_kt=16,
                 kln_n_phi=16, cc_fluctuation_model=0) if model != 5 else {}
    if model == "kln_direct":
        extra["kln_direct"] = 1
    p = H.glauber_params("PbPb", max_events=64, **extra)
    with _ctx(p) as g, _ctx(p) as e:
        calls = [(11, 0), (11, 64), (12, 5), (13, 1000), (11, 0)]
        for k, (seed, ev0) in enumerate(calls):
            rg = g.run_graph(seed, ev0, 64)
            re_ = e.run(seed, ev0, 64)
            assert np.array_equal(rg, re_), (k, seed, ev0)
            sg, dg, _ = g.density()
            se, de, _ = e.density()
            assert np.array_equal(sg, se) and np.array_equal(dg, de)
            if k == 2:  # eager work on the graph's context between replays
                g.run(99, 7, 16)
                g.synchronize()
        # another batch size: an eager call, a new capture, replays
        for seed in (21, 22, 23):
            assert np.array_equal(g.run_graph(seed, 0, 32), e.run(seed, 0, 32))
