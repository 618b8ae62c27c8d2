"""Failure paths of the device event loop, forced from the host so that
every kernel still ends normally (bounded loops, capacity checks):

  status bit 1  no accepted event within max_attempts (MakeDensity.cpp:
                2147-2162 retries forever; here the loop is bounded):
                Npmin above 2A with a small max_attempts; smc_collide then
                returns SMC_ERR_NO_EVENT;
  status bit 2  collision pairs beyond max_pairs: the first max_pairs pairs
                (reference order) are kept and smeared, the event is flagged
                and smc_collide returns SMC_ERR_CAPACITY;
  status bit 8  MC-KLN thickness beyond the dN/dy table (MCnucl.cpp:665-669
                exits): a table too small for a central event;
  fed positions beyond +-56 fm are refused by smc_set_positions
  (SMC_ERR_ARG) before anything runs, so no event can carry positions the
  integer arithmetic cannot hold.
"""
import numpy as np
import pytest

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu


def test_no_accepted_event_within_max_attempts():
    from supermc_amd import Context, SmcError
    from supermc_amd._native import SMC_ERR_NO_EVENT
    p = H.glauber_params("PbPb", Npmin=450, Npmax=500, max_attempts=4, max_events=16)
    with Context(p) as ctx:
        ev = ctx.run(seed=5, event0=0, nev=16)
        assert (ev[:, 57].astype(int) & 1).all()
        with pytest.raises(SmcError) as ei:
            ctx.collide()
        assert ei.value.code == SMC_ERR_NO_EVENT
    # the same context parameters with a reachable window accept every event
    q = H.glauber_params("PbPb", Npmin=2, Npmax=500, max_attempts=4000, max_events=16)
    with Context(q) as ctx:
        assert (ctx.run(seed=5, event0=0, nev=16)[:, 57] == 0).all()


def test_pair_capacity_exceeded_keeps_the_first_pairs():
    from supermc_amd import Context, SmcError
    from supermc_amd._native import SMC_ERR_CAPACITY
    mp = 64
    p = H.glauber_params("PbPb", max_pairs=mp, max_events=2, cc_fluctuation_model=0)
    xa, xb, bs = H.cpu_configs(p, 2, seed=8, bvals=np.array([0.0, 2.0]))
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        with pytest.raises(SmcError) as ei:
            ctx.collide()
        assert ei.value.code == SMC_ERR_CAPACITY
        npa, npb, nc, _ = ctx.collide(allow=(SMC_ERR_CAPACITY,))
        assert (nc > mp).all()
        s, _, _ = ctx.density()
        rec = ctx.moments()
        assert (rec[:, 57].astype(int) & 2).all()
        for e in range(2):
            nuc = ctx.nucleons(e)
            pa, pb, pf = ctx.pairs(e)
            pa, pb, pf = pa[:mp], pb[:mp], pf[:mp]
            col = oracle.collide(nuc["xy_proj"], nuc["xy_targ"], ctx.derived.width, ctx.derived.dsq)
            assert col["ncoll"] == nc[e]
            assert (pa == col["pair_a"][:mp]).all() and (pb == col["pair_b"][:mp]).all()
            # the stored pairs, and only they, are smeared
            ia = np.argsort(np.where(nuc["wound_proj"] >= 0, nuc["wound_proj"], 1 << 30))[: npa[e]]
            ib = np.argsort(np.where(nuc["wound_targ"] >= 0, nuc["wound_targ"], 1 << 30))[: npb[e]]
            g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
            mid = (nuc["xy_proj"][pa] + nuc["xy_targ"][pb]) / 2.0
            ref, _ = oracle.density_glauber(g, ctx.derived.width, p.alpha, nuc["xy_proj"][ia], np.ones(len(ia)),
                                            nuc["xy_targ"][ib], np.ones(len(ib)), mid, np.ones(mp))
            assert H.rel_linf(s[e], ref) <= 1e-5


def test_kln_table_overflow_is_flagged():
    from supermc_amd import Context
    # tmax 3, one subdivision: the table holds T < dT = 10 / sigma_inel, about
    # 1.6 fm^-2, below a central Pb+Pb thickness
    p = H.glauber_params("PbPb", which_mc_model=1, sub_model=7, ecm=2760.0, lambda_=0.138, tmax=3,
                         tmax_subdivision=1, kln_n_pt=8, kln_n_kt=8, kln_n_phi=8, cc_fluctuation_model=0,
                         max_events=2)
    xa, xb, bs = H.cpu_configs(p, 2, seed=3, bvals=np.array([0.0, 1.0]))
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide()
        ctx.density()
        rec = ctx.moments()
        assert (rec[:, 57].astype(int) & 8).all()


def test_positions_beyond_range_are_refused():
    from supermc_amd import Context, SmcError
    from supermc_amd._native import SMC_ERR_ARG
    p = H.glauber_params("AuAu", Aproj=1, Atarg=1, max_events=1)
    with Context(p) as ctx:
        with pytest.raises(SmcError) as ei:
            ctx.set_positions(np.array([[[57.0, 0.0]]]), np.array([[[0.0, 0.0]]]), np.array([0.0]))
        assert ei.value.code == SMC_ERR_ARG
        # the context still works afterwards
        ctx.set_positions(np.array([[[0.2, 0.0]]]), np.array([[[-0.2, 0.0]]]), np.array([0.4]))
        assert ctx.collide()[2][0] == 1
