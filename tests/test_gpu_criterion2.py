"""Gaussian collision criterion (collision_criterion 2, MCnucl.cpp:357-385,
GaussianNucleonsCal.cpp:59-67, 130-163) on the device.

The reference draws drand48 per tested pair; the device draws Philox, so
parity is in law (the draws cannot match):
- sigma_gg, the tested box and the criterion in effect equal the oracle's
  (sig_eff to 1e-10 relative; 0 selects the Gaussian test through
  shape_of_entropy 2 and the disk through 4..9; 3 and 4 are rejected);
- on fed Pb+Pb positions, every collision is a pair the reference's sweep
  tests (oracle.gauss_tested_pairs), and the hit counts per impact-parameter
  bin over 64 replicas of each configuration agree with sum P(b) within 5
  standard deviations (sum P (1 - P));
- sampled minimum-bias events with the reference parameter file's defaults
  (which_mc_model 7, shape_of_nucleons 2, criterion 2, 301^2) all pass the
  acceptance with Ncoll > 0.
"""
import numpy as np
import pytest

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu

KQ = 2.0 ** -24


def _params(**kw):
    kw.setdefault("collision_criterion", 2)
    return H.glauber_params("PbPb", **kw)


def test_derived_sigma_gg_box_and_criterion():
    from supermc_amd import Context, SmcError, default_params
    for kw in (dict(), dict(shape_of_nucleons=2, sigma_nn_mb=68.0), dict(gauss_nucl_width=0.812)):
        p = _params(max_events=4, **kw)
        with Context(p) as ctx:
            d = ctx.derived
        assert d.criterion == 2
        assert d.sigma_gg == pytest.approx(oracle.sig_eff(d.siginNN, d.width), rel=1e-10)
        assert d.box_lim == int(np.floor(8 * d.width / KQ))
    with Context(_params(max_events=4, collision_criterion=0)) as ctx:
        assert ctx.derived.criterion == 2  # shape_of_entropy 2 (MCnucl.cpp:376-383)
    with Context(H.glauber_params("PbPb", max_events=4)) as ctx:
        assert ctx.derived.criterion == 1
    for bad in (3, 4):
        with pytest.raises(SmcError):
            Context(_params(max_events=4, collision_criterion=bad))
    with Context(default_params(max_events=8)) as ctx:
        assert ctx.derived.criterion == 2
    # shape_of_entropy 4..9: the Gaussian entropy of 2, and the disk test
    # when no criterion is forced (MCnucl.cpp:376-383 falls to disk)
    with Context(_params(max_events=4, collision_criterion=0, shape_of_entropy=5)) as ctx:
        assert ctx.derived.criterion == 1
    for bad in (1, 3, 10):
        with pytest.raises(SmcError):
            Context(_params(max_events=4, shape_of_entropy=bad))


def test_gaussian_entropy_shapes_give_the_same_events():
    """shape_of_entropy 2 and 4..9 share the density and the fluctuation law
    (MCnucl.cpp:752-756, 853-859, 1278-1281): with the disk criterion forced,
    or chosen by the fallback for 4..9, the events are the same bit for bit."""
    from supermc_amd import Context
    recs = []
    for kw in (dict(shape_of_entropy=2, collision_criterion=1), dict(shape_of_entropy=6, collision_criterion=1),
               dict(shape_of_entropy=9, collision_criterion=0)):
        with Context(H.glauber_params("PbPb", max_events=64, **kw)) as ctx:
            recs.append(ctx.run(seed=8, event0=0, nev=64))
    assert np.array_equal(recs[0], recs[1], equal_nan=True)
    assert np.array_equal(recs[0], recs[2], equal_nan=True)


@pytest.mark.parametrize("lam", [4.14, 2.0])
def test_shape3_width_and_sigma_gg(lam):
    """shape_of_nucleons 3: width sqrt(0.1 sig / (4 pi lam ratio)) and
    sigma_gg = 0.1 sig / ratio, ratio = Ein(lam)/lam (GaussianNucleonsCal.cpp:41-46)."""
    from scipy import special
    from supermc_amd import Context
    p = _params(max_events=4, shape_of_nucleons=3, gaussian_lambda=lam, sigma_nn_mb=68.0)
    with Context(p) as ctx:
        d = ctx.derived
    ratio = (np.euler_gamma + np.log(lam) + special.exp1(lam)) / lam
    assert d.width == pytest.approx(np.sqrt(6.8 / (4 * np.pi * lam * ratio)), rel=1e-12)
    assert d.sigma_gg == pytest.approx(6.8 / ratio, rel=1e-12)


def test_collisions_follow_the_gaussian_law_on_fed_positions():
    from supermc_amd import Context
    reps = 64
    bvals = [2.0, 7.0, 12.0]
    p = _params(max_events=reps * len(bvals), cc_fluctuation_model=0)
    xa0, xb0, bs0 = H.cpu_configs(p, len(bvals), seed=321, bvals=np.array(bvals))
    xa = np.repeat(xa0, reps, axis=0)
    xb = np.repeat(xb0, reps, axis=0)
    bs = np.repeat(bs0, reps)
    edges = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 1e9])  # b / w
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=17)
        ctx.collide(allow=(4,))
        w, S = ctx.derived.width, ctx.derived.sigma_gg
        for c in range(len(bvals)):
            obs = np.zeros(len(edges) - 1)
            tested_ref = None
            for r in range(reps):
                e = c * reps + r
                nuc = ctx.nucleons(e)
                pa, pb, _ = ctx.pairs(e)
                if tested_ref is None:  # the same sorted nucleons in every replica
                    tested = oracle.gauss_tested_pairs(nuc["xy_proj"], nuc["xy_targ"], w)
                    tested_ref = (nuc["xy_proj"].copy(), nuc["xy_targ"].copy(), set(tested))
                    ti = np.array([t[0] for t in tested])
                    tj = np.array([t[1] for t in tested])
                    bt = np.hypot(*(nuc["xy_proj"][ti] - nuc["xy_targ"][tj]).T)
                    P = oracle.gauss_hit_prob(bt, w, S)
                    which = np.digitize(bt / w, edges) - 1
                    exp = np.bincount(which, weights=P, minlength=len(edges) - 1) * reps
                    var = np.bincount(which, weights=P * (1 - P), minlength=len(edges) - 1) * reps
                assert np.array_equal(nuc["xy_proj"], tested_ref[0]) and np.array_equal(nuc["xy_targ"], tested_ref[1])
                hits = set(zip(pa.tolist(), pb.tolist()))
                assert hits <= tested_ref[2], (c, r, sorted(hits - tested_ref[2])[:4])
                if len(pa):
                    bh = np.hypot(*(nuc["xy_proj"][pa] - nuc["xy_targ"][pb]).T)
                    obs += np.bincount(np.digitize(bh / w, edges) - 1, minlength=len(edges) - 1)
            for k in range(len(obs)):
                assert abs(obs[k] - exp[k]) <= 5 * np.sqrt(var[k]) + 1e-9, (bvals[c], k, obs[k], exp[k], var[k])


def test_sampled_minbias_with_reference_defaults():
    from supermc_amd import Context, default_params
    nev = 2000
    p = default_params(max_events=nev)
    with Context(p) as ctx:
        rec = ctx.run(seed=11, event0=0, nev=nev)
    assert (rec[:, 57] == 0).all()
    npart, ncoll = rec[:, 54], rec[:, 55]
    assert (ncoll > 0).all() and (npart >= 2).all() and (npart <= 416).all()
    assert np.isfinite(rec[:, :50]).all()
    # minimum bias over b <= 20 fm: central and peripheral events both present
    assert npart.max() > 350 and np.median(npart) < 150


def _table_npart(npart, cent):
    """The Npart column of the reference's centrality tables for sampled
    events (scripts/centrality_cut_h5.py:43-89): events sorted by Npart,
    descending; the row of upper edge c holds the smallest Npart among the
    int(N (c - c_prev) / 100) - 1 events from rank int(N c_prev / 100)."""
    s = np.sort(npart)[::-1]
    n = len(s)
    out = []
    for lo, hi in zip(cent[:-1], cent[1:]):
        off, cnt = int(n * lo / 100), int(n * (hi - lo) / 100) - 1
        out.append(s[off:off + cnt].min())
    return np.array(out)


@pytest.mark.parametrize("ecm", [5020, 2760])
def test_sampled_npart_centrality_matches_reference_tables(ecm):
    """Npart at the centrality edges of minimum-bias Pb+Pb with the
    reference parameter file (parameters.dat, the configuration the
    reference's tables were made with: operation 9, bmin 0, bmax 20, Gaussian
    collision criterion, d_min 0.9 fm, Regge sigma_NN) against
    scripts/centrality_cut_tables/iebe_centralityCut_Npart_MCGlbPbPb<ecm>_
    sigmaNN_gauss_d0.9_withMultFluct.dat (tests/golden/centrality_npart.json).

    Tolerance per edge c: the order statistic at rank r = N c / 100 has a
    binomial spread sqrt(N q (1 - q)) in rank (q = c / 100), i.e. sigma =
    half the Npart difference between ranks r -+ that spread in the sorted
    sample; the table, from its own finite sample, is taken to be no more
    precise, so the bar is 5 sqrt(2) sigma plus 1 for the integer Npart."""
    import json
    import os
    from supermc_amd import Context, default_params
    tab = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "centrality_npart.json")))[str(ecm)]
    rows = np.array(tab["rows"])
    cent, ref = rows[:, 0], rows[1:, 1]
    nb, batches = 25000, 4
    p = default_params(ecm=float(ecm), max_events=nb, max_pairs=4096)
    npart = []
    with Context(p) as ctx:
        for k in range(batches):
            ctx.sample(seed=20261203 + ecm, event0=k * nb, nev=nb)
            a, b, nc, _ = ctx.collide()
            assert (nc > 0).all()
            npart.append(a + b)
    npart = np.concatenate(npart).astype(float)
    n = len(npart)
    got = _table_npart(npart, cent)
    s = np.sort(npart)[::-1]
    checked = 0
    for i, c in enumerate(cent[1:]):
        if c < 1.0 or c > 95.0:
            continue  # the 0.1 % rows hold ~100 events; the last rows are Npart = 2
        q = c / 100
        r, k = int(n * q), int(np.ceil(np.sqrt(n * q * (1 - q))))
        sigma = 0.5 * (s[max(r - k, 0)] - s[min(r + k, n - 1)])
        tol = 5 * np.sqrt(2) * sigma + 1.0
        assert abs(got[i] - ref[i]) <= tol, (ecm, c, got[i], ref[i], tol)
        checked += 1
    assert checked >= 90
