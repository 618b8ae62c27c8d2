"""3D (x, y, eta) profiles of scripts/generate_3d_profiles/profile_3d.cpp on
the device (include/smc_p3d.h) against the FP64 oracle.

Bars (written here, checked below):
  eta table and derived constants: equal to the oracle's (relative 1e-15)
  smearing on the same source records: relative L-inf <= 1e-5 against the
    oracle (the device multiplies three FP32 exponentials, the reference
    takes one FP64 exponential of the sum)
  sampling (Philox, not GSL's stream): the laws -- eta from the tabulated
    distribution (KS p > 1e-3), widths uniform on sigma0 +- 0.3 (KS
    p > 1e-3), random_flag 0 values exact; the sampled records are then fed
    to the oracle for the grids
  generate = sample + smear (bit-exact), batch-independent draws
"""
import numpy as np
import pytest
from scipy import stats

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu

GRID_TOL = 1e-5
REF_GRID = dict(grid_nx=261, grid_ny=261, grid_neta=101, grid_dx=0.1, grid_dy=0.1, grid_deta=0.1)


def _p3d(**kw):
    from supermc_amd import Profile3D
    args = dict(REF_GRID, ecm=19.6, random_flag=1)
    args.update(kw)
    return Profile3D(**args)


def _oracle_rho(P, src):
    p = P.params
    return oracle.p3d_smear(p.grid_nx, p.grid_ny, p.grid_neta, p.grid_dx, p.grid_dy, p.grid_deta, src)


def _eta_cdf(t, table):
    """CDF of the density proportional to the linear interpolation of
    `table` on the eta nodes (the sampler's acceptance law)."""
    eta = t["eta"]
    h = np.diff(eta)
    seg = h * (table[:-1] + table[1:]) / 2
    cum = np.concatenate([[0.0], np.cumsum(seg)])

    def cdf(x):
        x = np.asarray(x, np.float64)
        k = np.clip(np.searchsorted(eta, x, side="right") - 1, 0, len(eta) - 2)
        f = (x - eta[k]) / h[k]
        return (cum[k] + h[k] * (table[k] * f + (table[k + 1] - table[k]) * f * f / 2)) / cum[-1]
    return cdf


def test_derived_constants_and_table_match_oracle():
    for ecm in (19.6, 200.0):
        with _p3d(ecm=ecm) as P:
            d, eta, proj, targ = P.derived()
        t = oracle.p3d_setup(ecm)
        assert d.y_beam == pytest.approx(t["y_beam"], rel=1e-15)
        assert d.sigma_inelastic == pytest.approx(t["sigma_inel"], rel=1e-15)
        assert d.sigma_x0 == pytest.approx(np.sqrt(t["sigma_inel"] / (8 * np.pi)), rel=1e-15)
        assert (d.eta0, d.sigma_eta0) == (2.0, 0.5)
        np.testing.assert_allclose(eta, t["eta"], rtol=1e-15, atol=0)
        np.testing.assert_allclose(proj, t["proj"], rtol=1e-15, atol=0)
        np.testing.assert_allclose(targ, t["targ"], rtol=1e-15, atol=0)
        assert d.proj_max == pytest.approx(t["proj_max"], rel=1e-15)


def _random_sources(rng, n, edge=False):
    src = np.column_stack([rng.uniform(-10, 10, n), rng.uniform(-10, 10, n), rng.uniform(-3, 3, n),
                           rng.uniform(0.06, 0.66, n), rng.uniform(0.06, 0.66, n), rng.uniform(0.2, 0.8, n)])
    if edge:  # windows clamped at every grid face, and one source off the grid
        src[:6, :3] = [[12.95, 0.3, 0.1], [-13.0, -12.9, 0.0], [0.0, 12.99, 4.95], [1.0, -1.0, -5.0],
                       [0.0, 0.0, 0.0], [20.0, 20.0, 9.0]]
    return src


def test_smear_matches_oracle_on_reference_grid():
    rng = np.random.default_rng(11)
    src = _random_sources(rng, 40, edge=True)
    with _p3d() as P:
        rho = P.smear([src])[0]
        ref = _oracle_rho(P, src)
    assert H.rel_linf(rho, ref) <= GRID_TOL, H.rel_linf(rho, ref)
    assert ((rho != 0) == (ref != 0)).all()  # the same cell windows


def test_smear_batch_ragged_and_empty_events():
    rng = np.random.default_rng(12)
    srcs = [_random_sources(rng, 9), np.zeros((0, 6)), _random_sources(rng, 70, edge=True), _random_sources(rng, 1)]
    # also a grid of unequal sizes and spacings in x and y
    with _p3d(grid_nx=81, grid_ny=91, grid_neta=31, grid_dy=0.12, grid_deta=0.2) as P:
        rho = P.smear(srcs)
        for e, s in enumerate(srcs):
            ref = _oracle_rho(P, s)
            if len(s) == 0:
                assert not rho[e].any()
                continue
            assert H.rel_linf(rho[e], ref) <= GRID_TOL, (e, H.rel_linf(rho[e], ref))


def _parts(n, ids, rng):
    return np.column_stack([rng.uniform(-6, 6, n), rng.uniform(-6, 6, n), ids])


def test_eta_law_and_widths_per_random_flag():
    rng = np.random.default_rng(13)
    n = 20000
    ids = np.where(np.arange(n) % 2 == 0, 1, 2)
    parts = [_parts(n, ids, rng)]
    bins = [np.zeros((0, 2))]
    t = oracle.p3d_setup(19.6)
    cdf_p, cdf_t = _eta_cdf(t, t["proj"]), _eta_cdf(t, t["targ"])
    sx0 = np.sqrt(t["sigma_inel"] / (8 * np.pi))
    for flag in (1, 2, 3):
        with _p3d(random_flag=flag) as P:
            ps, _ = P.sample(parts, bins, seed=99)
        s = ps[0]
        np.testing.assert_array_equal(s[:, :2], parts[0][:, :2])
        assert stats.kstest(s[ids == 1, 2], cdf_p).pvalue > 1e-3
        assert stats.kstest(s[ids == 2, 2], cdf_t).pvalue > 1e-3
        # the projectile weight (1 - eta/y_beam) of profile_3d.cpp:146 leans to
        # negative eta (random_flag 0 nevertheless puts projectiles at +2)
        assert s[ids == 1, 2].mean() < -0.3 and s[ids == 2, 2].mean() > 0.3
        if flag == 1:
            assert (s[:, 3] == sx0).all() and (s[:, 4] == sx0).all() and (s[:, 5] == 0.5).all()
        else:
            u = stats.uniform(loc=0.2, scale=0.6)
            assert stats.kstest(s[:, 5], u.cdf).pvalue > 1e-3
            ux = stats.uniform(loc=sx0 - 0.3, scale=0.6)
            assert stats.kstest(s[:, 3], ux.cdf).pvalue > 1e-3
            if flag == 2:
                assert (s[:, 3] == s[:, 4]).all()
            else:
                assert stats.kstest(s[:, 4], ux.cdf).pvalue > 1e-3
                assert abs(np.corrcoef(s[:, 3], s[:, 4])[0, 1]) < 0.05


def test_random_flag_0_and_binary_ids():
    rng = np.random.default_rng(14)
    parts = [_parts(5, [1, 2, 2, 1, 2], rng)]
    bins = [rng.uniform(-3, 3, (12, 2))]
    with _p3d(random_flag=0) as P:
        ps, bs = P.sample(parts, bins, seed=1)
        d = P.derived()[0]
    assert (ps[0][:, 2] == [2, -2, -2, 2, -2]).all()
    assert (bs[0][:, 2] == 2.0).all()  # the reference gives every binary collision +eta_0
    assert (ps[0][:, 3:] == [d.sigma_x0, d.sigma_x0, 0.5]).all()
    # sampled: binary i follows the law of participant i's id (wrapping)
    n = 20000
    parts = [_parts(2, [1, 2], rng)]
    bins = [rng.uniform(-3, 3, (n, 2))]
    t = oracle.p3d_setup(19.6)
    with _p3d(random_flag=1) as P:
        _, bs = P.sample(parts, bins, seed=2)
    eta = bs[0][:, 2]
    assert stats.kstest(eta[0::2], _eta_cdf(t, t["proj"])).pvalue > 1e-3
    assert stats.kstest(eta[1::2], _eta_cdf(t, t["targ"])).pvalue > 1e-3


def test_generate_is_sample_then_smear_and_batch_independent():
    rng = np.random.default_rng(15)
    parts = [_parts(k, rng.integers(1, 3, k), rng) for k in (30, 0, 55)]
    bins = [rng.uniform(-4, 4, (k, 2)) for k in (45, 3, 0)]
    with _p3d(random_flag=3, grid_nx=101, grid_ny=101, grid_neta=41) as P:
        rp, rb, ps, bs = P.generate(parts, bins, seed=7, event0=40)
        ps_s, bs_s = P.sample(parts, bins, seed=7, event0=40)
        np.testing.assert_array_equal(np.concatenate(ps_s), ps)
        np.testing.assert_array_equal(np.concatenate(bs_s), bs)
        assert np.array_equal(P.smear(ps_s), rp) and np.array_equal(P.smear(bs_s), rb)
        # event 2 alone, as event 42 of the same stream
        rp2, rb2, ps2, bs2 = P.generate(parts[2:], bins[2:], seed=7, event0=42)
        assert np.array_equal(rp2[0], rp[2]) and np.array_equal(ps2, ps_s[2])
        # another seed moves the draws
        ps3, _ = P.sample(parts, bins, seed=8, event0=40)
        assert not np.array_equal(ps3[0][:, 2], ps_s[0][:, 2])
        for e in range(3):
            assert H.rel_linf(rp[e], _oracle_rho(P, ps_s[e])) <= GRID_TOL if len(ps_s[e]) else not rp[e].any()


def test_full_chain_from_an_operation2_event():
    """Participants and binary collisions of a sampled Au+Au event (the
    operation-2 tables) through profile_3d's main configuration: 261 x 261 x
    101, ecm 19.6, random_flag 1."""
    from supermc_amd import Context
    p = H.glauber_params("AuAu", max_events=1)
    with Context(p) as ctx:
        ctx.run(seed=5, event0=3, nev=1)
        nuc = ctx.nucleons(0)
        pa, pb, _ = ctx.pairs(0)
    wa, wb = nuc["wound_proj"] >= 0, nuc["wound_targ"] >= 0
    parts = np.concatenate([np.column_stack([nuc["xy_proj"][wa], np.ones(wa.sum())]),
                            np.column_stack([nuc["xy_targ"][wb], 2 * np.ones(wb.sum())])])
    bins = (nuc["xy_proj"][pa] + nuc["xy_targ"][pb]) / 2
    with _p3d() as P:
        rp, rb, ps, bs = P.generate([parts], [bins], seed=123)
        assert H.rel_linf(rp[0], _oracle_rho(P, ps)) <= GRID_TOL
        assert H.rel_linf(rb[0], _oracle_rho(P, bs)) <= GRID_TOL
    # each participant carries unit weight, less what leaves the eta range
    dv = 0.1 ** 3
    inside = np.abs(ps[:, 2]) < 5.0 - 6 * 0.5
    assert inside.sum() * (1 - 1e-6) <= rp[0].sum() * dv <= len(parts) + 1e-3
