"""MC-KLN (which_mc_model 1, sub_model 7) on the HIP path against the FP64
oracle.

Bars (written here, checked below):
  k_kln_points (FP64 dN/dy at (T_A, T_B)): relative <= 1e-11 against the
    oracle's Gauss-Legendre restatement of KLNModel::getdNdy with the same
    nodes (only the summation order differs)
  dN/dy table (MCnucl::makeTable): row and column 0 exactly 0, sampled
    entries relative <= 1e-11
  table path density (sixPoint2dInterp): relative L-inf <= 1e-5 against the
    oracle fed FP64 T_A, T_B and the same table; dN/dy relative <= 1e-6
  direct per-cell FP32 quadrature: relative L-inf <= 1e-5 against the FP64
    quadrature on the same T_A, T_B
  eps_n (r^2 weight): absolute <= 1e-6
"""
import numpy as np
import pytest

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu

BVALS = np.array([0.0, 2.0, 4.5, 7.0, 9.5, 12.0])


def _kln_params(**kw):
    base = dict(which_mc_model=1, sub_model=7, max_events=len(BVALS), ecm=2760.0, lambda_=0.138,
                kln_n_pt=24, kln_n_kt=24, kln_n_phi=32)
    base.update(kw)
    return H.glauber_params("PbPb", **base)


def _ctx(p):
    from supermc_amd import Context
    return Context(p)


def _orc_kln(p):
    k = oracle.kln(p.ecm, p.lambda_)
    k.ptmin, k.ptmax = p.kln_pt_min, p.kln_pt_max
    return k


def test_kln_points_match_oracle_quadrature():
    p = _kln_params()
    k = _orc_kln(p)
    rng = np.random.default_rng(5)
    ta = np.concatenate([[0.0, 1e-3, 0.02, 0.5, 1.0, 3.0, 30.0, 90.0], rng.uniform(0, 20, 24)])
    tb = np.concatenate([[1.0, 1e-3, 0.5, 0.02, 1.0, 0.0, 60.0, 90.0], rng.uniform(0, 20, 24)])
    with _ctx(p) as ctx:
        got = ctx.kln_dndy(ta, tb)
    for a, b, g in zip(ta, tb, got):
        ref = oracle.kln_dndy_gl(k, a, b, p.kln_n_pt, p.kln_n_kt, p.kln_n_phi)
        if ref == 0.0:
            assert g == 0.0
        else:
            assert abs(g - ref) <= 1e-11 * abs(ref), (a, b, g, ref)


def test_kln_table_entries():
    p = _kln_params()
    k = _orc_kln(p)
    with _ctx(p) as ctx:
        d = ctx.derived
        t = ctx.kln_make_table()
    n = d.kln_tmax
    assert t.shape == (n, n)
    assert (t[0] == 0).all() and (t[:, 0] == 0).all()
    rng = np.random.default_rng(11)
    idx = [(1, 1), (n - 1, n - 1), (1, n - 1), (n - 1, 1)] + [tuple(v) for v in rng.integers(1, n, (24, 2))]
    for i, j in idx:
        ref = oracle.kln_dndy_gl(k, d.kln_dT * i, d.kln_dT * j, p.kln_n_pt, p.kln_n_kt, p.kln_n_phi)
        assert abs(t[i, j] - ref) <= 1e-11 * abs(ref), (i, j, t[i, j], ref)
    # dN/dy grows with either thickness
    assert (np.diff(t[1:, 1:], axis=0) > 0).all() and (np.diff(t[1:, 1:], axis=1) > 0).all()


@pytest.mark.parametrize("grid", [{}, dict(maxx=30.0, maxy=30.0), dict(maxy=9.0, dy=0.08)],
                         ids=["261", "601", "261x226"])
def test_kln_table_density_full_chain(grid):
    # 601^2: tiles and moments over grids wider than 576 columns; 261 x 226:
    # rectangular cells of 0.1 x 0.08 fm
    p = _kln_params(**grid)
    xa, xb, bs = H.cpu_configs(p, len(BVALS), seed=321, bvals=BVALS)
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        rec = ctx.moments()
        table = ctx.kln_make_table()
        for e in range(len(BVALS)):
            o = H.oracle_event(ctx, e, kln_table=table)
            if o["col"]["ncoll"] == 0:
                continue
            assert o["dndy"] > 0
            assert H.rel_linf(s[e], o["s"]) <= 1e-5, H.rel_linf(s[e], o["s"])
            assert dndy[e] == pytest.approx(o["dndy"], rel=1e-6)
            m = o["mom"]
            for n in range(p.ecc_from_order, p.ecc_to_order + 1):
                assert abs(rec[e, n] - m["mom_re"][n]) <= 1e-6
                assert abs(rec[e, 10 + n] - m["mom_im"][n]) <= 1e-6
            assert rec[e, 57] == 0


def test_kln_fused_density_is_the_table_of_its_own_thickness():
    """The one-launch smearing + table path against the oracle's
    MCnucl::setDensity table step fed the float T_A, T_B grids that the
    thickness call returns for the same events: the same FP64 arithmetic on
    the same inputs, so only rounding of the FP64 expression can differ
    (relative L-inf <= 1e-6)."""
    p = _kln_params()
    xa, xb, bs = H.cpu_configs(p, len(BVALS), seed=17, bvals=BVALS)
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        ta, tb = ctx.thickness()
        table = ctx.kln_make_table()
        g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
        for e in range(len(BVALS)):
            ref, ref_sum = oracle.density_kln_table(g, ta[e].astype(np.float64), tb[e].astype(np.float64), table,
                                                    ctx.derived.kln_dT)
            if ref_sum <= 0:
                continue
            assert H.rel_linf(s[e], ref) <= 1e-6, (e, H.rel_linf(s[e], ref))
            assert dndy[e] == pytest.approx(ref_sum, rel=1e-6)


def test_kln_direct_density_matches_fp64_quadrature():
    p = _kln_params(kln_direct=1, max_events=3)
    k = _orc_kln(p)
    xa, xb, bs = H.cpu_configs(p, 3, seed=99, bvals=np.array([0.0, 5.0, 10.0]))
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        ta, tb = ctx.thickness()
        for e in range(3):
            ref = ctx.kln_dndy(ta[e].ravel().astype(np.float64), tb[e].ravel().astype(np.float64))
            assert H.rel_linf(s[e].ravel(), ref) <= 1e-5, H.rel_linf(s[e].ravel(), ref)
            assert dndy[e] == pytest.approx(ref.sum(), rel=1e-5)
        # a few cells straight against the oracle on the oracle's own T_A, T_B
        o = H.oracle_event(ctx, 1)
        rng = np.random.default_rng(3)
        cells = np.argwhere((o["TA"] > 0) & (o["TB"] > 0))
        smax = s[1].max()
        for ix, iy in cells[rng.choice(len(cells), 12, replace=False)]:
            ref = oracle.kln_dndy_gl(k, o["TA"][ix, iy], o["TB"][ix, iy], p.kln_n_pt, p.kln_n_kt, p.kln_n_phi)
            assert abs(s[1][ix, iy] - ref) <= 1e-5 * smax


def test_kln_table_overflow_sets_status():
    # tmax = 3 without subdivision: the table ends at T = 10/sigma, below the
    # centre of a central Pb+Pb event (MCnucl.cpp:665-670 stops the run)
    p = _kln_params(tmax=3, tmax_subdivision=1, max_events=2)
    xa, xb, bs = H.cpu_configs(p, 2, seed=8, bvals=np.array([0.0, 1.0]))
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs)
        ctx.collide(allow=(4,))
        ctx.density()
        rec = ctx.moments()
    assert ((rec[:, 57].astype(int) & 8) != 0).all()


def test_kln_direct_chunked_batch_matches_small_batches():
    """The per-cell quadrature runs a batch in chunks of at most 2^28 cells
    (its active-cell list holds 32-bit indices): at 2048^2 that is 64 events,
    so 130 events take three chunks.  Events on both sides of each chunk
    edge give bit-identical records when the same events run alone in a
    batch of four (one chunk)."""
    kw = dict(kln_direct=1, maxx=20.47, maxy=20.47, dx=0.02, dy=0.02, kln_n_pt=4, kln_n_kt=4, kln_n_phi=8)
    with _ctx(_kln_params(max_events=130, **kw)) as ctx:
        assert ctx.nx == 2048 and ctx.ny == 2048
        big = ctx.run(seed=31, event0=0, nev=130)
    assert (big[:, 57] == 0).all()
    with _ctx(_kln_params(max_events=4, **kw)) as ctx:
        for e0 in (62, 126):
            small = ctx.run(seed=31, event0=e0, nev=4)
            assert np.array_equal(small, big[e0:e0 + 4]), e0
