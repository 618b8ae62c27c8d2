"""Parity of the HIP path against the FP64 oracle on the same nucleon
positions (and the same Gamma factors).

Bars (written here, checked below):
  Npart, Ncoll, participant sets, wound order, pair lists: bit-exact
  T_A, T_B, entropy grid: relative L-inf (max|d| / max|ref|) <= 1e-5
  eps_n (r^2 weight, n = 2..5) and eps'_n (r^n weight): absolute <= 1e-6
  <r^n>: relative <= 1e-5; dS/dy, grid sum: relative <= 1e-6
  centre of mass: absolute <= 1e-6 fm
on every path with FP32 (or better) products: the VALU tiles (smear_path
0), the FP32 MFMA contraction with the out-of-disk products subtracted (1),
the per-cell exponentials (2) and the BF16x6 MFMA contraction (4, each
product within 4 * 2^-24 relative).
BF16x3 (smear_path 3): each factor is split as bf16 hi + lo with
|f - hi - lo| <= 2^-16 |f|, and the dropped lo*lo product is below 2^-16 of
the term, so every term, and hence every cell (a sum of positive terms),
is within d = 3 * 2^-16 = 4.6e-5 relative.  Its bars follow from d:
  grids: relative L-inf <= 5e-5 (d)
  eps_n, eps'_n = N / D with |numerator weight| <= denominator weight:
    absolute <= 1e-4 (2 d)
  sums: relative <= 5e-5 (d); <r^n>: relative <= 2e-3 and centre of mass
    <= 1e-3 fm (2 d |x|max on the +-13 fm grid, then n times the
    relative shift of r)
"""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu

GRID_TOL = 1e-5
GRID_TOL_BF16X3 = 5e-5
EPS_TOL = 1e-6


def _grid_tol(p):
    return GRID_TOL_BF16X3 if p.smear_path == 3 else GRID_TOL


def _ctx(p):
    from supermc_amd import Context
    return Context(p)


def _check_collision(ctx, ev, o):
    nuc = o["nuc"]
    col = o["col"]
    pa, pb, _ = o["gpu_pairs"]
    # the HIP path already returns collision order
    assert (o["perm_a"] == np.arange(len(o["perm_a"]))).all()
    assert (o["perm_b"] == np.arange(len(o["perm_b"]))).all()
    assert (nuc["wound_proj"] >= 0).sum() == col["npart_a"]
    assert (nuc["wound_targ"] >= 0).sum() == col["npart_b"]
    assert ((nuc["wound_proj"] >= 0) == (col["wound_a"] == 1)).all()
    assert ((nuc["wound_targ"] >= 0) == (col["wound_b"] == 1)).all()
    assert len(pa) == col["ncoll"]
    assert (pa == col["pair_a"]).all() and (pb == col["pair_b"]).all()
    # wound order (Nucleus.cpp:325-331)
    assert (o["ia"] == col["order_a"]).all()
    assert (o["ib"] == col["order_b"]).all()


def _sum_tol(p):
    return 5e-5 if p.smear_path == 3 else 1e-6


def _check_event_record(rec, o, p):
    m = o["mom"]
    b3 = p.smear_path == 3
    tol = 1e-4 if b3 else EPS_TOL
    tol_rn = 2e-3 if b3 else 1e-5
    for n in range(p.ecc_from_order, p.ecc_to_order + 1):
        for key, off in (("mom_re", 0), ("mom_im", 10), ("momp_re", 20), ("momp_im", 30)):
            assert abs(rec[off + n] - m[key][n]) <= tol, (key, n, rec[off + n], m[key][n])
    for n in range(10):
        assert rec[40 + n] == pytest.approx(m["rn"][n], rel=tol_rn)
    tol_cm = 1e-3 if b3 else 1e-6
    assert abs(rec[50] - m["xc"]) <= tol_cm and abs(rec[51] - m["yc"]) <= tol_cm
    assert rec[52] == pytest.approx(m["total"], rel=_sum_tol(p))


# smear_path 0: separable VALU tiles (default)
#            1: FP32 MFMA contraction, out-of-disk products subtracted
#            2: per-cell exponentials
#            3: as 1 with BF16x3 split precision
#            4: as 1 with BF16x6 split precision
PATHS = [0, 1, 2, 3, 4]
MFMA_PATHS = (1, 3, 4)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("system", ["AuAu", "PbPb"])
def test_fed_positions_full_chain(system, path):
    p = H.glauber_params(system, max_events=12, smear_path=path)
    bvals = np.array([0.0, 1.0, 2.5, 4.0, 5.5, 7.0, 8.5, 10.0, 11.5, 12.5, 13.5, 14.5])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=1234 + p.Aproj, bvals=bvals)
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=99)
        npa, npb, nc, rc = ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        rec = ctx.moments()
        ta, tb = ctx.thickness()
        for e in range(len(bvals)):
            o = H.oracle_event(ctx, e)
            # positions come back exactly (quantised input)
            assert np.array_equal(np.sort(o["nuc"]["xy_proj"][:, 0]), np.sort(xa[e][:, 0]))
            assert npa[e] == o["col"]["npart_a"] and npb[e] == o["col"]["npart_b"]
            assert nc[e] == o["col"]["ncoll"]
            if nc[e] == 0:
                continue
            _check_collision(ctx, e, o)
            assert H.rel_linf(s[e], o["s"]) <= _grid_tol(p), H.rel_linf(s[e], o["s"])
            assert H.rel_linf(ta[e], o["TA"]) <= _grid_tol(p), H.rel_linf(ta[e], o["TA"])
            assert H.rel_linf(tb[e], o["TB"]) <= _grid_tol(p), H.rel_linf(tb[e], o["TB"])
            assert dndy[e] == pytest.approx(o["dndy"], rel=_sum_tol(p))
            _check_event_record(rec[e], o, p)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("system", ["AuAu", "PbPb"])
def test_ulrich_glauber_matches_oracle(system, path):
    """which_mc_model 5, sub_model 2 (MCnucl.cpp:713-717): no participant
    term, pair k weighted alpha + (1 - alpha) aw_k (MCnucl.cpp:345-348,
    750-755).  Pair weights exactly the oracle's rule on the reported pair
    lists; density grids on every smearing path, FP64 grids and eps_n
    against the oracle on fed positions."""
    p = H.glauber_params(system, max_events=8, smear_path=path, sub_model=2)
    bvals = np.array([0.0, 3.0, 6.0, 9.0, 11.0, 12.5, 13.5, 14.5])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=4321 + p.Aproj, bvals=bvals)
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=17)
        npa, npb, nc, rc = ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        rec = ctx.moments()
        s64 = ctx.density_f64() if path == 0 else None
        seen = set()
        for e in range(len(bvals)):
            if nc[e] == 0:
                continue
            o = H.oracle_event(ctx, e)
            _check_collision(ctx, e, o)
            _, kind, wsrc = ctx.sources(e)
            pa, pb, pf = o["gpu_pairs"]
            want = (pf.astype(np.float64) * (p.alpha + (1 - p.alpha) * o["aw"])).astype(np.float32)
            np.testing.assert_allclose(wsrc[kind == 2], want, rtol=2e-7)
            seen.update(o["aw"].tolist())
            assert H.rel_linf(s[e], o["s"]) <= _grid_tol(p), H.rel_linf(s[e], o["s"])
            assert dndy[e] == pytest.approx(o["dndy"], rel=_sum_tol(p))
            _check_event_record(rec[e], o, p)
            if s64 is not None:
                assert H.rel_linf(s64[e], o["s"]) <= 1e-12, H.rel_linf(s64[e], o["s"])
        assert seen == {0.0, 1.0, 2.0}  # central and peripheral events hold every weight


@pytest.mark.parametrize("orders,dx", [((1, 9), 0.1), ((1, 9), 0.05), ((3, 7), 0.1), ((1, 3), 0.1), ((3, 5), 0.05)])
def test_order_ranges_match_oracle(orders, dx):
    """eps_1..eps_9 (the parameter file's default range, with n = 1's r^3
    weight, MakeDensity.cpp:2408-2413) on the fixed-order moments kernel at
    261^2 and 521^2, a wider-than-four range on the same kernel (records
    keep the requested orders), and ranges of up to four orders on the
    general kernel, against the oracle on fed positions."""
    p = H.glauber_params("PbPb", max_events=5, ecc_from_order=orders[0], ecc_to_order=orders[1], dx=dx, dy=dx)
    bvals = np.array([0.0, 3.0, 6.5, 9.5, 12.0])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=77, bvals=bvals)
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=5)
        ctx.collide(allow=(4,))
        ctx.density()
        rec = ctx.moments()
        for e in range(len(bvals)):
            _check_event_record(rec[e], H.oracle_event(ctx, e), p)
        # orders outside the range stay zero
        for n in range(10):
            if not orders[0] <= n <= orders[1]:
                assert (rec[:, [n, 10 + n, 20 + n, 30 + n]] == 0).all()


@pytest.mark.parametrize("maxx,dx,orders,model", [(30.0, 0.1, (2, 5), 5), (13.0, 0.02, (2, 5), 5),
                                                  (13.0, 0.02, (1, 9), 5), (13.0, 0.02, (1, 9), 7)])
def test_wide_grids_match_oracle(maxx, dx, orders, model):
    """Grids wider than 576 columns, where the moment kernels take each
    event's rectangle in column segments of 576: 601^2 at 0.1 fm (one
    segment) and 1301^2 at 0.02 fm (rectangles over 576 columns wide), for
    orders 2..5 and 1..9, MC-Glauber and sqrt(T_A T_B) (the two-accumulator
    tile epilogue), against the oracle on fed positions."""
    p = H.glauber_params("PbPb", max_events=3, maxx=maxx, maxy=maxx, dx=dx, dy=dx, ecc_from_order=orders[0],
                         ecc_to_order=orders[1], which_mc_model=model)
    bvals = np.array([0.0, 8.0, 13.0])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=91, bvals=bvals)
    with _ctx(p) as ctx:
        assert ctx.ny > 576
        ctx.set_positions(xa, xb, bs, fluct_seed=3)
        ctx.collide(allow=(4,))
        s, _, _ = ctx.density()
        rec = ctx.moments()
        spans = []
        for e in range(len(bvals)):
            o = H.oracle_event(ctx, e)
            assert H.rel_linf(s[e], o["s"]) <= _grid_tol(p), H.rel_linf(s[e], o["s"])
            _check_event_record(rec[e], o, p)
            cols = np.nonzero(s[e].any(axis=0))[0]
            spans.append(cols[-1] - cols[0] + 1)
        if dx < 0.05:
            assert max(spans) > 576, spans


def test_largest_grid_runs():
    """2048 x 2048 cells, the largest grid a context accepts (the band mask
    of the moment kernel then fills 64 KB of LDS): sampled events run, and
    their records agree with the same events' records at 0.1 fm, to the
    resolution difference."""
    p = H.glauber_params("PbPb", max_events=4, maxx=20.47, maxy=20.47, dx=0.02, dy=0.02)
    q = H.glauber_params("PbPb", max_events=4, maxx=20.5, maxy=20.5, dx=0.1, dy=0.1)
    with _ctx(p) as ctx:
        assert ctx.nx == 2048 and ctx.ny == 2048
        fine = ctx.run(seed=12, event0=0, nev=4)
    with _ctx(q) as ctx:
        coarse = ctx.run(seed=12, event0=0, nev=4)
    assert (fine[:, 57] == 0).all()
    assert np.array_equal(fine[:, 54:57], coarse[:, 54:57])  # Npart, Ncoll, b: the same events
    assert np.all(np.isfinite(fine[:, :53]))
    np.testing.assert_allclose(fine[:, 2], coarse[:, 2], atol=2e-3)  # eps_2 (real part), grid resolution


@pytest.mark.parametrize("path", PATHS)
def test_sampled_events_match_oracle(path):
    p = H.glauber_params("PbPb", max_events=48, smear_path=path)
    with _ctx(p) as ctx:
        ev = ctx.run(seed=2026, event0=0, nev=48)
        assert (ev[:, 57] == 0).all()
        for e in range(0, 48, 3):
            o = H.oracle_event(ctx, e)
            _check_collision(ctx, e, o)
            rec = ev[e]
            assert rec[54] == o["col"]["npart_a"] + o["col"]["npart_b"]
            assert rec[55] == o["col"]["ncoll"]
            assert 2 <= rec[54] <= 500 and rec[55] > 0
            _check_event_record(rec, o, p)
        s, _, _ = ctx.density()
        for e in range(0, 48, 7):
            o = H.oracle_event(ctx, e)
            assert H.rel_linf(s[e], o["s"]) <= _grid_tol(p)


@pytest.mark.parametrize("path", PATHS)
def test_counter_based_events_are_batch_independent(path):
    p = H.glauber_params("PbPb", max_events=32, smear_path=path)
    with _ctx(p) as ctx:
        a = ctx.run(seed=7, event0=100, nev=32)
        b = ctx.run(seed=7, event0=110, nev=8)
        assert np.array_equal(a[10:18], b)
        c = ctx.run(seed=8, event0=100, nev=32)
        assert not np.array_equal(a[:, 54], c[:, 54])


def test_trento_model_density():
    p = H.glauber_params("PbPb", which_mc_model=7, max_events=6)
    xa, xb, bs = H.cpu_configs(p, 6, seed=77, bvals=np.array([0.0, 3.0, 6.0, 9.0, 11.0, 13.0]))
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=5)
        ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        rec = ctx.moments()
        for e in range(6):
            o = H.oracle_event(ctx, e)
            if o["col"]["ncoll"] == 0:
                continue
            assert H.rel_linf(s[e], o["s"]) <= GRID_TOL
            _check_event_record(rec[e], o, p)


def test_edge_cases_single_nucleons_and_no_collision():
    # p+p head-on (A=1 nucleons sit at +-b/2) and a miss
    p = H.glauber_params("AuAu", Aproj=1, Atarg=1, max_events=3)
    xa = np.array([[[0.25, 0.0]], [[3.0, 0.0]], [[0.0, 0.6]]])
    xb = np.array([[[-0.25, 0.0]], [[-3.0, 0.0]], [[0.0, -0.55]]])
    with _ctx(p) as ctx:
        ctx.set_positions(xa, xb, np.array([0.5, 6.0, 0.0]))
        npa, npb, nc, _ = ctx.collide()
        d = ctx.derived
        for e in range(3):
            dist2 = ((xa[e, 0] - xb[e, 0]) ** 2).sum()
            hit = np.sqrt(dist2) ** 2 <= d.dsq
            assert nc[e] == int(hit) and npa[e] == int(hit) and npb[e] == int(hit)


def test_gamma_factors_follow_gamma_law():
    from scipy import stats
    p = H.glauber_params("PbPb", max_events=256)
    with _ctx(p) as ctx:
        ctx.sample(seed=31, event0=0, nev=256)
        fp, fc = [], []
        for e in range(256):
            nuc = ctx.nucleons(e)
            fp += list(nuc["fluct_proj"][nuc["wound_proj"] >= 0])
            fc += list(ctx.pairs(e)[2])
    eps, th, a = 1e-8, 1.25, 0.14
    kp, sp = (1 - a + eps) / 2 / th, 2 / (1 - a + eps) * th
    kb, sb = (a + eps) / th, th / (a + eps)
    assert stats.kstest(fp, stats.gamma(kp, scale=sp).cdf).pvalue > 1e-3
    assert stats.kstest(fc, stats.gamma(kb, scale=sb).cdf).pvalue > 1e-3


def test_sampled_nuclei_follow_reference_sampler():
    """Two-sample KS of GPU-sampled nucleon radii (about the nucleus centre)
    and Npart against the oracle's restatement of Nucleus::populate."""
    from scipy import stats
    import oracle
    p = H.glauber_params("PbPb", bmin=0.0, bmax=0.0, Npmin=0, Npmax=1000, max_events=200)
    with _ctx(p) as ctx:
        ctx.sample(seed=4242, event0=0, nev=200)
        npa, npb, nc, _ = ctx.collide()
        rg = []
        for e in range(200):
            nuc = ctx.nucleons(e)
            rg += list(np.hypot(nuc["xy_proj"][:, 0], nuc["xy_proj"][:, 1]))
    rng = oracle.Rng(99)
    rc, npc = [], []
    dsq = oracle.lib().orc_dsq(64.0)
    for e in range(200):
        a = rng.populate(208, 6.62, 0.546, 0.4, 0.0)
        b = rng.populate(208, 6.62, 0.546, 0.4, 0.0)
        rc += list(np.hypot(a[:, 0], a[:, 1]))
        n, wa, wb = oracle.collide_bruteforce(a, b, dsq)
        npc.append(wa.sum() + wb.sum())
    assert stats.ks_2samp(rg, rc).pvalue > 1e-3
    assert abs(np.mean(npa + npb) - np.mean(npc)) < 4 * np.std(npc) / np.sqrt(200) * 1.5


def _oracle_loop_chunk(args):
    import oracle
    seed, nev = args
    cfg = oracle.OrcEventCfg(A=208, B=208, R_a=6.62, a_a=0.546, R_b=6.62, a_b=0.546, dmin=0.4, siginNN=64.0,
                             w=0.5, alpha=0.14, gamma_theta=1.25, fluct=1, bmin=0.0, bmax=20.0, npmin=2,
                             npmax=500, maxx=13.0, dx=0.1, from_order=2, to_order=5)
    n, o = oracle.run_events(cfg, seed, nev, want_eps=True)
    assert n == nev
    return o


def test_minbias_distributions_match_oracle_event_loop():
    """Two-sample KS (p > 1e-3) of accepted minimum-bias Pb+Pb events: 20000
    from the device's Philox sampling against 5000 from the oracle's
    restatement of the MakeDensity.cpp:2143-2224 event loop with its own
    streams (run over the host cores) -- impact parameter, Npart, Ncoll and
    |eps_2| (r^2 weight).  At these sizes a CDF shift of about 3 % fails."""
    import multiprocessing as mp
    from scipy import stats
    import oracle
    oracle.lib()  # built before the workers fork
    p = H.glauber_params("PbPb", max_events=20000)
    with _ctx(p) as ctx:
        ev = ctx.run(seed=31337, event0=0, nev=20000)
    assert (ev[:, 57] == 0).all()
    chunks = [(2718 + i, 250) for i in range(20)]
    with mp.get_context("fork").Pool(min(16, len(chunks))) as pool:
        o = np.concatenate(pool.map(_oracle_loop_chunk, chunks))
    assert len(o) == 5000
    gpu = dict(b=ev[:, 56], npart=ev[:, 54], ncoll=ev[:, 55], eps2=np.hypot(ev[:, 2], ev[:, 12]))
    orc = dict(b=o[:, 2], npart=o[:, 3], ncoll=o[:, 4], eps2=np.hypot(o[:, 0], o[:, 1]))
    for k in gpu:
        assert stats.ks_2samp(gpu[k], orc[k]).pvalue > 1e-3, k


@pytest.mark.parametrize("model", [5, 7, 1])
@pytest.mark.parametrize("system", ["PbPb", "AuAu"])
def test_mfma_paths_match_valu_tiles(system, model):
    """The MFMA contractions against the VALU tiles on the same sampled
    events: entropy grid (Glauber, one accumulator), sqrt(T_A T_B) (two
    accumulators combined in the epilogue), MC-KLN through the table, and
    T_A / T_B; FP32 and BF16x6 at the exact bar, BF16x3 at its bound."""
    kw = dict(which_mc_model=model)
    if model == 1:
        kw.update(sub_model=7, ecm=2760.0, lambda_=0.138, tmax=71, tmax_subdivision=3, kln_n_pt=12, kln_n_kt=12,
                  kln_n_phi=16, cc_fluctuation_model=0)
    grids = {}
    for path in (0, 1, 3, 4):
        p = H.glauber_params(system, max_events=16, smear_path=path, **kw)
        with _ctx(p) as ctx:
            ev = ctx.run(seed=31, event0=0, nev=16)
            s, _, _ = ctx.density()
            ta, tb = ctx.thickness()
            grids[path] = (s, ta, tb, ev)
    for path in (1, 3, 4):
        tol = GRID_TOL_BF16X3 if path == 3 else GRID_TOL
        for g0, g in zip(grids[0][:3], grids[path][:3]):
            for e in range(16):
                assert H.rel_linf(g[e], g0[e]) <= tol, (path, e, H.rel_linf(g[e], g0[e]))
        etol = 1e-4 if path == 3 else 2e-6
        assert np.abs(grids[path][3][:, :40] - grids[0][3][:, :40]).max() <= etol
    # split precision is not FP32 by accident
    assert not np.array_equal(grids[1][0], grids[3][0])


@pytest.mark.parametrize("model", [5, 7])
def test_mfma_paths_on_a_wide_grid(model):
    """601^2 (wider than 576 columns): the MFMA contractions (FP32, BF16x3,
    BF16x6) and the per-cell exponentials against the VALU tiles on the same
    sampled events, grids and records, at the bars of the 261^2
    comparison."""
    grids = {}
    for path in (0, 1, 2, 3, 4):
        p = H.glauber_params("PbPb", max_events=8, smear_path=path, which_mc_model=model, maxx=30.0, maxy=30.0)
        with _ctx(p) as ctx:
            assert ctx.ny > 576
            ev = ctx.run(seed=47, event0=0, nev=8)
            s, _, _ = ctx.density()
            grids[path] = (s, ev)
    assert (grids[0][1][:, 57] == 0).all()
    for path in (1, 2, 3, 4):
        tol = GRID_TOL_BF16X3 if path == 3 else GRID_TOL
        for e in range(8):
            assert H.rel_linf(grids[path][0][e], grids[0][0][e]) <= tol, (path, e)
        etol = 1e-4 if path == 3 else 2e-6
        assert np.abs(grids[path][1][:, :40] - grids[0][1][:, :40]).max() <= etol, path


@pytest.mark.parametrize("model", [5, 7])
@pytest.mark.parametrize("path", PATHS)
def test_rectangular_grid_and_cells_match_oracle(path, model):
    """maxx != maxy and dx != dy, which the reference keeps apart
    (MCnucl.cpp:39-40, 460; the hot-spot bound that divides a y extent by
    dx, MakeDensity.cpp:2357): 261 x 226 cells of 0.1 x 0.08 fm, Au+Au on
    fed positions, every smearing path, against the oracle."""
    p = H.glauber_params("AuAu", max_events=4, smear_path=path, which_mc_model=model, maxx=13.0, maxy=9.0, dx=0.1,
                         dy=0.08)
    bvals = np.array([0.0, 4.0, 8.0, 11.0])
    xa, xb, bs = H.cpu_configs(p, len(bvals), seed=63, bvals=bvals)
    with _ctx(p) as ctx:
        assert (ctx.nx, ctx.ny) == (261, 226)
        ctx.set_positions(xa, xb, bs)
        ctx.collide(allow=(4,))
        s, _, _ = ctx.density()
        rec = ctx.moments()
        checked = 0
        for e in range(len(bvals)):
            o = H.oracle_event(ctx, e)
            if o["col"]["ncoll"] == 0:
                continue
            assert H.rel_linf(s[e], o["s"]) <= _grid_tol(p), (e, H.rel_linf(s[e], o["s"]))
            _check_event_record(rec[e], o, p)
            checked += 1
        assert checked >= 3
