"""Parity of the HIP path against the FP64 oracle on the same nucleon
positions (and the same Gamma factors).

Bars (written here, checked below):
  Npart, Ncoll, participant sets, wound order, pair lists: bit-exact
  T_A, T_B, entropy grid: relative L-inf (max|d| / max|ref|) <= 1e-5
  exact-footprint paths (smear_path 0 and 2, the parity paths):
    eps_n (r^2 weight, n = 2..5) and eps'_n (r^n weight): absolute <= 1e-6
    <r^n>: relative <= 1e-5; dS/dy, grid sum: relative <= 1e-6
  MFMA paths (smear_path 1 FP32 and 3 BF16x3, opt-in, not parity paths):
  they smear over the
  square +-5w window, whose corners (r > 5w, each below e^-12.5 of the
  source's peak) the reference leaves out, so its moment bars are wider:
    eps_n: absolute <= 1e-5; eps'_n: absolute <= 5e-5 (the r^n weight
    grows toward the corners); <r^n>: relative <= 1e-2 (each corner
    cell carries up to e^-12.5 of a source's peak, and r^9 of a small
    peripheral event weights the corners most);
    dS/dy, grid sum: relative <= 1e-5; centre of mass: absolute <= 1e-5 fm
  centre of mass: absolute <= 1e-6 fm on the exact paths
  BF16x3 (smear_path 3): a factor split as bf16 hi + lo keeps 16 bits
  (|f - hi - lo| <= 2^-16 |f|), so its grids get relative L-inf <= 3e-5,
  against the oracle and against the FP32 MFMA strips on the same windows
"""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu

GRID_TOL = 1e-5
GRID_TOL_BF16X3 = 3e-5
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
    # the MFMA path keeps the square window's corners (each below e^-12.5 of
    # its source's peak), which moves grid sums by ~1e-6 relative
    return 1e-5 if p.smear_path in MFMA_PATHS and p.which_mc_model == 5 else 1e-6


def _check_event_record(rec, o, p):
    m = o["mom"]
    mfma = p.smear_path in MFMA_PATHS and p.which_mc_model == 5
    tol = 1e-5 if mfma else EPS_TOL
    tol_p = 5e-5 if mfma else EPS_TOL
    tol_rn = 1e-2 if mfma else 1e-5
    for n in range(p.ecc_from_order, p.ecc_to_order + 1):
        for key, off, t in (("mom_re", 0, tol), ("mom_im", 10, tol), ("momp_re", 20, tol_p), ("momp_im", 30, tol_p)):
            assert abs(rec[off + n] - m[key][n]) <= t, (key, n, rec[off + n], m[key][n])
    for n in range(10):
        assert rec[40 + n] == pytest.approx(m["rn"][n], rel=tol_rn)
    tol_cm = 1e-5 if mfma else 1e-6
    assert abs(rec[50] - m["xc"]) <= tol_cm and abs(rec[51] - m["yc"]) <= tol_cm
    assert rec[52] == pytest.approx(m["total"], rel=_sum_tol(p))


# smear_path 0: exact reference footprint, separable VALU tiles (default)
#            1: MFMA strips Gx.diag(w).Gy^T over the square window
#            2: exact reference footprint, per-cell exponentials
#            3: as 1 with BF16x3 split-precision MFMA
PATHS = [0, 1, 2, 3]
MFMA_PATHS = (1, 3)


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


def test_minbias_distributions_match_oracle_event_loop():
    """Two-sample KS (p > 1e-3) of accepted minimum-bias Pb+Pb events: the
    device's Philox sampling against the oracle's restatement of the
    MakeDensity.cpp:2143-2224 event loop with its own stream -- impact
    parameter, Npart, Ncoll and |eps_2| (r^2 weight)."""
    from scipy import stats
    import oracle
    p = H.glauber_params("PbPb", max_events=4000)
    with _ctx(p) as ctx:
        ev = ctx.run(seed=31337, event0=0, nev=4000)
    assert (ev[:, 57] == 0).all()
    cfg = oracle.OrcEventCfg(A=208, B=208, R_a=6.62, a_a=0.546, R_b=6.62, a_b=0.546, dmin=0.4, siginNN=64.0,
                             w=0.5, alpha=0.14, gamma_theta=1.25, fluct=1, bmin=0.0, bmax=20.0, npmin=2,
                             npmax=500, maxx=13.0, dx=0.1, from_order=2, to_order=5)
    n, o = oracle.run_events(cfg, 2718, 400, want_eps=True)
    assert n == 400
    gpu = dict(b=ev[:, 56], npart=ev[:, 54], ncoll=ev[:, 55], eps2=np.hypot(ev[:, 2], ev[:, 12]))
    orc = dict(b=o[:, 2], npart=o[:, 3], ncoll=o[:, 4], eps2=np.hypot(o[:, 0], o[:, 1]))
    for k in gpu:
        assert stats.ks_2samp(gpu[k], orc[k]).pvalue > 1e-3, k


@pytest.mark.parametrize("system", ["PbPb", "AuAu"])
def test_bf16x3_strips_match_fp32_strips(system):
    """The split-precision contraction against the FP32 one on the same
    sources and square windows: entropy grid (one accumulator) and T_A / T_B
    (two accumulators)."""
    grids = {}
    for path in (1, 3):
        p = H.glauber_params(system, max_events=16, smear_path=path)
        with _ctx(p) as ctx:
            ctx.run(seed=31, event0=0, nev=16)
            s, _, _ = ctx.density()
            ta, tb = ctx.thickness()
            grids[path] = (s, ta, tb)
    for g1, g3 in zip(grids[1], grids[3]):
        for e in range(16):
            assert H.rel_linf(g3[e], g1[e]) <= GRID_TOL_BF16X3, (e, H.rel_linf(g3[e], g1[e]))
        # and not the FP32 result by accident: split products differ in the last bits
        assert not np.array_equal(g1, g3)
