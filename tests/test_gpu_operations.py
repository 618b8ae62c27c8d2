"""Operations 2 and 3 of the reference beyond the eccentricity table.

Operation 2 grids (MakeDensity.cpp:148-500): smc_rho_binary
(MCnucl::calculate_rho_binary) and smc_spectator_density
(MCnucl::calculate_spectator_density) against the FP64 oracle's Glauber
collision term and thickness on the same nucleons: relative L-inf <= 1e-5.

Averaged, rotated profiles (operation 3, MakeDensity.cpp:736-1650,
pT-integrated): smc_average_profiles against the FP64 oracle's
rotated_profiles on the same nucleons and Gamma factors.

Bars: the sum over events of finalFactor * rho after the order-n recentring
and turn, and of the recentred-only (reaction-plane) profile: relative L-inf
(max|d| / max|ref|) <= 1e-5, as for single grids.  The device turns the
sources from its FP32 grid and keeps positions in half-quanta (2^-25 fm),
the oracle from its FP64 grid and FP64 positions; both are far below the bar.
"""
import numpy as np
import pytest

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle_sum(ctx, p, events, from_order, to_order, rp):
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    tot = np.zeros((to_order - from_order + 1, g.nx, g.ny))
    tot_rp = np.zeros_like(tot)
    for e in events:
        o = H.oracle_event(ctx, e)
        nuc = o["nuc"]
        pa, pb, pf = o["gpu_pairs"]
        xa, xb = nuc["xy_proj"], nuc["xy_targ"]
        ia, ib = o["ia"], o["ib"]
        mid = (xa[pa] + xb[pb]) / 2.0
        prof, rprof = oracle.rotated_profiles(g, ctx.derived.width, p.alpha, xa[ia], nuc["fluct_proj"][ia].astype(float),
                                              xb[ib], nuc["fluct_targ"][ib].astype(float), mid, pf.astype(float),
                                              from_order, to_order, p.finalFactor, reaction_plane=rp)
        tot += prof
        if rp:
            tot_rp += rprof
    return tot, tot_rp


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("grid", [{}, dict(maxy=9.0, dy=0.08)], ids=["261", "261x226"])
@pytest.mark.parametrize("rp", [False, True])
def test_rotated_average_matches_oracle(rp, grid):
    # 261 x 226: recentring and turning on rectangular cells (0.1 x 0.08 fm)
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=6, finalFactor=1.7, **grid)
    bvals = [0.0, 2.0, 4.5, 7.0, 9.5, 12.0]
    xa, xb, bs = H.cpu_configs(p, len(bvals), 41, bvals)
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=5)
        npa, npb, nc, _ = ctx.collide(allow=(4,))
        accept = (nc > 0).astype(np.uint8)
        accept[3] = 0  # an excluded event stays out of the sums
        events = [e for e in range(len(bvals)) if accept[e]]
        # the oracle works from the unturned nucleons: take them first
        ref, ref_rp = _oracle_sum(ctx, p, events, 2, 4, rp)
        out = ctx.average_profiles(2, 4, accept=accept, reaction_plane=rp)
    sums, sums_rp = out if rp else (out, None)
    for o in range(3):
        assert _rel(sums[o], ref[o]) <= TOL, (o, _rel(sums[o], ref[o]))
        if rp:
            assert _rel(sums_rp[o], ref_rp[o]) <= TOL, (o, _rel(sums_rp[o], ref_rp[o]))
    # turned profiles: the order-2 axis lies on y (the reference's convention)
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    x = g.xmin + np.arange(g.nx)[:, None] * g.dx
    y = g.ymin + np.arange(g.ny)[None, :] * g.dy
    w = sums[0] / sums[0].sum()
    assert ((y ** 2) * w).sum() > ((x ** 2) * w).sum()


def test_rotated_average_sampled_events_status_default():
    """Sampled events, accept = status 0 (NULL): finite, non-negative sums
    whose total equals the turned events' dS/dy up to the footprints that
    leave the grid."""
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=64, bmin=0.0, bmax=10.0)
    with Context(p) as ctx:
        rec = ctx.run(seed=77, event0=0, nev=64)
        sums = ctx.average_profiles(2, 2)
    assert np.isfinite(sums).all() and (sums >= 0).all()
    tot = sums[0].sum() * p.dx * p.dy
    assert tot == pytest.approx(rec[rec[:, 57] == 0, 53].sum(), rel=1e-3)


def test_average_sums_equal_host_sum_of_turned_grids():
    """The device sums (event segments, rectangle skipping) against a host
    FP64 sum of the same turned grids: 301 events (segments of 10: one group
    of eight plus a remainder), every seventh event excluded."""
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=301, finalFactor=1.3)
    with Context(p) as ctx:
        ctx.run(seed=9, event0=0, nev=301, fetch=False)
        accept = (np.arange(301) % 7 != 3).astype(np.uint8)
        sums = ctx.average_profiles(2, 2, accept=accept)
        rho, _, _ = ctx.density()  # the grids are left turned by the order-2 angle
    ref = p.finalFactor * rho[accept == 1].astype(np.float64).sum(axis=0)
    np.testing.assert_allclose(sums[0], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # no accepted event: zero sums
    with Context(p) as ctx:
        ctx.run(seed=9, event0=0, nev=301, fetch=False)
        s2 = ctx.average_profiles(2, 2, accept=np.zeros(301, np.uint8))
    assert not s2.any()


def test_rho_binary_and_spectators_match_oracle():
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=5)
    bvals = [0.0, 3.0, 6.5, 10.0, 13.5]
    xa, xb, bs = H.cpu_configs(p, len(bvals), 43, bvals)
    g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=2)
        npa, npb, nc, _ = ctx.collide(allow=(4,))
        rho = ctx.rho_binary()
        spa, spb = ctx.spectator_density()
        w = ctx.derived.width
        for e in range(len(bvals)):
            nuc = ctx.nucleons(e)
            pa, pb, _ = ctx.pairs(e)
            mid = (nuc["xy_proj"][pa] + nuc["xy_targ"][pb]) / 2.0
            no = np.zeros((0, 2))
            ref, _ = oracle.density_glauber(g, w, 1.0, no, np.zeros(0), no, np.zeros(0), mid, np.ones(len(mid)))
            if len(mid):
                assert _rel(rho[e], ref) <= TOL
            else:
                assert (rho[e] == 0).all()
            for got, key, wkey in ((spa[e], "xy_proj", "wound_proj"), (spb[e], "xy_targ", "wound_targ")):
                spec = nuc[key][nuc[wkey] < 0]
                ref = oracle.thickness(g, w, spec)
                assert _rel(got, ref) <= TOL
        # every nucleon is a participant or a spectator: the two grids add up
        # to the thickness of all nucleons
        e = 2
        nuc = ctx.nucleons(e)
        all_a = oracle.thickness(g, w, nuc["xy_proj"])
        ta, _ = ctx.thickness()
        assert _rel(spa[e] + ta[e], all_a) <= TOL


@pytest.mark.parametrize("turn", [False, True])
def test_recenter_and_rotate_sources_match_oracle(turn):
    """smc_recenter (MCnucl::recenterGrid / rotateGrid, MCnucl.cpp:1119-1174):
    the centre of mass and order-2 angle of the current density against
    the oracle's GlueDensity::calcCMAngle on the oracle density of the same
    event, and every source moved to R(angle) (x - c), rounded to the
    half-quantum; the re-smeared density then has its centre at the origin
    (to the grid's discretisation) and, turned, an order-2 angle of 0."""
    import oracle
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=6)
    with Context(p) as ctx:
        ctx.run(seed=404, event0=0, nev=6)
        before = [ctx.sources(e) for e in range(6)]
        ref = [H.oracle_event(ctx, e) for e in range(6)]
        rot = ctx.recenter(2, turn)
        after = [ctx.sources(e) for e in range(6)]
        s2, _, _ = ctx.density()
        g = ref[0]["grid"]
        for e in range(6):
            xc, yc, ang = oracle.cm_angle(g, ref[e]["s"], 2)
            assert abs(rot[e, 0] - xc) <= 1e-6 and abs(rot[e, 1] - yc) <= 1e-6
            assert abs(rot[e, 2] - (ang if turn else 0.0)) <= 1e-6
            (xy0, k0, w0), (xy1, k1, w1) = before[e], after[e]
            assert np.array_equal(k0, k1) and np.array_equal(w0, w1)
            phi = rot[e, 2]
            d = xy0 - rot[e, :2]
            want = np.column_stack([np.cos(phi) * d[:, 0] - np.sin(phi) * d[:, 1],
                                    np.sin(phi) * d[:, 0] + np.cos(phi) * d[:, 1]])
            assert np.abs(xy1 - want).max() <= 2.0 ** -25
            xc2, yc2, ang2 = oracle.cm_angle(g, s2[e].astype(np.float64), 2)
            assert abs(xc2) < 1e-3 and abs(yc2) < 1e-3
            if turn:
                assert abs(np.angle(np.exp(2j * ang2))) < 1e-3


def test_optional_averages_match_oracle():
    """smc_average_profiles_extra (operation 3's output_TATB,
    output_rho_binary, output_TA and output_spectator_density,
    MakeDensity.cpp:1290-1385): the sums over accepted events of T_A T_B,
    the binary-collision density, T_A and T_B after each order's turn and
    after the recentring of the reaction-plane pass, and of the spectator
    densities after the turn, against the oracle's restatement on the same
    nucleons (spectators moved with the sources, MCnucl.cpp:1139-1173):
    relative L-inf <= 1e-5 (1e-5 per factor for T_A T_B: 2e-5)."""
    from supermc_amd import Context
    p = H.glauber_params("PbPb", max_events=5)
    bvals = [1.0, 4.0, 7.0, 10.0, 13.0]
    xa, xb, bs = H.cpu_configs(p, len(bvals), 57, bvals)
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=9)
        npa, npb, nc, _ = ctx.collide(allow=(4,))
        accept = (nc > 0).astype(np.uint8)
        accept[1] = 0
        events = [e for e in range(len(bvals)) if accept[e]]
        g = oracle.grid(p.maxx, p.dx, p.maxy, p.dy)
        ref = {}
        for e in events:
            o = H.oracle_event(ctx, e)
            nuc = o["nuc"]
            pa, pb, pf = o["gpu_pairs"]
            xa_, xb_ = nuc["xy_proj"], nuc["xy_targ"]
            ia, ib = o["ia"], o["ib"]
            spec = (xa_[nuc["wound_proj"] < 0], xb_[nuc["wound_targ"] < 0])
            _, _, ex = oracle.rotated_profiles(g, ctx.derived.width, p.alpha, xa_[ia], nuc["fluct_proj"][ia].astype(float),
                                               xb_[ib], nuc["fluct_targ"][ib].astype(float), (xa_[pa] + xb_[pb]) / 2.0,
                                               pf.astype(float), 2, 3, p.finalFactor, reaction_plane=True,
                                               spectators=spec)
            for k, v in ex.items():
                ref[k] = ref.get(k, 0.0) + v
        _, _, got = ctx.average_profiles_extra(2, 3, accept=accept, reaction_plane=True)
    for k in Context.EXTRA_KEYS:
        rk = ("rp_" + k[:-3]) if k.endswith("_rp") else k
        tol = 2e-5 if k.startswith("tatb") else TOL
        for o in range(2):
            assert _rel(got[k][o], ref[rk][o]) <= tol, (k, o, _rel(got[k][o], ref[rk][o]))
