"""ctypes loader for the FP64 CPU oracle (smc_oracle.c) and, when built,
the reference harness oracle/_ref/libsmcref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libsmc_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsmcref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcGrid(C.Structure):
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int)]


class OrcKln(C.Structure):
    _fields_ = [("ecm", C.c_double), ("lambda_", C.c_double), ("ptmin", C.c_double), ("ptmax", C.c_double),
                ("g2hfac", C.c_double), ("lambda_qcd", C.c_double), ("alpha_max", C.c_double)]


class OrcRng(C.Structure):
    _fields_ = [("s", C.c_uint64)]


class OrcEventCfg(C.Structure):
    _fields_ = [("A", C.c_int), ("B", C.c_int), ("R_a", C.c_double), ("a_a", C.c_double), ("R_b", C.c_double),
                ("a_b", C.c_double), ("dmin", C.c_double), ("siginNN", C.c_double), ("w", C.c_double),
                ("alpha", C.c_double), ("gamma_theta", C.c_double), ("fluct", C.c_int), ("bmin", C.c_double),
                ("bmax", C.c_double), ("npmin", C.c_int), ("npmax", C.c_int), ("maxx", C.c_double),
                ("dx", C.c_double), ("from_order", C.c_int), ("to_order", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


_orc = None
_ref = None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        o = C.CDLL(ORACLE_SO)
        sig = {
            "orc_grid_init": (None, [C.POINTER(OrcGrid), C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_sigma_inel": (C.c_double, [C.c_double]),
            "orc_nucleon_width": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "orc_dsq": (C.c_double, [C.c_double]),
            "orc_particle_box": (None, [C.c_double, C.c_double, C.c_double, _dp]),
            "orc_sort_nucleus": (None, [C.c_int, _dp, _dp, C.c_double, _ip]),
            "orc_collide": (C.c_int, [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                      _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, C.c_int]),
            "orc_collide_bruteforce": (C.c_int, [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_double, _ip, _ip]),
            "orc_thickness": (None, [C.POINTER(OrcGrid), C.c_double, C.c_int, _dp, _dp, _dp]),
            "orc_add_density": (None, [C.POINTER(OrcGrid), C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_density_glauber": (C.c_double, [C.POINTER(OrcGrid), C.c_double, C.c_double, C.c_int,
                                                 C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                                 C.c_int, _dp, _dp, _dp, _dp, _dp]),
            "orc_density_trento": (C.c_double, [C.POINTER(OrcGrid), C.c_double, C.c_int, _dp, _dp, _dp,
                                                C.c_int, _dp, _dp, _dp, _dp]),
            "orc_moments": (C.c_int, [C.POINTER(OrcGrid), _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                      _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_kln_defaults": (None, [C.POINTER(OrcKln), C.c_double, C.c_double]),
            "orc_kln_func": (C.c_double, [C.POINTER(OrcKln), C.c_double, C.c_double, _dp]),
            "orc_kln_dndy_gl": (C.c_double, [C.POINTER(OrcKln), C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
            "orc_gauss_legendre01": (None, [C.c_int, _dp, _dp]),
            "orc_six_point": (C.c_double, [C.c_double] * 8),
            "orc_hotspot_mask": (C.c_int, [C.POINTER(OrcGrid), C.c_int, _dp, C.c_int, C.POINTER(C.c_ubyte)]),
            "orc_density_kln_table": (C.c_double, [C.POINTER(OrcGrid), _dp, _dp, _dp, C.c_int, C.c_double, _dp]),
            "orc_p3d_setup": (None, [C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_p3d_search": (C.c_long, [_dp, C.c_int, C.c_double]),
            "orc_p3d_smear": (None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                     _dp]),
            "orc_rng_seed": (None, [C.POINTER(OrcRng), C.c_uint64]),
            "orc_rng_uniform": (C.c_double, [C.POINTER(OrcRng)]),
            "orc_sample_ws": (None, [C.POINTER(OrcRng), C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _dp, _dp, _dp]),
            "orc_populate": (None, [C.POINTER(OrcRng), C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_double, C.c_double, _dp, _dp]),
            "orc_hulthen_cdf": (C.c_double, [C.c_double]),
            "orc_gamma": (C.c_double, [C.POINTER(OrcRng), C.c_double, C.c_double]),
            "orc_rotate": (None, [C.c_double, C.c_double, _dp, _dp, _dp]),
            "orc_ylm": (C.c_double, [C.c_int, C.c_double]),
            "orc_run_events": (C.c_int, [C.POINTER(OrcEventCfg), C.c_uint64, C.c_int, _dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(o, n)
            f.restype = r
            f.argtypes = a
        _orc = o
    return _orc


def ref():
    """The reference-harness library, or None when it is not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        r = C.CDLL(REF_SO)
        sig = {
            "ref_sigma_inel": (C.c_double, [C.c_double]),
            "ref_kln_dndy": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
            "ref_kln_func": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]),
            "ref_six_point": (C.c_double, [C.c_double] * 8),
            "ref_particle_box": (None, [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _dp]),
            "ref_overunion": (None, [_dp, _dp, _dp]),
            "ref_rotate": (None, [C.c_double, C.c_double, _dp]),
            "ref_hulthen_cdf": (C.c_double, [C.c_double]),
            "ref_hulthen_invcdf": (C.c_double, [C.c_double]),
            "ref_param_value": (C.c_double, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.c_char_p, _ip]),
        }
        for n, (rt, a) in sig.items():
            f = getattr(r, n)
            f.restype = rt
            f.argtypes = a
        _ref = r
    return _ref


def _p(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def grid(maxx=13.0, dx=0.1, maxy=None, dy=None):
    g = OrcGrid()
    lib().orc_grid_init(C.byref(g), maxx, maxx if maxy is None else maxy, dx, dx if dy is None else dy)
    return g


def particle_box(x, y, w):
    out = np.zeros(4)
    lib().orc_particle_box(x, y, w, _p(out))
    return out


def sort_nucleus(xy, w):
    xy = np.ascontiguousarray(xy, np.float64)
    x, y = np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1])
    perm = np.zeros(len(x), np.int32)
    lib().orc_sort_nucleus(len(x), _p(x), _p(y), w, _pi(perm))
    return perm


def collide(xy_a, xy_b, w, dsq):
    """Reference sweep; inputs must already be in sorted order."""
    xa, ya = np.ascontiguousarray(xy_a[:, 0]), np.ascontiguousarray(xy_a[:, 1])
    xb, yb = np.ascontiguousarray(xy_b[:, 0]), np.ascontiguousarray(xy_b[:, 1])
    na, nb = len(xa), len(xb)
    wa, wb = np.zeros(na, np.int32), np.zeros(nb, np.int32)
    oa, ob = np.zeros(na, np.int32), np.zeros(nb, np.int32)
    npa, npb = C.c_int(), C.c_int()
    mp = max(1, na * nb)
    pa, pb = np.zeros(mp, np.int32), np.zeros(mp, np.int32)
    nc = lib().orc_collide(na, _p(xa), _p(ya), nb, _p(xb), _p(yb), w, dsq, _pi(wa), _pi(wb), _pi(oa), _pi(ob),
                           C.byref(npa), C.byref(npb), _pi(pa), _pi(pb), mp)
    return dict(ncoll=nc, wound_a=wa, wound_b=wb, order_a=oa[:npa.value], order_b=ob[:npb.value],
                npart_a=npa.value, npart_b=npb.value, pair_a=pa[:nc], pair_b=pb[:nc])


def sig_eff(sigin_mb, w):
    """sigma_gg [fm^2] of the Gaussian collision test: Newton on
    sigma_in = int_0^{5w} 2 pi b db (1 - exp(-sigma_gg T_pp(b))),
    T_pp(b) = exp(-b^2/4w^2) / (4 pi w^2), with a 38-node Gauss-Legendre rule,
    started at 10 and stopped once a step is at most 1e-4
    (GaussianNucleonsCal.cpp:130-163)."""
    xg, wg = np.polynomial.legendre.leggauss(38)
    xg, wg = (xg + 1.0) / 2.0, wg / 2.0
    sigin, bmax, w4 = sigin_mb * 0.1, 5.0 * w, 4.0 * w * w
    s = 10.0
    while True:
        s0 = s
        b, db = xg * bmax, wg * bmax
        tpp = np.exp(-b * b / w4) / (np.pi * w4)
        total = float(np.sum(2 * np.pi * b * db * (1.0 - np.exp(-s * tpp))))
        dn = float(np.sum(2 * np.pi * b * db * tpp * np.exp(-s * tpp)))
        s -= (total - sigin) / dn
        if abs(s - s0) <= 1e-4:
            return s


def gauss_hit_prob(b, w, sigma_gg):
    """Collision probability of the Gaussian test at impact parameter b
    (GaussianNucleonsCal.cpp:59-67)."""
    b = np.asarray(b, np.float64)
    return 1.0 - np.exp(-sigma_gg * np.exp(-b * b / (4 * w * w)) / (4 * np.pi * w * w))


def gauss_tested_pairs(xy_a, xy_b, w):
    """Pairs (i, j) the reference's sweep hands to hit() under the Gaussian
    criterion (MCnucl.cpp:244-289): nucleons sorted by x, boxes of side 8w
    (Particle.cpp:20).  For each projectile the targets from the first whose
    box reaches the projectile's, while the previous target's box does --
    so one target past the box is tested after an in-range one -- and of
    those the ones whose boxes overlap in y."""
    h = 4.0 * w
    out = []
    start, nb = 0, len(xy_b)
    last = None  # the box most recently read, as the reference's targBox
    for i, (xp, yp) in enumerate(xy_a):
        pxl, pxr, pyl, pyr = xp - h, xp + h, yp - h, yp + h
        while start < nb:
            last = xy_b[start]
            if last[0] + h >= pxl:
                break
            start += 1
        j = start
        while j < nb and last is not None and pxr >= last[0] - h:
            xt, yt = xy_b[j]
            last = xy_b[j]
            if pyl <= yt + h and pyr >= yt - h:
                out.append((i, j))
            j += 1
    return out


def collide_bruteforce(xy_a, xy_b, dsq):
    xa, ya = np.ascontiguousarray(xy_a[:, 0]), np.ascontiguousarray(xy_a[:, 1])
    xb, yb = np.ascontiguousarray(xy_b[:, 0]), np.ascontiguousarray(xy_b[:, 1])
    wa, wb = np.zeros(len(xa), np.int32), np.zeros(len(xb), np.int32)
    nc = lib().orc_collide_bruteforce(len(xa), _p(xa), _p(ya), len(xb), _p(xb), _p(yb), dsq, _pi(wa), _pi(wb))
    return nc, wa, wb


def thickness(g, w, xy):
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    x, y = np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1])
    T = np.zeros((g.nx, g.ny))
    lib().orc_thickness(C.byref(g), w, len(x), _p(x), _p(y), _p(T))
    return T


def uli_weights(pair_a, pair_b, na, nb):
    """MCnucl.cpp:345-348 (which_mc_model 5, sub_model 2): each pair's
    additional_weight += 1/n for its projectile and its target nucleon, n
    that nucleon's number of collisions, in the reference's int arithmetic
    (Particle.h:64 returns an int), so a nucleon adds 1 exactly when it
    collided once."""
    pair_a = np.asarray(pair_a, np.int64)
    pair_b = np.asarray(pair_b, np.int64)
    na_coll = np.bincount(pair_a, minlength=na)
    nb_coll = np.bincount(pair_b, minlength=nb)
    return (1 // na_coll[pair_a] + 1 // nb_coll[pair_b]).astype(np.float64)


def density_glauber(g, w, alpha, xy_a, f_a, xy_b, f_b, xy_c, f_c, sub_model=1, aw=None):
    """MCnucl::setDensity, which_mc_model 5 (MCnucl.cpp:689-760): sub_model 1
    or 2, aw the pairs' additional weights (uli_weights) or None for 0."""
    def cols(xy):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        return len(xy), np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1])
    na, xa, ya = cols(xy_a)
    nb, xb, yb = cols(xy_b)
    nc, xc, yc = cols(xy_c)
    fa = np.ascontiguousarray(f_a, np.float64)
    fb = np.ascontiguousarray(f_b, np.float64)
    fc = np.ascontiguousarray(f_c, np.float64)
    awp = None if aw is None else np.ascontiguousarray(aw, np.float64)
    s = np.zeros((g.nx, g.ny))
    dndy = lib().orc_density_glauber(C.byref(g), w, alpha, int(sub_model), na, _p(xa), _p(ya), _p(fa), nb, _p(xb),
                                     _p(yb), _p(fb), nc, _p(xc), _p(yc), _p(fc), None if awp is None else _p(awp), _p(s))
    return s, dndy


def cm_angle(g, s, n):
    """GlueDensity::calcCMAngle (GlueDensity.cpp:87-144): centre of mass with
    weights rho dx dy over the full grid, then the order-n angle
    -atan2(-Im, -Re)/n of sum r^n e^{i n theta} rho about it (atan2, cos, sin
    and pow as the reference writes them)."""
    x = g.xmin + np.arange(g.nx)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    y = g.ymin + np.arange(g.ny)[None, :] * g.dy + 0.0 * np.arange(g.nx)[:, None]
    wei = s * g.dx * g.dy
    weight = wei.sum()
    xc, yc = (x * wei).sum() / weight, (y * wei).sum() / weight
    xr, yr = x - xc, y - yc
    th = np.arctan2(yr, xr)
    rwei = np.power(np.sqrt(xr * xr + yr * yr), n)
    re = (rwei * np.cos(n * th) * s).sum()
    im = (rwei * np.sin(n * th) * s).sum()
    return xc, yc, -np.arctan2(-im, -re) / n


def rotated_profiles(g, w, alpha, xy_a, f_a, xy_b, f_b, xy_c, f_c, from_order, to_order, ff, reaction_plane=False,
                     spectators=None):
    """One event of MakeDensity::generate_profile_average (operation 3,
    MakeDensity.cpp:1266-1345, Glauber, pT-integrated): for each order in
    turn, optionally recentre (MCnucl::recenterGrid, MCnucl.cpp:1119-1145)
    and record ff * rho, then recentre and turn the sources by the order-n
    angle (MCnucl::rotateGrid, MCnucl.cpp:1147-1174, Point3D::rotate with
    theta = 0) and record ff * rho.  The sources stay moved between orders.
    Returns (profiles [norder][nx][ny], recentred profiles or None).

    With spectators = (xy_A, xy_B) -- the nucleons without a collision,
    moved and turned with the sources (MCnucl.cpp:1139-1143, 1168-1173) --
    also returns a third item: the optional averages of MakeDensity.cpp:
    1290-1385 per pass, a dict of lists over orders: T_A T_B, rho_binary
    (MCnucl::calculate_rho_binary: the normalised Gaussians of the pairs,
    the thickness footprint), T_A and T_B after each pass ("rp_" keys for
    the recentred pass), and the spectator densities A, B after the turn
    only."""
    pts = [np.array(xy_a, np.float64).reshape(-1, 2), np.array(xy_b, np.float64).reshape(-1, 2),
           np.array(xy_c, np.float64).reshape(-1, 2)]
    if spectators is not None:
        pts += [np.array(spectators[0], np.float64).reshape(-1, 2), np.array(spectators[1], np.float64).reshape(-1, 2)]
    extra = {k: [] for k in ("tatb", "rho_binary", "ta", "tb", "spec_a", "spec_b",
                             "rp_tatb", "rp_rho_binary", "rp_ta", "rp_tb")}

    def dens():
        return density_glauber(g, w, alpha, pts[0], f_a, pts[1], f_b, pts[2], f_c)[0]

    def record(prefix, with_spec):
        if spectators is None:
            return
        ta, tb = thickness(g, w, pts[0]), thickness(g, w, pts[1])
        extra[prefix + "tatb"].append(ta * tb)
        extra[prefix + "rho_binary"].append(thickness(g, w, pts[2]))
        extra[prefix + "ta"].append(ta)
        extra[prefix + "tb"].append(tb)
        if with_spec:
            extra["spec_a"].append(thickness(g, w, pts[3]))
            extra["spec_b"].append(thickness(g, w, pts[4]))

    out, rp = [], []
    s = dens()
    for n in range(from_order, to_order + 1):
        if reaction_plane:
            xc, yc, _ = cm_angle(g, s, n)
            for q in pts:
                q[:, 0] -= xc
                q[:, 1] -= yc
            s = dens()
            rp.append(ff * s)
            record("rp_", False)
        xc, yc, ang = cm_angle(g, s, n)
        cp, sp = np.cos(ang), np.sin(ang)
        for q in pts:
            x0, y0 = q[:, 0] - xc, q[:, 1] - yc
            q[:, 0] = 1.0 * cp * x0 - sp * y0
            q[:, 1] = 1.0 * sp * x0 + cp * y0
        s = dens()
        out.append(ff * s)
        record("", True)
    res = (np.array(out), (np.array(rp) if reaction_plane else None))
    if spectators is None:
        return res
    return res + ({k: np.array(v) for k, v in extra.items() if v},)


def density_trento(g, w, xy_a, f_a, xy_b, f_b):
    def cols(xy):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        return len(xy), np.ascontiguousarray(xy[:, 0]), np.ascontiguousarray(xy[:, 1])
    na, xa, ya = cols(xy_a)
    nb, xb, yb = cols(xy_b)
    fa = np.ascontiguousarray(f_a, np.float64)
    fb = np.ascontiguousarray(f_b, np.float64)
    s = np.zeros((g.nx, g.ny))
    dndy = lib().orc_density_trento(C.byref(g), w, na, _p(xa), _p(ya), _p(fa), nb, _p(xb), _p(yb), _p(fb), _p(s))
    return s, dndy


def moments(g, density, boxes, ncoll, from_order, to_order):
    density = np.ascontiguousarray(density, np.float64)
    boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 4)
    out = {k: np.zeros(10) for k in ("mom_re", "mom_im", "momp_re", "momp_im", "rn")}
    cm = np.zeros(3)
    rc = lib().orc_moments(C.byref(g), _p(density), len(boxes), _p(boxes), ncoll, from_order, to_order,
                           _p(out["mom_re"]), _p(out["mom_im"]), _p(out["momp_re"]), _p(out["momp_im"]),
                           _p(out["rn"]), _p(cm))
    out["xc"], out["yc"], out["total"] = cm
    out["rc"] = rc
    return out


def hotspot_mask(g, boxes, ncoll):
    boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 4)
    m = np.zeros((g.nx, g.ny), np.uint8)
    lib().orc_hotspot_mask(C.byref(g), len(boxes), _p(boxes), ncoll, m.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return m.astype(bool)


def kln(ecm=2760.0, lam=0.138):
    k = OrcKln()
    lib().orc_kln_defaults(C.byref(k), ecm, lam)
    return k


def kln_func(k, ta, tb, u):
    u = np.ascontiguousarray(u, np.float64)
    return lib().orc_kln_func(C.byref(k), ta, tb, _p(u))


def kln_dndy_gl(k, ta, tb, n_pt, n_kt, n_phi):
    return lib().orc_kln_dndy_gl(C.byref(k), ta, tb, n_pt, n_kt, n_phi)


def gauss_legendre01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss_legendre01(n, _p(x), _p(w))
    return x, w


def six_point(*a):
    return lib().orc_six_point(*[float(v) for v in a])


def density_kln_table(g, ta, tb, table, dT):
    """MCnucl.cpp:654-687 over the grid g; returns (s, dndy), dndy < 0 when
    a cell leaves the table."""
    ta = np.ascontiguousarray(ta, np.float64)
    tb = np.ascontiguousarray(tb, np.float64)
    table = np.ascontiguousarray(table, np.float64)
    s = np.zeros_like(ta)
    dndy = lib().orc_density_kln_table(C.byref(g), _p(ta), _p(tb), _p(table), table.shape[0], dT, _p(s))
    return s, dndy


def p3d_setup(ecm, n=1000):
    """profile_3d.cpp:24-27, 133-155: dict(y_beam, sigma_inel (fm^2), eta,
    proj, targ, proj_max, targ_max)."""
    yb, sig = C.c_double(), C.c_double()
    eta, proj, targ, mx = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(2)
    lib().orc_p3d_setup(ecm, n, C.byref(yb), C.byref(sig), _p(eta), _p(proj), _p(targ), _p(mx))
    return dict(y_beam=yb.value, sigma_inel=sig.value, eta=eta, proj=proj, targ=targ, proj_max=mx[0],
                targ_max=mx[1])


def p3d_search(A, v):
    A = np.ascontiguousarray(A, np.float64)
    return lib().orc_p3d_search(_p(A), len(A), float(v))


def p3d_smear(nx, ny, neta, dx, dy, deta, src):
    """profile_3d.cpp:302-400 in FP64: rho [neta][nx][ny] from source rows
    (x, y, eta, sigma_x, sigma_y, sigma_eta)."""
    src = np.ascontiguousarray(np.asarray(src, np.float64).reshape(-1, 6))
    rho = np.zeros((neta, nx, ny))
    lib().orc_p3d_smear(nx, ny, neta, dx, dy, deta, len(src), _p(src), _p(rho))
    return rho


class Rng:
    def __init__(self, seed):
        self.r = OrcRng()
        lib().orc_rng_seed(C.byref(self.r), seed)

    def uniform(self):
        return lib().orc_rng_uniform(C.byref(self.r))

    def gamma(self, shape, scale):
        return lib().orc_gamma(C.byref(self.r), shape, scale)

    def populate(self, A, R, a, dmin, xcenter=0.0, beta2=0.0, beta4=0.0, deformed=0):
        x, y = np.zeros(A), np.zeros(A)
        lib().orc_populate(C.byref(self.r), A, R, a, beta2, beta4, deformed, dmin, xcenter, _p(x), _p(y))
        return np.stack([x, y], axis=1)


def hulthen_cdf(r):
    return lib().orc_hulthen_cdf(r)


def rotate(cth, phi, xyz):
    x, y, z = (C.c_double(v) for v in xyz)
    lib().orc_rotate(cth, phi, C.byref(x), C.byref(y), C.byref(z))
    return np.array([x.value, y.value, z.value])


def run_events(cfg, seed, nev, want_eps=False):
    """MakeDensity.cpp:2143-2224 event loop; with want_eps, per accepted event
    (Re eps_2, Im eps_2 (r^2 weight), b, Npart, Ncoll)."""
    eps = np.zeros(5 * nev) if want_eps else None
    n = lib().orc_run_events(C.byref(cfg), seed, nev, _p(eps) if eps is not None else None)
    return n, (eps.reshape(nev, 5) if eps is not None else None)
