"""The systems of the mixed-system workload: p+Pb, d+Au (Hulthen deuteron),
Cu+Cu, Au+Au on a 521^2 grid (dx = 0.05 fm) and deformed U+U.

Fed-position chains against the FP64 oracle (same bars as
test_gpu_parity: collisions bit-exact, grids 1e-5 relative L-inf, eps_n
1e-6 absolute), and the device samplers against closed forms / the
oracle's restatement of Nucleus::populate by two-sample KS tests:
  deformed U (beta2 0.28, beta4 0.093, random orientation): transverse
    radii and the event-by-event eps_2 distribution at b = 0
  deuteron: the p-n distance law of the Hulthen CDF (HulthenFunc.cpp:27-41)
    turned by Point3D::rotate
  proton: sits at the nucleus centre
"""
import math

import numpy as np
import pytest
from scipy import optimize, stats

import oracle
import parity_helpers as H

pytestmark = pytest.mark.gpu

Q = 2.0 ** -24

SYSTEMS = {
    # name: (Aproj, Atarg, overrides)
    "pPb": (1, 208, dict(targ_ws_R=6.62, targ_ws_a=0.546, sigma_nn_mb=70.0)),
    "dAu": (2, 197, dict(targ_ws_R=6.38, targ_ws_a=0.535, sigma_nn_mb=42.0)),
    "CuCu": (63, 63, dict(proj_ws_R=4.2, proj_ws_a=0.596, targ_ws_R=4.2, targ_ws_a=0.596, sigma_nn_mb=42.0)),
    "UU": (238, 238, dict(proj_ws_R=6.81, proj_ws_a=0.6, targ_ws_R=6.81, targ_ws_a=0.6, proj_deformed=1,
                          targ_deformed=1, proj_beta2=0.28, proj_beta4=0.093, targ_beta2=0.28, targ_beta4=0.093,
                          sigma_nn_mb=42.0)),
    "AuAu521": (197, 197, dict(proj_ws_R=6.38, proj_ws_a=0.535, targ_ws_R=6.38, targ_ws_a=0.535, sigma_nn_mb=42.0,
                               dx=0.05, dy=0.05)),
}


def _params(name, **kw):
    A, B, over = SYSTEMS[name]
    base = dict(Aproj=A, Atarg=B, cc_fluctuation_model=6, cc_fluctuation_Gamma_theta=1.25, max_events=8)
    base.update(over)
    base.update(kw)
    return H.glauber_params("AuAu", **base)


def _hulthen_inv(u):
    return optimize.brentq(lambda r: oracle.hulthen_cdf(r) - u, 1e-9, 200.0, xtol=1e-12)


def _nucleus(rng, A, R, a, beta2, beta4, deformed, xc):
    if A == 1:
        return np.array([[xc, 0.0]])
    if A == 2:
        d = _hulthen_inv(rng.uniform())
        c, ph = 1 - 2 * rng.uniform(), 2 * math.pi * rng.uniform()
        s = math.sqrt(1 - c * c)
        v = 0.5 * d * np.array([s * math.cos(ph), s * math.sin(ph)])
        return np.array([[xc + v[0], v[1]], [xc - v[0], -v[1]]])
    return rng.populate(A, R, a, 0.4, xc, beta2, beta4, deformed)


def _configs(p, nev, seed, bvals):
    rng = oracle.Rng(seed)
    npr = np.random.default_rng(seed)

    class R2:  # uniform() for the light nuclei, populate() for the heavy ones
        uniform = staticmethod(lambda: npr.uniform())
        populate = staticmethod(rng.populate)

    xa = np.zeros((nev, p.Aproj, 2))
    xb = np.zeros((nev, p.Atarg, 2))
    for e, b in enumerate(bvals):
        xa[e] = _nucleus(R2, p.Aproj, p.proj_ws_R, p.proj_ws_a, p.proj_beta2, p.proj_beta4, p.proj_deformed, b / 2)
        xb[e] = _nucleus(R2, p.Atarg, p.targ_ws_R, p.targ_ws_a, p.targ_beta2, p.targ_beta4, p.targ_deformed, -b / 2)
    return np.round(xa / Q) * Q, np.round(xb / Q) * Q, np.asarray(bvals, float)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_fed_chain(name):
    from supermc_amd import Context
    p = _params(name)
    bvals = [0.0, 1.5, 3.0, 4.5, 6.0, 8.0]
    xa, xb, bs = _configs(p, len(bvals), 17, bvals)
    with Context(p) as ctx:
        ctx.set_positions(xa, xb, bs, fluct_seed=3)
        npa, npb, nc, _ = ctx.collide(allow=(4,))
        s, dndy, _ = ctx.density()
        rec = ctx.moments()
        ta, tb = ctx.thickness()
        checked = 0
        for e in range(len(bvals)):
            o = H.oracle_event(ctx, e)
            assert npa[e] == o["col"]["npart_a"] and npb[e] == o["col"]["npart_b"] and nc[e] == o["col"]["ncoll"]
            if nc[e] == 0:
                continue
            pa, pb, _ = o["gpu_pairs"]
            assert (pa == o["col"]["pair_a"]).all() and (pb == o["col"]["pair_b"]).all()
            assert H.rel_linf(s[e], o["s"]) <= 1e-5
            assert H.rel_linf(ta[e], o["TA"]) <= 1e-5 and H.rel_linf(tb[e], o["TB"]) <= 1e-5
            assert dndy[e] == pytest.approx(o["dndy"], rel=1e-6)
            m = o["mom"]
            for n in range(2, 6):
                assert abs(rec[e, n] - m["mom_re"][n]) <= 1e-6 and abs(rec[e, 10 + n] - m["mom_im"][n]) <= 1e-6
                assert abs(rec[e, 20 + n] - m["momp_re"][n]) <= 1e-6
            checked += 1
        assert checked >= 3


def test_deformed_uranium_sampler():
    from supermc_amd import Context
    p = _params("UU", bmin=0.0, bmax=0.0, Npmin=0, Npmax=1000, max_events=200)
    with Context(p) as ctx:
        ev = ctx.run(seed=11, event0=0, nev=200)
        rg = []
        for e in range(60):
            nuc = ctx.nucleons(e)
            rg += list(np.hypot(nuc["xy_proj"][:, 0], nuc["xy_proj"][:, 1]))
    rng = oracle.Rng(5)
    rc = []
    for e in range(60):
        a = rng.populate(238, 6.81, 0.6, 0.4, 0.0, 0.28, 0.093, 1)
        rc += list(np.hypot(a[:, 0], a[:, 1]))
    assert stats.ks_2samp(rg, rc).pvalue > 1e-3
    # random orientations give a broad eps_2 distribution at b = 0 (tip-tip
    # to body-body); without multiplicity fluctuations its mean is well
    # above a spherical nucleus'
    e2m = []
    for defo in (1, 0):
        q = _params("UU", bmin=0.0, bmax=0.0, Npmin=0, Npmax=1000, max_events=400, cc_fluctuation_model=0,
                    proj_deformed=defo, targ_deformed=defo)
        with Context(q) as ctx:
            ev = ctx.run(seed=11, event0=0, nev=400)
        e2m.append(np.hypot(ev[:, 2], ev[:, 12]).mean())
    assert e2m[0] > 1.3 * e2m[1], e2m


def test_deuteron_sampler_follows_hulthen():
    from supermc_amd import Context
    p = _params("dAu", bmin=0.0, bmax=0.0, Npmin=0, Npmax=1000, max_events=2000)
    with Context(p) as ctx:
        ctx.sample_b(seed=21, event0=0, b=np.zeros(2000))
        dperp = []
        for e in range(2000):
            xy = ctx.nucleons(e)["xy_proj"]
            dperp.append(np.hypot(*(xy[0] - xy[1])))
    # (d/2, 0, 0) turned by Point3D::rotate(cos theta, phi) with cos theta
    # uniform (Nucleus.cpp:342-359, MathBasics.cpp:41-50): the transverse
    # separation is d |cos theta|
    npr = np.random.default_rng(9)
    ref = []
    for _ in range(4000):
        d = _hulthen_inv(npr.uniform())
        c = 1 - 2 * npr.uniform()
        ref.append(d * abs(c))
    assert stats.ks_2samp(dperp, ref).pvalue > 1e-3


def test_proton_sits_at_the_centre():
    from supermc_amd import Context
    p = _params("pPb", bmin=0.0, bmax=8.0, Npmin=2, Npmax=1000, max_events=64)
    with Context(p) as ctx:
        ev = ctx.run(seed=3, event0=0, nev=64)
        assert (ev[:, 57] == 0).all()
        for e in range(0, 64, 8):
            nuc = ctx.nucleons(e)
            b = ev[e, 56]
            assert np.allclose(nuc["xy_proj"][0], [b / 2, 0.0], atol=1e-6)
            assert (nuc["wound_proj"] >= 0).sum() == 1  # accepted events have a collision


def test_mixed_systems_in_one_batch_equal_single_system_runs():
    """One batch cycling p+Pb, d+Au, Cu+Cu, U+U and Au+Au (common sigma_NN,
    grid and fluctuations): every event equals the event a context made
    for its system alone gives for the same seed and index, record for
    record and nucleon for nucleon."""
    from supermc_amd import Context, SmcError

    def prm(name, **kw):
        A, B, over = SYSTEMS["AuAu521" if name == "AuAu" else name]
        over = {k: v for k, v in over.items() if k not in ("dx", "dy")}
        over["sigma_nn_mb"] = 42.0
        return _params(name if name != "AuAu" else "AuAu521", **{**over, **kw, "dx": 0.1, "dy": 0.1})

    names = ["pPb", "dAu", "CuCu", "UU", "AuAu"]
    systems = [prm(n, max_events=30) for n in names]
    pattern = [0, 1, 2, 3, 4, 3, 2]
    seed, ev0, nev = 77, 100, 30
    with Context(prm("UU", max_events=30)) as mix:
        mix.set_systems(systems, pattern)
        rec = mix.run(seed=seed, event0=ev0, nev=nev)
        assert (rec[:, 57] == 0).all()
        sys_of = [pattern[(ev0 + e) % len(pattern)] for e in range(nev)]
        assert [int(r) for r in rec[:, 58]] == sys_of
        assert [mix.event_system(e) for e in range(nev)] == sys_of
        mix_nuc = [mix.nucleons(e) for e in range(nev)]
        for s, p in enumerate(systems):
            with Context(p) as one:
                r1 = one.run(seed=seed, event0=ev0, nev=nev)
                for e in (e for e in range(nev) if sys_of[e] == s):
                    assert np.array_equal(rec[e, :58], r1[e, :58]), (names[s], e)
                    n1 = one.nucleons(e)
                    for side, A in (("proj", p.Aproj), ("targ", p.Atarg)):
                        assert np.array_equal(mix_nuc[e]["xy_" + side][:A], n1["xy_" + side])
                        assert np.isnan(mix_nuc[e]["xy_" + side][A:]).all()
                        assert np.array_equal(mix_nuc[e]["wound_" + side][:A], n1["wound_" + side])
        # the context's own system again
        mix.set_systems([], [])
        with Context(prm("UU", max_events=30)) as uu:
            assert np.array_equal(mix.run(seed=seed, event0=ev0, nev=nev)[:, :58],
                                  uu.run(seed=seed, event0=ev0, nev=nev)[:, :58])
        # systems must share everything but their nuclei and fit the context
        with pytest.raises(SmcError):
            mix.set_systems([prm("CuCu", sigma_nn_mb=64.0)], [0])
    with Context(prm("CuCu")) as small:
        with pytest.raises(SmcError):
            small.set_systems([prm("UU")], [0])
