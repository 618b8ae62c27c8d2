"""Pins the FP64 restatement of the differentiable batch (oracle/diff_oracle.py)
before it checks the HIP adjoint:
  Woods-Saxon radial quantile against scipy quad + brentq, its R and a
  derivatives against finite differences; the Gamma quantile's shape
  derivative against finite differences of scipy's gammaincinv; the
  hot-spot mask against the moments of orc_moments; the whole adjoint
  against central finite differences on a small synthetic topology.
"""
import numpy as np
import pytest
from scipy import integrate, optimize, special

import diff_oracle as D
import oracle


@pytest.mark.parametrize("R,a", [(6.62, 0.546), (3.0, 0.5), (6.38, 0.535)])
def test_ws_quantile_matches_quad_and_brent(R, a):
    us = np.array([1e-6, 0.01, 0.2, 0.5, 0.77, 0.99, 1 - 1e-9])
    r, rR, ra = D.ws_finv(us, R, a)
    rmax = R + 2.5
    N = lambda x: integrate.quad(lambda t: t * t / (1 + np.exp((t - R) / a)), 0, x, epsabs=0, epsrel=1e-13)[0]
    Z = N(rmax)
    for u, rv in zip(us, r):
        ref = optimize.brentq(lambda x: N(x) - u * Z, 0, rmax, xtol=1e-14, rtol=1e-14)
        assert rv == pytest.approx(ref, rel=1e-11, abs=1e-12)
    h = 1e-6
    rp, _, _ = D.ws_finv(us, R + h, a)
    rm, _, _ = D.ws_finv(us, R - h, a)
    np.testing.assert_allclose(rR, (rp - rm) / (2 * h), rtol=1e-6, atol=1e-8)
    rp, _, _ = D.ws_finv(us, R, a + h)
    rm, _, _ = D.ws_finv(us, R, a - h)
    np.testing.assert_allclose(ra, (rp - rm) / (2 * h), rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("s", [0.344, 0.112, 0.8, 1.7])
def test_gamma_quantile_shape_derivative(s):
    v = np.array([1e-8, 0.03, 0.3, 0.5, 0.9, 0.999])
    x, dx = D.gamma_quantile(v, np.full_like(v, s))
    np.testing.assert_allclose(special.gammainc(s, x), v, rtol=1e-12)
    h = 1e-6 * s
    fd = (special.gammaincinv(s + h, v) - special.gammaincinv(s - h, v)) / (2 * h)
    np.testing.assert_allclose(dx, fd, rtol=2e-6)
    dP = D.gamma_dP_ds(np.full_like(x, s), x)
    fdP = (special.gammainc(s + h, x) - special.gammainc(s - h, x)) / (2 * h)
    np.testing.assert_allclose(dP, fdP, rtol=1e-6, atol=1e-14)


def test_hotspot_mask_reproduces_orc_moments():
    g = oracle.grid(6.0, 0.1)
    rng = np.random.default_rng(4)
    pts = rng.normal(0, 1.2, (9, 2))
    boxes = np.array([oracle.particle_box(x, y, 0.5) for x, y in pts])
    dens = np.zeros((g.nx, g.ny))
    X = (g.xmin + g.dx * np.arange(g.nx))[:, None]
    Y = (g.ymin + g.dy * np.arange(g.ny))[None, :]
    for x, y in pts:
        dens += np.exp(-((X - x) ** 2 + (Y - y) ** 2) / 0.5)
    ref = oracle.moments(g, dens, boxes, 4, 2, 3)
    m = oracle.hotspot_mask(g, boxes, 4)
    xc, yc = ref["xc"], ref["yc"]
    z = (X - xc) + 1j * (Y - yc)
    r2 = np.abs(z) ** 2
    for n in (2, 3):
        num = (dens * z ** n * np.abs(z) ** (2 - n))[m].sum()
        e = -num / ((dens * r2)[m].sum() + 1e-15)
        assert e.real == pytest.approx(ref["mom_re"][n], abs=1e-13)
        assert e.imag == pytest.approx(ref["mom_im"][n], abs=1e-13)


def _synthetic_event(seed=11, A=24, R=3.2, a=0.5, w0=0.5, b=1.5):
    g = oracle.grid(6.0, 0.1)
    rng = np.random.default_rng(seed)
    uvw = [np.stack([rng.uniform(0, 1, A), rng.uniform(-1, 1, A), rng.uniform(0, 2 * np.pi, A)], 1)
           for _ in range(2)]
    ev0 = D.DiffEvent.__new__(D.DiffEvent)
    ev0.uvw, ev0.b = uvw, b
    pos = ev0.positions(R, a)
    xy0 = [np.round(p[0] / D.KQ) * D.KQ for p in pos]
    order = [oracle.sort_nucleus(xy, w0) for xy in xy0]
    uvw = [u[o] for u, o in zip(uvw, order)]
    xy0 = [xy[o] for xy, o in zip(xy0, order)]
    col = oracle.collide(xy0[0], xy0[1], w0, oracle.lib().orc_dsq(64.0))
    wa = np.full(A, -1)
    wb = np.full(A, -1)
    wa[col["order_a"]] = np.arange(col["npart_a"])
    wb[col["order_b"]] = np.arange(col["npart_b"])
    pairs = np.stack([col["pair_a"], col["pair_b"]], 1)
    K = col["npart_a"] + col["npart_b"] + col["ncoll"]
    gv = rng.uniform(0.02, 0.98, K)
    return D.DiffEvent(g, uvw[0], uvw[1], b, xy0[0], xy0[1], wa, wb, pairs, gv, w0, 1.0)


def test_adjoint_matches_central_differences():
    ev = _synthetic_event()
    assert len(ev.pairs) > 5
    th = np.array([0.5, 3.2, 0.5, 0.14, 0.8])
    cot = np.array([1.0, 1.0, 0.01])
    g = ev.adjoint(th, cot)
    fd = ev.fd_gradient(th, cot, h=1e-4)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())
    # away from theta0 as well (the topology stays frozen)
    th1 = np.array([0.62, 3.0, 0.56, 0.25, 1.3])
    np.testing.assert_allclose(ev.adjoint(th1, cot), ev.fd_gradient(th1, cot, h=1e-4), rtol=1e-5,
                               atol=1e-7 * np.abs(fd).max())
