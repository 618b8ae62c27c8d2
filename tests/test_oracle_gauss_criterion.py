"""Pins of the oracle's Gaussian collision criterion (collision_criterion 2,
MCnucl.cpp:357-385, GaussianNucleonsCal.cpp:59-67, 130-163).

- sig_eff against the closed form of its defining integral:
  int_0^{5w} 2 pi b (1 - exp(-S T_pp(b))) db = 4 pi w^2 [Ein(A) - Ein(A e^-6.25)],
  A = S / (4 pi w^2), Ein(z) = gamma + ln z + E1(z), solved by brentq;
- gauss_hit_prob: P(0) = 1 - e^-A, and its integral over the 5w disk
  reproduces sigma_in at S = sig_eff (adaptive quadrature);
- gauss_tested_pairs against brute-force box overlap (|dx|, |dy| <= 8w),
  plus exactly the reference loop's one extra target per projectile.
"""
import numpy as np
import pytest
from scipy import integrate, optimize, special

import oracle

CASES = [(68.0, np.sqrt(0.1 * 68.0 / np.pi) / np.sqrt(8)), (42.0, 0.5), (64.0, 0.812), (30.0, 0.3)]


def _ein(z):
    return np.euler_gamma + np.log(z) + special.exp1(z)


@pytest.mark.parametrize("sig,w", CASES)
def test_sig_eff_matches_closed_form(sig, w):
    c = 4 * np.pi * w * w

    def f(S):
        A = S / c
        return c * (_ein(A) - _ein(A * np.exp(-6.25))) - 0.1 * sig

    exact = optimize.brentq(f, 1e-6, 1e4, xtol=1e-14)
    got = oracle.sig_eff(sig, w)
    assert abs(got - exact) <= 1e-6 * exact, (got, exact)


@pytest.mark.parametrize("sig,w", CASES)
def test_hit_probability_reproduces_sigma_in(sig, w):
    S = oracle.sig_eff(sig, w)
    assert oracle.gauss_hit_prob(0.0, w, S) == pytest.approx(1 - np.exp(-S / (4 * np.pi * w * w)), rel=1e-15)
    b = np.linspace(0, 8 * w, 50)
    assert (np.diff(oracle.gauss_hit_prob(b, w, S)) < 0).all()
    area, _ = integrate.quad(lambda x: 2 * np.pi * x * oracle.gauss_hit_prob(x, w, S), 0, 5 * w, epsabs=0,
                             epsrel=1e-12, limit=200)
    assert area == pytest.approx(0.1 * sig, rel=1e-6)


def test_tested_pairs_are_box_overlaps_plus_one_extra_target():
    rng = np.random.default_rng(3)
    w = 0.5
    for trial in range(20):
        xa = rng.uniform(-8, 8, (rng.integers(20, 80), 2))
        xb = rng.uniform(-8, 8, (rng.integers(20, 80), 2))
        xa = xa[np.argsort(xa[:, 0], kind="stable")]
        xb = xb[np.argsort(xb[:, 0], kind="stable")]
        got = set(oracle.gauss_tested_pairs(xa, xb, w))
        box = {(i, j) for i in range(len(xa)) for j in range(len(xb))
               if abs(xb[j, 0] - xa[i, 0]) <= 8 * w and abs(xb[j, 1] - xa[i, 1]) <= 8 * w}
        assert box <= got
        # the rest: per projectile, the first target past the box in x after
        # an in-range one, when its y box overlaps
        expect_extra = set()
        for i in range(len(xa)):
            inx = [j for j in range(len(xb)) if abs(xb[j, 0] - xa[i, 0]) <= 8 * w]
            if inx and inx[-1] + 1 < len(xb):
                j = inx[-1] + 1
                if abs(xb[j, 1] - xa[i, 1]) <= 8 * w:
                    expect_extra.add((i, j))
        assert got - box == expect_extra


@pytest.mark.parametrize("sig,lam", [(68.0, 4.14), (42.0, 2.0), (64.0, 7.5)])
def test_shape3_width_and_sigma_gg_closed_form(sig, lam):
    """shape_of_nucleons 3 (GaussianNucleonsCal.cpp:41-46): ratio =
    (gamma + int_lam^{lam+100} e^-t/t dt + ln lam)/lam = Ein(lam)/lam up to
    E1(lam+100) < 1e-45; width = sqrt(0.1 sig / (4 pi lam ratio)),
    sigma_gg = 0.1 sig / ratio.  The device values (GPU test) are held to
    these; here the closed form is checked against direct quadrature."""
    ratio = _ein(lam) / lam
    quad_int, _ = integrate.quad(lambda t: np.exp(-t) / t, lam, lam + 100, epsabs=0, epsrel=1e-13, limit=400)
    assert (np.euler_gamma + quad_int + np.log(lam)) / lam == pytest.approx(ratio, rel=1e-12)
    # the Gaussian test at sigma_gg reproduces sigma_in over the whole plane:
    # int 2 pi b (1 - exp(-sigma_gg T_pp)) db = 4 pi w^2 Ein(A) with A = lam
    w = np.sqrt(0.1 * sig / (4 * np.pi * lam * ratio))
    S = 0.1 * sig / ratio
    assert S / (4 * np.pi * w * w) == pytest.approx(lam, rel=1e-12)
    assert 4 * np.pi * w * w * _ein(lam) == pytest.approx(0.1 * sig, rel=1e-12)
