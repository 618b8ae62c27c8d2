"""Two attempts per sampling pass (smc_set_attempts_per_pass): the
workgroup of a proton or deuteron projectile on a Woods-Saxon target draws
the targets of two consecutive attempts per pass, each from that attempt's
own counters.  The events must be bit-for-bit those of one attempt per
pass -- records (b, Npart, Ncoll, status, eps_n), nucleons, wound order
and pair lists -- for p+Pb and d+Au (both collision criteria), at the
max_attempts edges, in a mixed batch, and for systems the option does not
touch (Cu+Cu, Pb+Pb)."""
import numpy as np
import pytest

import parity_helpers as H
from test_gpu_systems import SYSTEMS, _params

pytestmark = pytest.mark.gpu


def _events(p, app, seed, nev, systems=None, pattern=None):
    from supermc_amd import Context
    with Context(p) as ctx:
        if systems:
            ctx.set_systems(systems, pattern)
        ctx.set_attempts_per_pass(app)
        assert ctx.attempts_per_pass() == app
        rec = ctx.run(seed=seed, event0=1000, nev=nev)
        nuc = [ctx.nucleons(e) for e in range(nev)]
        pairs = [ctx.pairs(e)[:2] for e in range(nev)]
    return rec, nuc, pairs


def _same(a, b):
    ra, na, pa = a
    rb, nb, pb = b
    assert np.array_equal(ra, rb, equal_nan=True)
    for x, y in zip(na, nb):
        for k in x:
            assert np.array_equal(x[k], y[k], equal_nan=True), k
    for x, y in zip(pa, pb):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("name,crit", [("pPb", 1), ("dAu", 1), ("pPb", 2), ("dAu", 2), ("CuCu", 1)])
def test_two_attempts_per_pass_give_the_same_events(name, crit):
    nev = 512
    p = _params(name, max_events=nev, collision_criterion=crit)
    one = _events(p, 1, 9, nev)
    two = _events(p, 2, 9, nev)
    _same(one, two)
    assert (one[0][:, 57] == 0).all()


@pytest.mark.parametrize("max_attempts", [1, 2, 3])
def test_attempt_limit_edges(max_attempts):
    """With few attempts allowed most p+Pb events fail (status bit 1); the
    failures and the accepted events match one attempt per pass."""
    nev = 256
    p = _params("pPb", max_events=nev, max_attempts=max_attempts)
    one = _events(p, 1, 4, nev)
    two = _events(p, 2, 4, nev)
    _same(one, two)
    st = one[0][:, 57].astype(int)
    assert (st & 1).any() and (st == 0).any()


def test_mixed_batch_and_untouched_systems():
    names = ["pPb", "dAu", "CuCu", "UU"]
    systems = [_params(n, max_events=64, sigma_nn_mb=42.0) for n in names]
    pattern = [0, 1, 2, 3, 0, 1]
    p = _params("UU", max_events=64, sigma_nn_mb=42.0)
    _same(_events(p, 1, 21, 64, systems, pattern), _events(p, 2, 21, 64, systems, pattern))
    q = H.glauber_params("PbPb", max_events=64)
    _same(_events(q, 1, 5, 64), _events(q, 2, 5, 64))


def test_refuses_other_counts():
    from supermc_amd import Context, SmcError
    with Context(_params("pPb")) as ctx:
        assert ctx.attempts_per_pass() == 2
        for n in (0, 3, -1):
            with pytest.raises(SmcError):
                ctx.set_attempts_per_pass(n)


def test_autotune_keeps_the_batch():
    """smc_autotune_attempts relaunches the event kernel on the batch's own
    seed with both settings; the events stay those of the batch."""
    from supermc_amd import Context
    p = _params("pPb", max_events=2048)
    with Context(p) as ctx:
        rec = ctx.run(seed=3, event0=0, nev=2048)
        nuc = [ctx.nucleons(e) for e in (0, 1000, 2047)]
        best, ms = ctx.autotune_attempts(reps=2)
        assert best in (1, 2) and ctx.attempts_per_pass() == best
        assert (ms > 0).all()
        for e, n in zip((0, 1000, 2047), nuc):
            m = ctx.nucleons(e)
            for k in n:
                assert np.array_equal(n[k], m[k], equal_nan=True)
        assert np.array_equal(ctx.moments(), rec, equal_nan=True)
        print(f"p+Pb k_event per 2048 events: one attempt per pass {ms[0]:.3f} ms, two {ms[1]:.3f} ms", flush=True)
