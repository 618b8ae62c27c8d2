"""The large mixed-system workload on the MFMA contraction: p+Pb, d+Au
(Hulthen deuteron), Cu+Cu, Au+Au and deformed U+U cycled event by event in
one sampled batch (smc_set_systems) on 521^2 cells at 0.05 fm.

Bars (the exact-footprint bars of test_gpu_parity):
  FP32 MFMA (smear_path 1) and BF16x6 MFMA (4) against the FP64 oracle on
    the same events: collisions bit-exact, grids relative L-inf <= 1e-5,
    eps_n and eps'_n absolute <= 1e-6, <r^n> relative <= 1e-5, centre of
    mass <= 1e-6 fm, grid sums relative <= 1e-6;
  BF16x3 (3) against the exact FP32 scatter path (the VALU tiles, smear_path
    0) on the same batch: grids <= 5e-5 and eps <= 1e-4 -- the bounds that
    follow from each term's 3 * 2^-16 relative error (test_gpu_parity).
"""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu

# (name, Aproj, Atarg, nuclei keys) -- the systems of the mixed workload
MIXED = [
    ("p+Pb", 1, 208, dict(targ_ws_R=6.62, targ_ws_a=0.546)),
    ("d+Au", 2, 197, dict(targ_ws_R=6.38, targ_ws_a=0.535)),
    ("Cu+Cu", 63, 63, dict(proj_ws_R=4.2, proj_ws_a=0.596, targ_ws_R=4.2, targ_ws_a=0.596)),
    ("Au+Au", 197, 197, dict(proj_ws_R=6.38, proj_ws_a=0.535, targ_ws_R=6.38, targ_ws_a=0.535)),
    ("U+U", 238, 238, dict(proj_ws_R=6.81, proj_ws_a=0.6, targ_ws_R=6.81, targ_ws_a=0.6, proj_deformed=1,
                           targ_deformed=1, proj_beta2=0.28, proj_beta4=0.093, targ_beta2=0.28, targ_beta4=0.093)),
]


def mixed_params(A, B, nuclei, **kw):
    base = dict(Aproj=A, Atarg=B, sigma_nn_mb=42.0, cc_fluctuation_model=6, cc_fluctuation_Gamma_theta=1.25,
                dx=0.05, dy=0.05, max_pairs=8192)
    base.update(nuclei)
    base.update(kw)
    return H.glauber_params("AuAu", **base)


def _run(path, nev, seed=4051):
    from supermc_amd import Context
    systems = [mixed_params(A, B, k, max_events=nev, smear_path=path) for _, A, B, k in MIXED]
    ctx = Context(mixed_params(238, 238, MIXED[-1][3], max_events=nev, smear_path=path))
    ctx.set_systems(systems, list(range(len(MIXED))))
    ev = ctx.run(seed=seed, event0=0, nev=nev)
    return ctx, ev


@pytest.mark.parametrize("path", [1, 4])
def test_mixed_521_mfma_matches_oracle(path):
    nev = 20
    ctx, ev = _run(path, nev)
    with ctx:
        assert (ev[:, 57] == 0).all()
        s, dndy, _ = ctx.density()
        assert ctx.nx == 521 and ctx.ny == 521
        seen = set()
        for e in range(10):  # two events of each system
            sysi = int(ev[e, 58])
            seen.add(sysi)
            _, A, B, _ = MIXED[sysi]
            o = H.oracle_event(ctx, e, sizes=(A, B))
            pa, pb, _ = o["gpu_pairs"]
            assert (pa == o["col"]["pair_a"]).all() and (pb == o["col"]["pair_b"]).all()
            assert ev[e, 54] == o["col"]["npart_a"] + o["col"]["npart_b"] and ev[e, 55] == o["col"]["ncoll"]
            assert H.rel_linf(s[e], o["s"]) <= 1e-5, (MIXED[sysi][0], H.rel_linf(s[e], o["s"]))
            assert dndy[e] == pytest.approx(o["dndy"], rel=1e-6)
            m = o["mom"]
            for n in range(2, 6):
                for key, off in (("mom_re", 0), ("mom_im", 10), ("momp_re", 20), ("momp_im", 30)):
                    assert abs(ev[e, off + n] - m[key][n]) <= 1e-6, (MIXED[sysi][0], key, n)
            for n in range(10):
                assert ev[e, 40 + n] == pytest.approx(m["rn"][n], rel=1e-5)
            assert abs(ev[e, 50] - m["xc"]) <= 1e-6 and abs(ev[e, 51] - m["yc"]) <= 1e-6
        assert len(seen) == len(MIXED)


def test_mixed_521_split_precision_against_exact_fp32_path():
    nev = 40
    out = {}
    for path in (0, 1, 3, 4):
        ctx, ev = _run(path, nev)
        with ctx:
            s, _, _ = ctx.density()
            out[path] = (s, ev)
    s0, ev0 = out[0]
    for path, gtol, etol in ((1, 1e-5, 2e-6), (4, 1e-5, 2e-6), (3, 5e-5, 1e-4)):
        s, ev = out[path]
        assert np.array_equal(ev[:, 54:58], ev0[:, 54:58])  # same events
        worst = max(H.rel_linf(s[e], s0[e]) for e in range(nev))
        assert worst <= gtol, (path, worst)
        assert np.abs(ev[:, :40] - ev0[:, :40]).max() <= etol, path


def test_million_event_stream_mixed_521():
    """10^6 mixed-system events on 521^2 in one smc_run_stream call through a
    5 x 10^4-event context on the BF16x3 MFMA path -- the large batch of the
    workload in bounded device memory: every event accepted, the system
    pattern in the records, and the first, a middle and the last events
    equal smc_run's for the same event indices (counter-based sampling)."""
    from supermc_amd import Context
    nev_total, cap = 1_000_000, 50_000
    systems = [mixed_params(A, B, k, max_events=cap, smear_path=3) for _, A, B, k in MIXED]
    with Context(mixed_params(238, 238, MIXED[-1][3], max_events=cap, smear_path=3)) as ctx:
        ctx.set_systems(systems, list(range(len(MIXED))))
        rec = ctx.run_stream(seed=99, event0=0, nev_total=nev_total)
        assert rec.shape == (nev_total, 64)
        assert (rec[:, 57] == 0).all()
        assert np.array_equal(rec[:, 58], np.arange(nev_total) % len(MIXED))
        for start in (0, 523_417, nev_total - 500):
            one = ctx.run(seed=99, event0=start, nev=500)
            assert np.array_equal(one[:, :58], rec[start:start + 500, :58]), start
