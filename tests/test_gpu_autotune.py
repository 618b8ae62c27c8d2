"""Tile shapes of the exact-footprint smearing (smc_set_tile_variant) and the
autotuner (smc_autotune): every shape gives bit-identical grids, so the
choice cannot move results; the event records differ at most by the order
of the centre-of-mass partial sums (relative 1e-12)."""
import numpy as np
import pytest

import parity_helpers as H

pytestmark = pytest.mark.gpu


def test_tile_variants_identical_and_autotune_picks_fastest():
    from supermc_amd import Context, SmcError
    p = H.glauber_params("PbPb", max_events=24)
    with Context(p) as ctx:
        ctx.run(seed=3, event0=0, nev=24, fetch=False)
        out = []
        for v in range(4):
            ctx.set_tile_variant(v)
            s, dndy, _ = ctx.density()
            ta, tb = ctx.thickness()
            out.append((s, dndy, ta, tb, ctx.moments()))
        for v in range(1, 4):
            for a, b in zip(out[v][:4], out[0][:4]):
                assert np.array_equal(a, b), v
            np.testing.assert_allclose(out[v][4][:, :54], out[0][4][:, :54], rtol=1e-12, atol=1e-14)
            assert np.array_equal(out[v][4][:, 54:58], out[0][4][:, 54:58])
        best, ms = ctx.autotune(reps=2)
        assert 0 <= best < 4 and (ms > 0).all() and ms[best] == ms.min()
        s2, _, _ = ctx.density()
        assert np.array_equal(s2, out[0][0])
        with pytest.raises(SmcError):
            ctx.set_tile_variant(4)
