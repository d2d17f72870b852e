"""snow_read_particles_noalloc, the read that allocates no device memory (it stages through a fixed
device array of the library), against snow_read_particles: the same bits, on two calls in a row,
the second with a smaller window of another scene; sizes beyond its limit are rejected."""
import numpy as np
import pytest

from snow_amd import SnowContext, scenes
from snow_amd._capi import SNOW_E_ARG, SNOW_E_CAPACITY, ptr

pytestmark = pytest.mark.gpu

FIELDS = (("x", 3), ("v", 3), ("fe", 9), ("fp", 9))


def read_window(ctx, scene, first, count):
    out = {k: np.full((count, w), np.nan, np.float32) for k, w in FIELDS}
    out["rho0"] = np.full(count, np.nan, np.float32)
    rc = ctx.lib.snow_read_particles_noalloc(ctx.h, scene, first, count, ptr(out["x"]), ptr(out["v"]),
                                             ptr(out["fe"]), ptr(out["fp"]), ptr(out["rho0"]))
    return rc, out


def test_noalloc_read_matches_read_particles():
    sc = [scenes.snowball_drop(), scenes.tiny_block(n_side=20, dims=32)]
    n0, n1 = len(sc[0].particles), len(sc[1].particles)
    assert n0 > 20000 and n1 == 8000
    with SnowContext(sc) as ctx:
        ctx.init()
        for _ in range(3):
            ctx.step(1e-4)
        rc, a = read_window(ctx, 0, 0, n0)
        assert rc == 0, ctx.lib.snow_last_error(ctx.h)
        ref0 = ctx.read_particles(0)
        for k in ref0:
            assert np.array_equal(a[k], ref0[k]), k
        ctx.step(1e-4)
        rc, b = read_window(ctx, 1, 1000, 3000)  # a smaller window of the other scene
        assert rc == 0, ctx.lib.snow_last_error(ctx.h)
        ref1 = ctx.read_particles(1)
        for k in ref1:
            assert np.array_equal(b[k], ref1[k][1000:4000]), k


def test_noalloc_read_limits():
    sc = scenes.tiny_block(n_side=42, dims=48)  # 74088 particles
    with SnowContext(sc) as ctx:
        ctx.init()
        rc, _ = read_window(ctx, 0, 0, 70000)
        assert rc == SNOW_E_CAPACITY
        assert b"65536" in ctx.lib.snow_last_error(ctx.h)
        rc, _ = read_window(ctx, 0, 74000, 100)
        assert rc == SNOW_E_ARG
        rc, w = read_window(ctx, 0, 74088 - 65536, 65536)
        assert rc == 0
        ref = ctx.read_particles(0)
        assert np.array_equal(w["x"], ref["x"][74088 - 65536:])
