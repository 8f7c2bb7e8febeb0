"""stater_calc_target_spp_noalloc (the AS map on caller planes with the
library's static scratch, no allocation) against stater_calc_target_spp on
the same values: bit-identical targets on two calls in a row, the second on
a smaller film, with caller planes of another pitch than the set's; films
above the stated maximum are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import stater_amd as sa
from stater_amd import hipmem

pytestmark = pytest.mark.gpu


def _inputs(rng, W, H, cap):
    n = rng.integers(1, 200, (H, W)).astype(np.int32)
    n[rng.random((H, W)) < 0.01] = cap - 1          # pixels near the cap
    varD = rng.lognormal(-2, 2, (3, H, W)).astype(np.float32)
    varD[:, rng.random((H, W)) < 0.2] = 0.0         # a flat part of the frame
    mean = rng.lognormal(-1, 1, (3, H, W)).astype(np.float32)
    return n, varD, mean


def _tight(arr, dtype):
    """A device plane with a tight row pitch (W * 4 bytes) holding arr [ch,H,W]."""
    ch, H, W = arr.shape
    buf = hipmem.DeviceBuffer(ch * H * W * 4)
    pl = sa.plane(buf.ptr, W, H, ch, W * 4, H * W * 4, dtype)
    hipmem.write_plane(pl, arr)
    return buf, pl


def _both(ctx, W, H, budget, cap, seed):
    rng = np.random.default_rng(seed)
    n, varD, mean = _inputs(rng, W, H, cap)
    bs = sa.BufferSet(ctx, W, H, (), 5)
    hipmem.write_plane(bs.view("n"), n[None])
    hipmem.write_plane(bs.view("t0-b0-film-varD"), varD)
    hipmem.write_plane(bs.view("t0-b0-film-mean"), mean)
    bs.calc_target_spp(budget, cap)
    ctx.synchronize()
    want = hipmem.read_plane(bs.view("targetspp"))[0]
    keep = [_tight(n[None], sa.STATER_I32), _tight(varD, sa.STATER_F32), _tight(mean, sa.STATER_F32),
            _tight(np.full((1, H, W), -7, np.int32), sa.STATER_I32)]
    (_, pn), (_, pv), (_, pm), (_, po) = keep
    lib = sa.load_library()
    sa.check(lib.stater_calc_target_spp_noalloc(ctx.h, W, H, C.byref(pn), C.byref(pv), C.byref(pm), budget, cap,
                                                C.byref(po), None))
    ctx.synchronize()
    got = hipmem.read_plane(po)[0]
    bs.close()
    return got, want


def test_noalloc_equals_the_set_entry_point_twice(ctx):
    got, want = _both(ctx, 700, 390, 32, 4096, 1)
    assert np.array_equal(got, want)
    assert int(got.sum()) == 32 * 700 * 390
    # the second call, on a smaller film, reuses the scratch the first one left
    got2, want2 = _both(ctx, 130, 47, 8, 300, 2)
    assert np.array_equal(got2, want2)
    assert int(got2.sum()) == 8 * 130 * 47


def test_noalloc_refuses_films_beyond_its_scratch(ctx):
    lib = sa.load_library()
    pl = sa.plane(0, 4097, 4096, 1, 4097 * 4, 0, sa.STATER_I32)
    rc = lib.stater_calc_target_spp_noalloc(ctx.h, 4097, 4096, C.byref(pl), C.byref(pl), C.byref(pl), 8, 64,
                                            C.byref(pl), None)
    assert rc == sa.STATER_E_ARG
    assert b"STATER_NOALLOC_MAX_PIXELS" in lib.stater_last_error()
