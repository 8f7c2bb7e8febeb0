"""The headline frame at its full size and sample count (1920 x 1080,
256 spp, mt5, r = 10, guarded samples and the pbrt Film as in the bench
step): the statistics and the Film bit for bit on a band of rows, the
prepass, the variance filter and the Welch filter against the oracle on an
interior window (whose outputs depend only on data inside the window plus
the radius) with every Welch decision of the window dumped and compared, and
size-independent properties on the whole frame."""
import ctypes as C

import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate, set_slots, STAT_NAMES
from stater_amd import hipmem
from stater_amd.workload import FEATURE_SDS, MT5_LAMBDAS

pytestmark = pytest.mark.gpu

W, H, SPP, R = 1920, 1080, 256, 10
FILTER_RTOL = 1e-4
DECISION_MISMATCH_MAX = 1e-4   # fraction of neighbour decisions that may differ (0.01%)


def _window(bs, name, x0, y0, w, h):
    return hipmem.read_plane(hipmem.sub_plane(bs.view(name), x0, y0, w, h))


def _stats_window(bs, x0, y0, w, h):
    out = []
    for s, j in enumerate(set_slots(MT5_LAMBDAS)):
        pre = f"t0-b{j}-film-" if s == 0 else f"t0-b{j}-"
        out.append(np.stack([_window(bs, pre + nm, x0, y0, w, h) for nm in STAT_NAMES]))
    return np.stack(out)


def _set_window(bs, name, x0, y0, w, h):
    return np.stack([_window(bs, (f"t0-b{j}-film-{name}" if s == 0 else f"t0-b{j}-{name}"), x0, y0, w, h)
                     for s, j in enumerate(set_slots(MT5_LAMBDAS))])


def _decision_rows(ctx, bs, p, g, film, y0, y1):
    """Every Welch decision of rows y0 <= y < y1 of the full-frame filter
    (stater_debug_film_decisions over the whole frame), [y1-y0, W, win, win, 3]."""
    win = 2 * R + 1
    row = W * win * win * 3
    buf = hipmem.DeviceBuffer(H * row)
    sa.check(sa.load_library().stater_debug_film_decisions(bs.h, C.byref(p), C.byref(g), C.byref(film), 0,
                                                           C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    out = buf.download(np.uint8, (y1 - y0) * row, offset=y0 * row).reshape(y1 - y0, W, win, win, 3)
    buf.free()
    return out


def test_full_frame(ctx):
    sc, chunk, feats = generate(ctx, 2, 2, W, H, SPP)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    bs.accumulate(chunk.desc(sanitize=True), sa.STATER_ACC_SEQUENTIAL | sa.STATER_ACC_FRESH)
    bs.film_resolve()
    p = sa.default_params(radius=R)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g, film=bs.view("film"))
    ctx.synchronize()

    # whole frame: every pixel took every sample, the outputs are finite and non-negative
    n = hipmem.read_plane(bs.view("n"))[0]
    assert np.all(n == SPP)
    filmD = hipmem.read_plane(bs.view("filmD"))
    film = hipmem.read_plane(bs.view("film"))
    assert np.all(np.isfinite(filmD)) and np.all(filmD >= 0)
    # a weighted mean of film values stays inside the film's range
    assert filmD.min() >= film.min() and filmD.max() <= film.max() * (1 + 1e-6)

    # a band of rows: statistics of the untransformed set and the Film bit for bit
    y0, y1 = 537, 545
    smp = chunk.to_numpy(rows=(y0, y1))
    n_ref, st_ref = oracle.accumulate(smp, MT5_LAMBDAS, 5, sanitize=True)
    st = _stats_window(bs, 0, y0, W, y1 - y0)
    assert np.array_equal(n[y0:y1], n_ref)
    assert np.array_equal(st[0].view(np.uint32), st_ref[0].view(np.uint32))
    film_ref = oracle.film_image(oracle.film_add(smp, sanitize=True))
    assert np.array_equal(film[:, y0:y1].view(np.uint32), film_ref.view(np.uint32))

    # an interior window: the filters against the oracle on the window's own
    # data, compared where the whole neighbourhood lies inside the window
    x0, yw, w, h = 1400, 300, 112, 80
    nw = n[yw:yw + h, x0:x0 + w]
    stw = _stats_window(bs, x0, yw, w, h)
    lnS2, se2, m2C = (_set_window(bs, k, x0, yw, w, h) for k in ("lnS2", "se2", "m2C"))
    jb, l64, s64, c64 = oracle.prepass(nw, stw, p.var_z)
    fin = np.isfinite(l64)
    assert np.max(np.abs(lnS2[fin] - l64[fin])) < 2e-5
    f = feats.to_numpy()
    fl = [(f["albedo"][:, yw:yw + h, x0:x0 + w], FEATURE_SDS[0]),
          (f["normal"][:, yw:yw + h, x0:x0 + w], FEATURE_SDS[1]),
          (f["depth"][:, yw:yw + h, x0:x0 + w], FEATURE_SDS[2])]
    vD, vCD, amb = oracle.denoise_vars(nw, stw[:, 1], lnS2, se2, m2C, R, p.ds_factor, p.var_z, fl,
                                       want_ambiguous=True)
    inner = (slice(None), slice(None), slice(R, h - R), slice(R, w - R))
    gD, gCD = _set_window(bs, "varD", x0, yw, w, h), _set_window(bs, "varCD", x0, yw, w, h)
    for gv, rv in ((gD, vD), (gCD, vCD)):
        ok = (amb == 0)[inner] & np.isfinite(rv[inner])
        assert np.count_nonzero(ok) > 0.9 * ok.size
        rel = np.abs(gv[inner][ok] - rv[inner][ok]) / np.maximum(np.abs(rv[inner][ok]), 1e-30)
        assert rel.max() < FILTER_RTOL
    ambf = amb[inner].sum() / (float(amb[inner].size) * (2 * R + 1) ** 2)
    assert ambf <= DECISION_MISMATCH_MAX, ambf
    ref = oracle.denoise_film(nw, stw[:, 0].copy(), gCD, film[:, yw:yw + h, x0:x0 + w], R, p.ds_factor, p.alpha,
                              feats=fl, want_decisions=True)
    dec = _decision_rows(ctx, bs, p, g, bs.view("film"), yw + R, yw + h - R)[:, x0 + R:x0 + w - R]
    diff = dec != ref["decisions"][R:h - R, R:w - R]
    frac = np.count_nonzero(diff) / diff.size
    assert frac <= DECISION_MISMATCH_MAX, frac
    # every pixel-channel whose decisions all agree is within the tolerance
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    got = filmD[:, yw + R:yw + h - R, x0 + R:x0 + w - R]
    want = ref["filmD"][:, R:h - R, R:w - R]
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-6)
    assert np.count_nonzero(same) > 0.99 * same.size
    assert rel[same].max() < FILTER_RTOL, rel[same].max()

    # the same window against the full-FP64 pipeline on its samples: the
    # end-to-end error of the FP32 path at the headline workload
    smw = chunk.to_numpy(rows=(yw, yw + h))[..., x0:x0 + w]
    ref64 = oracle.pipeline64(smw, MT5_LAMBDAS, R, p.ds_factor, p.var_z, p.alpha, feats=fl)
    win = 2 * R + 1
    decw = np.zeros((h, w, win, win, 3), np.uint8)
    decw[R:h - R, R:w - R] = dec
    gpu = {"n": nw, "lnS2": lnS2, "se2": se2, "varD": gD, "varCD": gCD,
           "filmD": filmD[:, yw:yw + h, x0:x0 + w], "decisions": decw}
    rep = oracle.fp64_error_report(gpu, ref64, p.var_z, R, crop=True)
    print("\n1080p 256 spp window vs FP64:", rep)
    for name in ("varD", "varCD", "filmD"):
        assert rep[name]["decisions_differ"] <= DECISION_MISMATCH_MAX, (name, rep[name])
        assert rep[name]["max_rel"] <= FILTER_RTOL, (name, rep[name])
