"""End-to-end error of the GPU's FP32 path against the full-FP64 pipeline
(FP64 recurrence with the FP64 Box-Cox, FP64 prepass, FP64 variance and
Welch filters, FP64 pbrt Film) on the same samples, per output buffer:

* the plain method at its full size (128 x 128, 64 spp, r = 10): meanD, varD;
* a whole mt5 frame with features (160 x 96, 64 spp, r = 10, the Film as the
  Welch filter's input): varD, varCD, filmD and the AS map's targetspp.

Bars (north star): <= 0.01 % of the accept/reject decisions differ, and every
pixel-channel whose decisions all agree is within 1e-4 relative of FP64.
The figures are printed (pytest -s)."""
import ctypes as C
import json

import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate, read_set_buf, read_stats
from stater_amd import hipmem
from stater_amd.workload import FEATURE_SDS, MT5_LAMBDAS

pytestmark = pytest.mark.gpu

DECISIONS_MAX = 1e-4
REL_MAX = 1e-4


def _dump(ctx, bs, p, g, film, plain, W, H, r):
    win = 2 * r + 1
    buf = hipmem.DeviceBuffer(W * H * win * win * 3)
    sa.check(sa.load_library().stater_debug_film_decisions(bs.h, C.byref(p), C.byref(g) if g else None,
                                                           C.byref(film) if film else None, 1 if plain else 0,
                                                           C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    return buf.download(np.uint8, W * H * win * win * 3).reshape(H, W, win, win, 3)


def test_plain_method_128_against_fp64(ctx):
    W = H = 128
    r, alpha = 10, 0.005
    _, chunk, _ = generate(ctx, 1, 1, W, H, 64)
    bs = sa.BufferSet(ctx, W, H, (), max_moment=3)
    bs.accumulate(chunk.desc())
    bs.denoise_plain(r, alpha)
    ctx.synchronize()
    smp = chunk.to_numpy()
    n, st = oracle.accumulate(smp, (), 3, fp64=True)
    s2 = np.where(n >= 2, st[0, 1] / np.maximum(n - 1, 1), 0.0)
    ref = oracle.denoise_film(n, st[0:1, 0].copy(), s2[None], st[0, 0].copy(), r, 0.0, alpha, want_varw=True,
                              want_decisions=True)
    dec = _dump(ctx, bs, sa.default_params(radius=r, alpha=alpha), None, None, True, W, H, r)
    diff = dec != ref["decisions"]
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    report = {"decisions_differ": np.count_nonzero(diff) / diff.size}
    for name, gpu, want in (("meanD", hipmem.read_plane(bs.view("t0-b0-film-meanD")), ref["filmD"]),
                            ("varD", hipmem.read_plane(bs.view("t0-b0-film-varD")), ref["varW"])):
        rel = np.abs(gpu[same] - want[same]) / np.maximum(np.abs(want[same]), 1e-30)
        report[name + "_max_rel"] = float(rel.max())
    print("\nplain 128x128 vs FP64:", json.dumps(report))
    assert report["decisions_differ"] <= DECISIONS_MAX
    assert report["meanD_max_rel"] <= REL_MAX and report["varD_max_rel"] <= REL_MAX


def test_mt5_frame_against_fp64(ctx):
    W, H, spp, r = 160, 96, 64, 10
    _, chunk, feats = generate(ctx, 2, 31, W, H, spp)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    bs.accumulate(chunk.desc(sanitize=True), sa.STATER_ACC_SEQUENTIAL | sa.STATER_ACC_FRESH)
    bs.film_resolve()
    p = sa.default_params(radius=r)
    g = feats.gbuffers()
    film = bs.view("film")
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g, film=film)
    bs.calc_target_spp(32, 4096)
    ctx.synchronize()
    f = feats.to_numpy()
    fl = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.pipeline64(chunk.to_numpy(), MT5_LAMBDAS, r, p.ds_factor, p.var_z, p.alpha, feats=fl)
    n, _ = read_stats(bs, MT5_LAMBDAS)
    gpu = {"n": n, "lnS2": read_set_buf(bs, MT5_LAMBDAS, "lnS2"), "se2": read_set_buf(bs, MT5_LAMBDAS, "se2"),
           "varD": read_set_buf(bs, MT5_LAMBDAS, "varD"), "varCD": read_set_buf(bs, MT5_LAMBDAS, "varCD"),
           "filmD": hipmem.read_plane(bs.view("filmD")), "targetspp": hipmem.read_plane(bs.view("targetspp"))[0],
           "decisions": _dump(ctx, bs, p, g, film, False, W, H, r)}
    rep = oracle.fp64_error_report(gpu, ref, p.var_z, r)
    print("\nmt5 160x96 vs FP64:", json.dumps(rep))
    for name in ("varD", "varCD", "filmD"):
        assert rep[name]["decisions_differ"] <= DECISIONS_MAX, (name, rep[name])
        assert rep[name]["max_rel"] <= REL_MAX, (name, rep[name])
    assert rep["targetspp"]["sum_gpu"] == rep["targetspp"]["sum_fp64"] == 32 * W * H


def test_bf16_feature_storage_against_fp64(ctx):
    """The serving variant that stores the filters' feature planes in bfloat16:
    its error against the full-FP64 pipeline (FP32 features there) is stated,
    not held to the FP32 path's bars: rounding an albedo to 8 significant bits
    moves it by up to 2e-3, a tenth of its range sd, so weights change by a
    few percent, the filtered variances with them, and the Welch decisions
    near the threshold follow those."""
    W, H, spp, r = 160, 96, 64, 10
    _, chunk, feats = generate(ctx, 2, 33, W, H, spp)
    ld = feats.ld
    b16 = hipmem.DeviceBuffer(7 * ld * H * 2)
    planes = []
    for k, (src, ch) in enumerate(((feats.albedo, 3), (feats.normal, 3), (feats.depth, 1))):
        off = (0, 3, 6)[k] * ld * H * 2
        dst = sa.plane(b16.ptr + off, W, H, ch, ld * 2, ld * H * 2, sa.STATER_BF16)
        sa.check(ctx.lib.stater_convert_plane(ctx.h, C.byref(src), C.byref(dst), None))
        planes.append(dst)
    g16 = sa.gbuffers(planes, list(FEATURE_SDS))
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    bs.accumulate(chunk.desc(sanitize=True), sa.STATER_ACC_SEQUENTIAL | sa.STATER_ACC_FRESH)
    bs.film_resolve()
    p = sa.default_params(radius=r)
    film = bs.view("film")
    bs.denoise_vars(p, g16)
    bs.denoise_film(p, g16, film=film)
    ctx.synchronize()
    f = feats.to_numpy()
    fl = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.pipeline64(chunk.to_numpy(), MT5_LAMBDAS, r, p.ds_factor, p.var_z, p.alpha, feats=fl)
    n, _ = read_stats(bs, MT5_LAMBDAS)
    gpu = {"n": n, "lnS2": read_set_buf(bs, MT5_LAMBDAS, "lnS2"), "se2": read_set_buf(bs, MT5_LAMBDAS, "se2"),
           "varD": read_set_buf(bs, MT5_LAMBDAS, "varD"), "varCD": read_set_buf(bs, MT5_LAMBDAS, "varCD"),
           "filmD": hipmem.read_plane(bs.view("filmD")), "decisions": _dump(ctx, bs, p, g16, film, False, W, H, r)}
    rep = oracle.fp64_error_report(gpu, ref, p.var_z, r)
    fd = np.asarray(gpu["filmD"], np.float64)
    rel_all = np.abs(fd - ref["filmD"]) / np.maximum(np.abs(ref["filmD"]), 1e-6)
    rep["filmD_all_pixels"] = {"median_rel": float(np.median(rel_all)), "p99_rel": float(np.percentile(rel_all, 99)),
                               "max_rel": float(rel_all.max())}
    print("\nbf16 features vs FP64:", json.dumps(rep))
    # measured: variance decisions 7.8e-5 and Welch decisions 1.1e-4 differ (the weights move varCD by up
    # to 5 %, which the cascaded Welch test then sees); filmD median 7e-5, p99 3.7e-2 relative
    assert rep["varD"]["decisions_differ"] <= 1e-3 and rep["filmD"]["decisions_differ"] <= 1e-3
    assert rep["filmD_all_pixels"]["median_rel"] < 1e-3 and rep["filmD_all_pixels"]["p99_rel"] < 0.1
