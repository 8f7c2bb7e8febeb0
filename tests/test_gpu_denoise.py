"""GPU parity of the denoiser: prepass, variance filter, Welch mean filter,
plain method, mean variance and the adaptive-sampling map, each against the
CPU oracle fed the GPU's own FP32 inputs (so a decision can only differ where
the FP32 statistic lies within rounding of the critical value).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate, read_set_buf, read_stats
from stater_amd import hipmem
from stater_amd.workload import FEATURE_SDS, MT5_LAMBDAS

pytestmark = pytest.mark.gpu

DECISION_MISMATCH_MAX = 1e-4   # fraction of neighbour decisions that may differ (0.01%)
FILTER_RTOL = 1e-4             # filtered outputs vs FP64 oracle


def _decisions(ctx, bs, params, gbufs, plain, W, H, r):
    win = 2 * r + 1
    buf = hipmem.DeviceBuffer(W * H * win * win * 3)
    lib = sa.load_library()
    sa.check(lib.stater_debug_film_decisions(bs.h, C.byref(params), C.byref(gbufs) if gbufs else None, None,
                                             1 if plain else 0, C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    return buf.download(np.uint8, W * H * win * win * 3).reshape(H, W, win, win, 3)


def test_plain_method_128(ctx):
    """The 128x128 plain case end to end: 64 spp, r = 10, alpha = 0.005, uniform weights."""
    W = H = 128
    r, alpha = 10, 0.005
    sc, chunk, _ = generate(ctx, 1, 1, W, H, 64)
    bs = sa.BufferSet(ctx, W, H, (), max_moment=3)
    bs.accumulate(chunk.desc())
    bs.denoise_plain(r, alpha)
    ctx.synchronize()
    n, st = read_stats(bs, ())
    meanD = hipmem.read_plane(bs.view("t0-b0-film-meanD"))
    varD = hipmem.read_plane(bs.view("t0-b0-film-varD"))
    s2 = (st[0, 1] / (n.astype(np.float32) - np.float32(1))).astype(np.float32)
    ref = oracle.denoise_film(n, st[0:1, 0], s2[None], st[0, 0], r, 0.0, alpha, want_varw=True,
                              want_decisions=True)
    dec = _decisions(ctx, bs, sa.default_params(radius=r, alpha=alpha), None, True, W, H, r)
    diff = dec != ref["decisions"]
    frac = np.count_nonzero(diff) / dec.size
    assert frac <= DECISION_MISMATCH_MAX, frac
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    np.testing.assert_allclose(meanD[same], ref["filmD"][same], rtol=FILTER_RTOL)
    np.testing.assert_allclose(varD[same], ref["varW"][same], rtol=FILTER_RTOL)
    # the x4 step edge at x = 64: pairs across it are accepted far less often
    # than pairs on the same side (the remainder involves firefly pixels,
    # whose variance hides the step at 64 spp)
    cross = [dec[:, x, :, 64 - x + r:, :] for x in range(54, 64)]
    same = [dec[:, x, :, :64 - x + r, :] for x in range(54, 64)]
    frac_cross = sum(int(c.sum()) for c in cross) / sum(c.size for c in cross)
    frac_same = sum(int(c.sum()) for c in same) / sum(c.size for c in same)
    assert frac_cross < 0.25 * frac_same, (frac_cross, frac_same)


@pytest.mark.parametrize("W,H,r", [(77, 45, 3), (130, 20, 12)])
def test_plain_method_shapes(ctx, W, H, r):
    """The plain method on ragged films and a radius above 10."""
    alpha = 0.005
    sc, chunk, _ = generate(ctx, 1, 4, W, H, 32)
    bs = sa.BufferSet(ctx, W, H, (), max_moment=3)
    bs.accumulate(chunk.desc())
    bs.denoise_plain(r, alpha)
    ctx.synchronize()
    n, st = read_stats(bs, ())
    meanD = hipmem.read_plane(bs.view("t0-b0-film-meanD"))
    varD = hipmem.read_plane(bs.view("t0-b0-film-varD"))
    s2 = (st[0, 1] / (n.astype(np.float32) - np.float32(1))).astype(np.float32)
    ref = oracle.denoise_film(n, st[0:1, 0], s2[None], st[0, 0], r, 0.0, alpha, want_varw=True,
                              want_decisions=True)
    dec = _decisions(ctx, bs, sa.default_params(radius=r, alpha=alpha), None, True, W, H, r)
    diff = dec != ref["decisions"]
    assert np.count_nonzero(diff) / dec.size <= DECISION_MISMATCH_MAX
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    assert np.count_nonzero(same) > 0.9 * same.size
    np.testing.assert_allclose(meanD[same], ref["filmD"][same], rtol=FILTER_RTOL)
    np.testing.assert_allclose(varD[same], ref["varW"][same], rtol=FILTER_RTOL)


def _mt5_setup(ctx, W=96, H=64, spp=64, seed=2, counts=None):
    """Accumulate spp synthetic samples per pixel (or counts[y, x] of them)."""
    sc, chunk, feats = generate(ctx, 2, seed, W, H, spp)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, max_moment=5)
    cplane = None
    if counts is not None:
        cbuf = hipmem.DeviceBuffer(W * H * 4)
        cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
        hipmem.write_plane(cplane, counts.astype(np.int32)[None])
    bs.accumulate(chunk.desc(counts=cplane))
    ctx.synchronize()
    return sc, chunk, feats, bs


def _ragged_counts(W, H, seed=11):
    """Columns x < 40 hold 0..20 samples (small n, n < 2 included), the rest 64:
    the 64-column tile at x >= 64 sees only n = 64 in its window."""
    rng = np.random.default_rng(seed)
    c = np.full((H, W), 64, np.int32)
    c[:, :40] = rng.integers(0, 21, (H, 40))
    return c


def test_prepass_and_variance_filter(ctx):
    W, H = 96, 64
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H)
    p = sa.default_params(radius=10)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    jb, l64, s64, c64 = oracle.prepass(n, st, p.var_z)
    lnS2 = read_set_buf(bs, MT5_LAMBDAS, "lnS2")
    se2 = read_set_buf(bs, MT5_LAMBDAS, "se2")
    m2C = read_set_buf(bs, MT5_LAMBDAS, "m2C")
    jbg = read_set_buf(bs, MT5_LAMBDAS, "jb")[:, 0]
    fin = np.isfinite(l64)
    assert np.array_equal(fin, np.isfinite(lnS2))
    assert np.max(np.abs(lnS2[fin] - l64[fin])) < 2e-5
    fin = np.isfinite(s64)
    assert np.max(np.abs(se2[fin] - s64[fin]) / np.abs(s64[fin])) < 1e-4
    fin = np.isfinite(c64) & (c64 < 1e37)
    assert np.max(np.abs(m2C[fin] - c64[fin]) / np.maximum(np.abs(c64[fin]), 1e-30)) < 1e-4
    fin = np.isfinite(jb)
    assert np.max(np.abs(jbg[fin] - jb[fin]) / np.maximum(np.abs(jb[fin]), 1e-3)) < 1e-3
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    vD, vCD, amb = oracle.denoise_vars(n, st[:, 1], lnS2, se2, m2C, 10, p.ds_factor, p.var_z, feats_list,
                                       want_ambiguous=True)
    # decisions within FP32 rounding of the threshold are not pinned: they must
    # be rare, and pixels that have one are left out of the value comparison
    tests = float(amb.size) * 440
    assert amb.sum() / tests <= DECISION_MISMATCH_MAX, amb.sum() / tests
    gD = read_set_buf(bs, MT5_LAMBDAS, "varD")
    gCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    ok = np.isfinite(vD) & (np.abs(vD) < 1e37) & (amb == 0)
    assert np.max(np.abs(gD[ok] - vD[ok]) / np.maximum(np.abs(vD[ok]), 1e-30)) < FILTER_RTOL
    ok = np.isfinite(vCD) & (np.abs(vCD) < 1e37) & (amb == 0)
    assert np.max(np.abs(gCD[ok] - vCD[ok]) / np.maximum(np.abs(vCD[ok]), 1e-30)) < FILTER_RTOL


@pytest.mark.parametrize("lambdas,r,ragged,weighted", [
    (MT5_LAMBDAS, 10, True, True),         # pixels with n < 4 keep their own variance; two 3-set groups
    (MT5_LAMBDAS, 12, False, True),        # r > 10: the wide-halo build
    ((0.5,), 3, True, True),               # one group of 2 sets
    ((0.25, 0.5, 3.0), 5, False, False),   # groups of 3 and 1 sets, no features and ds_factor 0
])
def test_variance_filter_cases(ctx, lambdas, r, ragged, weighted):
    W, H = 96, 64
    counts = _ragged_counts(W, H) if ragged else None
    sc, chunk, feats = generate(ctx, 2, 3, W, H, 64)
    bs = sa.BufferSet(ctx, W, H, lambdas, max_moment=5)
    cplane = None
    if counts is not None:
        cbuf = hipmem.DeviceBuffer(W * H * 4)
        cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
        hipmem.write_plane(cplane, counts[None])
    bs.accumulate(chunk.desc(counts=cplane))
    p = sa.default_params(radius=r)
    feats_list = None
    if not weighted:
        p.ds_factor = 0.0
    bs.denoise_vars(p, feats.gbuffers() if weighted else None)
    ctx.synchronize()
    n, st = read_stats(bs, lambdas)
    lnS2, se2, m2C = (read_set_buf(bs, lambdas, k) for k in ("lnS2", "se2", "m2C"))
    if weighted:
        f = feats.to_numpy()
        feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    vD, vCD, amb = oracle.denoise_vars(n, st[:, 1], lnS2, se2, m2C, r, p.ds_factor, p.var_z, feats_list,
                                       want_ambiguous=True)
    assert amb.sum() / (float(amb.size) * (2 * r + 1) ** 2) <= DECISION_MISMATCH_MAX
    gD, gCD = read_set_buf(bs, lambdas, "varD"), read_set_buf(bs, lambdas, "varCD")
    if ragged:  # pixels with fewer than 4 samples: their own s^2 and m2C/(n-1), unfiltered
        few = np.broadcast_to(n < 4, gD.shape)
        own = np.broadcast_to(n >= 2, gD.shape) & few
        assert np.count_nonzero(own) > 0
        assert np.array_equal(gD[few & ~own], np.zeros(np.count_nonzero(few & ~own), np.float32))
    for g, v in ((gD, vD), (gCD, vCD)):
        ok = np.isfinite(v) & (np.abs(v) < 1e37) & (amb == 0)
        assert np.max(np.abs(g[ok] - v[ok]) / np.maximum(np.abs(v[ok]), 1e-30)) < FILTER_RTOL


@pytest.mark.parametrize("mt_mode,ragged", [(2, False), (3, False), (2, True), (3, True)])
def test_welch_filter_mt5(ctx, mt_mode, ragged):
    W, H, r = 96, 64, 10
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, counts=_ragged_counts(W, H) if ragged else None)
    p = sa.default_params(radius=r, mt_mode=mt_mode)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    varCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    jb = read_set_buf(bs, MT5_LAMBDAS, "jb")[:, 0]
    film = st[0, 0]
    mean = st[:, 0].copy()
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.denoise_film(n, mean, varCD, film, r, p.ds_factor, p.alpha, mt_mode=mt_mode, jb=jb,
                              feats=feats_list, want_decisions=True)
    dec = _decisions(ctx, bs, p, g, False, W, H, r)
    diff = dec != ref["decisions"]
    frac = np.count_nonzero(diff) / dec.size
    assert frac <= DECISION_MISMATCH_MAX, frac
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)  # [3,H,W]: every decision agrees
    filmD = hipmem.read_plane(bs.view("filmD"))
    np.testing.assert_allclose(filmD[same], ref["filmD"][same], rtol=FILTER_RTOL)
    discr = read_set_buf(bs, MT5_LAMBDAS, "discr")
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(n >= 1, varCD / n.astype(np.float32), np.float32(0)).astype(np.float32)
    want[~np.isfinite(want)] = 0
    np.testing.assert_array_equal(discr, want)


@pytest.mark.parametrize("r", [10, 20])
def test_welch_filter_window_levels(ctx, r):
    """Tiles of every window level of the decision pass in one film: some
    n - 1 below 17 (x < 40), every n - 1 from 17 on (n = 64), every n - 1
    from 128 on with one n (n = 160 for 100 <= x < 224, which the 64-column
    tile at x >= 128 sees alone: the uniform-n form) and with several
    (n = 150 and 170 in alternate columns from x = 224 on, which the tiles
    from x = 192 on see)."""
    W, H = 320, 48
    counts = _ragged_counts(W, H)
    counts[:, 100:] = 160
    counts[:, 224::2] = 150
    counts[:, 225::2] = 170
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, spp=170, seed=7, counts=counts)
    p = sa.default_params(radius=r, mt_mode=2)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    assert np.array_equal(n, counts)
    varCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.denoise_film(n, st[:, 0].copy(), varCD, st[0, 0], r, p.ds_factor, p.alpha, mt_mode=2,
                              feats=feats_list, want_decisions=True)
    dec = _decisions(ctx, bs, p, g, False, W, H, r)
    diff = dec != ref["decisions"]
    for x0, x1 in ((0, 64), (64, 128), (128, 192), (192, W)):  # per tile column, each level on its own
        part = diff[:, x0:x1]
        assert np.count_nonzero(part) / part.size <= DECISION_MISMATCH_MAX, (x0, np.count_nonzero(part))
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    filmD = hipmem.read_plane(bs.view("filmD"))
    np.testing.assert_allclose(filmD[same], ref["filmD"][same], rtol=FILTER_RTOL)


@pytest.mark.parametrize("cascade,use_levc", [(0, 0), (1, 0)])
def test_welch_filter_variants(ctx, cascade, use_levc):
    """The reference's variant builds: "M" (CASCADE_MODE 0, USE_LEVC 0: the
    Welch test on the raw sample variances) and "MC" (USE_LEVC 0: on the
    denoised variance varD without the LEV correction),
    scripts/alternate-builds/build-asm+m.sh:6 and build-asm+mc.sh:5."""
    W, H, r = 96, 64, 10
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, counts=_ragged_counts(W, H))
    p = sa.default_params(radius=r, cascade=cascade, use_levc=use_levc)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    if cascade:
        var = read_set_buf(bs, MT5_LAMBDAS, "varD")
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            var = np.where(n >= 2, st[:, 1] / (n.astype(np.float32) - np.float32(1)), np.float32(0))
        var = var.astype(np.float32)
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.denoise_film(n, st[:, 0].copy(), var, st[0, 0], r, p.ds_factor, p.alpha, feats=feats_list,
                              want_decisions=True)
    dec = _decisions(ctx, bs, p, g, False, W, H, r)
    diff = dec != ref["decisions"]
    frac = np.count_nonzero(diff) / dec.size
    assert frac <= DECISION_MISMATCH_MAX, frac
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    filmD = hipmem.read_plane(bs.view("filmD"))
    np.testing.assert_allclose(filmD[same], ref["filmD"][same], rtol=FILTER_RTOL)
    # the variants really differ from the default "MCL" on this film
    assert np.count_nonzero(ref["decisions"] != oracle.denoise_film(
        n, st[:, 0].copy(), read_set_buf(bs, MT5_LAMBDAS, "varCD"), st[0, 0], r, p.ds_factor, p.alpha,
        feats=feats_list, want_decisions=True)["decisions"]) > 0


@pytest.mark.parametrize("W,H,r,weighted,mt_mode", [
    (77, 45, 1, True, 2),     # smallest window of the one-test-per-pair path, ragged tiles
    (77, 45, 3, False, 2),    # unweighted (no features, ds_factor 0)
    (130, 37, 7, True, 2),    # three 64-column tiles, the last one 2 wide
    (8, 5, 10, True, 2),      # image smaller than the window and than one tile
    (33, 20, 12, True, 2),    # r > 10: each pair tested once, in two bands of window rows
    (70, 30, 11, True, 2),    # r = 11: a second band of one row
    (90, 50, 20, True, 2),    # r = 20, the reference's default radius
    (200, 70, 20, True, 2),   # r = 20 over 4 x 18 tiles, the last column ragged
    (70, 45, 16, False, 2),   # r = 16 unweighted
    (70, 9, 0, True, 2),      # r = 0: the film itself
    (77, 45, 5, True, 3),     # mode 3 (the set with the smallest jb), ragged tiles
    (33, 20, 12, False, 3),   # mode 3, r > 10, unweighted
])
def test_welch_filter_shapes(ctx, W, H, r, weighted, mt_mode):
    """Window radii, ragged and tiny films, weighted and unweighted, both
    multi-transform modes: the decisions and filmD against the oracle."""
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, spp=32, seed=5)
    p = sa.default_params(radius=r, mt_mode=mt_mode)
    g, feats_list = None, None
    if weighted:
        g = feats.gbuffers()
        f = feats.to_numpy()
        feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    else:
        p.ds_factor = 0.0
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    varCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    jb = read_set_buf(bs, MT5_LAMBDAS, "jb")[:, 0] if mt_mode == 3 else None
    ref = oracle.denoise_film(n, st[:, 0].copy(), varCD, st[0, 0], r, p.ds_factor, p.alpha, mt_mode=mt_mode,
                              jb=jb, feats=feats_list, want_decisions=True)
    dec = _decisions(ctx, bs, p, g, False, W, H, r)
    diff = dec != ref["decisions"]
    assert np.count_nonzero(diff) / dec.size <= DECISION_MISMATCH_MAX
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    assert np.count_nonzero(same) > 0.9 * same.size
    filmD = hipmem.read_plane(bs.view("filmD"))
    np.testing.assert_allclose(filmD[same], ref["filmD"][same], rtol=FILTER_RTOL)
    if r == 0:
        assert np.array_equal(filmD, st[0, 0])


@pytest.mark.parametrize("W,H,r", [(77, 45, 1), (8, 5, 10), (130, 37, 20)])
def test_variance_filter_shapes(ctx, W, H, r):
    """Ragged and tiny films and the largest radius: varD and varCD against the oracle."""
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, spp=32, seed=6)
    p = sa.default_params(radius=r)
    bs.denoise_vars(p, feats.gbuffers())
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    lnS2, se2, m2C = (read_set_buf(bs, MT5_LAMBDAS, k) for k in ("lnS2", "se2", "m2C"))
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    vD, vCD, amb = oracle.denoise_vars(n, st[:, 1], lnS2, se2, m2C, r, p.ds_factor, p.var_z, feats_list,
                                       want_ambiguous=True)
    gD, gCD = read_set_buf(bs, MT5_LAMBDAS, "varD"), read_set_buf(bs, MT5_LAMBDAS, "varCD")
    for g, v in ((gD, vD), (gCD, vCD)):
        ok = np.isfinite(v) & (np.abs(v) < 1e37) & (amb == 0)
        why = [(s, c, int(np.count_nonzero(~np.isfinite(v[s, c]))), int(np.count_nonzero(amb[s, c])))
               for s in range(v.shape[0]) for c in range(3) if not ok[s, c].all()]
        assert np.count_nonzero(ok) > 0.9 * ok.size, why
        assert np.max(np.abs(g[ok] - v[ok]) / np.maximum(np.abs(v[ok]), 1e-30)) < FILTER_RTOL


def test_mean_vars_bitexact(ctx):
    W, H = 50, 30
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, spp=16)
    bs.calc_mean_vars()
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    mv = hipmem.read_plane(bs.view("t0-b0-film-meanVar"))
    assert np.array_equal(mv, oracle.mean_var(n, st[0, 1]))


@pytest.mark.parametrize("budget,cap", [(32, 4096), (32, 70), (8, 40)])
def test_target_spp_exact(ctx, budget, cap):
    W, H = 120, 72
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, spp=32)
    p = sa.default_params(radius=5)
    bs.denoise_vars(p, feats.gbuffers())
    bs.calc_target_spp(budget, cap)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    varD = read_set_buf(bs, MT5_LAMBDAS, "varD")[0]
    tgt = hipmem.read_plane(bs.view("targetspp"))[0]
    ref = oracle.target_spp(n, varD, st[0, 0], budget, cap)
    assert np.array_equal(tgt, ref)
    head = np.maximum(cap - n, 0)
    assert np.all(tgt >= 0) and np.all(tgt <= head)
    if head.sum() > budget * W * H:
        assert int(tgt.sum()) == budget * W * H
    else:
        assert np.array_equal(tgt, head)


def test_target_spp_as_mode_0_uses_the_sample_variance(ctx):
    """AS_MODE 0 ("ASN"): the map is built from each pixel's own sample
    variance m2/(n-1) instead of the denoised varD; ragged counts include
    pixels with n < 2 (variance 0)."""
    W, H, budget, cap = 120, 72, 16, 400
    sc, chunk, feats, bs = _mt5_setup(ctx, W, H, counts=_ragged_counts(W, H))
    bs.calc_target_spp(budget, cap, as_mode=0)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    with np.errstate(divide="ignore", invalid="ignore"):
        s2 = np.where(n >= 2, st[0, 1] / (n.astype(np.float32) - np.float32(1)), np.float32(0)).astype(np.float32)
    tgt = hipmem.read_plane(bs.view("targetspp"))[0]
    assert np.array_equal(tgt, oracle.target_spp(n, s2, st[0, 0], budget, cap))
    assert int(tgt.sum()) == budget * W * H


@pytest.mark.parametrize("budget,cap", [(32, 4096), (4, 64), (0, 4096)])
def test_target_spp_zero_error_pixels(ctx, budget, cap):
    """A flat background (varD = 0) beside noisy pixels one sample below the
    cap: the GPU map equals the oracle's and meets the budget exactly."""
    W, H = 150, 40
    rng = np.random.default_rng(12)
    bs = sa.BufferSet(ctx, W, H, (), 5)
    n = np.full((H, W), 16, np.int32)
    noisy = rng.random((H, W)) < 0.002
    n[noisy] = cap - 1
    varD = np.zeros((3, H, W), np.float32)
    varD[:, noisy] = rng.lognormal(0, 1, (3, int(noisy.sum()))).astype(np.float32)
    mean = np.full((3, H, W), 0.25, np.float32)
    hipmem.write_plane(bs.view("n"), n[None])
    hipmem.write_plane(bs.view("t0-b0-film-varD"), varD)
    hipmem.write_plane(bs.view("t0-b0-film-mean"), mean)
    bs.calc_target_spp(budget, cap)
    ctx.synchronize()
    tgt = hipmem.read_plane(bs.view("targetspp"))[0]
    ref = oracle.target_spp(n, varD, mean, budget, cap)
    assert np.array_equal(tgt, ref)
    assert int(tgt.sum()) == budget * W * H
    if budget:
        assert np.array_equal(tgt[noisy], (cap - n)[noisy])


def _write_stats(bs, lambdas, n, st):
    """Load n [H,W] and statistics [S,7,3,H,W] into a set through its views."""
    from gpu_helpers import STAT_NAMES, set_slots
    hipmem.write_plane(bs.view("n"), n[None].astype(np.int32))
    for s, j in enumerate(set_slots(lambdas)):
        pre = f"t0-b{j}-film-" if s == 0 else f"t0-b{j}-"
        for k, nm in enumerate(STAT_NAMES):
            hipmem.write_plane(bs.view(pre + nm), st[s, k].astype(np.float32))


def test_variance_filter_edge_statistics(ctx):
    """Statistics the sampler rarely makes, written straight into a set: an
    empty own Bonett interval (se2 < 0), a non-finite m2 and pixels with too
    few samples.  The pixel itself always counts with weight 1 and its exact
    s^2; everything else as the oracle on the same prepass values."""
    W, H, r = 40, 24, 4
    rng = np.random.default_rng(17)
    S = len(oracle.set_lambdas(MT5_LAMBDAS))
    n = np.full((H, W), 12, np.int32)
    n[3, 5] = 3                                        # not filtered: keeps its own s^2
    st = np.zeros((S, 7, 3, H, W), np.float32)
    st[:, 0] = rng.random((S, 3, H, W))
    st[:, 1] = rng.uniform(0.5, 2.0, (S, 3, H, W))
    st[:, 2] = rng.normal(0, 0.1, (S, 3, H, W))
    K = rng.uniform(1.5, 4.0, (S, 3, H, W))
    st[:, 3] = K * st[:, 1] ** 2 / 12.0
    st[:, 4:6] = rng.normal(0, 0.05, (S, 2, 3, H, W))
    st[:, 6] = rng.uniform(0.5, 1.0, (S, 3, H, W))
    st[:, 3, :, 10, 10] = 0.2 * st[:, 1, :, 10, 10] ** 2 / 12.0   # K = 0.2 < (n-3)/n: se2 < 0
    st[2, 1, 1, 20, 30] = np.inf                                  # a non-finite m2
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, max_moment=5)
    _write_stats(bs, MT5_LAMBDAS, n, st)
    sc, chunk, feats = generate(ctx, 2, 3, W, H, 4)
    p = sa.default_params(radius=r)
    bs.denoise_vars(p, feats.gbuffers())
    ctx.synchronize()
    lnS2, se2, m2C = (read_set_buf(bs, MT5_LAMBDAS, k) for k in ("lnS2", "se2", "m2C"))
    assert np.all(se2[:, :, 10, 10] < 0)
    assert np.isinf(lnS2[2, 1, 20, 30])
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    vD, vCD, amb = oracle.denoise_vars(n, st[:, 1], lnS2, se2, m2C, r, p.ds_factor, p.var_z, feats_list,
                                       want_ambiguous=True)
    gD, gCD = read_set_buf(bs, MT5_LAMBDAS, "varD"), read_set_buf(bs, MT5_LAMBDAS, "varCD")
    # the pixel with the empty own interval: finite and as the oracle
    assert np.allclose(gD[:, :, 10, 10], vD[:, :, 10, 10], rtol=FILTER_RTOL)
    assert np.allclose(gCD[:, :, 10, 10], vCD[:, :, 10, 10], rtol=FILTER_RTOL)
    # the pixel with the non-finite m2 keeps it (inf, as the oracle); where
    # it is an accepted neighbour the oracle's sums become inf, the filter's
    # take FLT_MAX times its weight (the record's clamp), so those pixels are
    # left out of the comparison below
    assert np.isinf(gD[2, 1, 20, 30]) and np.isinf(vD[2, 1, 20, 30])
    # the pixel with n < 4: its own s^2, unfiltered
    assert np.allclose(gD[:, :, 3, 5], st[:, 1, :, 3, 5] / 2.0, rtol=1e-6)
    ok = np.isfinite(vD) & (np.abs(vD) < 1e37) & (amb == 0)
    assert np.count_nonzero(ok) > 0.8 * ok.size
    assert np.max(np.abs(gD[ok] - vD[ok]) / np.maximum(np.abs(vD[ok]), 1e-30)) < FILTER_RTOL
    ok = np.isfinite(vCD) & (np.abs(vCD) < 1e37) & (amb == 0)
    assert np.max(np.abs(gCD[ok] - vCD[ok]) / np.maximum(np.abs(vCD[ok]), 1e-30)) < FILTER_RTOL


@pytest.mark.parametrize("n_bulk", [20, 200])
def test_welch_filter_edge_statistics(ctx, n_bulk):
    """Welch tests on statistics written straight into a set: zero variances
    (equal means accept, different means reject), a NaN mean (abstains),
    pixels with one or no sample, and a non-finite variance; with n = 200 on
    every pixel that has two samples or more, the uniform-n form decides."""
    W, H, r = 40, 24, 4
    rng = np.random.default_rng(23)
    S = len(oracle.set_lambdas(MT5_LAMBDAS))
    n = np.full((H, W), n_bulk, np.int32)
    n[2, 2], n[2, 3] = 0, 1
    if n_bulk == 20:
        n[5, 7] = 2
    st = np.zeros((S, 7, 3, H, W), np.float32)
    st[:, 0] = rng.normal(1.0, 0.05, (S, 3, H, W))
    st[:, 1] = rng.uniform(0.5, 2.0, (S, 3, H, W))
    st[:, 2] = rng.normal(0, 0.1, (S, 3, H, W))
    st[:, 3] = rng.uniform(1.5, 4.0, (S, 3, H, W)) * st[:, 1] ** 2 / n_bulk
    st[:, 4:6] = rng.normal(0, 0.05, (S, 2, 3, H, W))
    st[:, 6] = rng.uniform(0.5, 1.0, (S, 3, H, W))
    st[:, 1, :, 8:10, 8:12] = 0.0           # zero variance, equal means in pairs of columns
    st[:, 0, :, 8:10, 8:10] = 0.7
    st[:, 0, :, 8:10, 10:12] = 0.9          # and different means next to them
    st[:, 3, :, 8:10, 8:12] = 0.0
    st[1, 0, 2, 15, 20] = np.nan            # a NaN mean abstains
    st[3, 1, 0, 18, 30] = np.inf            # a non-finite variance
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, max_moment=5)
    _write_stats(bs, MT5_LAMBDAS, n, st)
    sc, chunk, feats = generate(ctx, 2, 3, W, H, 4)
    p = sa.default_params(radius=r, mt_mode=2)
    g = feats.gbuffers()
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    varCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    f = feats.to_numpy()
    feats_list = [(f["albedo"], FEATURE_SDS[0]), (f["normal"], FEATURE_SDS[1]), (f["depth"], FEATURE_SDS[2])]
    ref = oracle.denoise_film(n, st[:, 0].copy(), varCD, st[0, 0], r, p.ds_factor, p.alpha, mt_mode=2,
                              feats=feats_list, want_decisions=True)
    dec = _decisions(ctx, bs, p, g, False, W, H, r)
    diff = dec != ref["decisions"]
    assert np.count_nonzero(diff) / dec.size <= DECISION_MISMATCH_MAX
    # the zero-variance block: equal means accept each other, different means reject
    assert dec[8, 8, r, r + 1].all()
    assert not dec[8, 9, r, r + 1].any()
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    assert np.count_nonzero(same) > 0.9 * same.size
    filmD = hipmem.read_plane(bs.view("filmD"))
    fin = same & np.isfinite(ref["filmD"])
    np.testing.assert_allclose(filmD[fin], ref["filmD"][fin], rtol=FILTER_RTOL)
