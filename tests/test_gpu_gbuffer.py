"""Per-sample G-buffer means as filter features (src/statistics/statpath.cpp:
366-379 feeds one albedo / normal / depth value per camera sample into a
mean-only StatTile, estimator.h:226-232, whose film mean is the G-buffer).

Here each G-buffer is a buffer set with no lambdas and max_moment 1, fed
through stater_accumulate: its t0-b0-film-mean must equal the oracle's FP32
recurrence bit for bit, and the filters fed with those means must match the
oracle fed with the oracle's means."""
import ctypes as C

import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate, read_set_buf, read_stats
from stater_amd import hipmem
from stater_amd.workload import FEATURE_SDS, MT5_LAMBDAS, SampleChunk

pytestmark = pytest.mark.gpu

FILTER_RTOL = 1e-4
DECISION_MISMATCH_MAX = 1e-4   # fraction of neighbour decisions that may differ (0.01%)


def _gbuffer_samples(rng, spp, H, W):
    """[spp,5,H,W] per-sample albedo, normal and depth (depth in channel 0)."""
    base_alb = rng.random((3, H, W), dtype=np.float32)
    alb = np.clip(base_alb[None] + 0.05 * rng.standard_normal((spp, 3, H, W)).astype(np.float32), 0, 1)
    nrm = rng.standard_normal((spp, 3, H, W)).astype(np.float32) + np.array([0, 0, 3], np.float32)[None, :, None, None]
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    dep = (5.0 + 20.0 * rng.random((1, 1, H, W), dtype=np.float32)
           + rng.random((spp, 1, H, W), dtype=np.float32)).astype(np.float32)
    out = []
    for v in (alb, nrm, np.concatenate([dep, np.zeros((spp, 2, H, W), np.float32)], axis=1)):
        s = np.zeros((spp, 5, H, W), np.float32)
        s[:, :3] = v
        out.append(s)
    return out


def test_gbuffer_means_feed_the_filters(ctx):
    W, H, spp, r = 72, 40, 16, 5
    rng = np.random.default_rng(21)
    planes, ref_means, keep = [], [], []
    # albedo, normal and depth sets under the reference's stat types and names
    # (StatAlbedo 6, StatNormal 5, StatDepth 4: t6/t5/t4-b0-film-mean)
    for smp, ch, typ in zip(_gbuffer_samples(rng, spp, H, W), (3, 3, 1), (6, 5, 4)):
        chunk = SampleChunk(W, H, spp)
        chunk.from_numpy(smp)
        gb = sa.BufferSet(ctx, W, H, (), max_moment=1, stat_type=typ)
        gb.accumulate(chunk.desc())
        ctx.synchronize()
        n = hipmem.read_plane(gb.view("n"))[0]
        v = gb.view(f"t{typ}-b0-film-mean")
        assert v.channels == ch
        mean = hipmem.read_plane(v)
        n_ref, st_ref = oracle.accumulate(smp, (), 1)
        assert np.array_equal(n, n_ref)
        assert np.array_equal(mean.view(np.uint32), st_ref[0, 0][:ch].view(np.uint32))  # the mean, bit for bit
        with pytest.raises(sa.StaterError):
            gb.view("t0-b0-film-mean")  # the radiance names are not this set's
        planes.append(v)
        ref_means.append(st_ref[0, 0][:ch])
        keep += [chunk, gb]
    g = sa.gbuffers(planes, list(FEATURE_SDS))

    sc, chunk, _ = generate(ctx, 2, 7, W, H, 32)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    bs.accumulate(chunk.desc())
    p = sa.default_params(radius=r)
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g)
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    feats = [(ref_means[k], FEATURE_SDS[k]) for k in range(3)]
    lnS2, se2, m2C = (read_set_buf(bs, MT5_LAMBDAS, k) for k in ("lnS2", "se2", "m2C"))
    vD, vCD, amb = oracle.denoise_vars(n, st[:, 1], lnS2, se2, m2C, r, p.ds_factor, p.var_z, feats,
                                       want_ambiguous=True)
    gD = read_set_buf(bs, MT5_LAMBDAS, "varD")
    ok = np.isfinite(vD) & (np.abs(vD) < 1e37) & (amb == 0)
    assert np.count_nonzero(ok) > 0.9 * ok.size
    assert np.max(np.abs(gD[ok] - vD[ok]) / np.maximum(np.abs(vD[ok]), 1e-30)) < FILTER_RTOL
    varCD = read_set_buf(bs, MT5_LAMBDAS, "varCD")
    ambf = amb.sum() / (float(amb.size) * (2 * r + 1) ** 2)
    assert ambf <= DECISION_MISMATCH_MAX, ambf
    ref = oracle.denoise_film(n, st[:, 0].copy(), varCD, st[0, 0], r, p.ds_factor, p.alpha, feats=feats,
                              want_decisions=True)
    filmD = hipmem.read_plane(bs.view("filmD"))
    win = 2 * r + 1
    buf = hipmem.DeviceBuffer(W * H * win * win * 3)
    sa.check(sa.load_library().stater_debug_film_decisions(bs.h, C.byref(p), C.byref(g), None, 0,
                                                           C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    dec = buf.download(np.uint8, W * H * win * win * 3).reshape(H, W, win, win, 3)
    diff = dec != ref["decisions"]
    frac = np.count_nonzero(diff) / diff.size
    assert frac <= DECISION_MISMATCH_MAX, frac
    # every pixel-channel whose decisions all agree is within the tolerance
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    rel = np.abs(filmD - ref["filmD"]) / np.maximum(np.abs(ref["filmD"]), 1e-6)
    assert rel[same].max() < FILTER_RTOL, rel[same].max()
