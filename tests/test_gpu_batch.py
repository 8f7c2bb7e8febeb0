"""Batches of uneven films (multi-stream, tuned tile heights, graph replay):
every film's buffers must be bit-identical to running that film on its own."""
import numpy as np
import pytest

import stater_amd as sa
from gpu_helpers import read_set_buf, read_stats
from stater_amd import hipmem
from stater_amd.workload import FEATURE_SDS, Features, SampleChunk, make_scene, synth_features, synth_samples

pytestmark = pytest.mark.gpu

NONID = (0.01, 0.25, 0.5, 0.75, 3.0)
SHAPES = [(96, 64, 1, 40), (130, 70, 3, 23), (64, 40, 5, 64), (200, 48, 2, 17), (77, 90, 4, 50), (40, 40, 5, 33)]


def _image_lambdas(T):
    return NONID[:T] + (1.0,)


def _outputs(bs, lam):
    n, st = read_stats(bs, lam)
    out = {"n": n, "stats": st, "filmD": hipmem.read_plane(bs.view("filmD"))}
    for name in ("varD", "varCD", "lnS2", "discr"):
        out[name] = read_set_buf(bs, lam, name)
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_batch_matches_single_images(ctx):
    p = sa.default_params(radius=6)
    imgs, keep = [], []
    for i, (W, H, T, spp) in enumerate(SHAPES):
        sc = make_scene(2, 1000 + i, W, H, image_id=i)
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
        bs = sa.BufferSet(ctx, W, H, _image_lambdas(T), 5)
        g = feats.gbuffers(FEATURE_SDS)
        imgs.append((bs, sc, feats.albedo, g, spp))
        keep.append(feats)
    batch = sa.Batch(ctx, imgs, p, chunk_spp=16)
    films = batch.images()
    assert len(films) == len(SHAPES) and [f["sets"] for f in films] == [T + 1 for _, _, T, _ in SHAPES]
    assert len({f["branch"] for f in films}) == 4
    batch.run(use_graph=False)
    ctx.synchronize()
    got = [_outputs(bs, _image_lambdas(T)) for (bs, *_), (_, _, T, _) in zip(imgs, SHAPES)]

    # each image alone through the single-image entry points
    for i, ((bs, sc, alb, g, spp), feats, (W, H, T, _)) in enumerate(zip(imgs, keep, SHAPES)):
        lam = _image_lambdas(T)
        solo = sa.BufferSet(ctx, W, H, lam, 5)
        chunk = SampleChunk(W, H, spp)
        synth_samples(ctx, sc, chunk, 0, feats)
        solo.accumulate(chunk.desc())
        solo.denoise_vars(p, g)
        solo.denoise_film(p, g)
        ctx.synchronize()
        _same(got[i], _outputs(solo, lam))
        solo.close()

    # tuned tile heights, captured graph, replayed twice: same buffers
    batch.tune()
    for _ in range(2):
        batch.run(use_graph=True)
        ctx.synchronize()
        for i, ((bs, *_), (_, _, T, _)) in enumerate(zip(imgs, SHAPES)):
            _same(got[i], _outputs(bs, _image_lambdas(T)))
    print(batch.images())
    batch.close()


def test_batch_of_film_sets_runs_the_renderer_loop(ctx):
    """STATER_BUFSET_FILM sets: guarded samples, the resolved pbrt Film as the
    Welch filter's film; the same buffers as the single-image calls."""
    p = sa.default_params(radius=5)
    shapes = SHAPES[:3]
    imgs, keep = [], []
    for i, (W, H, T, spp) in enumerate(shapes):
        sc = make_scene(2, 2000 + i, W, H, image_id=i)
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
        bs = sa.BufferSet(ctx, W, H, _image_lambdas(T), 5, flags=sa.STATER_BUFSET_FILM)
        imgs.append((bs, sc, feats.albedo, feats.gbuffers(FEATURE_SDS), spp))
        keep.append(feats)
    batch = sa.Batch(ctx, imgs, p, chunk_spp=16)
    batch.run(use_graph=True)
    ctx.synchronize()
    for (bs, sc, alb, g, spp), feats, (W, H, T, _) in zip(imgs, keep, shapes):
        lam = _image_lambdas(T)
        got = _outputs(bs, lam)
        got["film"] = hipmem.read_plane(bs.view("film"))
        solo = sa.BufferSet(ctx, W, H, lam, 5, flags=sa.STATER_BUFSET_FILM)
        # one call with all the samples: the Film's sums run over the render,
        # so the batch's 16-sample passes leave the same Film
        chunk = SampleChunk(W, H, spp)
        synth_samples(ctx, sc, chunk, 0, feats)
        solo.accumulate(chunk.desc(sanitize=True))
        solo.denoise_vars(p, g)
        solo.film_resolve()
        solo.denoise_film(p, g, film=solo.view("film"))
        ctx.synchronize()
        want = _outputs(solo, lam)
        want["film"] = hipmem.read_plane(solo.view("film"))
        _same(got, want)
        # the Welch filter saw the Film, not the radiance mean
        assert not np.array_equal(want["film"], want["stats"][0, 0])
        solo.close()
    batch.close()


def test_batch_takes_caller_sample_chunks(ctx):
    """Films whose samples come from the caller (a renderer's chunks, one with
    ragged per-pixel counts) beside a film generated from its scene: each
    equals the single-image calls on the same samples."""
    p = sa.default_params(radius=5)
    shapes = [(90, 60, 3, 40, True, False), (150, 44, 2, 33, False, True), (64, 64, 5, 24, False, False)]
    imgs, solo_in = [], []
    for i, (W, H, T, spp, film, ragged) in enumerate(shapes):
        sc = make_scene(2, 3000 + i, W, H, image_id=i)
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
        g = feats.gbuffers(FEATURE_SDS)
        bs = sa.BufferSet(ctx, W, H, _image_lambdas(T), 5, flags=sa.STATER_BUFSET_FILM if film else 0)
        counts = None
        if i < 2:  # the caller's chunk
            chunk = SampleChunk(W, H, spp)
            synth_samples(ctx, sc, chunk, 0, feats)
            if ragged:
                cbuf = hipmem.DeviceBuffer(W * H * 4)
                counts = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
                hipmem.write_plane(counts, np.random.default_rng(i).integers(0, spp + 3, (1, H, W)).astype(np.int32))
                solo_in.append((chunk, cbuf))
            desc = chunk.desc(counts=counts, sanitize=film)
            imgs.append((bs, desc, None, g, spp))
            solo_in.append((sc, feats, chunk, desc))
        else:
            imgs.append((bs, sc, feats.albedo, g, spp))
            solo_in.append((sc, feats, None, None))
    batch = sa.Batch(ctx, imgs, p, chunk_spp=16)
    batch.run(use_graph=True)
    ctx.synchronize()
    entries = [e for e in solo_in if len(e) == 4]
    for (bs, _, _, g, spp), (sc, feats, chunk, desc), (W, H, T, _, film, _) in zip(imgs, entries, shapes):
        lam = _image_lambdas(T)
        got = _outputs(bs, lam)
        solo = sa.BufferSet(ctx, W, H, lam, 5, flags=sa.STATER_BUFSET_FILM if film else 0)
        if chunk is None:
            chunk = SampleChunk(W, H, spp)
            synth_samples(ctx, sc, chunk, 0, feats)
            desc = chunk.desc()
        solo.accumulate(desc)
        solo.denoise_vars(p, g)
        if film:
            solo.film_resolve()
        solo.denoise_film(p, g, film=solo.view("film") if film else None)
        ctx.synchronize()
        _same(got, _outputs(solo, lam))
        solo.close()
    batch.close()
