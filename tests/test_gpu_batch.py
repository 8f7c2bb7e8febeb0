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
        # the Film merges each accumulate call's sums (one pbrt tile merge per
        # call), so the single-image calls take the batch's 16-sample passes
        chunk = SampleChunk(W, H, 16)
        for off in range(0, spp, 16):
            k = min(16, spp - off)
            synth_samples(ctx, sc, chunk, off, feats, spp=k)
            solo.accumulate(chunk.desc(spp=k, sanitize=True))
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
