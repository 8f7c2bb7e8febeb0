"""The Film fed by stater_accumulate (STATER_BUFSET_FILM) against the
oracle's pbrt Film, bit for bit: one call, several calls with ragged counts,
the FP64-merge accumulation mode, and resolving into an interleaved plane."""
import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate
from stater_amd import hipmem
from stater_amd.workload import MT5_LAMBDAS

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("mode", [sa.STATER_ACC_SEQUENTIAL, sa.STATER_ACC_MERGE64])
def test_film_matches_pbrt_film(ctx, mode):
    W, H = 96, 64
    sc, chunk, feats = generate(ctx, 2, 7, W, H, 48)
    smp = chunk.to_numpy()
    rng = np.random.default_rng(4)
    counts = rng.integers(0, 30, (H, W)).astype(np.int32)
    cbuf = hipmem.DeviceBuffer(W * H * 4)
    cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
    hipmem.write_plane(cplane, counts[None])
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    # two calls: ragged counts over the first 30 samples, then the next 18 for every pixel
    bs.accumulate(chunk.desc(spp=30, sanitize=True, counts=cplane), mode)
    first = chunk.desc(spp=18, sanitize=True)
    first.data = chunk.desc().data + 30 * chunk.desc().sample_stride_bytes
    bs.accumulate(first, mode)
    bs.film_resolve()
    ctx.synchronize()
    ref = oracle.film_add(smp[:30], counts=counts, sanitize=True)
    ref = oracle.film_add(smp[30:48], ref, sanitize=True)
    film = hipmem.read_plane(bs.view("film"))
    assert np.array_equal(_bits(film), _bits(oracle.film_image(ref)))
    assert np.array_equal(hipmem.read_plane(bs.view("t0-b0-film-n"))[0], counts + 18)


def test_resolve_scale_interleaved_and_reset(ctx):
    W, H = 40, 24
    sc, chunk, feats = generate(ctx, 2, 8, W, H, 16)
    bs = sa.BufferSet(ctx, W, H, (), 4, flags=sa.STATER_BUFSET_FILM)
    bs.accumulate(chunk.desc(sanitize=True))
    buf = hipmem.DeviceBuffer(3 * 64 * 4 * H)
    inter = sa.plane(buf.ptr, W, H, 3, 3 * 64 * 4, 4, layout=sa.STATER_INTERLEAVED)
    bs.film_resolve(2.0, inter)
    ctx.synchronize()
    ref = oracle.film_image(oracle.film_add(chunk.to_numpy(), sanitize=True), 2.0)
    got = buf.download(np.float32, 3 * 64 * H).reshape(H, 3 * 64)[:, : 3 * W].reshape(H, W, 3).transpose(2, 0, 1)
    assert np.array_equal(_bits(got), _bits(ref))
    bs.reset()
    bs.film_resolve()
    ctx.synchronize()
    assert not np.any(hipmem.read_plane(bs.view("film")))
    # a set without the flag has no Film: "film" is the set-0 mean and resolving is refused
    plain = sa.BufferSet(ctx, W, H, (), 2)
    assert sa.load_library().stater_film_resolve(plain.h, 1.0, None, None) == sa.STATER_E_STATE


@pytest.mark.parametrize("mode", [sa.STATER_ACC_SEQUENTIAL, sa.STATER_ACC_MERGE64])
def test_frame_in_three_calls_equals_one_call(ctx, mode):
    """A Film fed as 16 + 16 + 16 samples per pixel is bit-identical to the
    one fed 48 in a single call, and to the oracle's pbrt Film (the tiles'
    sums run over the whole render: statpath.cpp:144,176,388)."""
    W, H = 72, 40
    _, chunk, _ = generate(ctx, 2, 12, W, H, 48)
    smp = chunk.to_numpy()
    one = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    three = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    one.accumulate(chunk.desc(sanitize=True), mode)
    base = chunk.desc()
    for k in range(3):
        d = chunk.desc(spp=16, sanitize=True)
        d.data = base.data + 16 * k * base.sample_stride_bytes
        three.accumulate(d, mode)
    one.film_resolve()
    three.film_resolve()
    ctx.synchronize()
    f1 = hipmem.read_plane(one.view("film"))
    f3 = hipmem.read_plane(three.view("film"))
    assert np.array_equal(_bits(f1), _bits(f3))
    assert np.array_equal(_bits(f1), _bits(oracle.film_image(oracle.film_add(smp, sanitize=True))))


@pytest.mark.parametrize("r", [7, 20])
def test_welch_filter_reads_and_writes_interleaved_films(ctx, r):
    """denoise_film on the reference's interleaved CV_32FC3 film, writing an
    interleaved filmD, equals the planar run bit for bit (no conversion pass),
    for the record-based filter (r <= 10) and the banded one (r = 20)."""
    from stater_amd.workload import FEATURE_SDS
    W, H = 88, 52
    _, chunk, feats = generate(ctx, 2, 13, W, H, 32)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5, flags=sa.STATER_BUFSET_FILM)
    bs.accumulate(chunk.desc(sanitize=True))
    bs.film_resolve()
    p = sa.default_params(radius=r)
    g = feats.gbuffers(FEATURE_SDS)
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g, film=bs.view("film"))
    ctx.synchronize()
    want = hipmem.read_plane(bs.view("filmD"))
    pitch = W * 12 + 64
    fin, fout = hipmem.DeviceBuffer(pitch * H), hipmem.DeviceBuffer(pitch * H)
    il_in = sa.plane(fin.ptr, W, H, 3, pitch, 4, layout=sa.STATER_INTERLEAVED)
    il_out = sa.plane(fout.ptr, W, H, 3, pitch, 4, layout=sa.STATER_INTERLEAVED)
    import ctypes as C
    film = bs.view("film")
    sa.check(ctx.lib.stater_convert_plane(ctx.h, C.byref(film), C.byref(il_in), None))
    bs.denoise_film(p, g, film=il_in, film_out=il_out)
    ctx.synchronize()
    got = fout.download(np.uint8, pitch * H).reshape(H, pitch)[:, :W * 12].copy().view(np.float32)
    got = got.reshape(H, W, 3).transpose(2, 0, 1)
    assert np.array_equal(_bits(got), _bits(want))
