"""PFM interchange on the device: plane round trips, and the reference's
offline --denoise flow (dump every buffer, read them into a fresh set,
denoise) against the in-memory run, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import stater_amd as sa
from gpu_helpers import generate
from stater_amd import hipmem, pfm
from stater_amd.workload import MT5_LAMBDAS

pytestmark = pytest.mark.gpu


def test_plane_round_trips(ctx, tmp_path):
    W, H = 37, 11
    lib = sa.load_library()
    rng = np.random.default_rng(2)
    a = rng.standard_normal((3, H, W)).astype(np.float32) * 100
    buf = hipmem.DeviceBuffer(4 * 64 * H * 3 * 4)
    planar = sa.plane(buf.ptr, W, H, 3, 64 * 4, 64 * 4 * H)
    hipmem.write_plane(planar, a)
    sa.check(lib.stater_pfm_write(str(tmp_path / "p.pfm").encode(), C.byref(planar), None))
    np.testing.assert_array_equal(pfm.read(tmp_path / "p.pfm"), a)
    # into an interleaved plane (the reference's CV_32FC3 layout) and back
    inter = sa.plane(buf.ptr + 3 * 64 * 4 * H, W, H, 3, 3 * 64 * 4, 4)
    inter.layout = sa.STATER_INTERLEAVED
    sa.check(lib.stater_pfm_read(str(tmp_path / "p.pfm").encode(), C.byref(inter), None))
    sa.check(lib.stater_pfm_write(str(tmp_path / "q.pfm").encode(), C.byref(inter), None))
    assert (tmp_path / "q.pfm").read_bytes() == (tmp_path / "p.pfm").read_bytes()
    # integer planes: stored as floats, rounded to nearest on reading
    cnt = np.array([[[0.0, 1.4, 2.6, 7.5]]], np.float32)
    pfm.write(tmp_path / "n.pfm", cnt)
    ib = hipmem.DeviceBuffer(64 * 4)
    ip = sa.plane(ib.ptr, 4, 1, 1, 64 * 4, 64 * 4, sa.STATER_I32)
    sa.check(lib.stater_pfm_read(str(tmp_path / "n.pfm").encode(), C.byref(ip), None))
    assert hipmem.read_plane(ip).ravel().tolist() == [0, 1, 3, 8]


def _set_views(bs, names):
    return {n: hipmem.read_plane(bs.view(n)) for n in names}


def test_dump_reload_denoise_matches_in_memory(ctx, tmp_path):
    W, H, spp = 96, 64, 64
    sc, chunk, feats = generate(ctx, 2, 4, W, H, spp)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, max_moment=5)
    bs.accumulate(chunk.desc())
    p = sa.default_params(radius=10)
    bs.denoise_vars(p, feats.gbuffers())
    bs.denoise_film(p, feats.gbuffers())
    ctx.synchronize()
    stem = tmp_path / "scene"
    n = pfm.write_set(bs, stem, spp, ".*")
    assert n == 2 + len(MT5_LAMBDAS) * (17 + 14)  # film, filmD; per slot 17 film-level + 14 transformed
    assert pfm.write_set(bs, tmp_path / "sel", spp) == 2  # the reference's default outputregex film.*
    f = feats.to_numpy()
    for name, t in (("albedo", 6), ("normal", 5), ("depth", 4)):
        pfm.write(pfm.file_name(stem, spp, f"t{t}-b0-film-mean"), f[name])

    bs2 = pfm.denoise_dumps(ctx, stem, spp, MT5_LAMBDAS, 5, params=p, out_stem=tmp_path / "out")
    names = ["t0-b0-film-n", "t0-b0-film-mean", "t0-b0-film-m4", "t0-b3-c2X", "t0-b5-m1", "t0-b1-film-varD",
             "t0-b2-varCD", "t0-b0-film-discr", "filmD"]
    a, b = _set_views(bs, names), _set_views(bs2, names)
    for k in names:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    np.testing.assert_array_equal(pfm.read(pfm.file_name(tmp_path / "out", spp, "filmD")), a["filmD"])


def test_gbuffer_set_files_carry_their_stat_type(ctx, tmp_path):
    """A normal G-buffer set (StatNormal, type 5) writes t5-b0-film-mean, the
    file the reference's loader reads as a feature buffer, and a fresh
    type-5 set reads it back bit for bit; a radiance set ignores the file."""
    W, H, spp = 40, 24, 8
    rng = np.random.default_rng(9)
    smp = np.zeros((spp, 5, H, W), np.float32)
    smp[:, :3] = rng.standard_normal((spp, 3, H, W)).astype(np.float32)
    from stater_amd.workload import SampleChunk
    chunk = SampleChunk(W, H, spp)
    chunk.from_numpy(smp)
    gb = sa.BufferSet(ctx, W, H, (), max_moment=1, flags=sa.STATER_BUFSET_NO_DERIVED, stat_type=5)
    gb.accumulate(chunk.desc())
    ctx.synchronize()
    stem = tmp_path / "scene"
    assert pfm.write_set(gb, stem, spp, "t5-b0-film-mean") == 1
    assert pfm.file_name(stem, spp, "t5-b0-film-mean").exists()
    back = sa.BufferSet(ctx, W, H, (), max_moment=1, flags=sa.STATER_BUFSET_NO_DERIVED, stat_type=5)
    assert pfm.read_set(back, stem, spp) == 1
    want = hipmem.read_plane(gb.view("t5-b0-film-mean"))
    assert np.array_equal(hipmem.read_plane(back.view("t5-b0-film-mean")).view(np.uint32), want.view(np.uint32))
    radiance = sa.BufferSet(ctx, W, H, (), max_moment=1, flags=sa.STATER_BUFSET_NO_DERIVED)
    assert pfm.read_set(radiance, stem, spp) == 0
