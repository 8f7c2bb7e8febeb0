"""The serving runtime: a job's buffers are bit-identical alone, co-scheduled
with other jobs, and on the single-image path (accumulate + filters, or the
adaptive-sampling rounds); admission keeps the device memory within budget in
arrival order; bf16 feature storage and the plane conversion at the boundary."""
import numpy as np
import pytest

import stater_amd as sa
from stater_amd import hipmem
from stater_amd.progressive import run_rounds
from stater_amd.serve import Server, job_lambdas
from stater_amd.workload import FEATURE_SDS, Features, SampleChunk, make_scene, synth_features, synth_samples

pytestmark = pytest.mark.gpu

MAXW, MAXH = 256, 192
OUT = ("n", "filmD", "t0-b0-film-varD", "t0-b0-film-m2", "t0-b1-m4", "targetspp")


def _server(ctx, slots, budget=8 << 30, chunk=16):
    return Server(ctx, max_bytes=budget, slots=slots, chunk_spp=chunk, max_width=MAXW, max_height=MAXH,
                  params=sa.default_params(radius=6))


def _read(srv, jid, names=OUT):
    return {n: hipmem.read_plane(srv.view(jid, n)) for n in names}


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


JOB_X = dict(width=160, height=96, lambdas=job_lambdas(3), spp=48, seed=5, image_id=7)


def test_job_alone_coscheduled_and_single_image_path_agree(ctx):
    srv = _server(ctx, slots=1)
    jid = srv.submit(Server.job(**JOB_X, keep=True))
    r = srv.wait(jid)
    assert r["status"] == 0 and r["samples"] == 48 * 160 * 96
    alone = _read(srv, jid)
    srv.release(jid)
    srv.close()

    # co-scheduled with jobs of other sizes, lambda counts and round counts
    srv = _server(ctx, slots=4)
    others = [Server.job(200, 150, job_lambdas(5), spp=40, seed=5, image_id=1),
              Server.job(96, 64, job_lambdas(1), spp=80, rounds=3, cap=400, seed=5, image_id=2),
              Server.job(256, 192, job_lambdas(2), spp=32, seed=5, image_id=3, bf16=True)]
    ids = [srv.submit(others[0]), srv.submit(others[1])]
    jid = srv.submit(Server.job(**JOB_X, keep=True))
    ids += [srv.submit(others[2]), srv.submit(Server.job(128, 128, job_lambdas(4), spp=24, seed=5, image_id=4))]
    res = [srv.wait(i) for i in ids]
    r = srv.wait(jid)
    assert r["status"] == 0 and all(x["status"] == 0 for x in res)
    _same(alone, _read(srv, jid))
    srv.release(jid)
    srv.close()

    # the single-image path with the same seeded inputs
    W, H, lam = JOB_X["width"], JOB_X["height"], JOB_X["lambdas"]
    sc = make_scene(2, 5, W, H, image_id=7)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    bs = sa.BufferSet(ctx, W, H, lam, 5)
    chunk = SampleChunk(W, H, 48)
    synth_samples(ctx, sc, chunk, 0, feats)
    bs.accumulate(chunk.desc())
    p = sa.default_params(radius=6)
    bs.denoise_vars(p, feats.gbuffers(FEATURE_SDS))
    bs.denoise_film(p, feats.gbuffers(FEATURE_SDS))
    ctx.synchronize()
    direct = {n: hipmem.read_plane(bs.view(n)) for n in OUT if n != "targetspp"}
    alone.pop("targetspp")
    _same(alone, direct)
    bs.close()


def test_film_job_runs_the_renderer_loop(ctx):
    """film=1: guarded samples, the set's pbrt Film resolved and used as the
    Welch filter's film, as the single-image calls do."""
    srv = _server(ctx, slots=2)
    jid = srv.submit(Server.job(**JOB_X, keep=True, film=True))
    other = srv.submit(Server.job(96, 64, job_lambdas(1), spp=40, rounds=2, cap=200, seed=5, image_id=2, film=True))
    r = srv.wait(jid)
    assert r["status"] == 0 and srv.wait(other)["status"] == 0
    got = _read(srv, jid, OUT[:-1] + ("film",))
    srv.release(jid)
    srv.close()

    W, H, lam = JOB_X["width"], JOB_X["height"], JOB_X["lambdas"]
    sc = make_scene(2, 5, W, H, image_id=7)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    bs = sa.BufferSet(ctx, W, H, lam, 5, flags=sa.STATER_BUFSET_FILM)
    chunk = SampleChunk(W, H, 48)  # one call: the server's 16-sample passes leave the same Film
    synth_samples(ctx, sc, chunk, 0, feats)
    bs.accumulate(chunk.desc(sanitize=True))
    p = sa.default_params(radius=6)
    g = feats.gbuffers(FEATURE_SDS)
    bs.denoise_vars(p, g)
    bs.film_resolve()
    bs.denoise_film(p, g, film=bs.view("film"))
    ctx.synchronize()
    _same(got, {n: hipmem.read_plane(bs.view(n)) for n in got})
    bs.close()


def test_adaptive_rounds_match_the_round_driver(ctx):
    W, H = 96, 64
    srv = _server(ctx, slots=2)
    jid = srv.submit(Server.job(W, H, job_lambdas(2), spp=16, rounds=3, cap=200, seed=5, image_id=9, keep=True))
    r = srv.wait(jid)
    got = _read(srv, jid, ("n", "filmD", "t0-b0-film-m2"))
    srv.release(jid)
    srv.close()
    assert r["samples"] == 3 * 16 * W * H  # every round spends exactly its budget

    sc = make_scene(2, 5, W, H, image_id=9)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    bs = sa.BufferSet(ctx, W, H, job_lambdas(2), 5)
    p = sa.default_params(radius=6)
    run_rounds(ctx, bs, sc, feats, p, feats.gbuffers(FEATURE_SDS), rounds=3, first_spp=16, budget_spp=16, cap=200,
               pass_spp=16, chunk=SampleChunk(W, H, 16))
    want = {n: hipmem.read_plane(bs.view(n)) for n in got}
    _same(got, want)
    bs.close()


def test_caller_samples_and_continuations_match_the_single_image_path(ctx):
    """A renderer's adaptive loop through the server: a kept job fed the
    caller's first chunk (and features) with next_spp, then a continuation
    whose chunk holds the samples its targetspp asked for (counts = the job's
    targetspp), each request equal bit for bit to the single-image calls."""
    W, H, lam, r = 120, 80, job_lambdas(3), 6
    sc = make_scene(2, 5, W, H, image_id=31)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    g = feats.gbuffers(FEATURE_SDS)
    first = SampleChunk(W, H, 24)
    synth_samples(ctx, sc, first, 0, feats)
    srv = _server(ctx, slots=2)
    jid = srv.submit(Server.job(W, H, lam, spp=24, keep=True, film=True, samples=first.desc(), gbufs=g,
                                next_spp=16, cap=300))
    other = srv.submit(Server.job(96, 64, job_lambdas(2), spp=32, seed=5, image_id=3))  # beside it
    res = srv.wait(jid)
    assert res["status"] == 0 and res["samples"] == 24 * W * H and srv.wait(other)["status"] == 0
    names = ("n", "filmD", "film", "t0-b0-film-varD", "t0-b1-m2", "targetspp")
    got1 = _read(srv, jid, names)

    # the single-image calls of the same request
    bs = sa.BufferSet(ctx, W, H, lam, 5, flags=sa.STATER_BUFSET_FILM)
    p = sa.default_params(radius=r)

    def request():
        bs.denoise_vars(p, g)
        bs.film_resolve()
        bs.denoise_film(p, g, film=bs.view("film"))
        bs.calc_target_spp(16, 300)
        ctx.synchronize()
        return {n: hipmem.read_plane(bs.view(n)) for n in names}

    bs.accumulate(first.desc(sanitize=True))
    _same(got1, request())

    # round 2: the renderer takes targetspp samples per pixel, continuing each pixel's sequence
    tgt = got1["targetspp"][0]
    more = SampleChunk(W, H, int(tgt.max()))
    tplane = srv.view(jid, "targetspp")
    synth_samples(ctx, sc, more, 0, feats, first=srv.view(jid, "n"), counts=tplane)
    ctx.synchronize()
    cid = srv.submit(Server.job(0, 0, lam, spp=1, samples=more.desc(counts=tplane), continue_id=jid,
                                next_spp=16, cap=300))
    res2 = srv.wait(cid)
    assert res2["status"] == 0 and res2["samples"] == int(tgt.sum())
    got2 = _read(srv, jid, names)
    assert np.array_equal(got2["n"][0], got1["n"][0] + tgt)
    srv.release(jid)
    srv.close()
    tbuf = hipmem.DeviceBuffer(W * H * 4)
    tp = sa.plane(tbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
    hipmem.write_plane(tp, tgt[None])
    bs.accumulate(more.desc(counts=tp, sanitize=True))
    _same(got2, request())
    bs.close()


def test_synthetic_job_continued_equals_one_longer_job(ctx):
    """A kept synthetic job of one round, continued by a request of two more
    adaptive rounds on its resident state, ends where a three-round job does."""
    W, H = 96, 64
    kw = dict(width=W, height=H, lambdas=job_lambdas(2), spp=16, cap=200, seed=5, image_id=12)
    srv = _server(ctx, slots=1)
    a = srv.submit(Server.job(**kw, rounds=3, keep=True))
    b = srv.submit(Server.job(**kw, rounds=1, keep=True))
    ra, rb = srv.wait(a), srv.wait(b)
    long = srv.submit(Server.job(256, 192, job_lambdas(5), spp=512, seed=5, image_id=13))  # holds the one slot
    c = srv.submit(Server.job(0, 0, kw["lambdas"], spp=16, rounds=2, cap=200, continue_id=b))
    with pytest.raises(sa.StaterError):
        srv.release(b)  # its continuation is queued behind the long job
    rc = srv.wait(c)
    assert srv.wait(long)["status"] == 0
    assert ra["status"] == rb["status"] == rc["status"] == 0
    assert rb["samples"] + rc["samples"] == ra["samples"] == 3 * 16 * W * H
    names = ("n", "filmD", "t0-b0-film-m2", "t0-b1-m4", "t0-b0-film-varCD")
    _same(_read(srv, a, names), _read(srv, b, names))
    with pytest.raises(sa.StaterError):
        srv.submit(Server.job(0, 0, kw["lambdas"], spp=16, continue_id=999))  # no such kept job
    srv.release(a)
    srv.release(b)
    srv.close()


def test_admission_stays_within_budget_in_arrival_order(ctx):
    jobs = [Server.job(192, 128, job_lambdas(5), spp=32, seed=5, image_id=20 + i) for i in range(6)]
    probe = _server(ctx, slots=1, budget=1 << 30)
    jb = probe.job_bytes(jobs[0])
    probe.close()
    srv = _server(ctx, slots=4, budget=int(2.5 * jb))  # room for two jobs at a time
    ids = [srv.submit(j) for j in jobs]
    res = [srv.wait(i) for i in ids]
    srv.close()
    assert all(r["status"] == 0 and r["bytes"] == jb for r in res)
    admits = [r["t_admit"] for r in res]
    assert admits == sorted(admits)  # FIFO
    events = sorted([(r["t_admit"], 1) for r in res] + [(r["t_done"], -1) for r in res], key=lambda e: (e[0], e[1]))
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= 2


def test_rejects_jobs_beyond_the_server(ctx):
    srv = _server(ctx, slots=1, budget=16 << 20)
    try:
        big = Server.job(256, 192, job_lambdas(5), spp=8)
        assert srv.job_bytes(big) > 16 << 20
        with pytest.raises(sa.StaterError, match="budget"):
            srv.submit(big)
        with pytest.raises(sa.StaterError, match="maximum"):
            srv.submit(Server.job(512, 64, job_lambdas(1), spp=8))
        with pytest.raises(sa.StaterError):
            srv.submit(Server.job(64, 64, job_lambdas(1), spp=0))
        # what fits still runs
        r = srv.wait(srv.submit(Server.job(64, 48, job_lambdas(1), spp=8)))
        assert r["status"] == 0
    finally:
        srv.close()


def _bf16_rne(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def test_convert_plane_bf16_and_interleaved(ctx):
    W, H = 70, 33
    rng = np.random.default_rng(3)
    x = rng.normal(0, 10, (3, H, W)).astype(np.float32)
    x[0, 0, :4] = [0.0, -0.0, np.inf, 1e-40]
    ld = 128
    src = hipmem.DeviceBuffer(3 * H * ld * 4)
    sp = sa.plane(src.ptr, W, H, 3, ld * 4, H * ld * 4)
    hipmem.write_plane(sp, x)
    b16 = hipmem.DeviceBuffer(3 * H * ld * 2)
    bp = sa.plane(b16.ptr, W, H, 3, ld * 2, H * ld * 2, sa.STATER_BF16)
    inter = hipmem.DeviceBuffer(H * W * 3 * 4)
    ip = sa.plane(inter.ptr, W, H, 3, W * 12, 0, sa.STATER_F32, sa.STATER_INTERLEAVED)
    lib = ctx.lib
    import ctypes as C
    sa.check(lib.stater_convert_plane(ctx.h, C.byref(sp), C.byref(bp), None))
    sa.check(lib.stater_convert_plane(ctx.h, C.byref(sp), C.byref(ip), None))
    ctx.synchronize()
    got = b16.download(np.uint16, 3 * H * ld).reshape(3, H, ld)[..., :W]
    assert np.array_equal(got, _bf16_rne(x))
    gi = inter.download(np.float32, H * W * 3).reshape(H, W, 3)
    assert np.array_equal(gi, x.transpose(1, 2, 0))


def test_bf16_features_stay_close_to_f32(ctx):
    srv = _server(ctx, slots=2)
    base = dict(width=192, height=128, lambdas=job_lambdas(5), spp=64, seed=5, image_id=11, keep=True)
    a = srv.submit(Server.job(**base))
    b = srv.submit(Server.job(**base, bf16=True))
    srv.wait(a)
    srv.wait(b)
    fa = hipmem.read_plane(srv.view(a, "filmD"))
    fb = hipmem.read_plane(srv.view(b, "filmD"))
    na = hipmem.read_plane(srv.view(a, "n"))
    nb = hipmem.read_plane(srv.view(b, "n"))
    srv.release(a)
    srv.release(b)
    srv.close()
    assert np.array_equal(na, nb)  # same samples: only the filters' feature storage differs
    rel = np.abs(fa.astype(np.float64) - fb) / np.maximum(np.abs(fa), 1e-3)
    print(f"bf16 features: filmD rel. difference median {np.median(rel):.2e}, p99 {np.percentile(rel, 99):.2e}")
    assert np.median(rel) < 1e-3 and np.percentile(rel, 99) < 5e-2
