"""The batching and serving configurations at their stated workloads.

* C4: the 48 uneven films (W ~ U[256, 2048], H ~ U[256, 1152],
  spp ~ LogU[16, 1024], T ~ U{1..5}) through one batch in the renderer's loop
  (guarded samples, pbrt Film), tuned and replayed as a captured graph:
  size-independent properties on every film, the untransformed statistics
  and the Film bit for bit on a band of rows of the largest and the smallest
  film, and the filters of the largest film against the oracle on an
  interior window.
* C5: a Poisson stream of the 720p / 1080p / 4K mix (T in {1, 3, 5}, spp in
  {64, 256, 1024}) through the server, with one 4K mt5 job kept and replayed
  on a band of rows through the oracle; every job completes with its full
  sample count, and latencies are ordered (submit <= admit <= done)."""
import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import STAT_NAMES, band_samples, set_slots, window_filters_vs_oracle
from stater_amd import hipmem
from stater_amd.serve import Mix, Server, job_lambdas
from stater_amd.workload import FEATURE_SDS, Features, c4_images, make_scene, synth_features

pytestmark = pytest.mark.gpu

R = 10


def _band_check(ctx, bs, lam, sc, feats, spp, y0, y1):
    """n, the untransformed statistics and the Film of rows y0..y1 against the
    oracle fed the same samples, bit for bit."""
    W, H = bs.width, bs.height
    smp = band_samples(ctx, sc, feats, W, H, spp, y0, y1)
    n_ref, st_ref = oracle.accumulate(smp, lam, 5, sanitize=True)
    n = hipmem.read_plane(hipmem.sub_plane(bs.view("n"), 0, y0, W, y1 - y0))[0]
    st0 = np.stack([hipmem.read_plane(hipmem.sub_plane(bs.view("t0-b0-film-" + nm), 0, y0, W, y1 - y0))
                    for nm in STAT_NAMES])
    assert np.array_equal(n, n_ref)
    assert np.array_equal(st0.view(np.uint32), st_ref[0].view(np.uint32))
    film = hipmem.read_plane(hipmem.sub_plane(bs.view("film"), 0, y0, W, y1 - y0))
    want = oracle.film_image(oracle.film_add(smp, sanitize=True))
    assert np.array_equal(film.view(np.uint32), want.view(np.uint32))


def test_c4_48_uneven_films(ctx):
    p = sa.default_params(radius=R)
    shapes = c4_images()
    imgs, keep = [], []
    for i, (W, H, spp, T) in enumerate(shapes):
        sc = make_scene(2, 1000 + i, W, H, image_id=i)
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
        bs = sa.BufferSet(ctx, W, H, job_lambdas(T), 5, flags=sa.STATER_BUFSET_FILM)
        imgs.append((bs, sc, feats.albedo, feats.gbuffers(FEATURE_SDS), spp))
        keep.append(feats)
    batch = sa.Batch(ctx, imgs, p, chunk_spp=64)
    batch.tune()
    batch.run(use_graph=True)
    batch.run(use_graph=True)  # a replay leaves the same buffers
    ctx.synchronize()
    print("\nc4 tile rows (variance, Welch):", [(f["tile_rows_var"], f["tile_rows_welch"]) for f in batch.images()])
    for (bs, *_), (W, H, spp, T) in zip(imgs, shapes):
        n = hipmem.read_plane(bs.view("n"))[0]
        assert np.all(n == spp)
        filmD, film = hipmem.read_plane(bs.view("filmD")), hipmem.read_plane(bs.view("film"))
        assert np.all(np.isfinite(filmD)) and np.all(filmD >= 0)
        assert filmD.min() >= film.min() and filmD.max() <= film.max() * (1 + 1e-6)
    area = [W * H for W, H, _, _ in shapes]
    for i in (int(np.argmax(area)), int(np.argmin(area))):
        bs, sc, _, g, spp = imgs[i]
        W, H, _, T = shapes[i]
        y0 = H // 2 - 2
        _band_check(ctx, bs, job_lambdas(T), sc, keep[i], spp, y0, y0 + 4)
    big = int(np.argmax(area))
    bs, sc, _, g, spp = imgs[big]
    W, H, _, T = shapes[big]
    amb, frac = window_filters_vs_oracle(ctx, bs, job_lambdas(T), keep[big].to_numpy(), p, g, bs.view("film"),
                                         W // 2, H // 3, 96, 64, FEATURE_SDS)
    print(f"c4 largest film {W}x{H} window: ambiguous variance decisions {amb:.2e}, Welch decisions differing "
          f"{frac:.2e}")
    batch.close()
    for bs, *_ in imgs:
        bs.close()


def test_c5_poisson_mix_with_a_4k_job(ctx):
    p = sa.default_params(radius=R)
    srv = Server(ctx, max_bytes=64 << 30, slots=4, chunk_spp=64, params=p)
    mix = Mix()
    rng = np.random.default_rng(55)
    jobs = [mix.draw(rng, i, 5000 + i) for i in range(14)]
    k4 = Server.job(3840, 2160, job_lambdas(5), spp=64, seed=5099, image_id=99, keep=True, film=True)
    jobs.insert(5, k4)
    gaps = rng.exponential(1.0 / 40.0, len(jobs))  # 40 jobs/s offered
    import time
    t0 = srv.clock()
    ids = []
    for job, ta in zip(jobs, t0 + np.cumsum(gaps)):
        dt = ta - srv.clock()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(job))
    res = [srv.wait(i) for i in ids]
    for job, r in zip(jobs, res):
        assert r["status"] == 0
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
        assert r["samples"] == job.spp * job.width * job.height
    lat = np.array([r["t_done"] - r["t_submit"] for r in res]) * 1e3
    print(f"\nc5 {len(jobs)} jobs: p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms")
    kid = ids[5]
    sc = make_scene(2, 5099, 3840, 2160, image_id=99)
    feats = Features(3840, 2160)
    synth_features(ctx, sc, feats)
    lam = job_lambdas(5)
    view = type("V", (), {"width": 3840, "height": 2160, "view": lambda self, n: srv.view(kid, n)})()
    _band_check(ctx, view, lam, sc, feats, 64, 1077, 1081)
    srv.release(kid)
    srv.close()
