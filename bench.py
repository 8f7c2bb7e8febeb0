"""Headline benchmark: denoised Mpixel/s of the statistics-and-denoise path.

Workload (one step = one frame through the hot path, inputs resident in HBM):
1920x1080 RGB, 256 samples per pixel, the mt5 lambda list
[.01 .25 .5 .75 1 3] (5 Box-Cox transforms + identity), M4C statistics;
per-pixel prepass (Jarque-Bera, Bonett, LEV), variance filter and
multi-transform Welch filter, radius 10, alpha 0.005, z 2.80703, spatial
sd 10, albedo/normal/depth range weights.  Synthetic seeded samples and
features (seed 2) are generated on the device before the timed region.
A step = accumulate from an empty set (STATER_ACC_FRESH: the result of a
reset followed by the call) + Film resolve + denoise_vars + denoise_film.

Prints ONE JSON line.  Per-kernel durations come from HIP events recorded
around every launch on the library's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# VALU peak: MI355X's published FP32 vector rate, 157.3 TFLOP/s, is 78.6 T
# lane-ops/s with an FMA counted as one lane-op (256 CU x 4 SIMD x 32 lanes x
# 2.4 GHz).  Non-packed FP32 reaches it: bench.py --valu-probe (16 independent
# FMA chains per lane, many waves) measures about 75 T.
PEAK_LANE_OPS = 78.6   # T lane-ops/s
PEAK_HBM = 8000.0      # GB/s
# Algorithmic work per unit (lane-ops), the per-unit figures of the path's
# scope table: a transcendental (rcp, sqrt, log2, exp2) counts 1, the rate at
# which it issues beside other VALU work on MI355X (it co-issues; on its own
# it takes about 2.7 issue slots).
WELCH_OPS_PER_TEST = 13    # dm, V, dm^2, shares, 1/nu, rcp, crit(u) in 6 FMAs, compare and reduce
VAR_OPS_PER_TEST = 8       # diff, square, add, multiply, compare, select, two accumulations
WEIGHT_OPS_PER_PAIR = 17   # spatial and 7 feature range terms, exp2 (once per ordered pair)
# Accumulation, per sample-pixel: the M4C update of one set-channel is 22
# FP32 lane-ops and one square root (m1); per channel one log2 for the
# Box-Cox transforms and one exp2 per transformed set.
ACC_OPS_PER_SET_CHANNEL = 22 + 1


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def window_pairs(W: int, H: int, r: int) -> int:
    """In-image neighbour pairs (p, q != p) of a (2r+1)^2 window."""
    sx = sum(W - abs(d) for d in range(-r, r + 1) if abs(d) < W)
    sy = sum(H - abs(d) for d in range(-r, r + 1) if abs(d) < H)
    return sx * sy - W * H


def acc_ops_per_sample(set_lambdas):
    """Algorithmic lane-ops of the M4C accumulation per sample-pixel with T = len(set_lambdas)
    Box-Cox sets: 3 channels x ((1 + T) set updates + a log2 + T exp2)."""
    T = len(set_lambdas)
    return 3 * ((1 + T) * ACC_OPS_PER_SET_CHANNEL + (1 if T else 0) + T)


def kernel_work(W, H, spp, S, r, set_lambdas):
    P = W * H
    pairs = window_pairs(W, H, r)
    tests = pairs * S * 3
    T = S - 1
    acc_ops = P * spp * acc_ops_per_sample(set_lambdas)
    return {
        # mode 2 tests each unordered pair once and weights every ordered pair; the variance
        # filter tests every ordered pair, and a pair's weight is the same for all its sets
        "welch_filter": {"ops": tests // 2 * WELCH_OPS_PER_TEST + pairs * WEIGHT_OPS_PER_PAIR, "bound": "valu"},
        "var_filter": {"ops": tests * VAR_OPS_PER_TEST + pairs * WEIGHT_OPS_PER_PAIR, "bound": "valu"},
        # samples read once; a fresh call writes the statistics (n + 7 per set-channel), the
        # reference samples and Box-Cox shifted means (3 + 3T planes) and the Film sums (4 planes)
        "accumulate": {"ops": acc_ops, "bytes": P * spp * 20 + P * (4 + S * 7 * 3 * 4) + P * 4 * (3 + 3 * T + 4),
                       "bound": "valu"},
    }


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(sc, W, H, spp, lambdas, r, p, region, nthreads, runs=5):
    """The CPU leg: the oracle (a C++ port of the reference's accumulation and
    of the filters) timed on a crop of the same frame, one warm-up run and
    then the median of `runs` runs."""
    import oracle

    x0, y0, w, h = region
    t0 = time.perf_counter()
    f = oracle.features(sc, region)
    smp = oracle.samples(sc, f["albedo"], 0, spp, nthreads=nthreads, region=region)
    log(f"cpu baseline inputs generated in {time.perf_counter() - t0:.1f}s")
    feats = [(f["albedo"], 0.02), (f["normal"], 0.1), (f["depth"], 0.5)]

    def one():
        t0 = time.perf_counter()
        n, st = oracle.accumulate(smp, lambdas, 5, sanitize=True, nthreads=nthreads)
        film = oracle.film_image(oracle.film_add(smp, sanitize=True))
        t1 = time.perf_counter()
        jb, lnS2, se2, m2C = oracle.prepass(n, st, p.var_z)
        vD, vCD = oracle.denoise_vars(n, st[:, 1], lnS2.astype(np.float32), se2.astype(np.float32),
                                      m2C.astype(np.float32), r, p.ds_factor, p.var_z, feats, nthreads=nthreads)
        oracle.denoise_film(n, st[:, 0], vCD.astype(np.float32), film, r, p.ds_factor, p.alpha, mt_mode=2,
                            jb=jb.astype(np.float32), feats=feats, nthreads=nthreads)
        return t1 - t0, time.perf_counter() - t1

    one()  # warm-up
    times = []
    for i in range(runs):
        times.append(one())
        log(f"cpu baseline run {i + 1}/{runs}: {sum(times[-1]):.2f}s")
    tot = sorted(sum(t) for t in times)[runs // 2]
    acc = sorted(t[0] for t in times)[runs // 2]
    filt = sorted(t[1] for t in times)[runs // 2]
    return {
        "value": round(w * h / tot / 1e6, 4),
        "unit": "Mpixel/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{w}x{h} crop at ({x0},{y0}) of the same 1080p frame, {spp} spp mt5; oracle accumulate "
                  f"(reference FP32 recurrence, 16x16 tiles, sanitised) and Film {acc:.2f}s + FP64 prepass/variance/"
                  f"Welch filters {filt:.2f}s on {nthreads} threads of {cpu_model()}; median of {runs} runs after "
                  f"one warm-up",
        "runs_s": [round(sum(t), 3) for t in times],
        "accumulate_gsample_s": round(w * h * spp / acc / 1e9, 4),
    }


def fp64_error(ctx, bs, chunk, feats, p, W, H, r, region):
    """Checker on the timed frame: a crop of its samples through the oracle's
    full-FP64 pipeline, against the GPU's FP32 buffers of the same frame (the
    variance outputs where their window lies in the crop, the Welch outputs
    and decisions where their neighbours' windows do too)."""
    import ctypes as C

    import oracle
    import stater_amd as sa
    from stater_amd import hipmem
    from stater_amd.workload import MT5_LAMBDAS

    x0, y0, w, h = region
    win = 2 * r + 1
    row = W * win * win * 3
    buf = hipmem.DeviceBuffer(H * row)
    film = bs.view("film")
    g = feats.gbuffers()
    sa.check(sa.load_library().stater_debug_film_decisions(bs.h, C.byref(p), C.byref(g), C.byref(film), 0,
                                                           C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    dec = buf.download(np.uint8, h * row, offset=y0 * row).reshape(h, W, win, win, 3)[:, x0:x0 + w]
    buf.free()
    lam = oracle.set_lambdas(MT5_LAMBDAS)
    slots = [0] + [list(np.float32(MT5_LAMBDAS)).index(np.float32(l)) for l in lam[1:]]

    def crop(name):
        return hipmem.read_plane(hipmem.sub_plane(bs.view(name), x0, y0, w, h))

    def sets(name):
        return np.stack([crop(f"t0-b{j}-film-{name}" if s == 0 else f"t0-b{j}-{name}") for s, j in enumerate(slots)])

    gpu = {"n": crop("n")[0], "lnS2": sets("lnS2"), "se2": sets("se2"), "varD": sets("varD"),
           "varCD": sets("varCD"), "filmD": crop("filmD"), "decisions": dec}
    fn = feats.to_numpy()
    fl = [(fn["albedo"][:, y0:y0 + h, x0:x0 + w], 0.02), (fn["normal"][:, y0:y0 + h, x0:x0 + w], 0.1),
          (fn["depth"][:, y0:y0 + h, x0:x0 + w], 0.5)]
    smp = chunk.to_numpy(rows=(y0, y0 + h))[..., x0:x0 + w]
    ref = oracle.pipeline64(smp, MT5_LAMBDAS, r, p.ds_factor, p.var_z, p.alpha, feats=fl)
    rep = oracle.fp64_error_report(gpu, ref, p.var_z, r, crop=True)
    rep["crop"] = f"{w}x{h} at ({x0},{y0}); variance outputs on the inner {w - 2 * r}x{h - 2 * r}, Welch outputs " \
                  f"on the inner {w - 4 * r}x{h - 4 * r}"
    rep["bars"] = {"decisions_differ": 1e-4, "max_rel": 1e-4}
    return rep


def bench_c4(args) -> int:
    """48 uneven images: the plain path (one image per launch, images one after
    another on one stream) against the batch (the same launches spread over 4
    streams, tile heights tuned per image, one hipGraph).  Both generate their
    samples on the device in chunks inside the timed region, run the
    renderer's loop (guarded samples, pbrt Film resolved and filtered) and end
    with bit-identical buffers."""
    import stater_amd as sa
    from stater_amd.serve import job_lambdas
    from stater_amd.workload import (FEATURE_SDS, Features, SampleChunk, c4_images, make_scene, synth_features,
                                     synth_samples)

    chunk_spp = args.chunk_spp
    ctx = sa.Context(0)
    p = sa.default_params(radius=args.radius)
    imgs = []
    for i, (W, H, spp, T) in enumerate(c4_images()):
        sc = make_scene(2, 1000 + i, W, H, image_id=i)
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
        bs = sa.BufferSet(ctx, W, H, job_lambdas(T), 5, flags=sa.STATER_BUFSET_FILM)
        imgs.append((bs, sc, feats, feats.gbuffers(FEATURE_SDS), spp, T))
    P = sum(im[0].width * im[0].height for im in imgs)
    samples = sum(im[0].width * im[0].height * im[4] for im in imgs)
    maxW = max(im[0].width for im in imgs)
    maxH = max(im[0].height for im in imgs)
    log(f"c4: {len(imgs)} images, {P / 1e6:.1f} Mpx, {samples / 1e9:.2f} G samples")
    big = SampleChunk(maxW, maxH, chunk_spp)

    def plain():
        for bs, sc, feats, g, spp, _ in imgs:
            ch = SampleChunk(bs.width, bs.height, chunk_spp, buf=big.buf)  # the largest chunk's memory
            for off in range(0, spp, chunk_spp):
                k = min(chunk_spp, spp - off)
                synth_samples(ctx, sc, ch, off, feats, spp=k)
                bs.accumulate(ch.desc(spp=k, sanitize=True), sa.STATER_ACC_FRESH if off == 0 else 0)
            bs.denoise_vars(p, g)
            bs.film_resolve()
            bs.denoise_film(p, g, film=bs.view("film"))

    def timed(fn, reps):
        fn()
        ctx.synchronize()
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ctx.synchronize()
            t.append(time.perf_counter() - t0)
        return min(t)

    t_plain = timed(plain, max(1, args.steps // 2)) if args.c4_mode != "batched" else None
    if t_plain is not None:
        log(f"c4 plain: {t_plain * 1e3:.1f} ms")
    if args.c4_mode == "plain":  # the one-image-per-launch path alone (for a kernel trace of it)
        print(json.dumps({"metric": "C4 denoised Mpixel/s, 48 uneven images, one image per launch",
                          "value": round(P / t_plain / 1e6, 3), "unit": "Mpixel/s", "n_gpus": 1,
                          "higher_is_better": True, "ms_plain": round(t_plain * 1e3, 2),
                          "runs": 1 + max(1, args.steps // 2)}), flush=True)
        ctx.close()
        return 0
    batch = sa.Batch(ctx, [(bs, sc, feats.albedo, g, spp) for bs, sc, feats, g, spp, _ in imgs], p,
                     chunk_spp=chunk_spp)
    batch.tune()
    t_batch = timed(lambda: batch.run(use_graph=True), max(1, args.steps // 2))
    log(f"c4 batched: {t_batch * 1e3:.1f} ms")
    if t_plain is None:
        t_plain = float("nan")
    out = {
        "metric": "C4 denoised Mpixel/s, 48 uneven images (batched: 4 streams, tuned tiles, hipGraph)",
        "value": round(P / t_batch / 1e6, 3), "unit": "Mpixel/s", "n_gpus": 1, "higher_is_better": True,
        "plain_mpixel_s": round(P / t_plain / 1e6, 3), "speedup_vs_plain": round(t_plain / t_batch, 3),
        "ms_batched": round(t_batch * 1e3, 2), "ms_plain": round(t_plain * 1e3, 2),
        "gsample_s_batched": round(samples / t_batch / 1e9, 3), "mpixel": round(P / 1e6, 3),
        "gsamples": round(samples / 1e9, 3), "dtype": "f32", "data": "synthetic, device-generated in the timed region",
        "config": {"workload": "48 images W~U[256,2048] H~U[256,1152] spp~LogU[16,1024] T~U{1..5}, r=%d, "
                               "chunk %d spp, sanitised samples + Film" % (args.radius, chunk_spp)},
        "tile_rows": [(f["tile_rows_var"], f["tile_rows_welch"]) for f in batch.images()],
    }
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()
    return 0


def bench_c3(args) -> int:
    """Progressive adaptive sampling at 1920x1080, mt5, r = 10: 8 rounds (32 spp
    per pixel, then 32 x pixels per round spread by the AS map, cap 4096) with
    the statistics resident on the device; each round generates its samples on
    the device inside the timed region, as the renderer would produce them.
    Timed for the sequential FP32 recurrence and for the FP64 per-pass merge."""
    import stater_amd as sa
    from stater_amd.progressive import run_rounds
    from stater_amd.workload import MT5_LAMBDAS, Features, SampleChunk, make_scene, synth_features

    W, H, rounds, cap, pass_spp = 1920, 1080, 8, 4096, args.pass_spp
    ctx = sa.Context(0)
    sc = make_scene(2, 4, W, H)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    p = sa.default_params(radius=args.radius)
    g = feats.gbuffers()
    chunk = SampleChunk(W, H, pass_spp)
    P = W * H
    modes = {}
    for name, mode in (("sequential", sa.STATER_ACC_SEQUENTIAL), ("merge64", sa.STATER_ACC_MERGE64)):
        bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
        best = None
        for rep in range(1 + max(1, args.steps // 10)):  # the first run is a warm-up
            bs.reset()
            ctx.synchronize()
            t0 = time.perf_counter()
            rl = run_rounds(ctx, bs, sc, feats, p, g, rounds=rounds, first_spp=32, budget_spp=32, cap=cap,
                            pass_spp=pass_spp, mode=mode, chunk=chunk)
            t = time.perf_counter() - t0
            if rep and (best is None or t < best[0]):
                best = (t, rl)
        t, rl = best
        samples = sum(r["samples"] for r in rl)
        modes[name] = {"mpixel_s": round(P / t / 1e6, 3), "ms_per_frame": round(t * 1e3, 2),
                       "gsample_s": round(samples / t / 1e9, 3),
                       "round_ms": [round(r["ms"], 2) for r in rl], "max_target": [r["max_target"] for r in rl]}
        log(f"c3 {name}: {t * 1e3:.1f} ms per 8-round frame, {samples / t / 1e9:.2f} Gsample/s")
        bs.close()
    seq = modes["sequential"]
    out = {
        "metric": "C3 adaptive-sampling Mpixel/s (1080p, mt5, 8 rounds, r=10)",
        "value": seq["mpixel_s"], "unit": "Mpixel/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_frame": seq["ms_per_frame"], "modes": modes, "dtype": "f32",
        "data": "synthetic, device-generated per round in the timed region",
        "config": {"workload": "1920x1080, mt5, 32 spp then 7 AS rounds of 32 spp x pixels (cap 4096), passes of "
                               "%d spp, variance filter + AS map every round, Welch filter after the last, r=%d"
                               % (pass_spp, args.radius)},
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def bench_c5(args) -> int:
    """Served jobs (mix of 720p / 1080p / 4K, T in {1,3,5}, spp in {64,256,1024})
    at Poisson arrivals, against the same jobs run one at a time."""
    import stater_amd as sa
    from stater_amd.serve import Mix, Server, run_poisson

    ctx = sa.Context(0)
    budget = int(args.budget_gb * 1e9)
    mix = Mix()
    rates = [float(x) for x in args.rates.split(",")]
    p = sa.default_params(radius=args.radius)

    one = Server(ctx, max_bytes=budget, slots=1, chunk_spp=args.chunk_spp, params=p)
    run_poisson(one, 1000.0, 4, mix, seed=1)  # warm-up
    plain = run_poisson(one, 1e6, args.jobs, mix, seed=5)  # all at once, one at a time
    one.close()
    log(f"c5 one at a time: {plain['mpixel_s']:.1f} Mpixel/s, {plain['jobs_s']:.2f} jobs/s")
    srv = Server(ctx, max_bytes=budget, slots=args.slots, chunk_spp=args.chunk_spp, params=p)
    run_poisson(srv, 1000.0, 4, mix, seed=1)
    sweep = []
    for rate in rates:
        r = run_poisson(srv, rate, args.jobs, mix, seed=5)
        sweep.append({k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()})
        log(f"c5 rate {rate}/s: {r['mpixel_s']:.1f} Mpixel/s p50 {r['p50_ms']:.0f} ms p99 {r['p99_ms']:.0f} ms")
    sat = run_poisson(srv, 1e6, args.jobs, mix, seed=5)
    srv.close()
    out = {
        "metric": "C5 served Mpixel/s with p50/p99 job latency (720p/1080p/4K mix, T 1/3/5, spp 64/256/1024)",
        "value": round(sat["mpixel_s"], 3), "unit": "Mpixel/s", "n_gpus": 1, "higher_is_better": True,
        "slots": args.slots, "chunk_spp": args.chunk_spp, "one_at_a_time_mpixel_s": round(plain["mpixel_s"], 3),
        "speedup_vs_one_at_a_time": round(sat["mpixel_s"] / plain["mpixel_s"], 3),
        "saturated": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in sat.items()},
        "sweep": sweep, "dtype": "f32", "data": "synthetic, device-generated per job",
        "config": {"workload": "Poisson arrivals, %d jobs per rate, admission budget %.0f GB, sanitised samples + "
                               "Film, r=%d" % (args.jobs, budget / 1e9, args.radius)},
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--radius", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-region", default="768,432,384,216", help="x0,y0,w,h of the CPU baseline's crop")
    ap.add_argument("--error-region", default="900,500,112,96",
                    help="x0,y0,w,h of the crop checked against the FP64 pipeline")
    ap.add_argument("--valu-probe", type=int, default=0,
                    help="only time this many launches of the VALU issue probe (stater_debug_valu_rate: 16 "
                         "independent FP32 FMA chains per lane) and print its rate")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="GiB: only copy a device buffer of this size three times (hipMemcpy, device to device), "
                         "the known-size copy that calibrates the HBM traffic counters")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5"], default="c2",
                    help="c2: the headline 1080p frame; c4: 48 uneven images, batched vs one image per launch; "
                         "c5: served Poisson arrivals vs one job at a time")
    ap.add_argument("--jobs", type=int, default=200, help="c5: jobs per arrival rate")
    ap.add_argument("--rates", default="20,40,100,200,400", help="c5: offered arrival rates, jobs/s")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="c5: admission budget (device memory), GB")
    ap.add_argument("--slots", type=int, default=4, help="c5: jobs in flight")
    ap.add_argument("--chunk-spp", type=int, default=64, help="c4, c5: samples per pixel of one accumulation pass")
    ap.add_argument("--pass-spp", type=int, default=256, help="c3: samples per pixel of one accumulation pass")
    ap.add_argument("--c4-mode", choices=["both", "plain", "batched"], default="both",
                    help="c4: time the plain path, the batch, or both")
    args = ap.parse_args()
    if args.valu_probe > 0:
        import ctypes as C

        import stater_amd as sa
        from stater_amd import hipmem

        ctx = sa.Context(0)
        blocks, iters = 256 * 64, 4096
        out = hipmem.DeviceBuffer(blocks * 256 * 4)
        lib = sa.load_library()

        def launch():
            sa.check(lib.stater_debug_valu_rate(ctx.h, blocks, iters, C.c_void_p(out.ptr), None))

        launch()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.valu_probe):
            launch()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        lane_ops = blocks * 256 * iters * 16 * args.valu_probe
        print(json.dumps({"valu_probe_tlane_op_s": round(lane_ops / dt / 1e12, 2), "launches": args.valu_probe,
                          "ms_per_launch": round(dt * 1e3 / args.valu_probe, 3), "peak_used": PEAK_LANE_OPS}),
              flush=True)
        ctx.close()
        return 0
    if args.copy_probe > 0:
        from stater_amd import hipmem

        nb = int(args.copy_probe * 2**30)
        a, b = hipmem.DeviceBuffer(nb), hipmem.DeviceBuffer(nb)
        a.zero()
        for _ in range(3):
            hipmem._ok(hipmem.hip().hipMemcpy(b.ptr, a.ptr, nb, hipmem.D2D), "hipMemcpy D2D")
        hipmem._ok(hipmem.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        print(json.dumps({"copy_probe_bytes": nb, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); on several GPUs the path "
              f"would shard by image and by row bands with an r-row halo, one process per GPU",
              file=sys.stderr)
        return 2

    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)

    import stater_amd as sa
    from stater_amd.workload import MT5_LAMBDAS, Features, SampleChunk, make_scene, synth_features, synth_samples

    W, H, spp, r = args.width, args.height, args.spp, args.radius
    P = W * H
    lambdas = MT5_LAMBDAS
    ctx = sa.Context(0, timing=True)
    sc = make_scene(2, 2, W, H)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    chunk = SampleChunk(W, H, spp)
    log(f"generating {P * spp / 1e6:.0f}M samples ({chunk.buf.nbytes / 2**30:.1f} GiB) on the device")
    synth_samples(ctx, sc, chunk, 0, feats)
    ctx.synchronize()
    # the reference's data flow: sanitised samples feed the statistics and the
    # Film (statpath.cpp:313-349); denoiseFilm filters the resolved Film
    bs = sa.BufferSet(ctx, W, H, lambdas, 5, flags=sa.STATER_BUFSET_FILM)
    S = len(set(float(x) for x in lambdas) | {1.0})
    p = sa.default_params(radius=r)
    g = feats.gbuffers()
    desc = chunk.desc(sanitize=True)
    film = bs.view("film")

    def step():
        bs.accumulate(desc, sa.STATER_ACC_SEQUENTIAL | sa.STATER_ACC_FRESH)  # = reset, then accumulate
        bs.film_resolve()
        bs.denoise_vars(p, g)
        bs.denoise_film(p, g, film=film)

    for i in range(args.warmup):
        step()
        ctx.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")
    ctx.timing_reset()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        ctx.synchronize()  # one job at a time: its latency is submit -> filmD ready
        lat.append(time.perf_counter() - ts)
    t1 = time.perf_counter()
    log(f"timed {args.steps} steps in {t1 - t0:.3f}s")
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = P * args.steps / (t1 - t0) / 1e6

    tim = ctx.timing()
    work = kernel_work(W, H, spp, S, r, [l for l in dict.fromkeys(float(x) for x in lambdas) if l != 1.0])
    kernels = {}
    for name, (cnt, ms) in tim.items():
        avg = ms / max(cnt, 1)
        k = {"launches": cnt, "avg_ms": round(avg, 4)}
        w = work.get(name, {})
        if "ops" in w:
            k["tlane_op_s"] = round(w["ops"] / (avg * 1e-3) / 1e12, 3)
            k["valu_frac"] = round(k["tlane_op_s"] / PEAK_LANE_OPS, 4)
        if "bytes" in w:
            k["gb_s"] = round(w["bytes"] / (avg * 1e-3) / 1e9, 1)
            k["hbm_frac"] = round(k["gb_s"] / PEAK_HBM, 4)
        kernels[name] = k
    dom = max(kernels, key=lambda n: kernels[n]["avg_ms"] * kernels[n]["launches"])
    dk = kernels[dom]
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / f"traffic_{dom}.json"
    tdata = json.loads(tfile.read_text()) if tfile.exists() else {}
    # the profile's figure only where it was taken on this workload
    if tdata.get("workload") == {"width": W, "height": H, "spp": spp, "radius": r, "sets": S}:
        traffic = tdata.get("bytes_per_launch")
        traffic_src = (f"profiles/{tfile.name}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this command, scaled "
                       f"by the same counters on a copy of known size (bench.py --copy-probe); not measured in this "
                       f"run")
    if "tlane_op_s" in dk:
        # VALU-bound: neither the HBM nor the MFMA roof describes it (no dense contraction on this
        # path); the same call's HBM figures are given beside it
        roofline = {"bound": "valu", "achieved": dk["tlane_op_s"], "peak": PEAK_LANE_OPS, "unit": "Tlane-op/s",
                    "frac": dk["valu_frac"], "traffic": traffic, "traffic_source": traffic_src, "kernel": dom}
        if "gb_s" in dk:
            roofline["hbm"] = {"achieved": dk["gb_s"], "peak": PEAK_HBM, "unit": "GB/s", "frac": dk["hbm_frac"]}
    else:
        roofline = {"bound": "hbm", "achieved": dk.get("gb_s"), "peak": PEAK_HBM, "unit": "GB/s",
                    "frac": dk.get("hbm_frac"), "traffic": traffic, "kernel": dom}
    acc = kernels.get("accumulate")
    gsample = round(P * spp / (acc["avg_ms"] * 1e-3) / 1e9, 3) if acc else None

    cpu, error = None, None
    if not args.no_cpu:
        region = tuple(int(v) for v in args.error_region.split(","))
        log(f"fp64 check of a {region[2]}x{region[3]} crop of the timed frame")
        error = fp64_error(ctx, bs, chunk, feats, p, W, H, r, region)
        region = tuple(int(v) for v in args.cpu_region.split(","))
        nthreads = max(1, min(os.cpu_count() or 1, 16))
        log(f"cpu baseline on a {region[2]}x{region[3]} crop, {nthreads} threads")
        cpu = cpu_baseline(sc, W, H, spp, lambdas, r, p, region, nthreads)

    out = {
        "metric": "denoised Mpixel/s (1080p, mt5, r=10) + p50/p99 job latency at fixed error",
        "value": round(value, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox sample model, device-generated before timing)",
        "config": {"workload": f"{W}x{H} RGB, {spp} spp sanitised, mt5 (6 lambda slots, 5 Box-Cox), M4C + "
                               f"Film, prepass + variance filter + Welch filter of the Film, r={r}, alpha=0.005, "
                               f"albedo/normal/depth weights", "width": W, "height": H, "spp": spp,
                   "radius": r, "sets": S},
        "gsample_s": gsample,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                       "what": "one 1080p job (accumulate from empty + film + denoise_vars + denoise_film) alone"},
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "error": error,
    }
    print(json.dumps(out), flush=True)
    bs.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
