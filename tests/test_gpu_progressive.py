"""Progressive adaptive-sampling rounds with device-resident state.

The union of all rounds gives pixel p its samples 0 .. n_p - 1 in order, so
the CPU oracle can replay the same sequence: the sequential mode must equal
the reference FP32 recurrence bit for bit on the untransformed set, and both
modes are compared with the FP64 recurrence per buffer.
"""
import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import read_stats
from stater_amd import hipmem
from stater_amd.progressive import run_rounds
from stater_amd.workload import MT5_LAMBDAS, Features, make_scene, synth_features
from test_gpu_accumulate import _scales
from test_gpu_fullsize import _stats_window

pytestmark = pytest.mark.gpu

NAMES = ["mean", "m2", "m3", "m4", "c2X", "c2Y", "m1"]


def _run(ctx, mode, W=48, H=32, rounds=4, cap=600):
    sc = make_scene(2, 3, W, H)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    p = sa.default_params(radius=5)
    log = run_rounds(ctx, bs, sc, feats, p, feats.gbuffers(), rounds=rounds, first_spp=32, budget_spp=32, cap=cap,
                     pass_spp=64, mode=mode)
    n, st = read_stats(bs, MT5_LAMBDAS)
    return sc, feats, n, st, log


def _errors(st, st64, n):
    worst = np.zeros((st.shape[0], 7))
    for s in range(st.shape[0]):
        scale = _scales(st64[s], n)
        fits = np.broadcast_to(np.all(np.abs(st64[s]) < 1e30, axis=0), st[s].shape)
        err = np.where(fits, np.abs(st[s].astype(np.float64) - st64[s]) / np.maximum(scale, 1e-30), 0.0)
        worst[s] = err.reshape(7, -1).max(axis=1)
    return worst


def test_progressive_rounds_sequential_and_merge64(ctx):
    sc, feats, n, st, log = _run(ctx, sa.STATER_ACC_SEQUENTIAL)
    total = sum(r["samples"] for r in log)
    assert int(n.sum()) == total == 32 * n.size * len(log)  # every round spends exactly its budget
    assert n.max() <= 600 and n.min() >= 32
    # replay: pixel p used samples 0 .. n_p - 1
    smp = oracle.samples(sc, feats.to_numpy()["albedo"], 0, int(n.max()))
    n32, st32 = oracle.accumulate(smp, MT5_LAMBDAS, 5, counts=n)
    _, st64 = oracle.accumulate(smp, MT5_LAMBDAS, 5, fp64=True, counts=n)
    assert np.array_equal(n, n32)
    assert np.array_equal(st[0], st32[0])  # resident state continued bit for bit across rounds and passes
    err_seq = _errors(st, st64, n)
    # the merge mode rounds differently, so its AS targets (and n) may differ
    _, _, n_m, st_m, log_m = _run(ctx, sa.STATER_ACC_MERGE64)
    assert int(n_m.sum()) == total
    smp = oracle.samples(sc, feats.to_numpy()["albedo"], 0, int(n_m.max()))
    _, st64m = oracle.accumulate(smp, MT5_LAMBDAS, 5, fp64=True, counts=n_m)
    err_m = _errors(st_m, st64m, n_m)
    for s in range(st.shape[0]):
        print("set", s, "sequential", " ".join(f"{NAMES[k]}={err_seq[s, k]:.1e}" for k in range(7)))
        print("set", s, "merge64   ", " ".join(f"{NAMES[k]}={err_m[s, k]:.1e}" for k in range(7)))
    assert err_seq.max() <= 1e-4
    assert err_m.max() <= 1e-4


def test_adaptive_rounds_full_frame(ctx):
    """1920 x 1080, mt5, r = 10, 8 rounds: 32 spp per pixel, then 32 x pixels
    per round spread by the AS map (cap 4096), statistics resident throughout.
    Whole frame: every round spends exactly its budget and the cap holds.  A
    band of rows: each pixel's samples 0 .. n_p - 1 replayed by the oracle give
    the untransformed set bit for bit and the Box-Cox sets within 1e-4 of FP64."""
    W, H, rounds, cap = 1920, 1080, 8, 4096
    sc = make_scene(2, 4, W, H)
    feats = Features(W, H)
    synth_features(ctx, sc, feats)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    p = sa.default_params(radius=10)
    log = run_rounds(ctx, bs, sc, feats, p, feats.gbuffers(), rounds=rounds, first_spp=32, budget_spp=32, cap=cap,
                     pass_spp=64)
    for r in log:
        print(f"round {r['round']}: {r['samples']} samples, max target {r['max_target']}, {r['ms']:.1f} ms")
    assert all(r["samples"] == 32 * W * H for r in log)
    n = hipmem.read_plane(bs.view("n"))[0]
    assert int(n.sum(dtype=np.int64)) == 32 * W * H * rounds
    assert n.min() >= 32 and n.max() <= cap
    filmD = hipmem.read_plane(bs.view("filmD"))
    assert np.all(np.isfinite(filmD))

    y0, y1 = 700, 702
    nb = n[y0:y1]
    alb = feats.to_numpy()["albedo"][:, y0:y1]
    smp = oracle.samples(sc, alb, 0, int(nb.max()), region=(0, y0, W, y1 - y0))
    n32, st32 = oracle.accumulate(smp, MT5_LAMBDAS, 5, counts=nb)
    _, st64 = oracle.accumulate(smp, MT5_LAMBDAS, 5, fp64=True, counts=nb)
    st = _stats_window(bs, 0, y0, W, y1 - y0)
    assert np.array_equal(n32, nb)
    assert np.array_equal(st[0].view(np.uint32), st32[0].view(np.uint32))
    err = _errors(st, st64, nb)
    print("band worst errors per set:", " ".join(f"{e:.1e}" for e in err.max(axis=1)))
    assert err.max() <= 1e-4
