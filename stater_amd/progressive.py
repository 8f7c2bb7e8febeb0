"""Progressive adaptive-sampling rounds with device-resident state.

Mirrors the reference's iteration driver (src/statistics/statpath.cpp:152-259
and :424-441): round 1 takes ``first_spp`` samples per pixel; every later
round runs DenoiseVars, CalcTargetSPPs and then lets each pixel take its
target number of further samples (the sampler's SetSPP/SetSampleNumber of
statpath.cpp:253-256), continuing its own sample sequence.  Here the samples
come from the seeded generator and the statistics never leave the device:
each round is consumed in passes of ``pass_spp`` samples through the
per-pixel ``counts``/``count_offset`` of the sample chunk.  Every pass reads
and writes the state of each pixel it gives samples, so passes are large by
default (256 spp: a 10.6 GB chunk at 1080p).
"""
from __future__ import annotations

import time

import numpy as np

from . import STATER_ACC_SEQUENTIAL, BufferSet, Context, DenoiseParams, GBuffers
from .hipmem import read_plane
from .workload import Features, SampleChunk, synth_samples


def run_rounds(ctx: Context, bs: BufferSet, scene, feats: Features | None, params: DenoiseParams,
               gbufs: GBuffers | None, rounds: int = 8, first_spp: int = 32, budget_spp: int = 32,
               cap: int = 4096, pass_spp: int = 256, mode: int = STATER_ACC_SEQUENTIAL,
               chunk: SampleChunk | None = None, denoise_last: bool = True, as_mode: int = 1) -> list[dict]:
    """Runs the rounds; returns per-round {round, samples, max_target, ms}.
    as_mode: the AS map's variance, 1 = denoised varD, 0 = sample variance
    (the reference's AS_MODE build flag)."""
    W, H = bs.width, bs.height
    n_view = bs.view("n")
    tgt_view = bs.view("targetspp")
    chunk = chunk or SampleChunk(W, H, pass_spp)
    log = []
    for r in range(rounds):
        t0 = time.perf_counter()
        if r == 0:
            counts, total, samples = None, first_spp, first_spp * W * H
        else:
            if as_mode:
                bs.denoise_vars(params, gbufs)
            bs.calc_target_spp(budget_spp, cap, as_mode=as_mode)
            ctx.synchronize()
            tgt = read_plane(tgt_view)[0]
            counts, total, samples = tgt_view, int(tgt.max()), int(tgt.sum(dtype=np.int64))
        for off in range(0, total, pass_spp):
            k = min(pass_spp, total - off)
            synth_samples(ctx, scene, chunk, 0, feats, spp=k, first=n_view, counts=counts, count_offset=off)
            bs.accumulate(chunk.desc(spp=k, counts=counts, count_offset=off), mode)
        ctx.synchronize()
        log.append({"round": r + 1, "samples": samples, "max_target": total,
                    "ms": (time.perf_counter() - t0) * 1e3})
    if denoise_last:
        bs.denoise_vars(params, gbufs)
        bs.denoise_film(params, gbufs)
        ctx.synchronize()
    return log
