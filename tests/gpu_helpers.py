"""Shared helpers of the GPU parity tests (not collected by pytest)."""
from __future__ import annotations

import numpy as np

import oracle
import stater_amd as sa
from stater_amd import hipmem
from stater_amd.workload import Features, SampleChunk, make_scene, synth_features, synth_samples

STAT_NAMES = ["mean", "m2", "m3", "m4", "c2X", "c2Y", "m1"]


def set_slots(lambdas) -> list[int]:
    """Slot index j that names each statistics set (set 0: the untransformed one)."""
    lam = oracle.set_lambdas(lambdas)
    slots = [0]
    for l in lam[1:]:
        slots.append([np.float32(v) for v in lambdas].index(np.float32(l)))
    return slots


def read_stats(bs: sa.BufferSet, lambdas) -> tuple[np.ndarray, np.ndarray]:
    """(n [H,W], stats [S,7,3,H,W]) read back through the named views."""
    slots = set_slots(lambdas)
    n = hipmem.read_plane(bs.view("n"))[0]
    out = []
    for s, j in enumerate(slots):
        pre = f"t0-b{j}-film-" if s == 0 else f"t0-b{j}-"
        out.append(np.stack([hipmem.read_plane(bs.view(pre + nm)) for nm in STAT_NAMES]))
    return n, np.stack(out)


def read_set_buf(bs: sa.BufferSet, lambdas, name: str) -> np.ndarray:
    """A derived per-set buffer (lnS2, se2, m2C, varD, varCD, discr, jb) as [S,ch,H,W]."""
    slots = set_slots(lambdas)
    out = []
    for s, j in enumerate(slots):
        pre = f"t0-b{j}-film-" if s == 0 else f"t0-b{j}-"
        out.append(hipmem.read_plane(bs.view(pre + name)))
    return np.stack(out)


def generate(ctx, model: int, seed: int, W: int, H: int, spp: int, first: int = 0, with_features: bool | None = None):
    """Scene + device samples (+ features).  Returns (scene, chunk, feats)."""
    sc = make_scene(model, seed, W, H)
    feats = None
    if with_features or (with_features is None and model == 2):
        feats = Features(W, H)
        synth_features(ctx, sc, feats)
    chunk = SampleChunk(W, H, spp)
    synth_samples(ctx, sc, chunk, first, feats)
    ctx.synchronize()
    return sc, chunk, feats


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
