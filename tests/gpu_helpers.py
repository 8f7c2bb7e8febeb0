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


def band_samples(ctx, sc, feats, W: int, H: int, spp: int, y0: int, y1: int, chunk_spp: int = 64) -> np.ndarray:
    """Samples 0 .. spp-1 of rows y0 <= y < y1 of a scene, [spp, 5, y1-y0, W],
    generated on the device in chunks (as the batch and the server do) so
    that large films never hold all their samples at once."""
    out = np.empty((spp, 5, y1 - y0, W), np.float32)
    chunk = SampleChunk(W, H, min(chunk_spp, spp))
    for off in range(0, spp, chunk_spp):
        k = min(chunk_spp, spp - off)
        synth_samples(ctx, sc, chunk, off, feats, spp=k)
        ctx.synchronize()
        out[off:off + k] = chunk.to_numpy(rows=(y0, y1))[:k]
    return out


def window_filters_vs_oracle(ctx, bs, lambdas, feats_np, p, g, film_plane, x0, y0, w, h, sds):
    """The variance and Welch filters of a whole film checked on an interior
    window against the oracle fed the window's own FP32 inputs (the GPU's
    statistics and prepass): variance decisions within 1e-5 of the threshold
    and differing Welch decisions each <= 0.01 %, every pixel-channel whose
    decisions agree within 1e-4.  Returns (ambiguous fraction, differing
    Welch decision fraction)."""
    import ctypes as C

    r = p.radius
    W, H = bs.width, bs.height
    slots = set_slots(lambdas)

    def win(name):
        return hipmem.read_plane(hipmem.sub_plane(bs.view(name), x0, y0, w, h))

    def sets(name):
        return np.stack([win(f"t0-b{j}-film-{name}" if s == 0 else f"t0-b{j}-{name}") for s, j in enumerate(slots)])

    nw = win("n")[0]
    stw = np.stack([np.stack([win((f"t0-b{j}-film-" if s == 0 else f"t0-b{j}-") + nm) for nm in STAT_NAMES])
                    for s, j in enumerate(slots)])
    lnS2, se2, m2C, jb = sets("lnS2"), sets("se2"), sets("m2C"), sets("jb")[:, 0]
    fl = [(feats_np[k][:, y0:y0 + h, x0:x0 + w], sd) for k, sd in zip(("albedo", "normal", "depth"), sds)]
    vD, vCD, amb = oracle.denoise_vars(nw, stw[:, 1], lnS2, se2, m2C, r, p.ds_factor, p.var_z, fl,
                                       want_ambiguous=True)
    inner = (slice(None), slice(None), slice(r, h - r), slice(r, w - r))
    ambf = amb[inner].sum() / (float(amb[inner].size) * (2 * r + 1) ** 2)
    assert ambf <= 1e-4, ambf
    gD, gCD = sets("varD"), sets("varCD")
    for gv, rv in ((gD, vD), (gCD, vCD)):
        ok = (amb == 0)[inner] & np.isfinite(rv[inner]) & (np.abs(rv[inner]) < 1e37)
        rel = np.abs(gv[inner][ok] - rv[inner][ok]) / np.maximum(np.abs(rv[inner][ok]), 1e-30)
        assert rel.max() < 1e-4, rel.max()
    film = hipmem.read_plane(hipmem.sub_plane(film_plane, x0, y0, w, h)) if film_plane is not None else stw[0, 0]
    mean = stw[:, 0].copy()
    mean[0] = film
    assert p.cascade, "the check covers the cascaded Welch test"
    var = gCD if p.use_levc else gD
    ref = oracle.denoise_film(nw, mean, var, film, r, p.ds_factor, p.alpha, mt_mode=p.mt_mode, jb=jb, feats=fl,
                              want_decisions=True)
    winw = 2 * r + 1
    row = W * winw * winw * 3
    dbuf = hipmem.DeviceBuffer(H * row)
    sa.check(sa.load_library().stater_debug_film_decisions(bs.h, C.byref(p), C.byref(g) if g else None,
                                                           C.byref(film_plane) if film_plane is not None else None,
                                                           0, C.c_void_p(dbuf.ptr), None))
    ctx.synchronize()
    dec = dbuf.download(np.uint8, (h - 2 * r) * row, offset=(y0 + r) * row)
    dbuf.free()
    dec = dec.reshape(h - 2 * r, W, winw, winw, 3)[:, x0 + r:x0 + w - r]
    diff = dec != ref["decisions"][r:h - r, r:w - r]
    frac = np.count_nonzero(diff) / diff.size
    assert frac <= 1e-4, frac
    same = ~np.any(diff, axis=(2, 3)).transpose(2, 0, 1)
    got = win("filmD")[:, r:h - r, r:w - r]
    want = ref["filmD"][:, r:h - r, r:w - r]
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-6)
    assert rel[same].max() < 1e-4, rel[same].max()
    return float(ambf), float(frac)
