"""ctypes/numpy wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, as the checker and the timed CPU port.  The
product (stater_amd) never imports it.  See oracle.cpp for what each function
restates (StatER file:line) and how it is pinned.

Array conventions: per-pixel planes are flattened row-major (p = y*W + x);
statistics are [S][7][3][P] (mean, m2, m3, m4, c2X, c2Y, m1); samples are
[spp][5][P] (R, G, B, u, v).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
    L.orc_scene_init.argtypes = [_vp, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_int]
    L.orc_features.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p]
    L.orc_samples.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, _f32p, C.c_int]
    for nm, st in (("orc_accumulate_f32", _f32p), ("orc_accumulate_f64", _f64p)):
        getattr(L, nm).argtypes = [C.c_int, C.c_int, C.c_int, _f32p, C.c_int, _f32p, C.c_int, _vp, C.c_int, _i32p,
                                   st, C.c_int]
    L.orc_moments_exact.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, C.c_int, _vp, C.c_int, _f64p]
    L.orc_merge_moments.argtypes = [_f64p, _f64p]
    L.orc_mean_var.argtypes = [C.c_int, _i32p, _f32p, _f32p]
    for nm in ("orc_betainc",):
        getattr(L, nm).argtypes = [C.c_double, C.c_double, C.c_double]
        getattr(L, nm).restype = C.c_double
    L.orc_t2_pvalue.argtypes = [C.c_double, C.c_double]
    L.orc_t2_pvalue.restype = C.c_double
    L.orc_t2_crit.argtypes = [C.c_double, C.c_double]
    L.orc_t2_crit.restype = C.c_double
    L.orc_prepass.argtypes = [C.c_int, C.c_int, _i32p, _f32p, C.c_double, _f64p, _f64p, _f64p, _f64p]
    L.orc_prepass64.argtypes = [C.c_int, C.c_int, _i32p, _f64p, C.c_double, _f64p, _f64p, _f64p, _f64p]
    L.orc_denoise_vars64.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p, _f64p, _f64p, C.c_int,
                                     C.c_double, C.c_double, C.c_int, _vp, _vp, _vp, _f64p, _f64p, _vp, _vp,
                                     C.c_int]
    L.orc_denoise_film64.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p, _vp, C.c_int, C.c_double,
                                     C.c_double, C.c_int, C.c_int, _vp, _vp, _vp, _f64p, _f64p, _vp, _vp, _vp,
                                     C.c_int]
    L.orc_var_decisions.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f32p, _f32p, C.c_int, C.c_double, _u8p]
    L.orc_var_decisions64.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p, C.c_int, C.c_double, _u8p]
    L.orc_film_image64.argtypes = [C.c_int, C.c_int, _f32p, C.c_int, _vp, C.c_int, _f64p]
    L.orc_denoise_vars.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f32p, _f32p, _f32p, _f32p, C.c_int,
                                   C.c_double, C.c_double, C.c_int, _vp, _vp, _vp, _f64p, _f64p, _vp, _vp, C.c_int]
    L.orc_denoise_film.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f32p, _f32p, _vp, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, _vp, _vp, _vp, _f32p, _f64p, _vp, _vp, _vp, C.c_int]
    L.orc_target_spp.argtypes = [C.c_int, _i32p, _f32p, _f32p, C.c_uint64, C.c_uint32, _i32p]
    L.orc_film_add.argtypes = [C.c_int, C.c_int, _f32p, C.c_int, _vp, C.c_int, _f32p]
    L.orc_film_xyz.argtypes = [C.c_int, _f32p, _f32p]
    L.orc_film_resolve.argtypes = [C.c_int, _f32p, C.c_float, _f32p]
    _lib = L
    return L


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().orc_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def scene_init(scene_struct_type, model: int, seed: int, image_id: int, W: int, H: int):
    s = scene_struct_type()
    rc = lib().orc_scene_init(C.addressof(s), model, seed, image_id, W, H)
    assert rc == 0
    return s


def features(scene, region=None) -> dict[str, np.ndarray]:
    """region = (x0, y0, w, h) or None for the whole frame."""
    x0, y0, W, H = region or (0, 0, scene.width, scene.height)
    a = np.zeros(3 * W * H, np.float32)
    n = np.zeros(3 * W * H, np.float32)
    d = np.zeros(W * H, np.float32)
    lib().orc_features(C.addressof(scene), x0, y0, W, H, a, n, d)
    return {"albedo": a.reshape(3, H, W), "normal": n.reshape(3, H, W), "depth": d.reshape(1, H, W)}


def samples(scene, albedo: np.ndarray | None, first: int, spp: int, nthreads: int = 8, region=None) -> np.ndarray:
    """[spp,5,h,w] samples of region (x0, y0, w, h) (default: whole frame);
    albedo is the region's [3,h,w] albedo (needed when the scene modulates)."""
    x0, y0, W, H = region or (0, 0, scene.width, scene.height)
    out = np.zeros(spp * 5 * W * H, np.float32)
    alb = None if albedo is None else np.ascontiguousarray(albedo, np.float32)
    lib().orc_samples(C.addressof(scene), None if alb is None else alb.ctypes.data, x0, y0, W, H, first, spp, out,
                      nthreads)
    return out.reshape(spp, 5, H, W)


def set_lambdas(lambdas) -> np.ndarray:
    """Distinct sets in the library's order: 1 (untransformed) then new lambdas != 1."""
    out = [1.0]
    for l in lambdas:
        if np.float32(l) not in [np.float32(v) for v in out]:
            out.append(float(l))
    return np.asarray(out, np.float32)


def accumulate(samples_arr: np.ndarray, lambdas, mm: int, fp64: bool = False, counts=None, sanitize=False,
               state=None, nthreads: int = 8):
    """Run the reference recurrence.  samples_arr: [spp,5,H,W].  Returns (n [H,W], stats [S,7,3,H,W])."""
    spp, _, H, W = samples_arr.shape
    lam = set_lambdas(lambdas)
    S = len(lam)
    P = W * H
    smp = np.ascontiguousarray(samples_arr.reshape(spp, 5, P), np.float32)
    if state is None:
        n = np.zeros(P, np.int32)
        st = np.zeros(S * 7 * 3 * P, np.float64 if fp64 else np.float32)
    else:
        n = np.ascontiguousarray(state[0].reshape(P), np.int32).copy()
        st = np.ascontiguousarray(state[1].reshape(-1), np.float64 if fp64 else np.float32).copy()
    cnt = None if counts is None else np.ascontiguousarray(counts.reshape(P), np.int32)
    fn = lib().orc_accumulate_f64 if fp64 else lib().orc_accumulate_f32
    fn(W, H, S, lam, mm, smp, spp, None if cnt is None else cnt.ctypes.data, int(sanitize), n, st, nthreads)
    return n.reshape(H, W), st.reshape(S, 7, 3, H, W)


def film_add(samples_arr: np.ndarray, sums: np.ndarray | None = None, counts=None, sanitize=False) -> np.ndarray:
    """pbrt Film running sums [4,H,W] (contribSum R, G, B and the weight sum)
    after adding one call's samples [spp,5,H,W] in sample order
    (src/core/film.h:152-190; the tiles live for the whole render,
    src/statistics/statpath.cpp:144)."""
    spp, _, H, W = samples_arr.shape
    P = W * H
    out = np.zeros(4 * P, np.float32) if sums is None else np.ascontiguousarray(sums, np.float32).reshape(-1).copy()
    cnt = None if counts is None else np.ascontiguousarray(counts.reshape(P), np.int32)
    lib().orc_film_add(W, H, np.ascontiguousarray(samples_arr.reshape(spp, 5, P), np.float32), spp,
                       None if cnt is None else cnt.ctypes.data, int(sanitize), out)
    return out.reshape(4, H, W)


def film_xyz(sums: np.ndarray) -> np.ndarray:
    """The Film's pixels [4,H,W] (X, Y, Z, weight) after Film::Clear and one
    MergeFilmTile (src/core/film.cpp:127-149)."""
    _, H, W = sums.shape
    out = np.zeros(4 * W * H, np.float32)
    lib().orc_film_xyz(W * H, np.ascontiguousarray(sums, np.float32).reshape(-1), out)
    return out.reshape(4, H, W)


def film_resolve(xyzw: np.ndarray, scale: float = 1.0) -> np.ndarray:
    """Film::UpdateImage (src/core/film.cpp:188-220) of the Film's pixels: RGB [3,H,W]."""
    _, H, W = xyzw.shape
    out = np.zeros(3 * W * H, np.float32)
    lib().orc_film_resolve(W * H, np.ascontiguousarray(xyzw, np.float32).reshape(-1), scale, out)
    return out.reshape(3, H, W)


def film_image64(samples_arr: np.ndarray, counts=None, sanitize=False) -> np.ndarray:
    """pbrt's Film image of the samples [spp,5,H,W] evaluated in FP64 [3,H,W]."""
    spp, _, H, W = samples_arr.shape
    P = W * H
    out = np.zeros(3 * P)
    cnt = None if counts is None else np.ascontiguousarray(counts.reshape(P), np.int32)
    lib().orc_film_image64(W, H, np.ascontiguousarray(samples_arr.reshape(spp, 5, P), np.float32), spp,
                           None if cnt is None else cnt.ctypes.data, int(sanitize), out)
    return out.reshape(3, H, W)


def film_image(sums: np.ndarray, scale: float = 1.0) -> np.ndarray:
    """The image of an iteration from the running sums: Clear, merge, UpdateImage."""
    return film_resolve(film_xyz(sums), scale)


def moments_exact(samples_arr: np.ndarray, lambdas, counts=None, sanitize=False) -> np.ndarray:
    spp, _, H, W = samples_arr.shape
    lam = set_lambdas(lambdas)
    S = len(lam)
    out = np.zeros(S * 4 * 3 * W * H, np.float64)
    cnt = None if counts is None else np.ascontiguousarray(counts.reshape(-1), np.int32)
    lib().orc_moments_exact(W, H, S, lam, np.ascontiguousarray(samples_arr.reshape(spp, 5, -1), np.float32), spp,
                            None if cnt is None else cnt.ctypes.data, int(sanitize), out)
    return out.reshape(S, 4, 3, H, W)


def prepass(n: np.ndarray, stats: np.ndarray, z: float):
    """FP32 statistics (the GPU's, or the reference recurrence's) or FP64
    ones (the full-FP64 pipeline: accumulate(..., fp64=True))."""
    S = stats.shape[0]
    H, W = n.shape
    P = W * H
    jb = np.zeros(S * P)
    l, s, m = (np.zeros(S * 3 * P) for _ in range(3))
    f64 = stats.dtype == np.float64
    fn = lib().orc_prepass64 if f64 else lib().orc_prepass
    fn(P, S, np.ascontiguousarray(n, np.int32).reshape(-1),
       np.ascontiguousarray(stats, np.float64 if f64 else np.float32).reshape(-1), z, jb, l, s, m)
    return jb.reshape(S, H, W), l.reshape(S, 3, H, W), s.reshape(S, 3, H, W), m.reshape(S, 3, H, W)


def _feat_args(feats):
    """feats: list of (array [ch,H,W] float32, sd)."""
    if not feats:
        return 0, None, None, None, []
    keep = [np.ascontiguousarray(a, np.float32) for a, _ in feats]
    ptrs = (C.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
    chs = (C.c_int * len(keep))(*[k.shape[0] for k in keep])
    drs = (C.c_double * len(keep))(*[-0.5 / (sd * sd) for _, sd in feats])
    return len(keep), C.cast(ptrs, C.c_void_p), C.cast(chs, C.c_void_p), C.cast(drs, C.c_void_p), [keep, ptrs, chs,
                                                                                                     drs]


def denoise_vars(n, m2, lnS2, se2, m2C, radius, ds_factor, z, feats=None, nthreads: int = 8, want_count=False,
                 want_ambiguous=False):
    """All per-set arrays [S,3,H,W], float32 (or all float64: the full-FP64
    pipeline).  Returns varD, varCD (FP64) [, accepted count] [, count of
    decisions within 1e-5 of the threshold]."""
    S = m2.shape[0]
    H, W = n.shape
    P = W * H
    ng, gp, gc, gd, keep = _feat_args(feats)
    vd, vcd = np.zeros(S * 3 * P), np.zeros(S * 3 * P)
    cnt = np.zeros(S * 3 * P, np.int32) if want_count else None
    amb = np.zeros(S * 3 * P, np.int32) if want_ambiguous else None
    f64 = np.asarray(m2).dtype == np.float64
    dt = np.float64 if f64 else np.float32
    f = lambda a: np.ascontiguousarray(a, dt).reshape(-1)
    fn = lib().orc_denoise_vars64 if f64 else lib().orc_denoise_vars
    fn(W, H, S, np.ascontiguousarray(n, np.int32).reshape(-1), f(m2), f(lnS2), f(se2), f(m2C),
                           radius, ds_factor, z, ng, gp, gc, gd, vd, vcd,
                           None if cnt is None else cnt.ctypes.data, None if amb is None else amb.ctypes.data,
                           nthreads)
    out = (vd.reshape(S, 3, H, W), vcd.reshape(S, 3, H, W))
    out = out + ((cnt.reshape(S, 3, H, W),) if want_count else ())
    return out + ((amb.reshape(S, 3, H, W),) if want_ambiguous else ())


def var_decisions(n, lnS2, se2, radius, z) -> np.ndarray:
    """The variance filter's decisions [S,3,H,W,win,win] (uint8, 1 = accepted)
    for FP32 or FP64 lnS2/se2 [S,3,H,W]; FP32 inputs are decided with the FP32
    path's own arithmetic (Z = z^2 se2 rounded, reject iff fma(dl, dl, -Z_q) >
    Z_p), so a pair within rounding of the threshold goes the way the kernel
    takes it."""
    S = lnS2.shape[0]
    H, W = n.shape
    win = 2 * radius + 1
    out = np.zeros(S * 3 * H * W * win * win, np.uint8)
    f64 = np.asarray(lnS2).dtype == np.float64
    dt = np.float64 if f64 else np.float32
    (lib().orc_var_decisions64 if f64 else lib().orc_var_decisions)(
        W, H, S, np.ascontiguousarray(n, np.int32).reshape(-1), np.ascontiguousarray(lnS2, dt).reshape(-1),
        np.ascontiguousarray(se2, dt).reshape(-1), radius, z, out)
    return out.reshape(S, 3, H, W, win, win)


def denoise_film(n, mean, var, film, radius, ds_factor, alpha, mt_mode=2, jb=None, feats=None, want_varw=False,
                 want_decisions=False, nthreads: int = 8):
    """mean/var [S,3,H,W] float32 (set 0 mean = film); film [3,H,W].  With
    float64 mean the inputs are taken in FP64 (the full-FP64 pipeline).
    Returns dict."""
    S = mean.shape[0]
    H, W = n.shape
    P = W * H
    ng, gp, gc, gd, keep = _feat_args(feats)
    filmD = np.zeros(3 * P)
    varw = np.zeros(3 * P) if want_varw else None
    win = 2 * radius + 1
    dec = np.zeros(P * win * win * 3, np.uint8) if want_decisions else None
    f64 = np.asarray(mean).dtype == np.float64
    dt = np.float64 if f64 else np.float32
    jbc = None if jb is None else np.ascontiguousarray(jb, dt).reshape(-1)
    f = lambda a: np.ascontiguousarray(a, dt).reshape(-1)
    (lib().orc_denoise_film64 if f64 else lib().orc_denoise_film)(W, H, S, np.ascontiguousarray(n, np.int32).reshape(-1), f(mean), f(var),
                           None if jbc is None else jbc.ctypes.data, radius, ds_factor, alpha, mt_mode, ng, gp, gc,
                           gd, f(film), filmD, None if varw is None else varw.ctypes.data,
                           None if dec is None else dec.ctypes.data, None, nthreads)
    out = {"filmD": filmD.reshape(3, H, W)}
    if want_varw:
        out["varW"] = varw.reshape(3, H, W)
    if want_decisions:
        out["decisions"] = dec.reshape(H, W, win, win, 3)
    return out


def target_spp(n, varD0, mean0, budget_spp: int, cap: int) -> np.ndarray:
    H, W = n.shape
    out = np.zeros(W * H, np.int32)
    lib().orc_target_spp(W * H, np.ascontiguousarray(n, np.int32).reshape(-1),
                         np.ascontiguousarray(varD0, np.float32).reshape(-1),
                         np.ascontiguousarray(mean0, np.float32).reshape(-1), budget_spp, cap, out)
    return out.reshape(H, W)


def mean_var(n, m2_0) -> np.ndarray:
    H, W = n.shape
    out = np.zeros(3 * W * H, np.float32)
    lib().orc_mean_var(W * H, np.ascontiguousarray(n, np.int32).reshape(-1),
                       np.ascontiguousarray(m2_0, np.float32).reshape(-1), out)
    return out.reshape(3, H, W)


def pipeline64(samples_arr: np.ndarray, lambdas, radius: int, ds_factor: float, z: float, alpha: float,
               feats=None, budget_spp: int = 32, cap: int = 4096, sanitize=True, film=True, mt_mode=2):
    """The whole path in FP64 from the same samples [spp,5,H,W]: FP64
    recurrence with the FP64 Box-Cox, FP64 prepass, variance filter and Welch
    filter (of the FP64 pbrt Film, or of the set-0 mean), and the AS map from
    the FP64 varD and mean (rounded to FP32 for its fixed FP32 metric)."""
    n, st = accumulate(samples_arr, lambdas, 5, fp64=True, sanitize=sanitize)
    jb, lnS2, se2, m2C = prepass(n, st, z)
    vD, vCD = denoise_vars(n, st[:, 1], lnS2, se2, m2C, radius, ds_factor, z, feats)
    fm = film_image64(samples_arr, sanitize=sanitize) if film else st[0, 0]
    mean = st[:, 0].copy()
    mean[0] = fm
    out = denoise_film(n, mean, vCD, fm, radius, ds_factor, alpha, mt_mode=mt_mode, jb=jb, feats=feats,
                       want_decisions=True)
    tgt = target_spp(n, vD[0].astype(np.float32), st[0, 0].astype(np.float32), budget_spp, cap)
    return {"n": n, "stats": st, "jb": jb, "lnS2": lnS2, "se2": se2, "m2C": m2C, "varD": vD, "varCD": vCD,
            "film": fm, "filmD": out["filmD"], "decisions": out["decisions"], "targetspp": tgt}


def fp64_error_report(gpu: dict, ref: dict, z: float, radius: int, crop: bool = False) -> dict:
    """Per-buffer error of the GPU's FP32 path against the full-FP64 pipeline
    (pipeline64) on the same pixels.

    gpu: n [H,W], lnS2, se2, varD, varCD [S,3,H,W], filmD [3,H,W], welch
    decisions [H,W,win,win,3] and optionally targetspp [H,W].  crop: the
    arrays are a crop of a larger frame, whose FP64 pipeline ran on the crop
    alone; then the variance outputs are compared where their window lies in
    the crop (a margin of radius) and the Welch outputs where the windows of
    their neighbours' variances do too (a margin of 2 radius: the Welch test
    reads the neighbours' filtered variances).  For each buffer: the fraction of
    accept/reject decisions that differ (variance filter: the decisions the
    oracle takes on the GPU's FP32 lnS2/se2, which the kernel reproduces, vs
    the FP64 ones), the largest and the 99.9th-percentile relative error
    over pixel-channels whose decisions all agree, and how many of those
    there are."""
    m = radius if crop else 0
    H, W = gpu["n"].shape
    sy, sx = slice(m, H - m), slice(m, W - m)
    out = {}
    d32 = var_decisions(gpu["n"], gpu["lnS2"], gpu["se2"], radius, z)[:, :, sy, sx]
    d64 = var_decisions(ref["n"], ref["lnS2"], ref["se2"], radius, z)[:, :, sy, sx]
    vdiff = d32 != d64
    vsame = ~np.any(vdiff, axis=(4, 5))
    for name in ("varD", "varCD"):
        g = np.asarray(gpu[name], np.float64)[:, :, sy, sx]
        r = np.asarray(ref[name], np.float64)[:, :, sy, sx]
        ok = vsame & np.isfinite(r) & (np.abs(r) < 1e30)
        rel = np.abs(g[ok] - r[ok]) / np.maximum(np.abs(r[ok]), 1e-30)
        out[name] = {"decisions_differ": float(np.count_nonzero(vdiff) / vdiff.size),
                     "max_rel": float(rel.max()) if rel.size else None,
                     "p999_rel": float(np.percentile(rel, 99.9)) if rel.size else None,
                     "agreeing": float(np.count_nonzero(ok) / ok.size)}
    sy, sx = slice(2 * m, H - 2 * m), slice(2 * m, W - 2 * m)
    wd = gpu["decisions"][sy, sx] != ref["decisions"][sy, sx]
    wsame = ~np.any(wd, axis=(2, 3)).transpose(2, 0, 1)
    g = np.asarray(gpu["filmD"], np.float64)[:, sy, sx]
    r = np.asarray(ref["filmD"], np.float64)[:, sy, sx]
    rel = np.abs(g[wsame] - r[wsame]) / np.maximum(np.abs(r[wsame]), 1e-6)
    out["filmD"] = {"decisions_differ": float(np.count_nonzero(wd) / wd.size),
                    "max_rel": float(rel.max()) if rel.size else None,
                    "p999_rel": float(np.percentile(rel, 99.9)) if rel.size else None,
                    "agreeing": float(np.count_nonzero(wsame) / wsame.size)}
    if "targetspp" in gpu:
        tg, tr = gpu["targetspp"][sy, sx], ref["targetspp"][sy, sx]
        out["targetspp"] = {"pixels_differ": float(np.count_nonzero(tg != tr) / tg.size),
                            "max_abs": int(np.max(np.abs(tg.astype(np.int64) - tr))),
                            "sum_gpu": int(tg.sum(dtype=np.int64)), "sum_fp64": int(tr.sum(dtype=np.int64))}
    return out
