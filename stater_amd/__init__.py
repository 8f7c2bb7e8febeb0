"""stater_amd — MI355X (gfx950) statistics-and-denoise path of StatER.

The product is ``libstater.so`` (HIP kernels behind the C-ABI declared in
``include/stater.h``).  This module binds that C-ABI with ctypes and offers a
small object layer (``Context``, ``BufferSet``) whose methods mirror the
reference's ``Estimator`` (src/statistics/estimator.h:333-520): accumulate,
``denoise_vars`` (DenoiseVars), ``denoise_film`` (DenoiseMeans),
``calc_target_spp`` (CalcTargetSPPs), ``calc_mean_vars`` (CalcMeanVars) and
``synchronize`` (Synchronize).  Device memory for sample chunks and feature
planes is supplied by the caller as raw device pointers (tests and the bench
use torch tensors for that).

There is no fallback: if ``libstater.so`` is missing, importing the binding
raises.
"""
from __future__ import annotations

import ctypes as C
import weakref
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libstater.so"

STATER_OK = 0
(STATER_E_ARG, STATER_E_HIP, STATER_E_NOMEM, STATER_E_STATE, STATER_E_UNSUPPORTED,
 STATER_E_NOTFOUND) = (-1, -2, -3, -4, -5, -6)
STATER_F32, STATER_I32, STATER_BF16 = 0, 1, 2
STATER_PLANAR, STATER_INTERLEAVED = 0, 1
STATER_BUFSET_NO_DERIVED, STATER_BUFSET_FILM = 1, 2
STATER_CTX_TIMING = 1
STATER_ACC_SEQUENTIAL, STATER_ACC_MERGE64, STATER_ACC_FRESH = 0, 1, 16
MAX_LAMBDAS, MAX_RADIUS, MAX_GBUFFERS, MAX_EDGES = 8, 20, 4, 16


class Plane(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("pitch_bytes", C.c_int64), ("channel_stride_bytes", C.c_int64),
                ("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32),
                ("dtype", C.c_int32), ("layout", C.c_int32), ("reserved", C.c_int32)]


class Samples(C.Structure):
    _fields_ = [("data", C.c_void_p), ("pitch_bytes", C.c_int64), ("comp_stride_bytes", C.c_int64),
                ("sample_stride_bytes", C.c_int64), ("spp", C.c_int32), ("sanitize", C.c_int32),
                ("counts", C.c_void_p), ("counts_pitch_bytes", C.c_int64), ("count_offset", C.c_int32),
                ("reserved", C.c_int32)]


class DenoiseParams(C.Structure):
    _fields_ = [("radius", C.c_int32), ("ds_factor", C.c_float), ("var_z", C.c_float), ("alpha", C.c_float),
                ("mt_mode", C.c_int32), ("cascade", C.c_int32), ("use_levc", C.c_int32), ("reserved", C.c_int32)]


class GBuffers(C.Structure):
    _fields_ = [("count", C.c_int32), ("reserved", C.c_int32), ("planes", Plane * MAX_GBUFFERS),
                ("dr_factors", C.c_float * MAX_GBUFFERS)]


class BatchImage(C.Structure):
    _fields_ = [("set", C.c_void_p), ("scene", C.c_void_p), ("albedo", C.c_void_p), ("gbufs", C.c_void_p),
                ("spp", C.c_int32), ("reserved", C.c_int32), ("samples", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("max_bytes", C.c_uint64), ("slots", C.c_int32), ("chunk_spp", C.c_int32),
                ("max_width", C.c_int32), ("max_height", C.c_int32), ("acc_mode", C.c_int32),
                ("feature_sds", C.c_float * 3), ("params", DenoiseParams), ("reserved", C.c_int32 * 4)]


class Job(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("n_lambdas", C.c_int32),
                ("lambdas", C.c_float * MAX_LAMBDAS), ("spp", C.c_int32), ("rounds", C.c_int32),
                ("cap_spp", C.c_uint32), ("scene_model", C.c_int32), ("seed", C.c_uint32), ("image_id", C.c_uint32),
                ("feature_dtype", C.c_int32), ("keep", C.c_int32), ("film", C.c_int32),
                ("next_spp", C.c_int32), ("continue_id", C.c_uint64), ("samples", C.c_void_p),
                ("gbufs", C.c_void_p), ("reserved", C.c_int32 * 2)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("slot", C.c_int32), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("samples", C.c_uint64), ("bytes", C.c_uint64)]


class Scene(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("seed", C.c_uint32), ("image_id", C.c_uint32),
                ("model", C.c_int32), ("n_edges", C.c_int32), ("base", C.c_double),
                ("amp", (C.c_double * 3) * 3), ("fx", C.c_double * 3), ("fy", C.c_double * 3),
                ("cos_phi", (C.c_double * 3) * 3), ("sin_phi", (C.c_double * 3) * 3),
                ("edge_vertical", C.c_int32 * MAX_EDGES), ("edge_pos", C.c_double * MAX_EDGES),
                ("edge_factor", C.c_double * MAX_EDGES), ("s_min", C.c_double), ("s_max", C.c_double),
                ("firefly_mult", C.c_double), ("firefly_thresh", C.c_uint32), ("albedo_modulation", C.c_int32),
                ("cell_size", C.c_int32), ("pad0", C.c_int32), ("normal_jitter", C.c_double),
                ("depth_min", C.c_double), ("depth_max", C.c_double)]


# (name, restype, argtypes) of every symbol include/stater.h declares.
_P = C.c_void_p
SIGNATURES = [
    ("stater_create", C.c_int, [C.c_int, C.c_uint, C.POINTER(_P)]),
    ("stater_destroy", C.c_int, [_P]),
    ("stater_ctx_stream", C.c_int, [_P, C.POINTER(_P)]),
    ("stater_last_error", C.c_char_p, []),
    ("stater_abi_version", C.c_int, []),
    ("stater_print_macro_configuration", None, []),
    ("stater_timing_reset", C.c_int, [_P]),
    ("stater_timing_count", C.c_int, [_P]),
    ("stater_timing_get", C.c_int, [_P, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_double)]),
    ("stater_bufset_create", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_uint,
                                       C.POINTER(_P)]),
    ("stater_bufset_create_in", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_uint,
                                          _P, C.c_uint64, _P, C.POINTER(_P)]),
    ("stater_bufset_destroy", C.c_int, [_P]),
    ("stater_convert_plane", C.c_int, [_P, C.POINTER(Plane), C.POINTER(Plane), _P]),
    ("stater_bufset_reset", C.c_int, [_P, _P]),
    ("stater_bufset_set_stat_type", C.c_int, [_P, C.c_int]),
    ("stater_bufset_view", C.c_int, [_P, C.c_char_p, C.POINTER(Plane)]),
    ("stater_bufset_bytes", C.c_int64, [_P]),
    ("stater_bufset_bytes_for", C.c_int64, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_uint]),
    ("stater_film_resolve", C.c_int, [_P, C.c_float, C.POINTER(Plane), _P]),
    ("stater_pfm_info", C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("stater_pfm_read_host", C.c_int, [C.c_char_p, _P, C.c_int, C.c_int, C.c_int]),
    ("stater_pfm_write_host", C.c_int, [C.c_char_p, _P, C.c_int, C.c_int, C.c_int]),
    ("stater_pfm_read", C.c_int, [C.c_char_p, C.POINTER(Plane), _P]),
    ("stater_pfm_write", C.c_int, [C.c_char_p, C.POINTER(Plane), _P]),
    ("stater_bufset_write_pfm", C.c_int, [_P, C.c_char_p, C.c_uint64, C.c_char_p, _P, C.POINTER(C.c_int)]),
    ("stater_bufset_read_pfm", C.c_int, [_P, C.c_char_p, C.c_uint64, _P, C.POINTER(C.c_int)]),
    ("stater_accumulate", C.c_int, [_P, C.POINTER(Samples), C.c_int, _P]),
    ("stater_default_params", None, [C.POINTER(DenoiseParams)]),
    ("stater_denoise_vars", C.c_int, [_P, C.POINTER(DenoiseParams), C.POINTER(GBuffers), _P]),
    ("stater_denoise_film", C.c_int, [_P, C.POINTER(DenoiseParams), C.POINTER(GBuffers), C.POINTER(Plane),
                                      C.POINTER(Plane), _P]),
    ("stater_denoise_plain", C.c_int, [_P, C.c_int, C.c_float, _P]),
    ("stater_calc_target_spp", C.c_int, [_P, C.c_uint64, C.c_uint32, C.POINTER(Plane), _P]),
    ("stater_calc_target_spp_ex", C.c_int, [_P, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(Plane), _P]),
    ("stater_calc_target_spp_noalloc", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Plane), C.POINTER(Plane),
                                                 C.POINTER(Plane), C.c_uint64, C.c_uint32, C.POINTER(Plane), _P]),
    ("stater_calc_mean_vars", C.c_int, [_P, _P]),
    ("stater_synchronize", C.c_int, [_P]),
    ("stater_scene_init", C.c_int, [C.POINTER(Scene), C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_int]),
    ("stater_synth_features", C.c_int, [_P, C.POINTER(Scene), C.POINTER(Plane), C.POINTER(Plane),
                                        C.POINTER(Plane), _P]),
    ("stater_synth_samples", C.c_int, [_P, C.POINTER(Scene), C.POINTER(Plane), C.c_uint32, C.POINTER(Plane),
                                       C.POINTER(Samples), _P]),
    ("stater_batch_create", C.c_int, [_P, C.c_int, C.POINTER(BatchImage), C.POINTER(DenoiseParams), C.c_int,
                                      C.POINTER(_P)]),
    ("stater_batch_tune", C.c_int, [_P, _P]),
    ("stater_batch_run", C.c_int, [_P, C.c_int, _P]),
    ("stater_batch_info", C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    ("stater_batch_destroy", C.c_int, [_P]),
    ("stater_debug_film_decisions", C.c_int, [_P, C.POINTER(DenoiseParams), C.POINTER(GBuffers),
                                              C.POINTER(Plane), C.c_int, _P, _P]),
    ("stater_debug_exact_forms", C.c_int, [_P, _P, _P]),
    ("stater_debug_quotient_sweep", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    ("stater_debug_valu_rate", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("stater_server_default_config", None, [C.POINTER(ServerConfig)]),
    ("stater_server_create", C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("stater_server_job_bytes", C.c_int64, [_P, C.POINTER(Job)]),
    ("stater_server_submit", C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    ("stater_server_wait", C.c_int, [_P, C.c_uint64, C.POINTER(JobResult)]),
    ("stater_server_view", C.c_int, [_P, C.c_uint64, C.c_char_p, C.POINTER(Plane)]),
    ("stater_server_release", C.c_int, [_P, C.c_uint64]),
    ("stater_server_drain", C.c_int, [_P]),
    ("stater_server_clock", C.c_double, [_P]),
    ("stater_server_destroy", C.c_int, [_P]),
]

_lib = None


class StaterError(RuntimeError):
    pass


def load_library(path: str | Path | None = None) -> C.CDLL:
    """Load libstater.so and bind every C-ABI symbol (raises if it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise StaterError(f"{p} not built: run `python -m stater_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != STATER_OK:
        msg = load_library().stater_last_error()
        raise StaterError(f"stater error {rc}: {msg.decode() if msg else ''}")


def scene(model: int, seed: int, image_id: int, width: int, height: int) -> Scene:
    s = Scene()
    check(load_library().stater_scene_init(C.byref(s), model, seed, image_id, width, height))
    return s


def plane(ptr: int, width: int, height: int, channels: int, pitch_bytes: int, channel_stride_bytes: int,
          dtype: int = STATER_F32, layout: int = STATER_PLANAR) -> Plane:
    return Plane(ptr, pitch_bytes, channel_stride_bytes, width, height, channels, dtype, layout, 0)


def default_params(**kw) -> DenoiseParams:
    p = DenoiseParams()
    load_library().stater_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gbuffers(planes: list[Plane], sds: list[float]) -> GBuffers:
    """Feature buffers with their range standard deviations (dr = -1/(2 sd^2))."""
    g = GBuffers()
    g.count = len(planes)
    for i, (pl, sd) in enumerate(zip(planes, sds)):
        g.planes[i] = pl
        g.dr_factors[i] = -0.5 / (sd * sd)
    return g


class Context:
    """A device and its stream (the reference's stat_denoiser::setup())."""

    def __init__(self, device: int = 0, timing: bool = False):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib.stater_create(device, STATER_CTX_TIMING if timing else 0, C.byref(h)))
        self.h = h
        s = C.c_void_p()
        check(self.lib.stater_ctx_stream(self.h, C.byref(s)))
        self.stream = s
        self._sets = weakref.WeakSet()
        self._users = weakref.WeakSet()  # batches and servers: closed before the sets they use

    def close(self) -> None:
        """Destroys the context, closing its remaining servers, batches and
        buffer sets first."""
        if self.h:
            for u in list(self._users):
                u.close()
            for bs in list(self._sets):
                bs.close()
            check(self.lib.stater_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(self.lib.stater_synchronize(self.stream))

    def timing(self) -> dict[str, tuple[int, float]]:
        out = {}
        n = self.lib.stater_timing_count(self.h)
        if n < 0:
            check(n)
        buf = C.create_string_buffer(128)
        for i in range(n):
            cnt, ms = C.c_uint64(), C.c_double()
            check(self.lib.stater_timing_get(self.h, i, buf, 128, C.byref(cnt), C.byref(ms)))
            out[buf.value.decode()] = (cnt.value, ms.value)
        return out

    def timing_reset(self) -> None:
        check(self.lib.stater_timing_reset(self.h))


class BufferSet:
    """Device-resident statistics buffers of one film (the Estimator's buffers)."""

    def __init__(self, ctx: Context, width: int, height: int, lambdas=(), max_moment: int = 5,
                 flags: int = 0, stat_type: int = 0):
        """stat_type: the reference's StatTypeIndex its buffer names carry
        (0 radiance; 4 depth, 5 normal, 6 albedo for G-buffer sets)."""
        self.ctx, self.lib = ctx, ctx.lib
        self.width, self.height = width, height
        self.lambdas = list(lambdas)
        arr = (C.c_float * max(1, len(self.lambdas)))(*self.lambdas)
        h = C.c_void_p()
        check(self.lib.stater_bufset_create(ctx.h, width, height, len(self.lambdas), arr, max_moment, flags,
                                            C.byref(h)))
        self.h = h
        ctx._sets.add(self)
        if stat_type:
            check(self.lib.stater_bufset_set_stat_type(self.h, stat_type))

    def close(self) -> None:
        if self.h:
            self.lib.stater_bufset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbytes(self) -> int:
        return self.lib.stater_bufset_bytes(self.h)

    def view(self, name: str) -> Plane:
        p = Plane()
        check(self.lib.stater_bufset_view(self.h, name.encode(), C.byref(p)))
        return p

    # Every call enqueues on `stream` (a hipStream_t handle; None: the context's stream).
    def reset(self, stream=None) -> None:
        check(self.lib.stater_bufset_reset(self.h, stream))

    def accumulate(self, samples: Samples, mode: int = STATER_ACC_SEQUENTIAL, stream=None) -> None:
        check(self.lib.stater_accumulate(self.h, C.byref(samples), mode, stream))

    def denoise_vars(self, params: DenoiseParams, gbufs: GBuffers | None = None, stream=None) -> None:
        check(self.lib.stater_denoise_vars(self.h, C.byref(params), C.byref(gbufs) if gbufs else None, stream))

    def denoise_film(self, params: DenoiseParams, gbufs: GBuffers | None = None, film: Plane | None = None,
                     film_out: Plane | None = None, stream=None) -> None:
        check(self.lib.stater_denoise_film(self.h, C.byref(params), C.byref(gbufs) if gbufs else None,
                                           C.byref(film) if film else None, C.byref(film_out) if film_out else None,
                                           stream))

    def denoise_plain(self, radius: int = 10, alpha: float = 0.005, stream=None) -> None:
        check(self.lib.stater_denoise_plain(self.h, radius, alpha, stream))

    def calc_target_spp(self, budget_spp: int, cap: int, out: Plane | None = None, stream=None,
                        as_mode: int = 1) -> None:
        """as_mode 1: from the denoised variance varD; 0: from the sample variance m2/(n-1)."""
        check(self.lib.stater_calc_target_spp_ex(self.h, budget_spp, cap, as_mode, C.byref(out) if out else None,
                                                 stream))

    def calc_mean_vars(self, stream=None) -> None:
        check(self.lib.stater_calc_mean_vars(self.h, stream))

    def film_resolve(self, scale: float = 1.0, out: Plane | None = None, stream=None) -> None:
        """Film::UpdateImage into out (None: the set's "film"); sets created with STATER_BUFSET_FILM."""
        check(self.lib.stater_film_resolve(self.h, scale, C.byref(out) if out else None, stream))


class Batch:
    """Uneven films processed together over several streams (optionally one hipGraph).

    images: list of (BufferSet, Scene | Samples, albedo Plane | None, GBuffers | None, spp): a
    seeded scene the batch generates samples from, or a caller's sample chunk
    (Samples, device memory kept valid while the batch runs)."""

    def __init__(self, ctx: Context, images, params: DenoiseParams, chunk_spp: int = 32):
        self.ctx, self.lib = ctx, ctx.lib
        self._keep = []
        arr = (BatchImage * len(images))()
        for i, (bs, src, alb, g, spp) in enumerate(images):
            self._keep += [src, alb, g]
            caller = isinstance(src, Samples)
            arr[i] = BatchImage(bs.h.value, None if caller else C.addressof(src),
                                C.addressof(alb) if alb is not None else None,
                                C.addressof(g) if g is not None else None, spp, 0,
                                C.addressof(src) if caller else None)
        h = C.c_void_p()
        check(self.lib.stater_batch_create(ctx.h, len(images), arr, C.byref(params), chunk_spp, C.byref(h)))
        self.h = h
        ctx._users.add(self)

    def tune(self) -> None:
        check(self.lib.stater_batch_tune(self.h, None))

    def run(self, use_graph: bool = True) -> None:
        check(self.lib.stater_batch_run(self.h, 1 if use_graph else 0, None))

    def images(self) -> list[dict]:
        """Per film: statistics sets, stream (branch) and the tuned tile heights."""
        n = self.lib.stater_batch_info(self.h, -1, None, None, None, None)
        out = []
        for i in range(n):
            s, br, tv, tw = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            check(self.lib.stater_batch_info(self.h, i, C.byref(s), C.byref(br), C.byref(tv), C.byref(tw)))
            out.append({"sets": s.value, "branch": br.value, "tile_rows_var": tv.value, "tile_rows_welch": tw.value})
        return out

    def close(self) -> None:
        if self.h:
            self.lib.stater_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
