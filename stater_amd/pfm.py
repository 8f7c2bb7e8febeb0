"""PFM buffer interchange in the reference's on-disk format, and the offline
denoise of dumped buffers.

Files are ``<stem>-<spp>-<name>.pfm`` with the reference's buffer names
(src/statistics/buffer.cpp:44-57 writes them; the --denoise mode of
src/statistics/statpath.cpp:486-560 reads them back, denoises and writes the
selected outputs).  G-buffers are the film means of their statistic types:
``t4-b0-film-mean`` (depth), ``t5-b0-film-mean`` (normal) and
``t6-b0-film-mean`` (albedo) (type numbering of src/statistics/statpath.h:28-37).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import (BufferSet, Context, check, default_params, gbuffers, load_library, plane)
from . import hipmem

# statistic type of each feature plane in the files, in the filters' order
GBUFFER_TYPES = (("albedo", 6, 3), ("normal", 5, 3), ("depth", 4, 1))


def info(path: str | Path) -> tuple[int, int, int]:
    """(width, height, channels) of a PFM file."""
    w, h, c = C.c_int(), C.c_int(), C.c_int()
    check(load_library().stater_pfm_info(str(path).encode(), C.byref(w), C.byref(h), C.byref(c)))
    return w.value, h.value, c.value


def read(path: str | Path) -> np.ndarray:
    """A PFM file as float32 [channels, H, W] (top row first)."""
    w, h, c = info(path)
    buf = np.empty((h, w, c), np.float32)
    check(load_library().stater_pfm_read_host(str(path).encode(), buf.ctypes.data, w, h, c))
    return np.ascontiguousarray(buf.transpose(2, 0, 1))


def write(path: str | Path, arr: np.ndarray) -> None:
    """Write float32 [channels, H, W] (1 or 3 channels, top row first)."""
    a = np.asarray(arr, np.float32)
    if a.ndim == 2:
        a = a[None]
    c, h, w = a.shape
    hwc = np.ascontiguousarray(a.transpose(1, 2, 0))
    check(load_library().stater_pfm_write_host(str(path).encode(), hwc.ctypes.data, w, h, c))


def write_set(bs: BufferSet, stem: str | Path, spp: int, regex: str = "film.*", stream=None) -> int:
    """Write the set's buffers whose names match `regex` (the reference's
    `outputregex`, default ``film.*``); returns the number of files."""
    n = C.c_int()
    check(load_library().stater_bufset_write_pfm(bs.h, str(stem).encode(), spp, regex.encode(), stream,
                                                 C.byref(n)))
    return n.value


def read_set(bs: BufferSet, stem: str | Path, spp: int, stream=None) -> int:
    """Read every ``<stem>-<spp>-<name>.pfm`` that names one of the set's
    buffers into it; returns the number of files read."""
    n = C.c_int()
    check(load_library().stater_bufset_read_pfm(bs.h, str(stem).encode(), spp, stream, C.byref(n)))
    return n.value


def file_name(stem: str | Path, spp: int, name: str) -> Path:
    return Path(f"{stem}-{spp}-{name}.pfm")


class Dumps:
    """Device planes for a film and its G-buffers read from dumped files
    (what StatPathIntegrator::Denoise finds next to the statistics)."""

    def __init__(self, stem: str | Path, spp: int):
        self.stem, self.spp = str(stem), spp
        f = file_name(stem, spp, "film")
        if not f.exists():
            f = file_name(stem, spp, "t0-b0-film-mean")
        self.width, self.height, _ = info(f)
        W, H = self.width, self.height
        self.ld = (W + 63) // 64 * 64
        self.plane_bytes = self.ld * 4 * H
        self.buf = hipmem.DeviceBuffer(self.plane_bytes * 10)
        p = self.buf.ptr
        self.film = plane(p, W, H, 3, self.ld * 4, self.plane_bytes)
        self.has_film = file_name(stem, spp, "film").exists()
        if self.has_film:
            check(load_library().stater_pfm_read(str(file_name(stem, spp, "film")).encode(), C.byref(self.film),
                                                 None))
        self.features = []
        off = 3
        for name, t, ch in GBUFFER_TYPES:
            path = file_name(stem, spp, f"t{t}-b0-film-mean")
            if path.exists():
                pl = plane(p + off * self.plane_bytes, W, H, ch, self.ld * 4, self.plane_bytes)
                check(load_library().stater_pfm_read(str(path).encode(), C.byref(pl), None))
                self.features.append((name, pl))
            off += ch

    def gbuffers(self, sds: dict[str, float]):
        """The filters' G-buffers from the files found, sds by feature name."""
        if not self.features:
            return None
        return gbuffers([pl for _, pl in self.features], [sds[n] for n, _ in self.features])


def denoise_dumps(ctx: Context, stem: str | Path, spp: int, lambdas=(), max_moment: int = 5, params=None,
                  sds: dict[str, float] | None = None, out_stem: str | Path | None = None,
                  out_regex: str = "filmD") -> BufferSet:
    """The reference's --denoise pass for one iteration: read the dumped
    statistics, film and G-buffers, run DenoiseVars and DenoiseMeans, and
    (with out_stem) write the buffers matching out_regex.  Returns the set."""
    from .workload import FEATURE_SDS
    d = Dumps(stem, spp)
    bs = BufferSet(ctx, d.width, d.height, lambdas, max_moment)
    read_set(bs, stem, spp)
    p = params or default_params()
    g = d.gbuffers(sds or {"albedo": FEATURE_SDS[0], "normal": FEATURE_SDS[1], "depth": FEATURE_SDS[2]})
    bs.denoise_vars(p, g)
    bs.denoise_film(p, g, film=d.film if d.has_film else None)
    ctx.synchronize()
    bs._dumps = d  # the film and feature planes live as long as the set
    if out_stem is not None:
        write_set(bs, out_stem, spp, out_regex)
    return bs
