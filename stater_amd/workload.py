"""Seeded synthetic workloads on the device (inputs of the hot path).

``SampleChunk`` is a device buffer in the layout ``stater_samples`` describes
(sample-major planes R, G, B, u, v with a 256-byte row pitch) and
``Features`` holds the albedo / normal / depth planes of a scene.  Both are
filled by the library's own generator kernels (stater_synth_*).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (STATER_F32, STATER_PLANAR, Context, Plane, Samples, check, gbuffers, load_library, plane,
               scene)
from .hipmem import DeviceBuffer

# Reference range standard deviations: albedo 0.02, normal 0.1
# (scenes/render-denoise*.pbrt:21-24); depth (cell values in [1, 100]) 0.5.
FEATURE_SDS = (0.02, 0.1, 0.5)
MT5_LAMBDAS = (0.01, 0.25, 0.5, 0.75, 1.0, 3.0)  # scenes/render-denoise-mt5.pbrt:26


def row_pitch(width: int) -> int:
    return (width + 63) // 64 * 64


class SampleChunk:
    def __init__(self, width: int, height: int, spp: int, buf: DeviceBuffer | None = None):
        """A chunk of spp samples for a width x height film; buf (optional)
        lends the memory of a larger chunk."""
        self.width, self.height, self.spp = width, height, spp
        self.ld = row_pitch(width)
        self.plane_bytes = self.ld * height * 4
        if buf is not None and buf.nbytes < self.plane_bytes * 5 * spp:
            raise ValueError("SampleChunk: lent buffer too small")
        self.buf = buf if buf is not None else DeviceBuffer(self.plane_bytes * 5 * spp)
        self.samples = Samples(self.buf.ptr, self.ld * 4, self.plane_bytes, 5 * self.plane_bytes, spp, 0, None, 0,
                               0, 0)

    def desc(self, spp: int | None = None, sanitize: bool = False, counts: Plane | None = None,
             count_offset: int = 0) -> Samples:
        s = Samples.from_buffer_copy(self.samples)
        if spp is not None:
            s.spp = spp
        s.sanitize = 1 if sanitize else 0
        if counts is not None:
            s.counts = counts.ptr
            s.counts_pitch_bytes = counts.pitch_bytes
            s.count_offset = count_offset
        return s

    def to_numpy(self, rows: tuple[int, int] | None = None) -> np.ndarray:
        """[spp, 5, H, W] float32, or only rows y0 <= y < y1 ([spp, 5, y1 - y0, W])."""
        if rows is None:
            flat = self.buf.download(np.float32, self.plane_bytes * 5 * self.spp // 4)
            return flat.reshape(self.spp, 5, self.height, self.ld)[..., : self.width].copy()
        y0, y1 = rows
        out = np.empty((self.spp, 5, y1 - y0, self.width), np.float32)
        for j in range(self.spp):
            for k in range(5):
                off = (j * 5 + k) * self.plane_bytes + y0 * self.ld * 4
                part = self.buf.download(np.float32, (y1 - y0) * self.ld, offset=off)
                out[j, k] = part.reshape(y1 - y0, self.ld)[:, : self.width]
        return out

    def from_numpy(self, arr: np.ndarray) -> None:
        a = np.zeros((self.spp, 5, self.height, self.ld), np.float32)
        a[..., : self.width] = arr
        self.buf.upload(a)


class Features:
    def __init__(self, width: int, height: int):
        self.width, self.height = width, height
        self.ld = row_pitch(width)
        self.plane_bytes = self.ld * height * 4
        self.buf = DeviceBuffer(self.plane_bytes * 7)
        p = self.buf.ptr
        self.albedo = plane(p, width, height, 3, self.ld * 4, self.plane_bytes)
        self.normal = plane(p + 3 * self.plane_bytes, width, height, 3, self.ld * 4, self.plane_bytes)
        self.depth = plane(p + 6 * self.plane_bytes, width, height, 1, self.ld * 4, self.plane_bytes)

    def gbuffers(self, sds=FEATURE_SDS):
        return gbuffers([self.albedo, self.normal, self.depth], list(sds))

    def to_numpy(self) -> dict[str, np.ndarray]:
        flat = self.buf.download(np.float32, self.plane_bytes * 7 // 4).reshape(7, self.height, self.ld)
        flat = flat[..., : self.width].copy()
        return {"albedo": flat[0:3], "normal": flat[3:6], "depth": flat[6:7]}


def make_scene(model: int, seed: int, width: int, height: int, image_id: int = 0):
    return scene(model, seed, image_id, width, height)


def synth_features(ctx: Context, sc, feats: Features) -> None:
    lib = load_library()
    check(lib.stater_synth_features(ctx.h, C.byref(sc), C.byref(feats.albedo), C.byref(feats.normal),
                                    C.byref(feats.depth), None))


def synth_samples(ctx: Context, sc, chunk: SampleChunk, first_sample: int = 0, feats: Features | None = None,
                  spp: int | None = None, first: Plane | None = None, counts: Plane | None = None,
                  count_offset: int = 0, stream=None) -> None:
    """Fill the chunk with samples first_sample + first[p] + j of each pixel
    (j < spp, or < the clamped per-pixel count when counts is given)."""
    lib = load_library()
    desc = chunk.desc(spp, counts=counts, count_offset=count_offset)
    check(lib.stater_synth_samples(ctx.h, C.byref(sc), C.byref(feats.albedo) if feats else None, first_sample,
                                   C.byref(first) if first else None, C.byref(desc), stream))


def c4_images(n: int = 48) -> list[tuple[int, int, int, int]]:
    """The uneven-image batch (W, H, spp, T): W ~ U[256, 2048], H ~ U[256, 1152],
    spp ~ LogU[16, 1024], T ~ U{1..5} (the first T non-identity mt5 lambdas
    plus the identity), image i drawn with seed 1000 + i."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(1000 + i)
        W = int(rng.integers(256, 2049))
        H = int(rng.integers(256, 1153))
        spp = int(round(float(np.exp(rng.uniform(np.log(16), np.log(1024))))))
        T = int(rng.integers(1, 6))
        out.append((W, H, spp, T))
    return out
