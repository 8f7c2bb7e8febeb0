"""Minimal HIP runtime plumbing over ctypes: device buffers and copies.

Used by the Python layer, the tests and the bench to hand device memory to
the C-ABI and to read planes back.  The library's streams do not synchronise
with the null stream, so every copy and clear here first waits for the whole
device and then runs the synchronous hipMemcpy / hipMemset (followed by a
device wait): plumbing for tests and set-up, not for timed paths.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _hip


def _sync() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    try:
        rc = hip().hipGetDeviceCount(C.byref(n))
    except OSError:
        return 0
    return n.value if rc == 0 else 0


class Stream:
    """An owned non-blocking HIP stream at the default priority."""

    def __init__(self):
        h = C.c_void_p()
        _ok(hip().hipStreamCreateWithFlags(C.byref(h), 1), "hipStreamCreateWithFlags")
        self.handle = h

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self) -> None:
        if self.handle:
            hip().hipStreamSynchronize(self.handle)
            hip().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """An owned device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 256)), "hipMalloc")
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        _sync()
        _ok(hip().hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "H2D")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _sync()
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes, D2H), "D2H")
        return out

    def zero(self) -> None:
        _sync()
        _ok(hip().hipMemset(C.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")
        _sync()


def sub_plane(pl, x0: int, y0: int, width: int, height: int):
    """The window [x0, x0 + width) x [y0, y0 + height) of a plane, as a plane
    over the same memory (same pitch and channel stride)."""
    if x0 < 0 or y0 < 0 or x0 + width > pl.width or y0 + height > pl.height:
        raise ValueError("sub_plane: window outside the plane")
    out = type(pl).from_buffer_copy(pl)
    out.ptr = pl.ptr + y0 * pl.pitch_bytes + x0 * 4
    out.width, out.height = width, height
    return out


def read_plane(pl, dtype=None) -> np.ndarray:
    """Copy a planar device plane (stater Plane) to a numpy array [channels, H, W]."""
    dt = dtype or (np.int32 if pl.dtype == 1 else np.float32)
    pitch_elems = pl.pitch_bytes // 4
    out = np.empty((pl.channels, pl.height, pl.width), dtype=dt)
    for c in range(pl.channels):
        nbytes = pl.pitch_bytes * (pl.height - 1) + pl.width * 4
        tmp = np.empty(nbytes // 4, dtype=dt)
        _sync()
        _ok(hip().hipMemcpy(tmp.ctypes.data_as(C.c_void_p), C.c_void_p(pl.ptr + c * pl.channel_stride_bytes),
                            nbytes, D2H), "D2H plane")
        full = np.empty(pitch_elems * pl.height, dtype=dt)
        full[: tmp.size] = tmp
        out[c] = full.reshape(pl.height, pitch_elems)[:, : pl.width]
    return out


def write_plane(pl, arr: np.ndarray) -> None:
    """Upload [channels, H, W] into a planar device plane (row by row pitch)."""
    dt = np.int32 if pl.dtype == 1 else np.float32
    a = np.asarray(arr, dtype=dt).reshape(pl.channels, pl.height, pl.width)
    pitch_elems = pl.pitch_bytes // 4
    for c in range(pl.channels):
        full = np.zeros((pl.height, pitch_elems), dtype=dt)
        full[:, : pl.width] = a[c]
        nbytes = pl.pitch_bytes * (pl.height - 1) + pl.width * 4
        flat = np.ascontiguousarray(full.reshape(-1)[: nbytes // 4])
        _sync()
        _ok(hip().hipMemcpy(C.c_void_p(pl.ptr + c * pl.channel_stride_bytes), flat.ctypes.data_as(C.c_void_p),
                            nbytes, H2D), "H2D plane")
