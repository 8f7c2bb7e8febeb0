"""The PFM codec against the format's definition (host side, no GPU).

The reference's files come from OpenCV's PFM encoder (a dependency absent
here, opencv_contrib 4.8.1 per README.md:24); the format itself pins them:
"PF" (RGB) / "Pf" (grey), width height, a scale whose sign is the byte order
(negative: little-endian), one white-space character, then float32 rows from
the bottom of the image up.  The known-answer files below are built byte by
byte from that definition, independently of the codec.
"""
import struct

import numpy as np
import pytest

import stater_amd as sa
from stater_amd import pfm


def _ref_bytes(arr, magic, scale_txt, endian="<"):
    """[C, H, W] float array -> PFM bytes built from the format definition."""
    c, h, w = arr.shape
    out = f"{magic}\n{w} {h}\n{scale_txt}\n".encode()
    for y in range(h - 1, -1, -1):  # bottom row first
        for x in range(w):
            for k in range(c):
                out += struct.pack(endian + "f", float(arr[k, y, x]))
    return out


def test_write_matches_the_format(tmp_path):
    a = np.arange(2 * 3 * 3, dtype=np.float32).reshape(3, 2, 3) * 0.5 - 1.0  # RGB, 2 rows x 3 columns
    pfm.write(tmp_path / "rgb.pfm", a)
    assert (tmp_path / "rgb.pfm").read_bytes() == _ref_bytes(a, "PF", "-1.000000")
    g = np.array([[[1.5, -2.0], [3.25, 1e-30]]], np.float32)
    pfm.write(tmp_path / "g.pfm", g)
    assert (tmp_path / "g.pfm").read_bytes() == _ref_bytes(g, "Pf", "-1.000000")


@pytest.mark.parametrize("magic,scale,endian", [("PF", "-1", "<"), ("PF", "1.0", ">"), ("Pf", "-0.5", "<"),
                                                ("Pf", "2", ">")])
def test_read_known_files(tmp_path, magic, scale, endian):
    c = 3 if magic == "PF" else 1
    rng = np.random.default_rng(3)
    a = rng.standard_normal((c, 4, 5)).astype(np.float32)
    (tmp_path / "k.pfm").write_bytes(_ref_bytes(a, magic, scale, endian))
    assert pfm.info(tmp_path / "k.pfm") == (5, 4, c)
    np.testing.assert_array_equal(pfm.read(tmp_path / "k.pfm"), a)


def test_round_trip_keeps_every_float(tmp_path):
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2**32, (3, 17, 9), dtype=np.uint64).astype(np.uint32)
    a = bits.view(np.float32)  # every pattern, NaNs and infinities included
    pfm.write(tmp_path / "r.pfm", a)
    assert np.array_equal(pfm.read(tmp_path / "r.pfm").view(np.uint32), bits)


def test_rejects_bad_files(tmp_path):
    lib = sa.load_library()
    (tmp_path / "bad.pfm").write_bytes(b"P6\n2 2\n255\n" + bytes(12))
    assert lib.stater_pfm_info(str(tmp_path / "bad.pfm").encode(), None, None, None) == sa.STATER_E_ARG
    a = np.ones((3, 2, 2), np.float32)
    (tmp_path / "short.pfm").write_bytes(_ref_bytes(a, "PF", "-1")[:-4])
    buf = np.empty(12, np.float32)
    assert lib.stater_pfm_read_host(str(tmp_path / "short.pfm").encode(), buf.ctypes.data, 2, 2, 3) == sa.STATER_E_ARG
    (tmp_path / "ok.pfm").write_bytes(_ref_bytes(a, "PF", "-1"))
    assert lib.stater_pfm_read_host(str(tmp_path / "ok.pfm").encode(), buf.ctypes.data, 2, 3, 3) == sa.STATER_E_ARG
    assert b"expected" in sa.load_library().stater_last_error()
    assert lib.stater_pfm_info(str(tmp_path / "missing.pfm").encode(), None, None, None) == sa.STATER_E_NOTFOUND
    assert lib.stater_pfm_write_host(str(tmp_path / "x.pfm").encode(), buf.ctypes.data, 2, 2, 2) == sa.STATER_E_ARG
