"""The C-ABI library on the CPU: it loads, exports every entry point that
include/stater.h declares, and its host-only paths (argument checks, layout
sizes, parameter defaults, scene set-up) behave as documented.  No call here
launches device work."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import stater_amd as sa

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "stater.h"


def _declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # drop comments
    names = re.findall(r"^\s*(?:const\s+)?(?:int|int64_t|void|char)\s*\**\s*(stater_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _define(name: str) -> int:
    m = re.search(rf"#define\s+{name}\s+(-?\d+)", HEADER.read_text())
    assert m, name
    return int(m.group(1))


@pytest.fixture(scope="module")
def lib():
    return sa.load_library()


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("stater_create", "stater_bufset_create", "stater_accumulate", "stater_denoise_vars",
                 "stater_denoise_film", "stater_calc_target_spp", "stater_calc_mean_vars", "stater_synchronize",
                 "stater_batch_create", "stater_batch_run"):
        assert must in names
    assert len(names) >= 32


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    if shutil.which("nm"):
        out = subprocess.run(["nm", "-D", "--defined-only", str(sa.LIB_PATH)], capture_output=True, text=True,
                             check=True).stdout
        exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
        assert not [n for n in _declared() if n not in exported]


def test_python_binding_covers_the_header():
    bound = {name for name, *_ in sa.SIGNATURES}
    assert set(_declared()) <= bound, set(_declared()) - bound


def test_abi_version_and_status_codes(lib):
    assert lib.stater_abi_version() == _define("STATER_ABI_VERSION")
    assert (_define("STATER_OK"), _define("STATER_E_ARG"), _define("STATER_E_HIP")) == (0, -1, -2)
    assert sa.STATER_E_ARG == _define("STATER_E_ARG")
    assert sa.STATER_E_UNSUPPORTED == _define("STATER_E_UNSUPPORTED")


def _bytes_for(lib, W, H, lambdas, mm=5, flags=0):
    arr = (C.c_float * max(len(lambdas), 1))(*lambdas)
    return lib.stater_bufset_bytes_for(W, H, len(lambdas), arr, mm, flags)


def _as_scratch(P):
    """Adaptive-sampling scratch after the planes: control block, 3 words per
    pixel, one int64 per 256-pixel block (each part 256-byte aligned)."""
    r = lambda b: (b + 255) // 256 * 256
    return 512 + 3 * r(P * 4) + r((P + 255) // 256 * 8)


def test_bufset_layout_size(lib):
    mt5 = (0.01, 0.25, 0.5, 0.75, 1.0, 3.0)
    S = 6                                  # untransformed + 5 distinct Box-Cox lambdas
    base = 1 + S * 7 * 3 + 3 + (S - 1) * 3  # n, moments, reference point, shifted means
    # jb, 6 per-set buffers, meanVar/meanD/filmD, targetspp, Welch bits, variance-filter records
    derived = S * (1 + 6 * 3) + 9 + 1 + 21 + 44
    ld = 1920                               # already a multiple of 64 floats
    zero = 256                              # the variance filter's zero record, last
    assert _bytes_for(lib, 1920, 1080, mt5) == (base + derived) * ld * 1080 * 4 + _as_scratch(1920 * 1080) + zero
    assert _bytes_for(lib, 1920, 1080, mt5, flags=sa.STATER_BUFSET_NO_DERIVED) == base * ld * 1080 * 4
    # rows padded to 256 bytes; repeated lambdas share one set; lambda 1 is the identity
    assert _bytes_for(lib, 100, 10, (1.0,), mm=2) == \
        (1 + 21 + 3 + 0 + 19 + 10 + 21 + 44) * 128 * 10 * 4 + _as_scratch(1000) + zero
    assert _bytes_for(lib, 100, 10, (0.5, 0.5, 1.0)) == _bytes_for(lib, 100, 10, (0.5,))


@pytest.mark.parametrize("W,H,lams,mm,msg", [
    (0, 10, (0.5,), 5, "width"),
    (10, 20000, (0.5,), 5, "width"),
    (10, 10, (-0.5,), 5, "lambda"),
    (10, 10, (0.5,), 6, "max_moment"),
    (10, 10, (0.1, 0.2, 0.3, 0.4, 0.6, 0.7), 5, "at most"),
])
def test_bufset_bytes_for_rejects(lib, W, H, lams, mm, msg):
    assert _bytes_for(lib, W, H, lams, mm) == -1
    assert msg in lib.stater_last_error().decode()


def test_default_params_match_reference_values(lib):
    p = sa.DenoiseParams()
    lib.stater_default_params(C.byref(p))
    # statpath.cpp:1071 (filterradius 20), scenes/render-denoise-mt5.pbrt:19
    assert p.radius == 20 and p.mt_mode == 2 and p.cascade == 1 and p.use_levc == 1
    assert np.isclose(p.var_z, 2.80703) and np.isclose(p.alpha, 0.005)
    assert np.isclose(p.ds_factor, -0.5 / 100.0)


def test_null_arguments_are_rejected_without_device_work(lib):
    E_ARG = _define("STATER_E_ARG")
    p = sa.DenoiseParams()
    lib.stater_default_params(C.byref(p))
    assert lib.stater_bufset_create(None, 8, 8, 0, None, 5, 0, C.byref(C.c_void_p())) == E_ARG
    assert lib.stater_accumulate(None, None, 0, None) == E_ARG
    assert lib.stater_denoise_vars(None, C.byref(p), None, None) == E_ARG
    assert lib.stater_denoise_film(None, C.byref(p), None, None, None, None) == E_ARG
    assert lib.stater_denoise_plain(None, 5, 0.005, None) == E_ARG
    assert lib.stater_calc_target_spp(None, 16, 64, None, None) == E_ARG
    assert lib.stater_calc_mean_vars(None, None) == E_ARG
    assert lib.stater_ctx_stream(None, None) == E_ARG
    assert lib.stater_timing_count(None) == E_ARG
    assert lib.stater_batch_create(None, 1, None, C.byref(p), 16, C.byref(C.c_void_p())) == E_ARG
    assert lib.stater_batch_run(None, 0, None) == E_ARG
    assert lib.stater_server_create(None, None, C.byref(C.c_void_p())) == E_ARG
    assert lib.stater_server_submit(None, None, None) == E_ARG
    assert lib.stater_server_wait(None, 1, None) == E_ARG
    assert lib.stater_server_destroy(None) == 0
    assert lib.stater_convert_plane(None, None, None, None) == E_ARG
    assert lib.stater_bufset_create_in(None, 8, 8, 0, None, 5, 0, None, 0, None, C.byref(C.c_void_p())) == E_ARG
    assert lib.stater_destroy(None) == 0 and lib.stater_bufset_destroy(None) == 0
    assert lib.stater_bufset_bytes(None) == -1
    assert "NULL" in lib.stater_last_error().decode() or "bad" in lib.stater_last_error().decode()


def test_server_default_config(lib):
    c = sa.ServerConfig()
    lib.stater_server_default_config(C.byref(c))
    assert c.slots >= 1 and c.chunk_spp >= 1 and c.max_bytes >= 1 << 30
    assert (c.max_width, c.max_height) == (3840, 2160)
    assert np.allclose(list(c.feature_sds), [0.02, 0.1, 0.5]) and c.params.radius == 20


def test_scene_init_is_deterministic(lib):
    a, b = sa.scene(2, 7, 0, 64, 48), sa.scene(2, 7, 0, 64, 48)
    assert bytes(a) == bytes(b)
    assert bytes(a) != bytes(sa.scene(2, 8, 0, 64, 48))
    assert a.width == 64 and a.height == 48
