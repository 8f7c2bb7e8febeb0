"""Pins the CPU oracle before it is trusted as the checker.

* Philox4x32-10 against the Random123 known-answer vectors.
* The FP64 restatement of the reference recurrence (estimator.h:226-305)
  against exact two-pass moments and closed forms; the FP32 restatement
  against the FP64 one; the Pebay merge against a single pass.
* Special functions and statistics against scipy fixtures
  (tests/golden/scipy_fixtures.json, written by tests/golden/make_golden.py).
* Filters against brute force and invariants; the AS map against its
  conservation and cap invariants.
A planted mistake (wrong sign, dropped term, wrong index) in any of these
functions fails at least one test here.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
import stater_amd as sa

FX = json.loads((Path(__file__).parent / "golden" / "scipy_fixtures.json").read_text())
MT5 = (0.01, 0.25, 0.5, 0.75, 1.0, 3.0)


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


# ------------------------------------------------------------ accumulation
def _rand_samples(spp, H, W, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    s = np.empty((spp, 5, H, W), np.float32)
    s[:, :3] = (rng.lognormal(0, 0.8, (spp, 3, H, W)) * scale).astype(np.float32)
    s[:, 3:] = rng.random((spp, 2, H, W), dtype=np.float32)
    return s


def test_fp64_recurrence_matches_two_pass_moments():
    smp = _rand_samples(50, 4, 5, 1)
    n, st = oracle.accumulate(smp, MT5, 5, fp64=True)
    ex = oracle.moments_exact(smp, MT5)
    assert np.all(n == 50)
    for k in range(4):
        np.testing.assert_allclose(st[:, k], ex[:, k], rtol=1e-10, atol=1e-12 * np.abs(ex[:, k]).max())


def test_fp32_recurrence_close_to_fp64():
    smp = _rand_samples(64, 6, 7, 2)
    _, s32 = oracle.accumulate(smp, (), 5)
    _, s64 = oracle.accumulate(smp, (), 5, fp64=True)
    for k in range(7):
        np.testing.assert_allclose(s32[0, k], s64[0, k], rtol=2e-4, atol=1e-5 * np.abs(s64[0, k]).max())


def test_closed_forms():
    # constant input: zero central moments, m1 = 0
    smp = np.zeros((10, 5, 1, 1), np.float32)
    smp[:, :3] = 2.5
    smp[:, 3:] = 0.3
    n, st = oracle.accumulate(smp, (), 5)
    assert n[0, 0] == 10
    assert np.all(st[0, 0] == 2.5) and not np.any(st[0, 1:4]) and not np.any(st[0, 6])
    # two samples a, b: m2 = (a-b)^2/2, m3 = 0, m4 = (a-b)^4/8
    a, b = 1.0, 4.0
    smp = np.zeros((2, 5, 1, 1), np.float32)
    smp[0, :3], smp[1, :3] = a, b
    _, st = oracle.accumulate(smp, (), 4, fp64=True)
    assert st[0, 1, 0, 0, 0] == pytest.approx((a - b) ** 2 / 2)
    assert st[0, 2, 0, 0, 0] == pytest.approx(0.0, abs=1e-12)
    assert st[0, 3, 0, 0, 0] == pytest.approx((a - b) ** 4 / 8)
    # m1 of the reference recurrence: sum of |d| sqrt(1 - 1/n)
    _, st = oracle.accumulate(smp, (), 5, fp64=True)
    assert st[0, 6, 0, 0, 0] == pytest.approx(abs(b - a) * np.sqrt(0.5))


def test_c2x_linear_signal_is_explained():
    """x exactly linear in (u - 1/2) -> the LEV-corrected m2C is a small
    remainder (it uses the known in-pixel mean 1/2 and variance 1/12, so the
    sampling error of u's own moments is left over)."""
    rng = np.random.default_rng(3)
    spp = 4000
    u = rng.random(spp).astype(np.float32)
    v = rng.random(spp).astype(np.float32)
    smp = np.zeros((spp, 5, 1, 1), np.float32)
    smp[:, 0, 0, 0] = 1.0 + 0.5 * (u - 0.5)
    smp[:, 1, 0, 0] = 1.0 + 0.2 * (v - 0.5)
    smp[:, 2, 0, 0] = 1.0
    smp[:, 3, 0, 0], smp[:, 4, 0, 0] = u, v
    n, st = oracle.accumulate(smp, (), 5)
    jb, lnS2, se2, m2C = oracle.prepass(n, st, 2.80703)
    m2 = st[0, 1, :, 0, 0]
    assert m2C[0, 0, 0, 0] < 0.06 * m2[0]
    assert m2C[0, 1, 0, 0] < 0.06 * m2[1]
    assert m2C[0, 2, 0, 0] == 0.0  # constant channel


def test_boxcox_identity_and_limits():
    smp = _rand_samples(8, 2, 2, 4)
    smp[0, :3, 0, 0] = 0.0
    _, st = oracle.accumulate(smp, (1.0, 0.5), 1, fp64=True)
    assert st.shape[0] == 2  # lambda 1 maps onto the untransformed set
    x = smp[:, :3].astype(np.float64)
    bc = (np.power(x, 0.5) - 1) / 0.5
    np.testing.assert_allclose(st[1, 0], bc.mean(axis=0), rtol=1e-12)
    np.testing.assert_allclose(st[0, 0], x.mean(axis=0), rtol=1e-12)


def test_boxcox_negative_samples_follow_pow():
    """boxCox(x, l) = (pow(x, l) - 1)/l (estimator.h:56) for a negative x:
    NaN for a non-integer lambda, the real power for an integer one."""
    smp = _rand_samples(6, 1, 1, 3)
    smp[2, 0, 0, 0] = -0.5
    smp[4, 1, 0, 0] = -2.0
    _, st = oracle.accumulate(smp, (0.5, 3.0), 1, fp64=True)
    assert np.isnan(st[1, 0, 0, 0, 0]) and np.isnan(st[1, 0, 1, 0, 0])  # lambda 0.5: NaN
    assert np.isfinite(st[1, 0, 2, 0, 0])
    x = smp[:, :3].astype(np.float64)
    np.testing.assert_allclose(st[2, 0], ((x ** 3 - 1) / 3).mean(axis=0), rtol=1e-12)


def test_sanitize_counts_black_samples():
    smp = _rand_samples(4, 1, 3, 5)
    smp[1, 0, 0, 0] = np.nan
    smp[2, 1, 0, 1] = -1.0
    smp[3, 2, 0, 2] = np.inf
    n, st = oracle.accumulate(smp, (), 1, sanitize=True)
    assert np.all(n == 4)
    ref = smp[:, :3].astype(np.float64).copy()
    ref[1, :, 0, 0] = 0
    ref[2, :, 0, 1] = 0
    ref[3, :, 0, 2] = 0
    np.testing.assert_allclose(st[0, 0], ref.mean(axis=0), rtol=1e-6)


def test_pebay_merge_equals_single_pass():
    rng = np.random.default_rng(6)
    x = rng.lognormal(0, 1, 300)
    def mom(v):
        m = v.mean(); d = v - m
        return np.array([len(v), m, (d ** 2).sum(), (d ** 3).sum(), (d ** 4).sum()], np.float64)
    a = mom(x[:113]); b = mom(x[113:])
    oracle.lib().orc_merge_moments(a, b)
    np.testing.assert_allclose(a, mom(x), rtol=1e-11)


def test_mean_var_formula():
    n = np.array([[4, 10]], np.int32)
    m2 = np.arange(6, dtype=np.float32).reshape(3, 1, 2)
    mv = oracle.mean_var(n, m2)
    np.testing.assert_allclose(mv[:, 0, 0], m2[:, 0, 0] / 12.0, rtol=1e-7)
    np.testing.assert_allclose(mv[:, 0, 1], m2[:, 0, 1] / 90.0, rtol=1e-7)


# -------------------------------------------------------- special functions
def test_betainc_against_scipy():
    L = oracle.lib()
    for a, b, x, v in FX["betainc"]:
        assert L.orc_betainc(a, b, x) == pytest.approx(v, rel=1e-10, abs=1e-300)


def test_t_tail_and_critical_values_against_scipy():
    L = oracle.lib()
    for t2, nu, v in FX["t2_tail"]:
        assert L.orc_t2_pvalue(t2, nu) == pytest.approx(v, rel=1e-9)
    for nu, v in FX["t2_crit_0.005"]:
        assert L.orc_t2_crit(nu, 0.005) == pytest.approx(v, rel=1e-9)
    z = FX["z_0.9975"]
    assert L.orc_t2_crit(float("inf"), 0.005) == pytest.approx(z * z, rel=1e-9)
    assert z == pytest.approx(2.80703, abs=1e-5)  # the reference's varZValue default


def test_welch_filter_decisions_match_scipy_welch():
    """Two-pixel images: the oracle's decision equals scipy's Welch p > alpha."""
    alpha = 0.005
    for m1, s1, n1, m2, s2, n2, t, pval in FX["welch"]:
        n = np.array([[n1, n2]], np.int32)
        mean = np.array([m1, m2], np.float32).reshape(1, 1, 1, 2).repeat(3, 1)
        var = np.array([s1 * s1, s2 * s2], np.float32).reshape(1, 1, 1, 2).repeat(3, 1)
        film = mean[0]
        out = oracle.denoise_film(n, mean, var, film, 1, 0.0, alpha, want_decisions=True)
        d = out["decisions"][0, 0, 1, 2, 0]  # pixel 0, neighbour dx = +1
        if abs(pval - alpha) > 1e-6:
            assert bool(d) == (pval > alpha), (pval, d)


def test_jarque_bera_against_scipy():
    for case in FX["jb"]:
        x = np.asarray(case["x"], np.float32)
        smp = np.zeros((len(x), 5, 1, 1), np.float32)
        smp[:, :3, 0, 0] = x[:, None]
        smp[:, 3:] = 0.5
        n, st = oracle.accumulate(smp, (), 4, fp64=True)
        jb, _, _, _ = oracle.prepass(n, st.astype(np.float32), 2.80703)
        assert jb[0, 0, 0] == pytest.approx(case["jb"], rel=2e-4)


# -------------------------------------------------------------- filters
def _brute_film(n, mean, var, film, r, alpha):
    from scipy import stats
    H, W = n.shape
    out = np.zeros((3, H, W))
    for y in range(H):
        for x in range(W):
            for c in range(3):
                num = den = 0.0
                for dy in range(-r, r + 1):
                    for dx in range(-r, r + 1):
                        yq, xq = y + dy, x + dx
                        if not (0 <= yq < H and 0 <= xq < W):
                            continue
                        acc = (dy == 0 and dx == 0)
                        if not acc:
                            vp = np.float32(var[0, c, y, x] / np.float32(n[y, x]))
                            vq = np.float32(var[0, c, yq, xq] / np.float32(n[yq, xq]))
                            sp = np.sqrt(float(vp) * n[y, x]); sq = np.sqrt(float(vq) * n[yq, xq])
                            res = stats.ttest_ind_from_stats(mean[0, c, y, x], sp, n[y, x], mean[0, c, yq, xq],
                                                             sq, n[yq, xq], equal_var=False)
                            acc = res.pvalue > alpha
                        if acc:
                            num += float(film[c, yq, xq]); den += 1
                out[c, y, x] = num / den
    return out


def test_welch_filter_brute_force_small():
    rng = np.random.default_rng(7)
    H, W, r = 6, 7, 2
    n = rng.integers(8, 40, (H, W)).astype(np.int32)
    mean = rng.normal(1, 0.2, (1, 3, H, W)).astype(np.float32)
    var = rng.uniform(0.2, 1.5, (1, 3, H, W)).astype(np.float32)
    out = oracle.denoise_film(n, mean, var, mean[0], r, 0.0, 0.005)
    np.testing.assert_allclose(out["filmD"], _brute_film(n, mean, var, mean[0], r, 0.005), rtol=1e-12)


def test_constant_image_accepts_everything():
    H, W, r = 8, 9, 3
    n = np.full((H, W), 64, np.int32)
    mean = np.ones((2, 3, H, W), np.float32)
    var = np.full((2, 3, H, W), 0.5, np.float32)
    out = oracle.denoise_film(n, mean, var, mean[0] * 2, r, 0.0, 0.005, want_decisions=True)
    d = out["decisions"]
    ys, xs = np.mgrid[0:H, 0:W]
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            inb = (ys + dy >= 0) & (ys + dy < H) & (xs + dx >= 0) & (xs + dx < W)
            assert np.all(d[:, :, dy + r, dx + r, 0][inb] == 1)
            assert np.all(d[:, :, dy + r, dx + r, 0][~inb] == 0)
    np.testing.assert_allclose(out["filmD"], 2.0)


def test_variance_filter_uniform_and_rejection():
    H, W, r = 5, 6, 2
    n = np.full((H, W), 100, np.int32)
    m2 = np.full((1, 3, H, W), 99.0, np.float32)         # s^2 = 1
    m2[:, :, :, 3:] = 99.0 * 100                          # s^2 = 100 on the right half
    lnS2 = np.log(m2 / 99.0).astype(np.float32)
    se2 = np.full_like(m2, 0.02)
    vD, vCD, cnt = oracle.denoise_vars(n, m2, lnS2, se2, m2, r, 0.0, 2.80703, want_count=True)
    np.testing.assert_allclose(vD[0, :, :, :3], 1.0)      # the two halves never mix
    np.testing.assert_allclose(vD[0, :, :, 3:], 100.0)
    assert cnt[0, 0, 2, 2] == 5 * 3                       # 5 rows x columns 0..2 within r
    # with a huge standard error every neighbour is accepted: plain mean
    vD2, _ = oracle.denoise_vars(n, m2, lnS2, np.full_like(m2, 1e6), m2, r, 0.0, 2.80703)
    ref = np.zeros((H, W))
    s2 = m2[0, 0].astype(np.float64) / 99.0
    for y in range(H):
        for x in range(W):
            ref[y, x] = s2[max(0, y - r):y + r + 1, max(0, x - r):x + r + 1].mean()
    np.testing.assert_allclose(vD2[0, 0], ref, rtol=1e-12)


# ----------------------------------------------------------------- AS map
@pytest.mark.parametrize("budget,cap", [(32, 4096), (32, 60), (4, 33)])
def test_target_spp_invariants(budget, cap):
    rng = np.random.default_rng(8)
    H, W = 20, 30
    n = rng.integers(16, 64, (H, W)).astype(np.int32)
    varD = rng.lognormal(0, 1, (3, H, W)).astype(np.float32)
    mean = rng.lognormal(0, 1, (3, H, W)).astype(np.float32)
    t = oracle.target_spp(n, varD, mean, budget, cap)
    head = np.maximum(cap - n, 0)
    assert np.all(t >= 0) and np.all(t <= head)
    if head.sum() > budget * W * H:
        assert t.sum() == budget * W * H
        # monotone in the error: a pixel with a larger error never gets fewer samples unless capped
        e = (np.sqrt(np.maximum(varD, 0) / n) / np.maximum(np.abs(mean), 1e-3)).mean(0)
        order = np.argsort(e.ravel())
        tt, hh = t.ravel()[order], head.ravel()[order]
        free = tt < hh
        assert np.all(np.diff(tt[free]) >= -1)
    else:
        assert np.array_equal(t, head)


@pytest.mark.parametrize("n_noisy,n_flat,budget,cap", [(4, 60, 32, 4096), (10, 300, 4, 64), (0, 50, 8, 100),
                                                         (3, 40, 0, 4096)])
def test_target_spp_zero_error_pixels(n_noisy, n_flat, budget, cap):
    """Pixels with zero error (a flat background, varD = 0) beside noisy
    pixels close to the cap: the noisy pixels' headroom is smaller than the
    budget, so they take all of it and the flat pixels share the rest evenly
    (their targets differ by at most 1); the budget is met exactly."""
    rng = np.random.default_rng(11)
    P = n_noisy + n_flat
    H, W = 1, P
    n = np.full((H, W), 16, np.int32)
    varD = np.zeros((3, H, W), np.float32)
    mean = np.full((3, H, W), 0.5, np.float32)
    noisy = rng.permutation(P)[:n_noisy]
    n[0, noisy] = cap - 1
    varD[:, 0, noisy] = rng.lognormal(0, 1, (3, n_noisy)).astype(np.float32)
    t = oracle.target_spp(n, varD, mean, budget, cap)[0]
    head = np.maximum(cap - n[0], 0)
    assert np.all(t >= 0) and np.all(t <= head)
    assert int(t.sum()) == min(budget * P, int(head.sum()))
    assert np.array_equal(t[noisy], head[noisy]) or budget == 0
    flat = np.setdiff1d(np.arange(P), noisy)
    free = flat[t[flat] < head[flat]]
    if free.size:
        assert t[free].max() - t[free].min() <= 1


def test_scene_struct_matches_library():
    """The library's scene parameters (host code) equal the oracle's; the
    trigonometric fields may differ by an ulp between the two compilers' libm
    calls, everything else is identical."""
    import ctypes as C
    for model in (1, 2):
        a = oracle.scene_init(sa.Scene, model, 5, 3, 200, 120)
        b = sa.scene(model, 5, 3, 200, 120)
        for name, typ in sa.Scene._fields_:
            va = np.ctypeslib.as_array(getattr(a, name)) if hasattr(getattr(a, name), "_length_") else getattr(a, name)
            vb = np.ctypeslib.as_array(getattr(b, name)) if hasattr(getattr(b, name), "_length_") else getattr(b, name)
            if name in ("cos_phi", "sin_phi"):
                np.testing.assert_allclose(np.asarray(va, float), np.asarray(vb, float), rtol=4e-16, atol=4e-16)
            else:
                assert np.array_equal(np.asarray(va), np.asarray(vb)), name
