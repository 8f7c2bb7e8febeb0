"""GPU parity of the sample generator and the moment accumulation.

The untransformed statistics must be bit-identical to the oracle's FP32
restatement of StatTile (src/statistics/estimator.h:226-324) on the same
samples; the Box-Cox statistics are compared with the FP64 evaluation of the
same recurrence at the relative tolerance stated below.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import stater_amd as sa
from gpu_helpers import generate, read_stats
from stater_amd import hipmem
from stater_amd.workload import MT5_LAMBDAS, SampleChunk

pytestmark = pytest.mark.gpu

# Box-Cox statistics against FP64, relative to each statistic's natural scale.
BOXCOX_RTOL = 1e-4


def _scales(st64, n):
    """Natural magnitude per statistic [7,3,H,W] for relative errors.  A sum's
    rounding error scales with the sum of the magnitudes of its terms; for m3
    that is sum |d|^3 <= sqrt(m2 m4) (Cauchy-Schwarz), which for heavy-tailed
    pixels is much larger than |m3|."""
    nn = np.maximum(n, 1)[None].astype(np.float64)
    m2 = np.abs(st64[1])
    sd = np.sqrt(m2 / nn)
    return np.stack([
        np.maximum(np.abs(st64[0]), sd),            # mean
        m2,                                         # m2
        np.sqrt(m2 * np.abs(st64[3])),              # m3: bound of sum |d|^3
        np.abs(st64[3]),                            # m4
        np.sqrt(m2 * nn / 12.0),                    # c2X ~ sqrt(n/12) sd sqrt(n)
        np.sqrt(m2 * nn / 12.0),                    # c2Y
        np.abs(st64[6]),                            # m1
    ])


def test_samples_match_cpu_generator(ctx):
    sc, chunk, feats = generate(ctx, 2, 7, 80, 48, 6)
    f_cpu = oracle.features(sc)
    f_gpu = feats.to_numpy()
    for k in ("albedo", "normal", "depth"):
        mism = np.count_nonzero(f_cpu[k] != f_gpu[k])
        assert mism <= 2, (k, mism)
        np.testing.assert_allclose(f_gpu[k], f_cpu[k], rtol=2e-7, atol=0)
    s_cpu = oracle.samples(sc, f_gpu["albedo"], 0, 6)
    s_gpu = chunk.to_numpy()
    mism = np.count_nonzero(s_cpu != s_gpu)
    assert mism <= max(2, s_cpu.size // 100000), mism
    np.testing.assert_allclose(s_gpu, s_cpu, rtol=2.5e-7, atol=0)


def test_accumulate_plain_bitexact(ctx):
    """128x128, 64 spp, identity transform, n/mean/M2/M3 (the plain case)."""
    W = H = 128
    sc, chunk, _ = generate(ctx, 1, 1, W, H, 64)
    bs = sa.BufferSet(ctx, W, H, (), max_moment=3)
    bs.accumulate(chunk.desc())
    ctx.synchronize()
    n, st = read_stats(bs, ())
    n_ref, st_ref = oracle.accumulate(chunk.to_numpy(), (), 3)
    assert np.array_equal(n, n_ref)
    for k in range(3):
        assert np.array_equal(st[0, k], st_ref[0, k]), k


def test_accumulate_mt5_m4c(ctx):
    W, H, spp = 96, 64, 64
    sc, chunk, _ = generate(ctx, 2, 2, W, H, spp)
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, max_moment=5)
    bs.accumulate(chunk.desc())
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    smp = chunk.to_numpy()
    n_ref, st32 = oracle.accumulate(smp, MT5_LAMBDAS, 5)
    _, st64 = oracle.accumulate(smp, MT5_LAMBDAS, 5, fp64=True)
    assert np.array_equal(n, n_ref)
    # untransformed set: the reference's FP32 numbers, bit for bit
    assert np.array_equal(st[0], st32[0])
    # every identity slot of the lambda list maps onto set 0
    for s in range(1, st.shape[0]):
        scale = _scales(st64[s], n)
        fits = np.all(np.abs(st64[s]) < 1e30, axis=0)  # exclude FP32 overflow of lambda = 3
        fits = np.broadcast_to(fits, st[s].shape)
        err = np.abs(st[s].astype(np.float64) - st64[s]) / np.maximum(scale, 1e-30)
        worst = float(np.max(np.where(fits, err, 0.0)))
        print(f"Box-Cox set {s}: worst error {worst:.2e} of the natural scale")
        assert worst <= BOXCOX_RTOL, (s, worst)
        assert fits.mean() > 0.9


def _edge_samples(spp, W, seed):
    """[spp, 5, 8, W] samples, one regime per row: ordinary, a tiny first
    sample, a zero first sample, zeros and negatives among the samples, a
    nearly constant pixel, large values, twenty decades of range, a huge
    first sample."""
    rng = np.random.default_rng(seed)
    H = 8
    x = rng.lognormal(0.0, 1.0, (spp, 3, H, W))
    x[0, :, 1] = 1e-12
    x[0, :, 2] = 0.0
    x[:, :, 3] *= rng.random((spp, 3, W)) > 0.3
    x[5, 0, 3] = -0.25                        # a negative channel sample
    x[9, 2, 3, ::3] = -3.0
    x[:, :, 4] = 1.0 + 1e-5 * rng.standard_normal((spp, 3, W))
    x[:, :, 5] *= 1e10
    x[:, :, 6] = 10.0 ** rng.uniform(-20, 10, (spp, 3, W))
    x[0, :, 7] = 1e15
    smp = np.zeros((spp, 5, H, W), np.float32)
    smp[:, :3] = x
    smp[:, 3:] = rng.random((spp, 2, H, W))
    return smp


@pytest.mark.parametrize("lambdas", [MT5_LAMBDAS, (0.5,), (0.5, 0.75), (0.25, 1.0), (0.25, 0.5, 0.75, 1.0),
                                     (0.3, 2.0)])
def test_boxcox_edge_samples_against_fp64(ctx, lambdas):
    """Every transform form (the reference's lambda lists take the ratio form
    for 1/4, 1/2, 3/4, 3 and the log form for 0.01; another list takes the
    log form for all) on edge samples: the FP64 statistics wherever they fit
    FP32, NaN where a negative sample makes pow NaN, set 0 bit for bit."""
    W, spp = 64, 48
    smp = _edge_samples(spp, W, 17)
    chunk = SampleChunk(W, 8, spp)
    chunk.from_numpy(smp)
    bs = sa.BufferSet(ctx, W, 8, lambdas, max_moment=5)
    bs.accumulate(chunk.desc())
    ctx.synchronize()
    n, st = read_stats(bs, lambdas)
    n_ref, st32 = oracle.accumulate(smp, lambdas, 5)
    _, st64 = oracle.accumulate(smp, lambdas, 5, fp64=True)
    assert np.array_equal(n, n_ref)
    assert np.array_equal(st[0], st32[0], equal_nan=True)
    for s in range(1, st.shape[0]):
        nan64 = np.isnan(st64[s])
        fits = np.all(np.isfinite(st64[s]) & (np.abs(st64[s]) < 1e30), axis=0)
        fits = np.broadcast_to(fits, st[s].shape)
        assert np.all(np.isnan(st[s])[nan64]), s                  # the NaN pixels of pow ...
        bad = fits & ~np.isfinite(st[s])
        if bad.any():
            idx = np.argwhere(bad)[:8]
            print("non-finite where FP64 fits (stat, channel, row, column):", idx.tolist(),
                  [float(st[s][tuple(i)]) for i in idx], [float(st64[s][tuple(i)]) for i in idx])
        assert not bad.any(), s                                   # ... and every pixel that fits FP32 finite
        lam = float(oracle.set_lambdas(lambdas)[s])
        assert nan64.any() == (lam != round(lam))                 # the samples hold negatives
        scale = _scales(np.where(np.isfinite(st64[s]), st64[s], 0.0), n)
        err = np.abs(st[s].astype(np.float64) - st64[s]) / np.maximum(scale, 1e-300)
        worst = float(np.max(np.where(fits, err, 0.0)))
        rows = [float(np.max(np.where(fits, err, 0.0)[..., r, :])) for r in range(8)]
        print(f"lambdas {lambdas} set {s}: worst error {worst:.2e}; per row " + " ".join(f"{v:.1e}" for v in rows))
        assert worst <= BOXCOX_RTOL, (s, worst)


def _extreme_set0_samples(spp, W, seed):
    """Rows whose deviations x - mean reach every branch of set 0's exact
    update: below 2^-40 (the IEEE division and sqrtf), subnormal quotients
    and exact midpoints among them, radicands around 2^-104, deviations on
    both sides of 2^-40, values whose squares overflow, and zeros."""
    rng = np.random.default_rng(seed)
    H = 8
    tiny = np.float32(2.0 ** -149)
    x = np.empty((spp, 3, H, W), np.float64)
    x[:, :, 0] = 1e-14 * rng.lognormal(0.0, 1.0, (spp, 3, W))
    x[:, :, 1] = 1e-30 * rng.lognormal(0.0, 1.0, (spp, 3, W))
    x[:, :, 2] = 1e-40 * rng.lognormal(0.0, 1.0, (spp, 3, W))            # subnormal samples
    x[:, :, 3] = tiny * rng.integers(0, 8, (spp, 3, W))                  # quotients on midpoints (3 x 2^-149 / 2)
    x[:, :, 4] = 2.0 ** -52 * (1.0 + rng.uniform(-1.0, 1.0, (spp, 3, W)))  # radicands near 2^-104
    x[:, :, 5] = 2.0 ** -39 * rng.uniform(0.0, 2.0, (spp, 3, W))        # deviations around 2^-40
    x[:, :, 6] = 10.0 ** rng.uniform(18, 38, (spp, 3, W))                # d^2 overflows
    x[:, :, 7] = rng.lognormal(0.0, 1.0, (spp, 3, W)) * (rng.random((spp, 3, W)) > 0.5)
    smp = np.zeros((spp, 5, H, W), np.float32)
    smp[:, :3] = x.astype(np.float32)
    smp[:, 3:] = rng.random((spp, 2, H, W))
    return smp


@pytest.mark.parametrize("mm", [1, 2, 3, 4, 5])
def test_set0_exact_at_extreme_deviations(ctx, mm):
    """Set 0 takes d / n from an FP64 reciprocal of n and the root of m1's
    increment from a neighbour-checked hardware square root, with the IEEE
    division and sqrtf where a deviation is below 2^-40: bit for bit the
    reference's FP32 update on samples that reach every one of those forms,
    for every max_moment (estimator.h:226-305)."""
    W, spp = 64, 40
    smp = _extreme_set0_samples(spp, W, 31)
    chunk = SampleChunk(W, 8, spp)
    chunk.from_numpy(smp)
    bs = sa.BufferSet(ctx, W, 8, (), max_moment=mm)
    bs.accumulate(chunk.desc())
    ctx.synchronize()
    n, st = read_stats(bs, ())
    n_ref, st_ref = oracle.accumulate(smp, (), mm)
    assert np.array_equal(n, n_ref)
    k = min(mm, 4) if mm < 5 else 7
    got, want = st[0, :k], st_ref[0, :k]
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    if not same.all():
        idx = np.argwhere(~same)[:8]
        print("differing (stat, channel, row, column):", idx.tolist(),
              [(float(got[tuple(i)]), float(want[tuple(i)])) for i in idx])
    assert same.all()


def test_exact_forms_on_every_float(ctx):
    """The neighbour-checked square root equals sqrtf on every non-negative
    float outside (0, 2^-104), -0 and NaN included; the quotient through the
    correctly rounded reciprocal of n equals the IEEE d / n for every finite
    float d outside 0 < |d| < 2^-40 (with a pseudo-random n per d); and that
    reciprocal equals 1.f / m for every float m in [1, 2^32]: the ranges on
    which set 0's update uses them."""
    buf = hipmem.DeviceBuffer(24)
    sa.check(sa.load_library().stater_debug_exact_forms(ctx.h, C.c_void_p(buf.ptr), None))
    ctx.synchronize()
    counts = buf.download(np.uint64, 3)
    assert counts.tolist() == [0, 0, 0], counts


def test_quotient_on_every_pair_of_significands(ctx):
    """The quotient form equals the IEEE x / m for every significand x in
    [1, 2) and every integer m in [2^23, 2^24]: with the reciprocal check
    above and the exact scaling by powers of two, every quotient set 0's
    update forms."""
    lib = sa.load_library()
    buf = hipmem.DeviceBuffer(8)
    # argument checks
    assert lib.stater_debug_quotient_sweep(ctx.h, 0, 5, C.c_void_p(buf.ptr), None) != 0
    assert lib.stater_debug_quotient_sweep(ctx.h, 9, 9, C.c_void_p(buf.ptr), None) != 0
    assert lib.stater_debug_quotient_sweep(ctx.h, 1, (1 << 24) + 2, C.c_void_p(buf.ptr), None) != 0
    step = 1 << 20
    for lo in range(1 << 23, (1 << 24) + 1, step):
        hi = min(lo + step, (1 << 24) + 1)
        sa.check(lib.stater_debug_quotient_sweep(ctx.h, lo, hi, C.c_void_p(buf.ptr), None))
        ctx.synchronize()
        bad = int(buf.download(np.uint64, 1)[0])
        print(f"m in [{lo}, {hi}): {bad} differing", flush=True)
        assert bad == 0


def test_accumulate_in_chunks_continues_sequence(ctx):
    W, H = 70, 33
    sc, chunk, _ = generate(ctx, 2, 3, W, H, 40)
    full = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    full.accumulate(chunk.desc())
    parts = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    d = chunk.desc(spp=25)
    parts.accumulate(d)
    d2 = chunk.desc(spp=15)
    d2.data = chunk.buf.ptr + 25 * 5 * chunk.plane_bytes
    parts.accumulate(d2)
    ctx.synchronize()
    a = read_stats(full, MT5_LAMBDAS)
    b = read_stats(parts, MT5_LAMBDAS)
    # a second call continues the recurrence bit for bit, for every set
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    full.reset()
    ctx.synchronize()
    n0, st0 = read_stats(full, MT5_LAMBDAS)
    assert not n0.any() and not st0.any()


def test_count_factor_paths(ctx):
    """The per-sample factors of the sample count come from a table shared by
    the wavefront where every lane starts from the same count, and from each
    lane's own count otherwise.  A fresh call with ragged counts (pixels
    without samples, a partial last wavefront, more than 64 samples, so that
    the table is refilled by fewer lanes each time) takes the first; a call
    continuing those unequal counts takes the second.  Set 0 stays bit-exact
    with the oracle and the Box-Cox sets within their tolerance of FP64
    through both."""
    W, H, spp1, spp2 = 100, 3, 150, 70
    rng = np.random.default_rng(23)

    def samples(spp):
        smp = rng.lognormal(0.0, 1.0, (spp, 5, H, W)).astype(np.float32)
        smp[:, 3:] = rng.random((spp, 2, H, W), dtype=np.float32)
        return smp

    smp1, smp2 = samples(spp1), samples(spp2)
    counts = rng.integers(-1, spp1 + 1, (H, W)).astype(np.int32)
    counts[:, 10:20] = 0
    counts[0] = spp1
    c1, c2 = SampleChunk(W, H, spp1), SampleChunk(W, H, spp2)
    c1.from_numpy(smp1)
    c2.from_numpy(smp2)
    cbuf = hipmem.DeviceBuffer(W * H * 4)
    cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
    hipmem.write_plane(cplane, counts[None])
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    ref32 = ref64 = None
    for smp, desc in ((smp1, c1.desc(counts=cplane)), (smp2, c2.desc())):
        bs.accumulate(desc)
        ctx.synchronize()
        n, st = read_stats(bs, MT5_LAMBDAS)
        kw = {"counts": counts} if ref32 is None else {}
        ref32 = oracle.accumulate(smp, MT5_LAMBDAS, 5, state=ref32, **kw)
        ref64 = oracle.accumulate(smp, MT5_LAMBDAS, 5, fp64=True, state=ref64, **kw)
        assert np.array_equal(n, ref32[0])
        assert np.array_equal(st[0].view(np.uint32), ref32[1][0].view(np.uint32))
        for s in range(1, st.shape[0]):
            scale = _scales(ref64[1][s], n)
            err = np.abs(st[s].astype(np.float64) - ref64[1][s]) / np.maximum(scale, 1e-30)
            worst = float(np.max(np.where(n[None, None] > 0, err, 0.0)))
            assert worst <= BOXCOX_RTOL, (s, worst)
    assert len(set(ref32[0][0])) == 1 and len(set(ref32[0][1])) > 10  # uniform, then unequal counts


def test_sanitize_and_counts(ctx):
    W, H, spp = 40, 24, 16
    rng = np.random.default_rng(5)
    smp = rng.lognormal(0, 1, (spp, 5, H, W)).astype(np.float32)
    smp[:, 3:] = rng.random((spp, 2, H, W), dtype=np.float32)
    bad = rng.random((spp, H, W)) < 0.05
    kinds = rng.integers(0, 4, (spp, H, W))
    for k, val in enumerate((np.nan, np.inf, -3.0, -np.inf)):
        m = bad & (kinds == k)
        smp[:, 1][m] = val
    counts = rng.integers(0, spp + 1, (H, W)).astype(np.int32)
    chunk = SampleChunk(W, H, spp)
    chunk.from_numpy(smp)
    cbuf = hipmem.DeviceBuffer(W * H * 4)
    cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
    hipmem.write_plane(cplane, counts[None])
    bs = sa.BufferSet(ctx, W, H, MT5_LAMBDAS, 5)
    bs.accumulate(chunk.desc(sanitize=True, counts=cplane))
    ctx.synchronize()
    n, st = read_stats(bs, MT5_LAMBDAS)
    n_ref, st_ref = oracle.accumulate(smp, MT5_LAMBDAS, 5, counts=counts, sanitize=True)
    assert np.array_equal(n, n_ref)
    assert np.array_equal(st[0], st_ref[0], equal_nan=True)


@pytest.mark.parametrize("mode,mm", [(sa.STATER_ACC_SEQUENTIAL, 5), (sa.STATER_ACC_MERGE64, 5),
                                     (sa.STATER_ACC_SEQUENTIAL, 3)])
def test_fresh_call_equals_reset_then_call(ctx, mode, mm):
    """STATER_ACC_FRESH on a set holding other statistics and Film sums gives
    bit for bit what a cleared set gives, pixels without samples included."""
    W, H, spp = 70, 20, 24
    _, chunk, _ = generate(ctx, 2, 9, W, H, spp)
    _, other, _ = generate(ctx, 2, 10, W, H, spp)
    counts = np.random.default_rng(3).integers(-2, spp + 1, (H, W)).astype(np.int32)
    counts[:, :5] = 0
    cbuf = hipmem.DeviceBuffer(W * H * 4)
    cplane = sa.plane(cbuf.ptr, W, H, 1, W * 4, W * H * 4, sa.STATER_I32)
    hipmem.write_plane(cplane, counts[None])
    lam = MT5_LAMBDAS if mm == 5 else (0.5,)
    a = sa.BufferSet(ctx, W, H, lam, mm, flags=sa.STATER_BUFSET_FILM)
    b = sa.BufferSet(ctx, W, H, lam, mm, flags=sa.STATER_BUFSET_FILM)
    a.accumulate(other.desc(sanitize=True))
    a.accumulate(chunk.desc(sanitize=True, counts=cplane), mode | sa.STATER_ACC_FRESH)
    b.accumulate(chunk.desc(sanitize=True, counts=cplane), mode)
    a.film_resolve()
    b.film_resolve()
    ctx.synchronize()
    na, sta = read_stats(a, lam)
    nb, stb = read_stats(b, lam)
    assert np.array_equal(na, nb) and not na[:, :5].any()
    assert np.array_equal(sta.view(np.uint32), stb.view(np.uint32))
    fa, fb = hipmem.read_plane(a.view("film")), hipmem.read_plane(b.view("film"))
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    # a later plain call continues both the same way
    a.accumulate(other.desc(spp=7, sanitize=True), mode)
    b.accumulate(other.desc(spp=7, sanitize=True), mode)
    ctx.synchronize()
    assert np.array_equal(read_stats(a, lam)[1].view(np.uint32), read_stats(b, lam)[1].view(np.uint32))
    # no samples at all: the set is cleared
    a.accumulate(chunk.desc(spp=0), mode | sa.STATER_ACC_FRESH)
    ctx.synchronize()
    n0, st0 = read_stats(a, lam)
    assert not n0.any() and not st0.any()
