// denoise.hip — the denoiser of the statistics path on gfx950.
//
// Replaces cv::cuda::stat_denoiser::{denoiseVars, denoiseFilm, calcTargetSPP}
// called at src/statistics/estimator.cpp:374-414, 421-457 and 502-511.  The
// reference's kernel source (stat_denoiser.cu, an external submodule) is not
// part of the reference checkout, so the arithmetic follows the published
// methods the call sites and buffer names point to:
//   prepass   Jarque-Bera (1980) per channel, mean over RGB -> jb;
//             Bonett (2006) log-variance: c = n/(n-z), lnS2 = ln(c s^2),
//             se2 = c^2 (g4 - (n-3)/n)/(n-1), g4 = n m4/m2^2;
//             LEV: m2C = max(0, m2 - 12 (c2X^2 + c2Y^2)/n), the variance not
//             explained by a linear fit on the in-pixel position (var 1/12).
//   varD      weighted mean of s^2_q over neighbours q whose Bonett intervals
//             agree, (lnS2_p - lnS2_q)^2 <= z^2 (se2_p + se2_q); varCD the same
//             mean of m2C_q/(n_q-1).
//   filmD     weighted mean of the film over neighbours whose means pass a
//             two-sided Welch t-test at level alpha in every lambda set
//             (mode 2) or in the set with the smallest jb (mode 3).
//   weights   exp(ds_factor |dq|^2 + sum_k dr_k |g_p - g_q|^2) (estimator.h:353,
//             estimator.cpp:10-18); uniform for the plain method.
// A test whose operands are not finite abstains (accepts).
//
// Both filters stream a ring of TY halo rows through LDS: a 64x4 output tile
// walks the 2r+1 neighbour rows, and after each row the oldest halo row is
// replaced, so every neighbour pair sees all its operands in LDS while the
// per-workgroup footprint stays a few rows wide.  Weights are evaluated once
// per pair for all sets and channels.

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

#include "exact.hpp"
#include "kernels.hpp"

namespace stater {

constexpr int TX = 64, kTY = 4;  // output tile: 64 columns x kTY rows (one wavefront per row)
constexpr int kDecideTY = 4;     // rows of the Welch decision pass's tiles (2, 6 and 8 measured slower)
constexpr float kLog2e = 1.4426950408889634f;


FeatArgs make_feats(const stater_gbuffers *g) {
    FeatArgs f{};
    if (!g) return f;
    for (int k = 0; k < g->count; ++k)
        for (int c = 0; c < g->planes[k].channels; ++c) {
            f.ptr[f.nf] = (const float *)((const char *)g->planes[k].ptr + c * g->planes[k].channel_stride_bytes);
            f.pitch[f.nf] = g->planes[k].pitch_bytes;
            f.scale[f.nf] = std::sqrt(-g->dr_factors[k] * kLog2e);
            if (g->planes[k].dtype == STATER_BF16) f.bf16 |= 1u << f.nf;
            f.nf++;
        }
    return f;
}

constexpr int kNF = 7;  // feature channels per pair weight (unused ones stay zero)

// The kNF features of pixel (x, y), 0 beyond the set's nf.  Every load is
// issued before any value is used (a bf16 value is widened afterwards), so
// a pixel's features cost one memory latency, not one per feature.
__device__ __forceinline__ void feats_at(const FeatArgs &f, int x, int y, float (&out)[kNF]) {
    uint32_t raw[kNF];
#pragma unroll
    for (int k = 0; k < kNF; ++k) {
        raw[k] = 0u;
        if (k < f.nf) {
            const char *row = (const char *)f.ptr[k] + (int64_t)y * f.pitch[k];
            raw[k] = (f.bf16 >> k & 1u) ? (uint32_t)*(const uint16_t *)(row + (int64_t)x * 2)
                                        : *(const uint32_t *)(row + (int64_t)x * 4);
        }
    }
#pragma unroll
    for (int k = 0; k < kNF; ++k)
        out[k] = __uint_as_float((f.bf16 >> k & 1u) ? raw[k] << 16 : raw[k]) * f.scale[k];
}

// ================================================================ prepass
// One pixel-channel of one set: lnS2, se2 and m2C, and the channel's
// Jarque-Bera term (S^2 + (K - 3)^2 / 4; +inf when m2 or m4 is not finite,
// 0 without a term).  A pixel with n < 4 gets zeros; its jb is 0.
struct PreOut {
    float lnS2, se2, m2C, jbt;
};


// ======================================================= variance filter
// Two passes per group of up to 3 sets.  The pack pass writes one record per
// pixel: (ok, f0, f1, f2), (f3 .. f6) with features, then (L, Z, V, C) =
// (lnS2, z^2 se2, s^2, m2C/(n-1)) per set-channel, padded to an odd number of
// 16-byte items.  The filter pass stages rows of records into an LDS ring with
// direct global->LDS loads issued by a dedicated loader wavefront, so the
// loads of the next row are in flight while the other wavefronts test pairs.
//
// The default filter (k_var_filter1) runs one pixel per lane and eight
// filtering wavefronts (seven for r > 10) per workgroup, one tile row each:
// the ring of their live rows plus the rows in flight fills one CU's LDS, and
// two filtering wavefronts per SIMD hide each other's LDS and barrier waits
// (one per SIMD issued VALU work on 0.70 of the cycles; LDS bandwidth is not
// the limit: a CU serves about 210 bytes per clock of such reads, and the
// filter asks for a third of that).  The 4-row tiles the batch tuner may pick
// for small films run k_var_filter, whose lanes handle two vertically
// adjacent pixels, A (row 2 ty of the tile) and B (row 2 ty + 1): every
// staged neighbour row a wavefront reads serves both.  There wavefront ty
// reads row 2 ty + k at step k, so 2 TY - 1 consecutive rows are live.  In
// LDS a record keeps its odd 16-byte stride, so the 16-byte reads of 8 (or
// 16) consecutive lanes fall in distinct banks, and one base address serves
// all of a pair's reads as immediate offsets.

// Waves per SIMD that the LDS of a block allows (160 KiB per CU, 4 SIMDs),
// rounded up.  The filter kernels size their dynamic LDS at compile time, so
// the compiler is told this occupancy and schedules for latency hiding
// (LDS reads issued ahead) instead of for occupancy it can never reach.
constexpr int lds_waves_per_simd(int lds_bytes, int threads) {
    const int blocks = (160 * 1024) / lds_bytes;
    const int w = (blocks * (threads / 64) + 3) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <int G, bool HASW>
struct VarRec {
    static constexpr int SINGLE4 = HASW ? 2 : 1;                     // float4 items of ok + features
    static constexpr int ITEMS = SINGLE4 + 3 * G;                    // float4 items per record
    static constexpr int FLOAT4S = (ITEMS % 2) ? ITEMS : ITEMS + 1;  // odd
    static_assert(FLOAT4S * 4 <= kVarRecFloats, "record larger than its planes");
};

// Ring slot: the halo row's HWC records, rounded up to whole 64-item loads
// (the loader fills everything past the halo with zeros).  The ring holds the
// live rows (TY for one-row wavefronts, 2 TY - 1 for two-row ones) and AHEAD
// rows in flight: two for r <= 10, one for the wider windows, whose steps
// (2r + 1 columns) outlast a row's loads many times over and whose wider halo
// rows leave room for seven one-row wavefronts only.
constexpr int kVarAhead = 2;
template <int G, bool HASW, int HWC>
constexpr int var_loads_per_row() { return (HWC * VarRec<G, HASW>::FLOAT4S + 63) / 64; }
template <int G, bool HASW, int HWC>
constexpr int var_slot_float4s() { return var_loads_per_row<G, HASW, HWC>() * 64; }
template <int TY, int AHEAD = kVarAhead>
constexpr int var_ring_slots() { return 2 * TY - 1 + AHEAD; }
template <int G, bool HASW, int HWC, int TY, int AHEAD>
constexpr int var_lds_bytes() { return var_ring_slots<TY, AHEAD>() * var_slot_float4s<G, HASW, HWC>() * 16; }

// s_waitcnt immediates that wait for vector-memory count <= v, or LDS / scalar
// count <= v, only.
constexpr int waitcnt_vm(int v) { return (v & 15) | ((v >> 4) << 14) | (7 << 4) | (15 << 8); }
constexpr int waitcnt_lgkm(int v) { return 15 | (3 << 14) | (7 << 4) | ((v > 15 ? 15 : v) << 8); }

__device__ __forceinline__ float finite_or_max(float v) { return __builtin_isfinite(v) ? v : FLT_MAX; }

// Pack pass, with the prepass of the group's sets: one thread per pixel,
// four rows of 64 pixels per workgroup.  Each plane load of a wavefront is one
// coalesced 256-byte row segment; the pixel's record goes out item by item
// (the 16-byte stores of consecutive lanes fill consecutive records).  The
// prepass's factors of n are formed once per pixel and its quotients, roots
// and logarithm take the hardware reciprocal, reciprocal square root and
// log2 (1 ulp; the prepass is held to FP64 within 2e-5 absolute for lnS2 and
// 1e-4 relative for se2 and m2C), where the first version spent about 160
// instructions per set-channel on IEEE quotients and roots.  The record's
// s^2 and m2C/(n-1) keep the IEEE quotient 1/(n-1).
struct PreN {
    float nf, sqn, cs2, se2f, nm3n, in12, jbw;  // n, sqrt n, c/(n-1), c^2/(n-1), (n-3)/n, 12/n, n/6
};

__device__ __forceinline__ PreN pre_n(int n, float z) {
    PreN q;
    q.nf = (float)n;
    q.sqn = __builtin_amdgcn_sqrtf(q.nf);
    const float cb = q.nf * __builtin_amdgcn_rcpf(q.nf - z);
    const float inm1 = __builtin_amdgcn_rcpf(q.nf - 1.f);
    q.cs2 = cb * inm1;
    q.se2f = cb * cb * inm1;
    const float in = __builtin_amdgcn_rcpf(q.nf);
    q.nm3n = (q.nf - 3.f) * in;
    q.in12 = 12.f * in;
    q.jbw = q.nf / 6.f;
    return q;
}

// Per set-channel: lnS2 = ln(c s^2), se2 = c^2 (K - (n-3)/n)/(n-1), the LEV
// m2C and the Jarque-Bera term S^2 + (K-3)^2/4 (S = sqrt(n) m3/m2^1.5,
// K = n m4/m2^2, scale-safe: m2 enters as 1/m2 and 1/sqrt(m2)).  n < 4 is
// degenerate (zeros); a non-finite m2 or m4 gives lnS2 = se2 = +inf (the
// tests abstain); m2 <= 0 gives lnS2 = ln 1e-30 and zeros.
__device__ __forceinline__ PreOut prepass_px(int n, const PreN &q, float m2, float m3, float m4, float c2x, float c2y,
                                             bool lev) {
    PreOut o{0.f, 0.f, 0.f, 0.f};
    if (n < 4) return o;
    if (!(__builtin_isfinite(m2) && __builtin_isfinite(m4))) return PreOut{INFINITY, INFINITY, m2, INFINITY};
    if (m2 > 0.f) {
        const float im2 = __builtin_amdgcn_rcpf(m2), rsq = __builtin_amdgcn_rsqf(m2);
        const float S3 = q.sqn * (m3 * im2) * rsq;
        const float K = q.nf * (m4 * im2) * im2;
        const float k3 = K - 3.f;
        o.jbt = __builtin_fmaf(0.25f * k3, k3, S3 * S3);
        o.lnS2 = __builtin_amdgcn_logf(fmaxf(q.cs2 * m2, 1e-30f)) * 0.69314718055994531f;
        o.se2 = q.se2f * (K - q.nm3n);
        float m2c = m2;
        if (lev) {
            const float cx = c2x * rsq, cy = c2y * rsq;
            m2c = m2 * fmaxf(0.f, __builtin_fmaf(-q.in12, __builtin_fmaf(cx, cx, cy * cy), 1.f));
        }
        o.m2C = m2c;
    } else {
        o.lnS2 = -69.07755278982137f;  // ln 1e-30
    }
    return o;
}

template <int G, bool HASW>
__global__ void __launch_bounds__(256) k_var_pack(VarArgs a, float4 *zero) {
    using R = VarRec<G, HASW>;
    // each wavefront's 64 records pass through LDS, so the global stores are
    // contiguous (one store at a 176-byte lane stride touches 64 lines)
    __shared__ float4 recs[4][64 * R::FLOAT4S];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < R::FLOAT4S) zero[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = blockIdx.x * 64, x = x0 + lane, y = blockIdx.y * 4 + wave;
    if (y >= a.H) return;  // wavefront-uniform
    const bool act = x < a.W;
    const size_t pix = (size_t)y * a.ld + (act ? x : x0);
    float4 *lrec = recs[wave] + lane * R::FLOAT4S;
    if (act) {
        const int n = a.n[pix];
        float f[kNF] = {};
        if (HASW) feats_at(a.f, x, y, f);
        const PreN q = pre_n(n, a.z);
        const float inv = n >= 2 ? 1.f / (float)(n - 1) : 0.f;
        const bool lev = a.lev != 0;
        // every input of the group first: the outputs could alias them, so loads
        // after a store would each wait for the one before
        float in[G][3][5];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const float *st = a.stats + ((size_t)(a.g0 + g) * N_STATS * 3) * a.plane + pix;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                in[g][c][0] = st[((size_t)ST_M2 * 3 + c) * a.plane];
                in[g][c][1] = st[((size_t)ST_M3 * 3 + c) * a.plane];
                in[g][c][2] = st[((size_t)ST_M4 * 3 + c) * a.plane];
                in[g][c][3] = lev ? st[((size_t)ST_C2X * 3 + c) * a.plane] : 0.f;
                in[g][c][4] = lev ? st[((size_t)ST_C2Y * 3 + c) * a.plane] : 0.f;
            }
        }
        lrec[0] = make_float4(n >= 4 ? 1.f : 0.f, f[0], f[1], f[2]);
        if (HASW) lrec[1] = make_float4(f[3], f[4], f[5], f[6]);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float jbs = 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float m2 = in[g][c][0];
                const PreOut o = prepass_px(n, q, m2, in[g][c][1], in[g][c][2], in[g][c][3], in[g][c][4], lev);
                const size_t off = ((size_t)(a.g0 + g) * 3 + c) * a.plane + pix;
                a.lnS2[off] = o.lnS2;
                a.se2[off] = o.se2;
                a.m2C[off] = o.m2C;
                jbs = __builtin_fmaf(q.jbw, o.jbt, jbs);
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (n >= 4) { v.x = o.lnS2; v.y = a.z2 * o.se2; }
                if (n >= 2) {
                    v.z = finite_or_max(m2 * inv);
                    v.w = finite_or_max(o.m2C * inv);
                }
                lrec[R::SINGLE4 + g * 3 + c] = v;
            }
            a.jb[(size_t)(a.g0 + g) * a.plane + pix] = n < 4 ? 0.f : jbs / 3.f;
        }
        if (R::FLOAT4S > R::ITEMS) lrec[R::ITEMS] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // the wavefront's own records: no barrier beyond its own LDS writes
    __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(0));
    __builtin_amdgcn_wave_barrier();
    const int npx = min(64, a.W - x0);
    float4 *grec = a.rec + ((size_t)y * a.ld + x0) * R::FLOAT4S;
    for (int k = lane; k < npx * R::FLOAT4S; k += 64) grec[k] = recs[wave][k];
}

// One output pixel's state: its own statistics and the running sums.
template <int G>
struct VarPix {
    float Lp[G][3], Zp[G][3], den[G][3], num[G][3], numC[G][3];
    float gp[kNF];
};

// The pixel's own operands come from its record (all items issued at once,
// one memory latency per tile instead of one per dependent load); an
// inactive lane reads the zero record.  The sums start at zero: the pixel
// itself is one of the window's columns (dx = dy = 0), where its weight is
// exactly 1 and its test compares its interval with itself.
template <int G, bool HASW>
__device__ __forceinline__ void var_pix_init(const VarArgs &a, int x, int y, bool active, VarPix<G> &P) {
    using R = VarRec<G, HASW>;
    const size_t pix = (size_t)y * a.ld + x;
    const float4 *q = active ? a.rec + pix * R::FLOAT4S : a.zero;
    float4 v[R::ITEMS];
#pragma unroll
    for (int i = 0; i < R::ITEMS; ++i) v[i] = q[i];
    const float gp[kNF] = {v[0].y, v[0].z, v[0].w, v[HASW ? 1 : 0].x, v[HASW ? 1 : 0].y, v[HASW ? 1 : 0].z,
                           v[HASW ? 1 : 0].w};
#pragma unroll
    for (int k = 0; k < kNF; ++k) P.gp[k] = HASW ? gp[k] : 0.f;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float4 it = v[R::SINGLE4 + g * 3 + c];
            P.Lp[g][c] = it.x;
            P.Zp[g][c] = it.y;
            P.num[g][c] = P.numC[g][c] = P.den[g][c] = 0.f;
        }
}

template <int G, bool HASW>
struct VarCol {
    float4 s0, s1, v[G][3];
};

template <int G, bool HASW>
__device__ __forceinline__ void var_col_load(const float4 *q, VarCol<G, HASW> &c) {
    using R = VarRec<G, HASW>;
    c.s0 = q[0];
    c.s1 = HASW ? q[1] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int k = 0; k < 3; ++k) c.v[g][k] = q[R::SINGLE4 + g * 3 + k];
}

template <bool HASW>
__device__ __forceinline__ float var_weight(float sp_scale, const float (&gp)[kNF], float okq, const float (&fq)[kNF],
                                            float lw0, float m, float dx2) {
    float w = okq;
    if (HASW) {
        float lw = __builtin_fmaf(sp_scale, dx2, lw0);
#pragma unroll
        for (int f = 0; f < kNF; ++f) {
            const float d = gp[f] - fq[f];
            lw = __builtin_fmaf(-d, d, lw);
        }
        w *= __builtin_amdgcn_exp2f(lw);
    } else {
        w *= m;
    }
    return w;
}

__device__ __forceinline__ void var_test(float Lp, float Zp, float4 q, float w, float &den, float &num, float &numC) {
    const float dl = Lp - q.x;
    // reject iff dl^2 - z^2 se2_q > z^2 se2_p; NaN/inf operands abstain (accept)
    const float wa = (__builtin_fmaf(dl, dl, -q.y) > Zp) ? 0.f : w;
    den += wa;
    num = __builtin_fmaf(wa, q.z, num);
    numC = __builtin_fmaf(wa, q.w, numC);
}

// The pairs of neighbour column dx with A and B: both weights, then every
// set-channel's log-variance test for A and for B side by side.  A row
// outside a pixel's window comes with lw = -inf (weighted) or m = 0
// (unweighted), so its weight is exactly 0 and every sum stays as it is.
template <int G, bool HASW>
__device__ __forceinline__ void var_col_test(const VarArgs &a, const VarCol<G, HASW> &c, float dx2, float lwA,
                                             float lwB, float mA, float mB, VarPix<G> &A, VarPix<G> &B) {
    const float fq[kNF] = {c.s0.y, c.s0.z, c.s0.w, c.s1.x, c.s1.y, c.s1.z, c.s1.w};
    const float wA = var_weight<HASW>(a.sp_scale, A.gp, c.s0.x, fq, lwA, mA, dx2);
    const float wB = var_weight<HASW>(a.sp_scale, B.gp, c.s0.x, fq, lwB, mB, dx2);
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            var_test(A.Lp[g][k], A.Zp[g][k], c.v[g][k], wA, A.den[g][k], A.num[g][k], A.numC[g][k]);
            var_test(B.Lp[g][k], B.Zp[g][k], c.v[g][k], wB, B.den[g][k], B.num[g][k], B.numC[g][k]);
        }
}

// One pixel per lane: its pair with neighbour column dx.
template <int G, bool HASW>
__device__ __forceinline__ void var_col_test1(const VarArgs &a, const VarCol<G, HASW> &c, float dx2, float lw,
                                              VarPix<G> &A) {
    const float fq[kNF] = {c.s0.y, c.s0.z, c.s0.w, c.s1.x, c.s1.y, c.s1.z, c.s1.w};
    const float w = var_weight<HASW>(a.sp_scale, A.gp, c.s0.x, fq, lw, 1.f, dx2);
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int k = 0; k < 3; ++k) var_test(A.Lp[g][k], A.Zp[g][k], c.v[g][k], w, A.den[g][k], A.num[g][k], A.numC[g][k]);
}

// The 2r + 1 columns of one staged row for one pixel per lane, read ahead as
// in var_cols.
template <int G, bool HASW>
__device__ __forceinline__ void var_cols1(const VarArgs &a, const float4 *slot, int tx, float lw, VarPix<G> &A) {
    constexpr int F4 = VarRec<G, HASW>::FLOAT4S;
    constexpr int ITEMS = VarRec<G, HASW>::ITEMS;
    const float4 *q = slot + tx * F4;
    VarCol<G, HASW> X, Y;
    var_col_load<G, HASW>(q, X);
    for (int dx = -a.r; dx < a.r; dx += 2, q += 2 * F4) {
        var_col_load<G, HASW>(q + F4, Y);
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(ITEMS));
        __builtin_amdgcn_sched_barrier(0);
        var_col_test1<G, HASW>(a, X, (float)(dx * dx), lw, A);
        var_col_load<G, HASW>(q + 2 * F4, X);
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(ITEMS));
        __builtin_amdgcn_sched_barrier(0);
        var_col_test1<G, HASW>(a, Y, (float)((dx + 1) * (dx + 1)), lw, A);
    }
    var_col_test1<G, HASW>(a, X, (float)(a.r * a.r), lw, A);
}

// The 2r + 1 columns of one staged row: the LDS reads of column dx + 1 are
// issued before the tests of column dx, in two register sets that
// alternate; pairs of columns from dx = -r, then the last column alone.
template <int G, bool HASW>
__device__ __forceinline__ void var_cols(const VarArgs &a, const float4 *slot, int tx, float lwA, float lwB,
                                         float mA, float mB, VarPix<G> &A, VarPix<G> &B) {
    constexpr int F4 = VarRec<G, HASW>::FLOAT4S;
    constexpr int ITEMS = VarRec<G, HASW>::ITEMS;
    const float4 *q = slot + tx * F4;  // column of dx = -r
    VarCol<G, HASW> X, Y;
    var_col_load<G, HASW>(q, X);
    for (int dx = -a.r; dx < a.r; dx += 2, q += 2 * F4) {
        var_col_load<G, HASW>(q + F4, Y);
        // X (issued one column earlier) has landed once at most Y's reads are outstanding
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(ITEMS));
        __builtin_amdgcn_sched_barrier(0);
        var_col_test<G, HASW>(a, X, (float)(dx * dx), lwA, lwB, mA, mB, A, B);
        var_col_load<G, HASW>(q + 2 * F4, X);
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(ITEMS));
        __builtin_amdgcn_sched_barrier(0);
        var_col_test<G, HASW>(a, Y, (float)((dx + 1) * (dx + 1)), lwA, lwB, mA, mB, A, B);
    }
    var_col_test<G, HASW>(a, X, (float)(a.r * a.r), lwA, lwB, mA, mB, A, B);
}

// Step k of the tile: the staged row at dy = k - r from A (k <= 2r) and
// k - r - 1 from B (k >= 1).
template <int G, bool HASW>
__device__ __forceinline__ void var_row(const VarArgs &a, const float4 *slot, int tx, int k, VarPix<G> &A,
                                        VarPix<G> &B) {
    const int dyA = k - a.r, dyB = k - a.r - 1;
    const bool offA = k > 2 * a.r, offB = k < 1;  // row outside the pixel's window
    const float lwA = offA ? -INFINITY : a.sp_scale * (float)(dyA * dyA);
    const float lwB = offB ? -INFINITY : a.sp_scale * (float)(dyB * dyB);
    const float mA = offA ? 0.f : 1.f, mB = offB ? 0.f : 1.f;
    var_cols<G, HASW>(a, slot, tx, lwA, lwB, mA, mB, A, B);
}

// Whether the pixel's own column may need var_own_fix: the loop rejected it
// (its own interval is empty), or its lnS2 is not finite, which the prepass
// sets whenever m2 (and so the record's s^2 or m2C/(n-1)) is not finite.
__device__ __forceinline__ bool var_own_suspect(float Lp, float Zp) {
    const float dl = Lp - Lp;
    return __builtin_fmaf(dl, dl, -Zp) > Zp || !__builtin_isfinite(Lp);
}

// The pixel's own column as the pair loop left it, made what the filter
// defines: the pixel itself always counts with weight 1 (the loop rejects it
// only when its own interval is empty, z^2 se2 < 0), and with its exact s^2
// (the record holds FLT_MAX for a non-finite one).  Rare: only pixel-channels
// that var_own_suspect flags come here.
template <int G, bool HASW>
__device__ __forceinline__ void var_own_fix(const VarArgs &a, size_t pix, int g, int c, float Lp, float Zp, float &den,
                                            float &num, float &numC) {
    using R = VarRec<G, HASW>;
    const float dl = Lp - Lp;
    const bool rejected = __builtin_fmaf(dl, dl, -Zp) > Zp;
    const float4 own = a.rec[pix * R::FLOAT4S + R::SINGLE4 + g * 3 + c];
    if (!rejected && own.z != FLT_MAX && own.w != FLT_MAX) return;
    if (rejected) {
        den += 1.f;
        num += own.z;
        numC += own.w;
    }
    const float inv = 1.f / (float)(a.n[pix] - 1);
    if (own.z == FLT_MAX) num = (num - FLT_MAX) + a.m2[(size_t)(a.g0 + g) * a.set_stride + c * a.plane + pix] * inv;
    if (own.w == FLT_MAX) numC = (numC - FLT_MAX) + a.m2C[((size_t)(a.g0 + g) * 3 + c) * a.plane + pix] * inv;
}

// A pixel with fewer than 4 samples is not filtered: it keeps its own
// variance, read again here so that the pair loop runs for it like for
// every other lane.  Both pixels of the lane are written in one pass over the
// set-channels, with plain per-pixel values (no selection between the two
// pixels' state).
template <int G, bool HASW>
__device__ __forceinline__ void var_store1(const VarArgs &a, int x, int y, VarPix<G> &A) {
    const size_t pA = (size_t)y * a.ld + x;
    const int nA = a.n[pA];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t off = ((size_t)(a.g0 + g) * 3 + c) * a.plane;
            const size_t m2off = (size_t)(a.g0 + g) * a.set_stride + c * a.plane;
            if (nA >= 4 && var_own_suspect(A.Lp[g][c], A.Zp[g][c]))
                var_own_fix<G, HASW>(a, pA, g, c, A.Lp[g][c], A.Zp[g][c], A.den[g][c], A.num[g][c], A.numC[g][c]);
            float VA = A.num[g][c] / A.den[g][c], CA = A.numC[g][c] / A.den[g][c];
            if (nA < 4) {
                const float inv = nA >= 2 ? 1.f / (float)(nA - 1) : 0.f;
                VA = nA >= 2 ? a.m2[m2off + pA] * inv : 0.f;
                CA = nA >= 2 ? a.m2C[off + pA] * inv : 0.f;
            }
            a.varD[off + pA] = VA;
            a.varCD[off + pA] = CA;
        }
}

template <int G, bool HASW>
__device__ __forceinline__ void var_store(const VarArgs &a, int x, int yA, bool actA, bool actB, VarPix<G> &A,
                                          VarPix<G> &B) {
    const size_t pA = (size_t)yA * a.ld + x, pB = pA + a.ld;
    const int nA = actA ? a.n[pA] : 4, nB = actB ? a.n[pB] : 4;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t off = ((size_t)(a.g0 + g) * 3 + c) * a.plane;
            const size_t m2off = (size_t)(a.g0 + g) * a.set_stride + c * a.plane;
            if (actA && nA >= 4 && var_own_suspect(A.Lp[g][c], A.Zp[g][c]))
                var_own_fix<G, HASW>(a, pA, g, c, A.Lp[g][c], A.Zp[g][c], A.den[g][c], A.num[g][c], A.numC[g][c]);
            if (actB && nB >= 4 && var_own_suspect(B.Lp[g][c], B.Zp[g][c]))
                var_own_fix<G, HASW>(a, pB, g, c, B.Lp[g][c], B.Zp[g][c], B.den[g][c], B.num[g][c], B.numC[g][c]);
            float VA = A.num[g][c] / A.den[g][c], CA = A.numC[g][c] / A.den[g][c];
            float VB = B.num[g][c] / B.den[g][c], CB = B.numC[g][c] / B.den[g][c];
            if (nA < 4) {
                const float inv = nA >= 2 ? 1.f / (float)(nA - 1) : 0.f;
                VA = nA >= 2 ? a.m2[m2off + pA] * inv : 0.f;
                CA = nA >= 2 ? a.m2C[off + pA] * inv : 0.f;
            }
            if (nB < 4) {
                const float inv = nB >= 2 ? 1.f / (float)(nB - 1) : 0.f;
                VB = nB >= 2 ? a.m2[m2off + pB] * inv : 0.f;
                CB = nB >= 2 ? a.m2C[off + pB] * inv : 0.f;
            }
            if (actA) { a.varD[off + pA] = VA; a.varCD[off + pA] = CA; }
            if (actB) { a.varD[off + pB] = VB; a.varCD[off + pB] = CB; }
        }
}

typedef __attribute__((address_space(3))) void lds_void;

// Loader wavefront: row yq of a tile's halo (columns x0 - r .. x0 + 63 + r)
// of F4-item records into one ring slot, with IPR direct 64-item loads for
// every row, so that a count of outstanding loads names a number of rows.
// Items of columns and rows outside the image, and every item past the halo,
// read the zero record.
template <int F4, int IPR>
__device__ __forceinline__ void ring_load_row(const float4 *rec, const float4 *zero, int W, int H, int ld, int r,
                                              float4 *dst, int yq, int x0, int lane) {
    const int nch = (TX + 2 * r) * F4;
    const bool row_in = yq >= 0 && yq < H;
    const float4 *row = rec + (size_t)(row_in ? yq : 0) * ld * F4;
#pragma unroll
    for (int i = 0; i < IPR; ++i) {
        const int ch = i * 64 + lane;
        const int col = ch / F4, item = ch - col * F4;
        const int xq = x0 - r + col;
        const bool real = ch < nch && row_in && xq >= 0 && xq < W;
        const float4 *src = real ? row + (size_t)xq * F4 + item : zero + item;
        __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)(dst + i * 64), 16, 0, 0);
    }
}

template <int G, bool HASW, int HWC>
__device__ __forceinline__ void var_load_row(const VarArgs &a, float4 *dst, int yq, int x0, int lane) {
    ring_load_row<VarRec<G, HASW>::FLOAT4S, var_loads_per_row<G, HASW, HWC>()>(a.rec, a.zero, a.W, a.H, a.ld, a.r,
                                                                                dst, yq, x0, lane);
}

// One tile of 64 x 2 TY pixels per workgroup: TY filtering wavefronts (two
// rows each) and one loader wavefront.  Wavefront ty reads row 2 ty + k at
// step k, so rows k .. k + P - 1 (P = 2 TY - 1) are live during step k while
// the loader keeps AHEAD rows in flight; row j lives in slot j mod
// (P + AHEAD).  Both roles pass the same 2r + 3 barriers: the first once
// rows 0 .. P - 1 have landed, then one at the end of every step k, by which
// row k + P has landed (it replaces nothing live: its slot held row k - 1 or
// was never used).  Built for the 4-row tiles of the r <= 10
// feature-weighted filter, which the batch tuner may choose per film.
template <int G, bool HASW, int HWC, int TY, int AHEAD>
__global__ void __launch_bounds__(64 * (TY + 1))
    __attribute__((amdgpu_waves_per_eu(1, lds_waves_per_simd(var_lds_bytes<G, HASW, HWC, TY, AHEAD>(),
                                                             64 * (TY + 1))))) k_var_filter(VarArgs a) {
    extern __shared__ float4 vring[];
    constexpr int P = 2 * TY - 1, NS = var_ring_slots<TY, AHEAD>();
    constexpr int SLOT = var_slot_float4s<G, HASW, HWC>();
    constexpr int AHEAD_LOADS = var_loads_per_row<G, HASW, HWC>() * (AHEAD - 1);
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + 2 * TY - 1) / (2 * TY)));
    if (t < 0) return;
    const int x0 = (t % tiles_x) * TX, y0 = (t / tiles_x) * 2 * TY;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int last = 2 * a.r + 1;  // steps 0 .. last

    // The loader meets the filtering wavefronts' __syncthreads() with a bare
    // s_barrier: a workgroup fence there would wait for the rows still in
    // flight too.  Its own counter wait covers the rows the next step reads.
    if (wave == TY) {  // loader
        for (int j = 0; j < P + AHEAD - 1; ++j)
            var_load_row<G, HASW, HWC>(a, vring + j * SLOT, y0 - a.r + j, x0, lane);
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));  // rows 0 .. P - 1 have landed
        __builtin_amdgcn_s_barrier();
        for (int k = 0; k <= last; ++k) {
            const int j = k + P + AHEAD - 1;
            var_load_row<G, HASW, HWC>(a, vring + (j % NS) * SLOT, y0 - a.r + j, x0, lane);
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));  // row k + P has landed
            __builtin_amdgcn_s_barrier();
        }
        __builtin_amdgcn_s_waitcnt(0);  // nothing may land in LDS after the workgroup ends
        return;
    }

    const int x = x0 + lane, yA = y0 + 2 * wave;
    const bool actA = x < a.W && yA < a.H, actB = x < a.W && yA + 1 < a.H;
    VarPix<G> A, B;
    var_pix_init<G, HASW>(a, x, yA, actA, A);
    var_pix_init<G, HASW>(a, x, yA + 1, actB, B);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));  // the own records have landed: no counter waits in the pair loop
    __syncthreads();
    for (int k = 0; k <= last; ++k) {
        const float4 *slot = vring + ((2 * wave + k) % NS) * SLOT;
        if (actA) var_row<G, HASW>(a, slot, lane, k, A, B);
        __syncthreads();
    }
    var_store<G, HASW>(a, x, yA, actA, actB, A, B);
}

// One pixel per lane: TY filtering wavefronts of one row each (wavefront ty
// reads row ty + k at step k, k = 0 .. 2r) and the loader.  Per output the
// LDS reads double against the two-row lanes, but the ring of TY + AHEAD rows
// holds eight filtering wavefronts in one CU's LDS, two per SIMD, so one
// wavefront's waits are filled by the other's arithmetic.
template <int G, bool HASW, int HWC, int TY, int AHEAD>
__global__ void __launch_bounds__(64 * (TY + 1))
    __attribute__((amdgpu_waves_per_eu(1, 2))) k_var_filter1(VarArgs a) {
    extern __shared__ float4 vring[];
    constexpr int P = TY, NS = TY + AHEAD;
    constexpr int SLOT = var_slot_float4s<G, HASW, HWC>();
    constexpr int AHEAD_LOADS = var_loads_per_row<G, HASW, HWC>() * (AHEAD - 1);
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    const int x0 = (t % tiles_x) * TX, y0 = (t / tiles_x) * TY;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int last = 2 * a.r;  // steps 0 .. last

    if (wave == TY) {  // loader
        for (int j = 0; j < P + AHEAD - 1; ++j)
            var_load_row<G, HASW, HWC>(a, vring + j * SLOT, y0 - a.r + j, x0, lane);
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));  // rows 0 .. P - 1 have landed
        __builtin_amdgcn_s_barrier();
        for (int k = 0; k <= last; ++k) {
            const int j = k + P + AHEAD - 1;
            var_load_row<G, HASW, HWC>(a, vring + (j % NS) * SLOT, y0 - a.r + j, x0, lane);
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));  // row k + P has landed
            __builtin_amdgcn_s_barrier();
        }
        __builtin_amdgcn_s_waitcnt(0);
        return;
    }

    const int x = x0 + lane, y = y0 + wave;
    const bool act = x < a.W && y < a.H;
    VarPix<G> A;
    var_pix_init<G, HASW>(a, x, y, act, A);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    __syncthreads();
    for (int k = 0; k <= last; ++k) {
        const float4 *slot = vring + ((wave + k) % NS) * SLOT;
        const int dy = k - a.r;
        if (act) var_cols1<G, HASW>(a, slot, lane, a.sp_scale * (float)(dy * dy), A);
        __syncthreads();
    }
    if (act) var_store1<G, HASW>(a, x, y, A);
}

template <int G, bool HASW, int HWC, int TY, int AHEAD>
static int var_launch1(const VarArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_var_pack<G, HASW>), dim3((a.W + 63) / 64, (a.H + 3) / 4), dim3(256), 0, st, a,
                       const_cast<float4 *>(a.zero));
    STATER_HIP(hipGetLastError());
    constexpr size_t lds = (size_t)(TY + AHEAD) * var_slot_float4s<G, HASW, HWC>() * 16;
    static_assert(lds <= 160 * 1024, "ring larger than a CU's LDS");
    STATER_HIP(hipFuncSetAttribute((const void *)k_var_filter1<G, HASW, HWC, TY, AHEAD>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY)));
    hipLaunchKernelGGL((k_var_filter1<G, HASW, HWC, TY, AHEAD>), grid, dim3(64 * (TY + 1)), lds, st, a);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

template <int G, bool HASW, int HWC, int TY, int AHEAD = 2>
static int var_launch4(const VarArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_var_pack<G, HASW>), dim3((a.W + 63) / 64, (a.H + 3) / 4), dim3(256), 0, st, a,
                       const_cast<float4 *>(a.zero));
    STATER_HIP(hipGetLastError());
    const size_t lds = (size_t)var_lds_bytes<G, HASW, HWC, TY, AHEAD>();
    static_assert(var_lds_bytes<G, HASW, HWC, TY, AHEAD>() <= 160 * 1024, "ring larger than a CU's LDS");
    STATER_HIP(hipFuncSetAttribute((const void *)k_var_filter<G, HASW, HWC, TY, AHEAD>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + 2 * TY - 1) / (2 * TY))));
    hipLaunchKernelGGL((k_var_filter<G, HASW, HWC, TY, AHEAD>), grid, dim3(64 * (TY + 1)), lds, st, a);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

// rows: output rows per tile (4 or 8; two per filtering wavefront).
template <int G>
static int var_launch(const VarArgs &a, bool hasw, int rows, hipStream_t st) {
    if (a.r > 10)
        return hasw ? var_launch1<G, true, TX + 40, 7, 1>(a, st) : var_launch1<G, false, TX + 40, 7, 1>(a, st);
    if (!hasw) return var_launch1<G, false, TX + 20, 8, 2>(a, st);
    return rows == 8 ? var_launch1<G, true, TX + 20, 8, 2>(a, st) : var_launch4<G, true, TX + 20, 2>(a, st);
}

VarArgs make_var_args(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g, int g0) {
    VarArgs a{};
    a.n = bs->n;
    a.m2 = bs->stat(0, ST_M2, 0);
    a.lnS2 = bs->lnS2; a.se2 = bs->se2; a.m2C = bs->m2C; a.jb = bs->jb;
    a.stats = bs->stats;
    a.varD = bs->varD; a.varCD = bs->varCD;
    a.f = make_feats(g);
    a.W = bs->W; a.H = bs->H; a.ld = bs->ld; a.r = p->radius; a.g0 = g0;
    a.plane = bs->plane;
    a.set_stride = (size_t)N_STATS * 3 * bs->plane;
    a.z = p->var_z;
    a.z2 = p->var_z * p->var_z;
    a.lev = bs->mm >= 5 ? 1 : 0;
    a.sp_scale = p->ds_factor * kLog2e;
    a.rec = (float4 *)bs->var_rec;
    a.zero = (const float4 *)bs->var_zero;
    return a;
}

int launch_var_filter(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g, hipStream_t st,
                      int ty) {
    VarArgs a = make_var_args(bs, p, g, 0);
    const bool hasw = a.f.nf > 0 || p->ds_factor != 0.f;
    LaunchTimer t(bs->ctx, st, "var_filter");
    for (int g0 = 0; g0 < bs->S;) {
        const int left = bs->S - g0;
        const int G = left >= 3 ? 3 : left;
        a.g0 = g0;
        int rc = G == 3 ? var_launch<3>(a, hasw, ty, st) : G == 2 ? var_launch<2>(a, hasw, ty, st)
                                                                   : var_launch<1>(a, hasw, ty, st);
        if (rc) return rc;
        g0 += G;
    }
    return STATER_OK;
}

// ========================================================== Welch filter
// LDS ring layout (floats), per slot: single rows (ok, n-1, padded features,
// film[3], plain: s^2[3]) then one float2 row (M, V) per set-channel.

// FP64 tail for pairs near the critical value (rare; kept out of line).
__device__ __noinline__ bool welch_exact(float mp, float vp, float nm1p, float mq, float vq, float nm1q, double alpha) {
    const double V = (double)vp + (double)vq, dm = (double)mp - (double)mq;
    if (!(V > 0.0)) return dm == 0.0;
    const double t2 = dm * dm / V;
    const double Wd = (double)vp * vp / (double)nm1p + (double)vq * vq / (double)nm1q;
    const double nu = Wd > 0.0 ? V * V / Wd : INFINITY;
    return t2_tail(t2, nu) > alpha;
}

constexpr float kUBig = 281474976710656.f;  // 2^48: pushes e below zero for u > umax

// Per-pair factors of u = 1/nu.  With ap = vp/V and aq = 1 - ap,
// u = rp ap^2 + rq aq^2 = (A ap + B) ap + C, A = rp + rq, B = -2 rq, C = rq
// (r = 1/(n-1)).  Its absolute rounding error stays below ~4 rq ulp(1),
// negligible against the band for rq <= 1/17; windows holding a smaller n take
// the scale-free form with aq = vq/V.
struct PairR {
    float rp, rq, A, B, C;
};

__device__ __forceinline__ PairR pair_r(float rp, float rq) { return PairR{rp, rq, rp + rq, -2.f * rq, rq}; }

// t2 = dm^2 / V and the fitted critical value at u = 1/nu (clamped to umax)
// of one FP32 Welch test.
template <bool SMALLN>
__device__ __forceinline__ void welch_t2_crit(const TCritPoly &tp, float mp, float vp, float mq, float vq,
                                              const PairR &r, float &t2, float &crit, float &u) {
    const float V = vp + vq;
    const float iV = __builtin_amdgcn_rcpf(V);
    const float dm = mp - mq;
    t2 = (dm * iV) * dm;
    const float ap = vp * iV;
    if (SMALLN) {
        const float aq = vq * iV;
        u = __builtin_fmaf(ap * ap, r.rp, (aq * aq) * r.rq);
    } else {
        u = __builtin_fmaf(__builtin_fmaf(r.A, ap, r.B), ap, r.C);
    }
    const float uc = fminf(u, tp.umax);  // NaN (V = 0) -> umax
    crit = tp.c[kTCritDeg];
#pragma unroll
    for (int k = kTCritDeg - 1; k >= 0; --k) crit = __builtin_fmaf(crit, uc, tp.c[k]);
}

// Decision margins of one FP32 Welch test, t2 = dm^2/V against crit(u):
//   e > 0: t2 > (1 + band) crit(u), u <= umax -> rejected beyond doubt;
//   f < 0: t2 < (1 - band) crit(u)            -> accepted beyond doubt
// (crit grows with u, so f < 0 stays valid for u > umax).  A NaN mean
// (abstain) or V = 0 with equal means gives NaN margins, which the max
// reductions skip (accept); V = 0 with unequal means gives t2 = inf, e = inf.
// SMALLN: some n - 1 of the window may be below 1/umax; otherwise u <= umax
// holds for every pair and the range term is left out.
template <bool SMALLN>
__device__ __forceinline__ void welch_margin(const TCritPoly &tp, float mp, float vp, float mq, float vq,
                                             const PairR &r, float &e, float &f) {
    float t2, crit, u;
    welch_t2_crit<SMALLN>(tp, mp, vp, mq, vq, r, t2, crit, u);
    e = __builtin_fmaf(crit, tp.neg_hi, t2);
    f = __builtin_fmaf(crit, tp.neg_lo, t2);
    if (SMALLN) e = __builtin_fmaf(-kUBig, fmaxf(u - tp.umax, 0.f), e);
}

// Both margins of the large-n case as one number, g = t2 - crit(u), with
// the degree-D fit c valid on u <= 1/m where every n - 1 of the window is at
// least m: g > band_abs (band times the largest crit of the fit's range)
// implies e > 0 and g < -band_abs implies f < 0.  The fma forms t2 - crit
// with one rounding, so g is exact to within an ulp of itself.  The weight
// share ap = vp/V takes the legacy multiply (0 x inf = 0), so V = 0 (both
// variances zero; they are never negative) gives u = rq, inside the fit's
// range, without a clamp: t2 is then +inf (rejected) for unequal means and
// NaN (skipped by the max reduction: accepted) for equal ones.
// The legacy multiply (v_mul_legacy_f32: 0 x anything = 0), an LLVM
// intrinsic without a clang builtin.
extern "C" __device__ float stater_fmul_legacy(float, float) __asm("llvm.amdgcn.fmul.legacy");

template <int D>
__device__ __forceinline__ float welch_gap(const float (&c)[D + 1], float mp, float vp, float mq, float vq,
                                           const PairR &r) {
    const float V = vp + vq;
    const float iV = __builtin_amdgcn_rcpf(V);
    const float dm = mp - mq;
    const float ap = stater_fmul_legacy(vp, iV);
    const float u = __builtin_fmaf(__builtin_fmaf(r.A, ap, r.B), ap, r.C);
    float crit = c[D];
#pragma unroll
    for (int k = D - 1; k >= 0; --k) crit = __builtin_fmaf(crit, u, c[k]);
    return __builtin_fmaf(dm * iV, dm, -crit);
}

// The fast margins of one pair per channel, max-reduced over the NT tested
// sets: E[c] > 0 rejected beyond doubt, F[c] < 0 accepted beyond doubt.  With
// every n - 1 of the window large, one running maximum of g per channel with
// the band applied once (welch_gap); otherwise both margins (welch_margin).
// The decision pass keeps its own copy of this loop, which schedules two
// instructions per column shorter there.
template <int NT, bool SMALLN>
__device__ __forceinline__ void welch_fast(const TCritPoly &tp, const float (&Mp)[NT][3], const float (&Vp)[NT][3],
                                           const float2 (&mv)[NT][3], const PairR &pr, float (&E)[3],
                                           float (&F)[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) E[c] = F[c] = -INFINITY;
    if (SMALLN) {
#pragma unroll
        for (int s = 0; s < NT; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float e, f;
                welch_margin<SMALLN>(tp, Mp[s][c], Vp[s][c], mv[s][c].x, mv[s][c].y, pr, e, f);
                E[c] = fmaxf(E[c], e);
                F[c] = fmaxf(F[c], f);
            }
    } else {
#pragma unroll
        for (int s = 0; s < NT; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                E[c] = fmaxf(E[c], welch_gap<kTCritDeg>(tp.c, Mp[s][c], Vp[s][c], mv[s][c].x, mv[s][c].y, pr));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            F[c] = E[c] + tp.band_abs;
            E[c] = E[c] - tp.band_abs;
        }
    }
}

// Per-pixel test operands: M (mean; NaN = abstain) and V = var/n (discr).
__device__ __forceinline__ void welch_load(const WelchArgs &a, int xq, int yq, int s, int c, float &M, float &V) {
    const size_t pix = (size_t)yq * a.ld + xq;
    const int n = a.n[pix];
    M = s == 0 ? *(const float *)((const char *)a.film + (int64_t)yq * a.film_pitch + (int64_t)xq * a.film_xs + c * a.film_cs)
               : a.mean[(size_t)s * a.set_stride + c * a.plane + pix];
    float var;
    if (a.var_is_m2) var = n >= 2 ? a.var[(size_t)s * a.set_stride + c * a.plane + pix] / (float)(n - 1) : 0.f;
    else var = a.var[((size_t)s * 3 + c) * a.plane + pix];
    V = n >= 1 ? var / (float)n : 0.f;
    if (!(__builtin_isfinite(M) && __builtin_isfinite(V))) { M = NAN; V = 0.f; }
}

template <int S, bool HASW, bool PLAIN>
struct WelchLayout {
    static constexpr int FEAT = 2;                        // after ok, n-1
    static constexpr int FILM = 2 + (HASW ? kNF : 0);
    static constexpr int S2 = FILM + 3;                   // plain: s^2 [3]
    static constexpr int SINGLES = FILM + 3 + (PLAIN ? 3 : 0);
    static constexpr int FLOATS = SINGLES + 6 * S;        // per column
};

template <int S, bool HASW, bool PLAIN, int HWC>
__device__ __forceinline__ void welch_stage_row(const WelchArgs &a, float *slot, int yq, int x0) {
    using LY = WelchLayout<S, HASW, PLAIN>;
    const int HW = TX + 2 * a.r;
    const int items = HW * (1 + 3 * S);
    float2 *pairs = (float2 *)(slot + LY::SINGLES * HWC);
    for (int i = threadIdx.x; i < items; i += blockDim.x) {
        const int kind = i / HW, col = i - kind * HW;
        const int xq = x0 - a.r + col;
        const bool in = xq >= 0 && xq < a.W && yq >= 0 && yq < a.H;
        if (kind == 0) {
            float ok = 0.f, nm1 = 1.f;
            int n = 0;
            if (in) {
                n = a.n[(size_t)yq * a.ld + xq];
                ok = n >= 2 ? 1.f : 0.f;
                nm1 = n >= 2 ? (float)(n - 1) : 1.f;
            }
            slot[col] = ok;
            slot[HWC + col] = nm1;
            if (HASW) {
                float fv[kNF] = {};
                if (in) feats_at(a.f, xq, yq, fv);
#pragma unroll
                for (int k = 0; k < kNF; ++k) slot[(LY::FEAT + k) * HWC + col] = fv[k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                slot[(LY::FILM + c) * HWC + col] =
                    in ? *(const float *)((const char *)a.film + (int64_t)yq * a.film_pitch + (int64_t)xq * a.film_xs +
                                          c * a.film_cs)
                       : 0.f;
            if (PLAIN)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float s2 = 0.f;
                    if (in && n >= 2) s2 = a.var[c * a.plane + (size_t)yq * a.ld + xq] / (float)(n - 1);
                    slot[(LY::S2 + c) * HWC + col] = finite_or_max(s2);
                }
        } else {
            const int sc = kind - 1, s = sc / 3, c = sc - s * 3;
            float M = 0.f, V = 0.f;
            if (in) welch_load(a, xq, yq, s, c, M, V);
            pairs[sc * HWC + col] = make_float2(M, V);
        }
    }
}

// Below this n - 1 a pair may have nu < 1/umax (u <= max(1/(np-1), 1/(nq-1))).
constexpr int kWelchLargeNm1 = 17;
// From this n - 1 on, every pair has u <= kTCritUmaxShort.
constexpr int kWelchShortNm1 = 128;

// The tile's window level (block-uniform), over the pixels of its window
// region with n >= 2: 0 if some n - 1 < kWelchLargeNm1, 2 if every
// n - 1 >= kWelchShortNm1, 1 otherwise.  UNIFORM: level 3 instead of 2 when
// all those pixels have the same n, which is returned in *nuni.
template <bool UNIFORM = false>
__device__ __forceinline__ int welch_window_level(const WelchArgs &a, int x0, int y0, int TYv, int *nuni = nullptr) {
    __shared__ int s_lo, s_hi;
    if (UNIFORM) {
        if (threadIdx.x == 0) {
            s_lo = INT_MAX;
            s_hi = 0;
        }
        __syncthreads();
    }
    const int xa = max(x0 - a.r, 0), xb = min(x0 + TX + a.r, a.W);
    const int ya = max(y0 - a.r, 0), yb = min(y0 + TYv + a.r, a.H);
    const int w = xb - xa, cnt = w * (yb - ya);
    int small = 0, mid = 0, lo = INT_MAX, hi = 0;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int yy = ya + i / w, xx = xa + i % w;
        const int n = a.n[(size_t)yy * a.ld + xx];
        small |= (n >= 2 && n - 1 < kWelchLargeNm1) ? 1 : 0;
        mid |= (n >= 2 && n - 1 < kWelchShortNm1) ? 1 : 0;
        if (UNIFORM && n >= 2) {
            lo = min(lo, n);
            hi = max(hi, n);
        }
    }
    if (UNIFORM && lo <= hi) {
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
    }
    if (__syncthreads_or(small)) return 0;
    if (__syncthreads_or(mid)) return 1;
    if (UNIFORM && s_lo == s_hi) {
        *nuni = s_lo;
        return 3;
    }
    return 2;
}

// The test on a window whose pixels (those with n >= 2) all have the same n,
// n - 1 >= kWelchShortNm1.  With r = 1/(n - 1) and the weight shares
// ap = vp/V, aq = vq/V (ap + aq = 1), 1/nu = r (ap^2 + aq^2) = r (1 - 2P)
// with P = ap aq in [0, 1/4], so the short fit crit(u) = cs0 + cs1 u + cs2 u^2
// is a quadratic a0 + a1 P + a2 P^2 in P.  Its minimax line on [0, 1/4]
// (P^2 -> P/4 - 1/128, off by at most 1/128) gives
//     g = t2 - crit ~ t2 - alpha - beta ap aq,
// three instructions after t2 where the general form takes five; the band
// widens by |a2|/128 (5e-5 at n = 256, against crit ~ 7.9) and two ulp of
// alpha for its rounding.  nbeta = -beta.
struct WelchUniform {
    float alpha, nbeta, band;
};

static_assert(kTCritDegShort == 2, "the uniform-n test expands the short fit as a quadratic");

__device__ __forceinline__ WelchUniform welch_uniform(const TCritPoly &tp, int n) {
    const double r = 1.0 / (double)(n - 1);
    const double c0 = tp.cs[0], c1 = tp.cs[1], c2 = tp.cs[2];
    const double a0 = c0 + r * (c1 + r * c2), a1 = -2.0 * r * (c1 + 2.0 * r * c2), a2 = 4.0 * r * r * c2;
    const double alpha = a0 - a2 / 128.0;
    WelchUniform U;
    U.alpha = (float)alpha;
    U.nbeta = (float)-(a1 + 0.25 * a2);
    U.band = (float)((double)tp.band_abs_short + fabs(a2) / 128.0 + fabs(alpha) * 0x1p-22);
    return U;
}

// g of the uniform-n form; vpn = nbeta vp.  The legacy multiplies keep V = 0
// (both variances zero) and V = inf as welch_gap takes them: the shares are 0
// there, so g = t2 - alpha (+inf for unequal means with V = 0: rejected; NaN
// for equal ones: skipped by the max reduction, accepted).
__device__ __forceinline__ float welch_gap_uniform(const WelchUniform &U, float mp, float vp, float vpn, float mq,
                                                   float vq) {
    const float V = vp + vq;
    const float iV = __builtin_amdgcn_rcpf(V);
    const float dm = mp - mq;
    const float h = __builtin_fmaf(dm * iV, dm, -U.alpha);
    return __builtin_fmaf(stater_fmul_legacy(vpn, iV), stater_fmul_legacy(vq, iV), h);
}

// One window row of the Welch filter for the pixel of this lane.
// MODE3: only the set of minimum Jarque-Bera statistic (per pixel) is tested.
template <int S, bool HASW, bool PLAIN, bool MODE3, bool SMALLN, int HWC>
__device__ __forceinline__ void welch_row(const WelchArgs &a, const float *slot, int tx, int x, int y, int dy,
                                          const float (&Mp)[MODE3 ? 1 : S][3], const float (&Vp)[MODE3 ? 1 : S][3],
                                          const float (&gp)[kNF], float okp, float nm1p, float rp, int bset,
                                          float (&den)[3], float (&num)[3], float (&numV)[3]) {
    using LY = WelchLayout<S, HASW, PLAIN>;
    constexpr int NT = MODE3 ? 1 : S;  // sets tested per pair
    const int win = 2 * a.r + 1;
    const float spy = a.sp_scale * (float)(dy * dy);
    for (int dx = -a.r; dx <= a.r; ++dx) {
        if (dx == 0 && dy == 0) continue;
        const float *q = slot + tx + a.r + dx;
        // Operands of set s, channel c of q: pq[(s*3 + c) * HWC]; MODE3 reads
        // the lane's own set.  Every LDS read of the pair is issued before any
        // use, so the wave waits once per pair rather than once per test.
        const float2 *pq = (const float2 *)(slot + LY::SINGLES * HWC) + tx + a.r + dx + (MODE3 ? bset * 3 * HWC : 0);
        float fq[kNF], filmq[3];
        float2 mv[NT][3];
        const float okq = q[0];
        const float nm1q = q[HWC];
#pragma unroll
        for (int f = 0; f < (HASW ? kNF : 0); ++f) fq[f] = q[(LY::FEAT + f) * HWC];
#pragma unroll
        for (int s = 0; s < NT; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) mv[s][c] = pq[(s * 3 + c) * HWC];
#pragma unroll
        for (int c = 0; c < 3; ++c) filmq[c] = q[(LY::FILM + c) * HWC];
        __builtin_amdgcn_sched_barrier(0);
        const PairR pr = pair_r(rp, __builtin_amdgcn_rcpf(nm1q));
        const float okpq = okq * okp;
        float w = okpq;
        if (HASW) {
            float lw = __builtin_fmaf(a.sp_scale, (float)(dx * dx), spy);
#pragma unroll
            for (int f = 0; f < kNF; ++f) {
                const float d = gp[f] - fq[f];
                lw = __builtin_fmaf(-d, d, lw);
            }
            w *= __builtin_amdgcn_exp2f(lw);
        }
        float E[3], F[3];
        welch_fast<NT, SMALLN>(a.tp, Mp, Vp, mv, pr, E, F);
        // A channel neither rejected nor accepted beyond doubt by its fast
        // tests is decided by the exact tails of its tests.
        bool near = false;
#pragma unroll
        for (int c = 0; c < 3; ++c) near |= !(E[c] > 0.f) && F[c] >= 0.f;
        if (near && okpq > 0.f) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (E[c] > 0.f || F[c] < 0.f) continue;
                bool pass = true;
                for (int s = 0; s < NT; ++s) {
                    const float2 mv = pq[(s * 3 + c) * HWC];
                    float e, f;
                    welch_margin<SMALLN>(a.tp, Mp[s][c], Vp[s][c], mv.x, mv.y, pr, e, f);
                    if (e > 0.f) { pass = false; break; }
                    if (f >= 0.f && !welch_exact(Mp[s][c], Vp[s][c], nm1p, mv.x, mv.y, nm1q, a.alpha)) {
                        pass = false;
                        break;
                    }
                }
                E[c] = pass ? -1.f : 1.f;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float wa = E[c] > 0.f ? 0.f : w;
            den[c] += wa;
            num[c] = __builtin_fmaf(wa, filmq[c], num[c]);
            if (PLAIN) numV[c] = __builtin_fmaf(wa, q[(LY::S2 + c) * HWC], numV[c]);
        }
        if (a.dec) {
            const size_t kk = ((size_t)(y * a.W + x) * win * win + (size_t)(dy + a.r) * win + (dx + a.r)) * 3;
            const int xq = x + dx, yq = y + dy;
            const bool in = xq >= 0 && xq < a.W && yq >= 0 && yq < a.H;
            for (int c = 0; c < 3; ++c) a.dec[kk + c] = (in && okpq > 0.f && !(E[c] > 0.f)) ? 1 : 0;
        }
    }
}

template <int S, bool HASW, bool PLAIN, bool MODE3, int HWC, int TY>
__device__ __forceinline__ void welch_tile(const WelchArgs a, int bx, int by, float *ring) {
    using LY = WelchLayout<S, HASW, PLAIN>;
    constexpr int SLOT = LY::FLOATS * HWC;
    constexpr int NT = MODE3 ? 1 : S;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = bx * TX, y0 = by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < a.W && y < a.H;

    float Mp[NT][3], Vp[NT][3];
    float gp[kNF];
    float okp = 0.f, nm1p = 1.f, rp = 0.f;
    float den[3], num[3], numV[3];
    int bset = 0;
#pragma unroll
    for (int s = 0; s < NT; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) Mp[s][c] = Vp[s][c] = 0.f;
#pragma unroll
    for (int k = 0; k < kNF; ++k) gp[k] = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) den[c] = 1.f, num[c] = numV[c] = 0.f;
    if (active) {
        const size_t pix = (size_t)y * a.ld + x;
        const int n = a.n[pix];
        okp = n >= 2 ? 1.f : 0.f;
        nm1p = n >= 2 ? (float)(n - 1) : 1.f;
        rp = __builtin_amdgcn_rcpf(nm1p);
        if (HASW) feats_at(a.f, x, y, gp);
        if (MODE3) {
            float bj = a.jb[pix];
            for (int s = 1; s < S; ++s) {
                const float j = a.jb[(size_t)s * a.plane + pix];
                if (j < bj) { bj = j; bset = s; }
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float M, V;
                welch_load(a, x, y, s, c, M, V);
                a.discr[((size_t)s * 3 + c) * a.plane + pix] = V;
                if (!MODE3) Mp[s][c] = M, Vp[s][c] = V;
                else if (s == bset) Mp[0][c] = M, Vp[0][c] = V;
            }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            num[c] = *(const float *)((const char *)a.film + (int64_t)y * a.film_pitch + (int64_t)x * a.film_xs + c * a.film_cs);
            if (PLAIN) numV[c] = n >= 2 ? a.var[c * a.plane + pix] / (float)(n - 1) : 0.f;
        }
        if (a.dec) {
            const int win = 2 * a.r + 1;
            const size_t k0 = ((size_t)(y * a.W + x) * win * win + (size_t)a.r * win + a.r) * 3;
            a.dec[k0] = a.dec[k0 + 1] = a.dec[k0 + 2] = 1;
        }
    }
    const bool large_n = welch_window_level(a, x0, y0, TY) > 0;

    for (int t = 0; t < TY; ++t) welch_stage_row<S, HASW, PLAIN, HWC>(a, ring + t * SLOT, y0 - a.r + t, x0);
    __syncthreads();
    for (int k = 0; k <= 2 * a.r; ++k) {
        const float *slot = ring + ((ty + k) % TY) * SLOT;
        if (active) {
            if (large_n)
                welch_row<S, HASW, PLAIN, MODE3, false, HWC>(a, slot, tx, x, y, k - a.r, Mp, Vp, gp, okp, nm1p, rp,
                                                             bset, den, num, numV);
            else
                welch_row<S, HASW, PLAIN, MODE3, true, HWC>(a, slot, tx, x, y, k - a.r, Mp, Vp, gp, okp, nm1p, rp,
                                                            bset, den, num, numV);
        }
        if (k == 2 * a.r) break;
        __syncthreads();
        welch_stage_row<S, HASW, PLAIN, HWC>(a, ring + (k % TY) * SLOT, y0 - a.r + k + TY, x0);
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        *(float *)((char *)a.out + (int64_t)y * a.out_pitch + (int64_t)x * a.out_xs + c * a.out_cs) = num[c] / den[c];
        if (PLAIN) a.varW[c * a.plane + (size_t)y * a.ld + x] = numV[c] / den[c];
    }
}

// ====================================================== symmetric Welch filter
// The mode-2 decision of a pair (every set's Welch test, per channel) is
// symmetric in p and q, so each unordered pair is tested once.  Pass 1 tests
// every pixel against its forward half-window (dy > 0, or dy = 0 and dx > 0)
// and stores one bit per channel and offset (1: accepted, both pixels valid);
// pass 2 weights the full window and takes a backward neighbour's bit from
// that neighbour's record.  Each pair's decision is the one the one-pass form
// takes: the fast test outside the exact-tail band is correct either way
// round, and the FP64 tail is symmetric.

// Index of forward offset (dx, dy) among the (2r+1)^2/2 forward offsets.
__device__ __forceinline__ int fwd_index(int dx, int dy, int r) {
    return dy == 0 ? dx - 1 : r + (dy - 1) * (2 * r + 1) + dx + r;
}

// Pass 1 LDS slot: ok (-1 where n >= 2, so that its sign bit is the flag; +1
// otherwise), n - 1, then per column its 3S (M, V) pairs side by side, so a
// lane reads a neighbour's operands at immediate offsets from one address
// (a column is 24 S bytes: for S = 6, nine 16-byte units, an odd count, and
// the lanes' reads are bank-conflict free).
template <int S>
struct DecideLayout {
    static constexpr int FLOATS = 2 + 6 * S;
};

// The decision pass's (M, V) of a pixel: welch_load with V = var/n through
// the correctly rounded reciprocal rn of nf = max(n, 1), which gives the IEEE
// quotient bit for bit (div_exact_by_n).
__device__ __forceinline__ void decide_load(const WelchArgs &a, size_t pix, int xq, int yq, int s, int c, int n,
                                            float nf, float rn, float &M, float &V) {
    M = s == 0 ? *(const float *)((const char *)a.film + (int64_t)yq * a.film_pitch + (int64_t)xq * a.film_xs + c * a.film_cs)
               : a.mean[(size_t)s * a.set_stride + c * a.plane + pix];
    float var;
    if (a.var_is_m2) var = n >= 2 ? a.var[(size_t)s * a.set_stride + c * a.plane + pix] / (float)(n - 1) : 0.f;
    else var = a.var[((size_t)s * 3 + c) * a.plane + pix];
    V = n >= 1 ? div_exact_by_n(var, rn, nf) : 0.f;
    if (!(__builtin_isfinite(M) && __builtin_isfinite(V))) { M = NAN; V = 0.f; }
}

// Stages image row yq of the tile's halo: lane l takes columns l and l + 64,
// wavefront w the set-channels w, w + (waves), ... of them, so n, its
// reciprocal and the pixel address are formed once per column and the
// global loads of a set-channel plane stay coalesced.
template <int S, int HWC>
__device__ __forceinline__ void decide_stage_row(const WelchArgs &a, float *slot, int yq, int x0) {
    const int HW = TX + 2 * a.r;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    float2 *pairs = (float2 *)(slot + 2 * HWC);
    for (int col = lane; col < HW; col += 64) {
        const int xq = x0 - a.r + col;
        const bool in = xq >= 0 && xq < a.W && yq >= 0 && yq < a.H;
        const size_t pix = in ? (size_t)yq * a.ld + xq : 0;
        const int n = in ? a.n[pix] : 0;
        if (wave == 0) {
            slot[col] = n >= 2 ? -1.f : 1.f;
            slot[HWC + col] = n >= 2 ? (float)(n - 1) : 1.f;
        }
        const float nf = (float)max(n, 1);
        const float rn = recip_rn(nf);
        for (int sc = wave; sc < 3 * S; sc += waves) {
            const int s = sc / 3, c = sc - 3 * s;
            float M = 0.f, V = 0.f;
            if (in) decide_load(a, pix, xq, yq, s, c, n, nf, rn, M, V);
            pairs[col * (3 * S) + sc] = make_float2(M, V);
        }
    }
}

// Stores the word group that ends at forward index idx (bits shifted in
// from below, the group's first offset highest): reversed and shifted down
// so that bit i is the group's i-th offset, masked by both pixels' ok flags;
// to the weighting pass's record as one (c0, c1, c2, 0) group, or to the bit
// planes.
template <bool BANDED>
__device__ __forceinline__ void flush_words(const WelchArgs &a, int idx, bool okp, uint32_t (&word)[3],
                                            uint32_t &okw, size_t pix) {
    const uint32_t sh = 31u - (uint32_t)(idx & 31);
    const uint32_t ok = okp ? __builtin_bitreverse32(okw) >> sh : 0u;
    uint32_t w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c] = (__builtin_bitreverse32(word[c]) >> sh) & ok;
    if (!BANDED && a.rec_f4) {
        ((uint4 *)(a.rec + pix * a.rec_f4 + a.rec_head))[idx >> 5] = make_uint4(w[0], w[1], w[2], 0u);
    } else {
        const int nwords = BANDED ? a.bits_words : kBitsWords;
#pragma unroll
        for (int c = 0; c < 3; ++c) a.bits[((size_t)c * nwords + (idx >> 5)) * a.plane + pix] = w[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) word[c] = 0u;
    okw = 0u;
}

// Pass 1, one window row dy >= 0 of the lane's pixel: dx from dx0 to r.
// LEVEL is the tile's window level (welch_window_level): 0 tests both margins
// per set with the range term, 1 one gap per set with the fit on u <= 1/16,
// 2 with the short fit on u <= 1/128, 3 with the uniform-n form of the short
// fit (welch_uniform).  Each column leaves per channel a number R whose sign
// is the decision (negative: accepted): the reject margin for level 0, the
// largest gap minus the band otherwise, +-1 where the exact path decided.
// Its sign bit is shifted into word[c] from below (alignbit), and the
// neighbour's ok flag (the sign of its staged ok) into okw, so a word holds
// its offsets in reverse order until flush_words turns it round.
template <int S, int LEVEL, int HWC, bool BANDED>
__device__ __forceinline__ void decide_row(const WelchArgs &a, const float *slot, int tx, int dy, int dx0,
                                           const float (&Mp)[S][3], const float (&Vp)[S][3], bool okp, float nm1p,
                                           float rp, const WelchUniform &U, uint32_t (&word)[3], uint32_t &okw,
                                           size_t pix) {
    const int base = BANDED ? a.band_base : 0;
    const int count = BANDED ? a.band_count : 2 * a.r * a.r + 2 * a.r;
    float Vn[S][3];  // level 3: nbeta vp
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) Vn[s][c] = LEVEL == 3 ? U.nbeta * Vp[s][c] : 0.f;
    // The fit's coefficients and bands in VGPRs for this row (kernel arguments
    // sit in SGPRs, and an FMA that reads an SGPR issues at half the rate).
    TCritPoly tp = a.tp;
    if (LEVEL < 3) {
#pragma unroll
        for (int k = 0; k <= kTCritDeg; ++k) asm volatile("" : "+v"(tp.c[k]));
#pragma unroll
        for (int k = 0; k <= kTCritDegShort; ++k) asm volatile("" : "+v"(tp.cs[k]));
        asm volatile("" : "+v"(tp.umax), "+v"(tp.neg_hi), "+v"(tp.neg_lo), "+v"(tp.band_abs), "+v"(tp.band_abs_short));
    }
    // The ring as an LDS pointer: 32-bit address arithmetic per column, where
    // the generic pointer took two 64-bit multiply-adds.
    typedef float v2f __attribute__((ext_vector_type(2)));
    typedef float v4f __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) const float lds_float;
    typedef __attribute__((address_space(3))) const v2f lds_float2;
    typedef __attribute__((address_space(3))) const v4f lds_float4;
    // byte addresses in LDS of the first column's ok flag and operands, stepped per column
    const uint32_t col0 = (uint32_t)(tx + a.r + dx0);
    uint32_t aok = (uint32_t)(uintptr_t)(lds_float *)slot + col0 * 4u;
    uint32_t aop = aok + (2u * HWC + col0 * (6u * S - 1u)) * 4u;
    for (int dx = dx0; dx <= a.r; ++dx, aok += 4u, aop += 24u * S) {
        lds_float2 *pq = (lds_float2 *)(uintptr_t)aop;
        float2 mv[S][3];
        const float okq = *(lds_float *)(uintptr_t)aok;
        const float nm1q = *(lds_float *)(uintptr_t)(aok + 4u * HWC);
        if ((3 * S) % 2 == 0) {  // whole 16-byte units per column (slot, 2 HWC floats and the column 16-byte aligned)
            lds_float4 *p4 = (lds_float4 *)pq;
#pragma unroll
            for (int k = 0; k < 3 * S / 2; ++k) {
                const v4f t = p4[k];
                mv[(2 * k) / 3][(2 * k) % 3] = make_float2(t.x, t.y);
                mv[(2 * k + 1) / 3][(2 * k + 1) % 3] = make_float2(t.z, t.w);
            }
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const v2f t = pq[s * 3 + c];
                    mv[s][c] = make_float2(t.x, t.y);
                }
        }
        __builtin_amdgcn_sched_barrier(0);
        // level 3's gaps need no 1/(n_q - 1): there only a near-threshold test reads it
        PairR pr{};
        if (LEVEL != 3) pr = pair_r(rp, __builtin_amdgcn_rcpf(nm1q));
        float R[3];
        bool nc[3];
        if (LEVEL == 0) {
            float E[3], F[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) E[c] = F[c] = -INFINITY;
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float e, f;
                    welch_margin<true>(tp, Mp[s][c], Vp[s][c], mv[s][c].x, mv[s][c].y, pr, e, f);
                    E[c] = fmaxf(E[c], e);
                    F[c] = fmaxf(F[c], f);
                }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                nc[c] = !(E[c] > 0.f) && F[c] >= 0.f;
                R[c] = E[c];  // E <= F: outside the band E > 0 rejects, E <= F < 0 accepts
            }
        } else {  // one running maximum per channel, the band applied once
            float G[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) G[c] = -INFINITY;
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    G[c] = fmaxf(G[c], LEVEL == 3   ? welch_gap_uniform(U, Mp[s][c], Vp[s][c], Vn[s][c], mv[s][c].x,
                                                                        mv[s][c].y)
                                       : LEVEL == 2 ? welch_gap<kTCritDegShort>(tp.cs, Mp[s][c], Vp[s][c],
                                                                                mv[s][c].x, mv[s][c].y, pr)
                                                    : welch_gap<kTCritDeg>(tp.c, Mp[s][c], Vp[s][c], mv[s][c].x,
                                                                           mv[s][c].y, pr));
            const float band = LEVEL == 3 ? U.band : LEVEL == 2 ? tp.band_abs_short : tp.band_abs;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                nc[c] = fabsf(G[c]) <= band;
                R[c] = G[c] - band;  // outside the band: > band rejects, < -band accepts
            }
        }
        if ((nc[0] || nc[1] || nc[2]) && okq < 0.f && okp) {
            if (LEVEL == 3) {
                float nq = nm1q;
                asm volatile("" : "+v"(nq));  // keeps the reciprocal in this branch (the compiler hoisted it)
                pr = pair_r(rp, __builtin_amdgcn_rcpf(nq));
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (!nc[c]) continue;
                bool pass = true;
                for (int s = 0; s < S; ++s) {
                    const v2f m = pq[s * 3 + c];
                    float e, f;
                    welch_margin<LEVEL == 0>(a.tp, Mp[s][c], Vp[s][c], m.x, m.y, pr, e, f);
                    if (e > 0.f) { pass = false; break; }
                    if (f >= 0.f && !welch_exact(Mp[s][c], Vp[s][c], nm1p, m.x, m.y, nm1q, a.alpha)) {
                        pass = false;
                        break;
                    }
                }
                R[c] = pass ? -1.f : 1.f;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) word[c] = __builtin_amdgcn_alignbit(word[c], __float_as_uint(R[c]), 31);
        okw = __builtin_amdgcn_alignbit(okw, __float_as_uint(okq), 31);
        const int idx = fwd_index(dx, dy, a.r) - base;
        if ((idx & 31) == 31 || idx == count - 1) flush_words<BANDED>(a, idx, okp, word, okw, pix);
    }
}

// The head of a pixel's weighting-pass record (head4 items): film (3) and,
// with weights, the 7 features: (film0, film1, film2, f0) (f1 .. f4) (f5, f6,
// 0, 0), or (film0, film1, film2, 0) without.
__device__ __forceinline__ void welch_record_head(const WelchArgs &a, int x, int y, float4 *rec, int head4) {
    float film[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        film[c] = *(const float *)((const char *)a.film + (int64_t)y * a.film_pitch + (int64_t)x * a.film_xs + c * a.film_cs);
    if (head4 == 1) {
        rec[0] = make_float4(film[0], film[1], film[2], 0.f);
        return;
    }
    float f[kNF];
    feats_at(a.f, x, y, f);
    rec[0] = make_float4(film[0], film[1], film[2], f[0]);
    rec[1] = make_float4(f[1], f[2], f[3], f[4]);
    rec[2] = make_float4(f[5], f[6], 0.f, 0.f);
}

template <int S, int HWC, int TY, bool BANDED>
__device__ __forceinline__ void decide_tile(const WelchArgs a, int bx, int by, float *ring) {
    constexpr int SLOT = DecideLayout<S>::FLOATS * HWC;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = bx * TX, y0 = by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < a.W && y < a.H;
    const size_t pix = (size_t)y * a.ld + x;

    float Mp[S][3], Vp[S][3];
    bool okp = false;
    float nm1p = 1.f, rp = 0.f;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) Mp[s][c] = Vp[s][c] = 0.f;
    if (active) {
        const int n = a.n[pix];
        okp = n >= 2;
        nm1p = n >= 2 ? (float)(n - 1) : 1.f;
        rp = __builtin_amdgcn_rcpf(nm1p);
        const float nf = (float)max(n, 1);
        const float rn = recip_rn(nf);
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                decide_load(a, pix, x, y, s, c, n, nf, rn, Mp[s][c], Vp[s][c]);
                if (!BANDED || a.band_dy0 == 0) a.discr[((size_t)s * 3 + c) * a.plane + pix] = Vp[s][c];
            }
        if (!BANDED && a.rec_f4) welch_record_head(a, x, y, a.rec + pix * a.rec_f4, a.rec_head);
        if (BANDED && a.head) welch_record_head(a, x, y, a.head + pix * a.head4, a.head4);
    }
    int nuni = 0;
    const int level = welch_window_level<true>(a, x0, y0, TY, &nuni);
    WelchUniform U{0.f, 0.f, 0.f};
    if (level == 3) U = welch_uniform(a.tp, nuni);

    // window rows dy = d0 .. d1; staged row j is image row y0 + d0 + j
    const int d0 = BANDED ? a.band_dy0 : 0, d1 = BANDED ? a.band_dy1 : a.r;
    for (int t = 0; t < TY; ++t) decide_stage_row<S, HWC>(a, ring + t * SLOT, y0 + d0 + t, x0);
    __syncthreads();
    uint32_t word[3] = {0u, 0u, 0u}, okw = 0u;
    for (int k = d0; k <= d1; ++k) {  // dy = k
        const float *slot = ring + ((ty + k - d0) % TY) * SLOT;
        const int dx0 = k == 0 ? 1 : -a.r;
        if (active) {
            if (level == 3)
                decide_row<S, 3, HWC, BANDED>(a, slot, tx, k, dx0, Mp, Vp, okp, nm1p, rp, U, word, okw, pix);
            else if (level == 2)
                decide_row<S, 2, HWC, BANDED>(a, slot, tx, k, dx0, Mp, Vp, okp, nm1p, rp, U, word, okw, pix);
            else if (level == 1)
                decide_row<S, 1, HWC, BANDED>(a, slot, tx, k, dx0, Mp, Vp, okp, nm1p, rp, U, word, okw, pix);
            else
                decide_row<S, 0, HWC, BANDED>(a, slot, tx, k, dx0, Mp, Vp, okp, nm1p, rp, U, word, okw, pix);
        }
        if (k == d1) break;
        __syncthreads();
        decide_stage_row<S, HWC>(a, ring + ((k - d0) % TY) * SLOT, y0 + k + TY, x0);
        __syncthreads();
    }
}

template <int S, int HWC, int TY, bool BANDED = false>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(
    1, lds_waves_per_simd(TY * DecideLayout<S>::FLOATS * HWC * 4, 64 * TY)))) k_welch_decide(WelchArgs a) {
    extern __shared__ float4 ring4[];  // float4: the dynamic ring starts 16-byte aligned after the static LDS
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    decide_tile<S, HWC, TY, BANDED>(a, t % tiles_x, t / tiles_x, (float *)ring4);
}

// Pass 2 LDS slot: features, film[3], then the 3 x NW decision words of the
// window's radius (NW = bits_words(r) for the largest r the tile width holds).
template <bool HASW, int HWC>
struct AccumLayout {
    static constexpr int FILM = HASW ? kNF : 0;
    static constexpr int WORDS = FILM + 3;
    static constexpr int NW = bits_words((HWC - TX) / 2);
    static constexpr int FLOATS = WORDS + 3 * NW;
};

template <bool HASW, int HWC>
__device__ __forceinline__ void accum_stage_row(const WelchArgs &a, float *slot, int yq, int x0) {
    using LY = AccumLayout<HASW, HWC>;
    const int HW = TX + 2 * a.r;
    const int items = HW * (1 + 3 * LY::NW);
    uint32_t *words = (uint32_t *)(slot + LY::WORDS * HWC);
    for (int i = threadIdx.x; i < items; i += blockDim.x) {
        const int kind = i / HW, col = i - kind * HW;
        const int xq = x0 - a.r + col;
        const bool in = xq >= 0 && xq < a.W && yq >= 0 && yq < a.H;
        if (kind == 0) {
            if (HASW) {
                float fv[kNF] = {};
                if (in) feats_at(a.f, xq, yq, fv);
#pragma unroll
                for (int k = 0; k < kNF; ++k) slot[k * HWC + col] = fv[k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
                slot[(LY::FILM + c) * HWC + col] =
                    in ? *(const float *)((const char *)a.film + (int64_t)yq * a.film_pitch + (int64_t)xq * a.film_xs +
                                          c * a.film_cs)
                       : 0.f;
        } else {
            const int w = kind - 1, c = w / LY::NW, k = w - c * LY::NW;  // c * NW + k
            words[w * HWC + col] = in ? a.bits[((size_t)c * a.bits_words + k) * a.plane + (size_t)yq * a.ld + xq] : 0u;
        }
    }
}

// Pass 2, one window row: weights and the decisions of pass 1.
template <bool HASW, int HWC>
__device__ __forceinline__ void accum_row(const WelchArgs &a, const float *slot, int tx, int x, int y, int dy,
                                          const float (&gp)[kNF], float (&den)[3], float (&num)[3], size_t pix) {
    using LY = AccumLayout<HASW, HWC>;
    const int win = 2 * a.r + 1;
    const float spy = a.sp_scale * (float)(dy * dy);
    const uint32_t *words = (const uint32_t *)(slot + LY::WORDS * HWC);
    // this pixel's own words for the forward offsets of the row: its 2r + 1
    // consecutive indices span at most three words (two while 2r + 1 <= 33)
    uint32_t own0[3] = {0u, 0u, 0u}, own1[3] = {0u, 0u, 0u}, own2[3] = {0u, 0u, 0u};
    int k0 = 0;
    if (dy >= 0) {
        const int i0 = fwd_index(dy == 0 ? 1 : -a.r, dy, a.r), i1 = fwd_index(a.r, dy, a.r);
        k0 = i0 >> 5;
        const int k1 = i1 >> 5;
        if (a.r > 0)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                own0[c] = a.bits[((size_t)c * a.bits_words + k0) * a.plane + pix];
                own1[c] = k1 > k0 ? a.bits[((size_t)c * a.bits_words + k0 + 1) * a.plane + pix] : 0u;
                own2[c] = k1 > k0 + 1 ? a.bits[((size_t)c * a.bits_words + k0 + 2) * a.plane + pix] : 0u;
            }
    }
    for (int dx = -a.r; dx <= a.r; ++dx) {
        if (dx == 0 && dy == 0) continue;
        const float *q = slot + tx + a.r + dx;
        float fq[kNF], filmq[3];
        uint32_t wq[3];
        const bool fwd = dy > 0 || (dy == 0 && dx > 0);
        const int idx = fwd ? fwd_index(dx, dy, a.r) : fwd_index(-dx, -dy, a.r);
#pragma unroll
        for (int f = 0; f < (HASW ? kNF : 0); ++f) fq[f] = q[f * HWC];
#pragma unroll
        for (int c = 0; c < 3; ++c) filmq[c] = q[(LY::FILM + c) * HWC];
        if (!fwd)
#pragma unroll
            for (int c = 0; c < 3; ++c) wq[c] = words[(c * LY::NW + (idx >> 5)) * HWC + tx + a.r + dx];
        __builtin_amdgcn_sched_barrier(0);
        float w = 1.f;
        if (HASW) {
            float lw = __builtin_fmaf(a.sp_scale, (float)(dx * dx), spy);
#pragma unroll
            for (int f = 0; f < kNF; ++f) {
                const float d = gp[f] - fq[f];
                lw = __builtin_fmaf(-d, d, lw);
            }
            w = __builtin_amdgcn_exp2f(lw);
        }
        bool acc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int kw = (idx >> 5) - k0;
            const uint32_t wd = fwd ? (kw == 0 ? own0[c] : kw == 1 ? own1[c] : own2[c]) : wq[c];
            acc[c] = (wd >> (idx & 31)) & 1u;
            const float wa = acc[c] ? w : 0.f;
            den[c] += wa;
            num[c] = __builtin_fmaf(wa, filmq[c], num[c]);
        }
        if (a.dec) {
            const size_t kk = ((size_t)(y * a.W + x) * win * win + (size_t)(dy + a.r) * win + (dx + a.r)) * 3;
            for (int c = 0; c < 3; ++c) a.dec[kk + c] = acc[c] ? 1 : 0;
        }
    }
}

template <bool HASW, int HWC, int TY>
__device__ __forceinline__ void accum_tile(const WelchArgs a, int bx, int by, float *ring) {
    constexpr int SLOT = AccumLayout<HASW, HWC>::FLOATS * HWC;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = bx * TX, y0 = by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < a.W && y < a.H;
    const size_t pix = (size_t)y * a.ld + x;
    float gp[kNF], den[3], num[3];
#pragma unroll
    for (int k = 0; k < kNF; ++k) gp[k] = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) den[c] = 1.f, num[c] = 0.f;  // the pixel itself, weight 1
    if (active) {
        if (HASW) feats_at(a.f, x, y, gp);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            num[c] = *(const float *)((const char *)a.film + (int64_t)y * a.film_pitch + (int64_t)x * a.film_xs + c * a.film_cs);
        if (a.dec) {
            const int win = 2 * a.r + 1;
            const size_t k0 = ((size_t)(y * a.W + x) * win * win + (size_t)a.r * win + a.r) * 3;
            a.dec[k0] = a.dec[k0 + 1] = a.dec[k0 + 2] = 1;
        }
    }
    for (int t = 0; t < TY; ++t) accum_stage_row<HASW, HWC>(a, ring + t * SLOT, y0 - a.r + t, x0);
    __syncthreads();
    for (int k = 0; k <= 2 * a.r; ++k) {
        const float *slot = ring + ((ty + k) % TY) * SLOT;
        if (active) accum_row<HASW, HWC>(a, slot, tx, x, y, k - a.r, gp, den, num, pix);
        if (k == 2 * a.r) break;
        __syncthreads();
        accum_stage_row<HASW, HWC>(a, ring + (k % TY) * SLOT, y0 - a.r + k + TY, x0);
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        *(float *)((char *)a.out + (int64_t)y * a.out_pitch + (int64_t)x * a.out_xs + c * a.out_cs) = num[c] / den[c];
}

template <bool HASW, int HWC, int TY>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(
    1, lds_waves_per_simd(TY * AccumLayout<HASW, HWC>::FLOATS * HWC * 4, 64 * TY)))) k_welch_accum(WelchArgs a) {
    extern __shared__ float ring[];
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    accum_tile<HASW, HWC, TY>(a, t % tiles_x, t / tiles_x, ring);
}

// ------------------------------------------------------------------------
// Pass 2 from records (the path without a decision dump): film (3), features
// (7, weighted filter), then pass 1's decision words grouped by word index,
// (c0, c1, c2, 0) per group, so that one 16-byte read fetches a neighbour's
// three channel words.  Pass 1 writes the records (the head when it loads the
// pixel's own operands, each word group when it is complete); pass 2 stages
// rows of them with a loader wavefront as in the variance filter, one pixel
// per lane and eight filtering wavefronts (k_welch_accum1).  A forward
// offset's decision comes from the pixel's own words (in registers), a
// backward offset's from the neighbour's record, at a word index that is the
// same for the whole wavefront.  The column code is written for two pixels
// per lane (A, B) in modes of their own; the one-pixel pass runs B off.
template <bool HASW>
struct AccRec {
    static constexpr int HEAD4 = HASW ? 3 : 1;                  // float4 items of film (+ features)
    static constexpr int RAW4 = HEAD4 + kRecWords;              // + one item per word group
    static constexpr int FLOAT4S = RAW4 % 2 ? RAW4 : RAW4 + 1;  // 11 or 9: odd
    static_assert(FLOAT4S * 4 <= kVarRecFloats, "record larger than its planes");
};
template <bool HASW, int HWC>
constexpr int acc_loads_per_row() { return (HWC * AccRec<HASW>::FLOAT4S + 63) / 64; }

template <bool HASW>
struct AccPix {
    float gp[kNF], den[3], num[3];
    uint32_t own[3][kRecWords];
};

// The pixel's own film, features and decision words come from its record in
// one batch of loads; an inactive lane reads the zero record.
template <bool HASW>
__device__ __forceinline__ void acc_pix_init(const WelchArgs &a, int x, int y, bool active, AccPix<HASW> &P) {
    using R = AccRec<HASW>;
    const float4 *q = active ? a.rec + ((size_t)y * a.ld + x) * R::FLOAT4S : a.zero;
    float4 h[R::HEAD4];
    uint4 wd[kRecWords];
#pragma unroll
    for (int i = 0; i < R::HEAD4; ++i) h[i] = q[i];
#pragma unroll
    for (int w = 0; w < kRecWords; ++w) wd[w] = ((const uint4 *)q)[R::HEAD4 + w];
    const float gp[kNF] = {h[0].w, h[HASW ? 1 : 0].x, h[HASW ? 1 : 0].y, h[HASW ? 1 : 0].z,
                           h[HASW ? 1 : 0].w, h[HASW ? 2 : 0].x, h[HASW ? 2 : 0].y};
#pragma unroll
    for (int k = 0; k < kNF; ++k) P.gp[k] = HASW ? gp[k] : 0.f;
    const float film[3] = {h[0].x, h[0].y, h[0].z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        P.den[c] = 1.f;  // the pixel itself, weight 1
        P.num[c] = film[c];
#pragma unroll
        for (int w = 0; w < kRecWords; ++w) P.own[c][w] = c == 0 ? wd[w].x : c == 1 ? wd[w].y : wd[w].z;
    }
}

// How a pixel finds the decisions of one window row: all forward offsets
// (dy > 0: its own words, turned into a per-row mask), all backward offsets
// (dy < 0: the neighbour's word, at an index that falls by one per column),
// the centre row (dy = 0, both kinds, per column), or outside its window.
enum AccMode { kAccFwd, kAccBwd, kAccGen, kAccOff };

template <bool HASW>
struct AccCol {
    float4 f0, f1, f2;  // film 0..2, features 0..6 (features only with HASW)
    uint4 wa, wb;       // the neighbour's word group for A and for B
    int sha, shb;       // bit of the neighbour's word (backward offsets)
};

// Row-constant part of a pixel's step.  The own words are consumed front
// first: rows run in order of increasing forward index, so the pixel's words
// are shifted down (base counts the shifts) until the row's first word is at
// index 0; o0, o1 are then the row's two own words (the row's 2r + 1
// consecutive forward indices span at most two).
struct AccRowP {
    int dy, k0, i0;  // i0: forward index of column dx = -r (own for dy > 0, the neighbour's for dy < 0)
    float lw;
    uint32_t o0[3], o1[3], m[3];  // m: dy > 0, bit dx + r = decision of column dx
};

template <bool HASW>
__device__ __forceinline__ AccRowP acc_row_p(const WelchArgs &a, AccPix<HASW> &P, int &base, int dy, bool off) {
    const int r = a.r;
    AccRowP R;
    R.dy = dy;
    R.lw = a.sp_scale * (float)(dy * dy);
    R.i0 = dy > 0 ? fwd_index(-r, dy, r) : (dy < 0 ? fwd_index(r, -dy, r) : 0);
    R.k0 = dy > 0 && !off ? R.i0 >> 5 : (dy == 0 ? 0 : base);
    for (; base < R.k0; ++base)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
            for (int j = 0; j + 1 < kRecWords; ++j) P.own[ch][j] = P.own[ch][j + 1];
            P.own[ch][kRecWords - 1] = 0u;
        }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        R.o0[ch] = P.own[ch][0];
        R.o1[ch] = P.own[ch][1];
        R.m[ch] = __builtin_amdgcn_alignbit(R.o1[ch], R.o0[ch], (uint32_t)(R.i0 & 31));
    }
    return R;
}

// Word group and bit of the neighbour's decision for pixel mode M at column dx.
template <int M>
__device__ __forceinline__ void acc_qsel(const AccRowP &R, int dx, int r, int &wq, int &sh) {
    if (M == kAccBwd) {
        const int idx = R.i0 - (dx + r);
        wq = idx >> 5;  // idx >= i0 - 2r >= r
        sh = idx & 31;
    } else if (M == kAccGen) {
        const int idx = dx < 0 ? -dx - 1 : 0;  // centre row: backward offsets dx < 0
        wq = idx >> 5;
        sh = idx & 31;
    } else {
        wq = 0;
        sh = 0;
    }
}

template <bool HASW, int MA, int MB>
__device__ __forceinline__ void acc_col_load(const float4 *q, const AccRowP &RA, const AccRowP &RB, int dx, int r,
                                             AccCol<HASW> &c) {
    c.f0 = q[0];
    if (HASW) {
        c.f1 = q[1];
        c.f2 = q[2];
    }
    const uint4 *g = (const uint4 *)(q + AccRec<HASW>::HEAD4);
    int wA, wB;
    acc_qsel<MA>(RA, dx, r, wA, c.sha);
    acc_qsel<MB>(RB, dx, r, wB, c.shb);
    if (MA == kAccBwd || MA == kAccGen) c.wa = g[wA];
    if (MB == kAccBwd || MB == kAccGen) c.wb = g[wB];
}

// The weight and the three channel decisions of column dx (j = dx + r) for
// one pixel in mode M; the pixel itself weighs 0.
template <bool HASW, int M>
__device__ __forceinline__ void acc_pix_col(const WelchArgs &a, const AccCol<HASW> &c, const float (&fq)[kNF],
                                            const float (&film)[3], const uint4 &qw, int sh, int dx, const AccRowP &R,
                                            AccPix<HASW> &P) {
    if (M == kAccOff) return;
    const bool z = M == kAccGen && dx == 0;
    const float w = var_weight<HASW>(a.sp_scale, P.gp, 1.f, fq, z ? -INFINITY : R.lw, z ? 0.f : 1.f, (float)(dx * dx));
    const uint32_t q3[3] = {qw.x, qw.y, qw.z};
    const int j = dx + a.r;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        // the decision bit sign-extended to a mask (all ones: accepted), and
        // the weight masked with it: two instructions where a compare and a
        // select take three
        int mask;
        if (M == kAccFwd) {
            mask = __builtin_amdgcn_sbfe((int)R.m[ch], j, 1);
        } else if (M == kAccBwd) {
            mask = __builtin_amdgcn_sbfe((int)q3[ch], sh, 1);
        } else {  // centre row: dx > 0 own index dx - 1 (word 0), dx < 0 the neighbour's
            mask = dx > 0 ? __builtin_amdgcn_sbfe((int)R.o0[ch], dx - 1, 1) : __builtin_amdgcn_sbfe((int)q3[ch], sh, 1);
        }
        const float wa = __uint_as_float(__float_as_uint(w) & (uint32_t)mask);
        P.den[ch] += wa;
        P.num[ch] = __builtin_fmaf(wa, film[ch], P.num[ch]);
    }
}

template <bool HASW, int MA, int MB>
__device__ __forceinline__ void acc_col_test(const WelchArgs &a, const AccCol<HASW> &c, int dx, const AccRowP &RA,
                                             const AccRowP &RB, AccPix<HASW> &A, AccPix<HASW> &B) {
    const float fq[kNF] = {c.f0.w, c.f1.x, c.f1.y, c.f1.z, c.f1.w, c.f2.x, c.f2.y};
    const float film[3] = {c.f0.x, c.f0.y, c.f0.z};
    acc_pix_col<HASW, MA>(a, c, fq, film, c.wa, c.sha, dx, RA, A);
    acc_pix_col<HASW, MB>(a, c, fq, film, c.wb, c.shb, dx, RB, B);
}

// One window row for A and B in modes MA, MB: the reads of column dx + 1
// issued before the weights of column dx; pairs of columns from dx = -r, then
// the last column alone.
template <bool HASW, int MA, int MB>
__device__ __forceinline__ void acc_cols(const WelchArgs &a, const float4 *q, const AccRowP &RA, const AccRowP &RB,
                                         AccPix<HASW> &A, AccPix<HASW> &B) {
    constexpr int F4 = AccRec<HASW>::FLOAT4S;
    // LDS reads per column: film and features, and a word group per pixel that reads its neighbour's
    constexpr int NL = AccRec<HASW>::HEAD4 + (MA == kAccBwd || MA == kAccGen) + (MB == kAccBwd || MB == kAccGen);
    const int r = a.r;
    AccCol<HASW> X, Y;
    acc_col_load<HASW, MA, MB>(q, RA, RB, -r, r, X);
    for (int dx = -r; dx < r; dx += 2, q += 2 * F4) {
        acc_col_load<HASW, MA, MB>(q + F4, RA, RB, dx + 1, r, Y);
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(NL));  // X, issued one column earlier, has landed
        __builtin_amdgcn_sched_barrier(0);
        acc_col_test<HASW, MA, MB>(a, X, dx, RA, RB, A, B);
        acc_col_load<HASW, MA, MB>(q + 2 * F4, RA, RB, dx + 2, r, X);
        __builtin_amdgcn_s_waitcnt(waitcnt_lgkm(NL));
        __builtin_amdgcn_sched_barrier(0);
        acc_col_test<HASW, MA, MB>(a, Y, dx + 1, RA, RB, A, B);
    }
    acc_col_test<HASW, MA, MB>(a, X, r, RA, RB, A, B);
}

// One pixel per lane: TY filtering wavefronts of one row each and the loader,
// as k_var_filter1; the ring of TY + 2 rows of records fits one CU's LDS with
// eight filtering wavefronts, two per SIMD.  Wavefront ty reads row ty + k at
// step k (dy = k - r): backward rows, the centre row, forward rows.
template <bool HASW, int HWC, int TY>
__global__ void __launch_bounds__(64 * (TY + 1)) __attribute__((amdgpu_waves_per_eu(1, 2))) k_welch_accum1(WelchArgs a) {
    extern __shared__ float4 aring[];
    constexpr int F4 = AccRec<HASW>::FLOAT4S, IPR = acc_loads_per_row<HASW, HWC>();
    constexpr int P = TY, NS = TY + kVarAhead, SLOT = IPR * 64;
    constexpr int AHEAD_LOADS = IPR * (kVarAhead - 1);
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    const int x0 = (t % tiles_x) * TX, y0 = (t / tiles_x) * TY;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int last = 2 * a.r;

    if (wave == TY) {  // loader
        for (int j = 0; j < P + kVarAhead - 1; ++j)
            ring_load_row<F4, IPR>(a.rec, a.zero, a.W, a.H, a.ld, a.r, aring + j * SLOT, y0 - a.r + j, x0, lane);
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));
        __builtin_amdgcn_s_barrier();
        for (int k = 0; k <= last; ++k) {
            const int j = k + P + kVarAhead - 1;
            ring_load_row<F4, IPR>(a.rec, a.zero, a.W, a.H, a.ld, a.r, aring + (j % NS) * SLOT, y0 - a.r + j, x0,
                                   lane);
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(AHEAD_LOADS));
            __builtin_amdgcn_s_barrier();
        }
        __builtin_amdgcn_s_waitcnt(0);
        return;
    }

    const int x = x0 + lane, y = y0 + wave;
    const bool act = x < a.W && y < a.H;
    AccPix<HASW> A;
    acc_pix_init<HASW>(a, x, y, act, A);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    int base = 0;
    __syncthreads();
    for (int k = 0; k <= last; ++k) {
        const float4 *q = aring + ((wave + k) % NS) * SLOT + lane * F4;  // column of dx = -r
        if (act) {
            const int dy = k - a.r;
            const AccRowP R = acc_row_p<HASW>(a, A, base, dy, false);
            if (dy < 0) acc_cols<HASW, kAccBwd, kAccOff>(a, q, R, R, A, A);
            else if (dy == 0) acc_cols<HASW, kAccGen, kAccOff>(a, q, R, R, A, A);
            else acc_cols<HASW, kAccFwd, kAccOff>(a, q, R, R, A, A);
        }
        __syncthreads();
    }
    if (!act) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        *(float *)((char *)a.out + (int64_t)y * a.out_pitch + (int64_t)x * a.out_xs + c * a.out_cs) = A.num[c] / A.den[c];
}

template <bool HASW>
static int accum1_launch(const WelchArgs &a, hipStream_t st) {
    constexpr int HWC = TX + 2 * kRecBitsRadius, TY = 8;
    constexpr size_t lds = (size_t)(TY + kVarAhead) * acc_loads_per_row<HASW, HWC>() * 64 * 16;
    static_assert(lds <= 160 * 1024, "ring larger than a CU's LDS");
    STATER_HIP(hipFuncSetAttribute((const void *)k_welch_accum1<HASW, HWC, TY>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY)));
    hipLaunchKernelGGL((k_welch_accum1<HASW, HWC, TY>), grid, dim3(64 * (TY + 1)), lds, st, a);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}


// ------------------------------------------------------------------------
// Banded once-per-pair form, 10 < r <= 20.  The forward half-window is split
// into two bands of rows, dy = 0 .. 10 and 11 .. r, whose decision words (at
// most 14 per channel: 21 r + 10 <= 430 and 10 (2r + 1) <= 410 offsets) fit
// the set's variance-filter record planes.  Each band runs the decision pass
// over its rows and then a weighting pass over the pixel's pairs at dy in
// [-d1, -d0] and [d0, d1].  The first band's sums (den, num per channel) carry
// over to the second in planes of the set's decision-word area, which also
// holds the weighting records (film, features) the first band's decision
// pass writes.
//
// The weighting pass stages, per row step, only the word groups the step
// reads: the 2r + 1 offsets of a window row span at most three words.
// Wavefront ty (image row y0 + ty) at step dy reads the words of image row
// y0 + ty + min(dy, 0), a backward neighbour's (dy < 0, or dy = 0 and dx < 0)
// at the neighbour's column, its own (forward) at its own column, as
// (c0, c1, c2, 0) groups; and the neighbour row's records from a ring of
// TY + 1 rows.  The next step's words and ring row are loaded into registers
// before the current step's pairs and written to LDS after them, so one
// barrier per step separates the two.
constexpr int kBandRows0 = 10;     // the first band holds dy = 0 .. 10
constexpr int kBandWords = 14;     // words per channel of a band (r <= 20)
constexpr int kBandWordSlots = 3;  // word groups of one window row
static_assert(kBandRows0 == kRecBitsRadius, "radii above the record path's take two bands");
static_assert(21 * 20 + 10 <= 32 * kBandWords && 10 * 41 <= 32 * kBandWords, "band words for r <= 20");
static_assert(3 * kBandWords <= kVarRecFloats, "band words must fit the variance-filter record planes");
static_assert(6 + 4 * 3 <= 3 * kBitsWords, "band sums and weighting records must fit the decision-word planes");

template <bool HASW>
constexpr int band_head4() { return HASW ? 3 : 1; }

template <bool HASW, int HWC, int TY>
struct BandStage {
    static constexpr int F4 = band_head4<HASW>();
    static constexpr int HITEMS = HWC * F4;                   // float4 items of a ring row
    static constexpr int WITEMS = TY * kBandWordSlots * HWC;  // word groups of a step
    static constexpr int HPER = (HITEMS + 64 * TY - 1) / (64 * TY);
    static constexpr int WPER = (WITEMS + 64 * TY - 1) / (64 * TY);
    static constexpr int LDS_BYTES = ((TY + 1) * HITEMS + 2 * WITEMS) * 16;
    float4 h[HPER];
    uint4 w[WPER];
};

// First word of the band that a step at window row dy reads.
__device__ __forceinline__ int band_w0(const WelchArgs &a, int dy) {
    const int ady = dy < 0 ? -dy : dy;
    return ((ady == 0 ? 0 : fwd_index(-a.r, ady, a.r)) - a.band_base) >> 5;
}

template <bool HASW, int HWC, int TY>
__device__ __forceinline__ void band_load_head(const WelchArgs &a, BandStage<HASW, HWC, TY> &S, int yq, int x0) {
    using B = BandStage<HASW, HWC, TY>;
    const bool row_in = yq >= 0 && yq < a.H;
#pragma unroll
    for (int i = 0; i < B::HPER; ++i) {
        const int it = threadIdx.x + i * 64 * TY;
        const int col = it / B::F4, item = it - col * B::F4;
        const int xq = x0 - a.r + col;
        const bool in = row_in && it < B::HITEMS && col < TX + 2 * a.r && xq >= 0 && xq < a.W;
        S.h[i] = in ? a.head[((size_t)yq * a.ld + xq) * B::F4 + item] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

template <bool HASW, int HWC, int TY>
__device__ __forceinline__ void band_store_head(const BandStage<HASW, HWC, TY> &S, float4 *slot) {
    using B = BandStage<HASW, HWC, TY>;
#pragma unroll
    for (int i = 0; i < B::HPER; ++i) {
        const int it = threadIdx.x + i * 64 * TY;
        if (it < B::HITEMS) slot[it] = S.h[i];
    }
}

template <bool HASW, int HWC, int TY>
__device__ __forceinline__ void band_load_words(const WelchArgs &a, BandStage<HASW, HWC, TY> &S, int dy, int x0,
                                                int y0) {
    using B = BandStage<HASW, HWC, TY>;
    const int w0 = band_w0(a, dy);
#pragma unroll
    for (int i = 0; i < B::WPER; ++i) {
        const int it = threadIdx.x + i * 64 * TY;
        const int t = it / (kBandWordSlots * HWC), rem = it - t * (kBandWordSlots * HWC);
        const int j = rem / HWC, col = rem - j * HWC;
        const int ys = y0 + t + (dy < 0 ? dy : 0), xq = x0 - a.r + col, wi = w0 + j;
        const bool in = it < B::WITEMS && col < TX + 2 * a.r && ys >= 0 && ys < a.H && xq >= 0 && xq < a.W &&
                        wi < a.bits_words;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (in) {
            const uint32_t *b = a.bits + (size_t)wi * a.plane + (size_t)ys * a.ld + xq;
            const size_t cs = (size_t)a.bits_words * a.plane;
            v = make_uint4(b[0], b[cs], b[2 * cs], 0u);
        }
        S.w[i] = v;
    }
}

template <bool HASW, int HWC, int TY>
__device__ __forceinline__ void band_store_words(const BandStage<HASW, HWC, TY> &S, uint4 *buf) {
    using B = BandStage<HASW, HWC, TY>;
#pragma unroll
    for (int i = 0; i < B::WPER; ++i) {
        const int it = threadIdx.x + i * 64 * TY;
        if (it < B::WITEMS) buf[it] = S.w[i];
    }
}

// One column dx of a window row: the weight and the three channel bits
// (bit 0 of q3[c] >> sh); DEC writes the pair's decisions.
template <bool HASW, bool DEC>
__device__ __forceinline__ void band_col(const WelchArgs &a, const float4 *hq, uint32_t q0, uint32_t q1, uint32_t q2,
                                         int sh, int dx, int dy, int x, int y, float spy, const float (&gp)[kNF],
                                         float (&den)[3], float (&num)[3]) {
    const float4 h0 = hq[0];
    const float4 h1 = HASW ? hq[1] : h0, h2 = HASW ? hq[2] : h0;
    __builtin_amdgcn_sched_barrier(0);
    float w = 1.f;
    if (HASW) {
        const float fq[kNF] = {h0.w, h1.x, h1.y, h1.z, h1.w, h2.x, h2.y};
        float lw = __builtin_fmaf(a.sp_scale, (float)(dx * dx), spy);
#pragma unroll
        for (int f = 0; f < kNF; ++f) {
            const float d = gp[f] - fq[f];
            lw = __builtin_fmaf(-d, d, lw);
        }
        w = __builtin_amdgcn_exp2f(lw);
    }
    const uint32_t q3[3] = {q0, q1, q2};
    const float film[3] = {h0.x, h0.y, h0.z};
    bool acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[c] = (q3[c] >> sh) & 1u;
        const float wa = acc[c] ? w : 0.f;
        den[c] += wa;
        num[c] = __builtin_fmaf(wa, film[c], num[c]);
    }
    if (DEC) {
        const int win = 2 * a.r + 1;
        const size_t kk = ((size_t)(y * a.W + x) * win * win + (size_t)(dy + a.r) * win + (dx + a.r)) * 3;
        for (int c = 0; c < 3; ++c) a.dec[kk + c] = acc[c] ? 1 : 0;
    }
}

// One window row dy of the lane's pixel, specialised by kind: all columns
// backward (dy < 0: the neighbour's word group, at a bit index that falls by
// one per column), the centre row (dy = 0) and all columns forward (dy > 0:
// the pixel's own words, turned into two 32-bit masks per channel for the
// row's 2r + 1 consecutive bits).  hrow holds the neighbour row's records
// (column dx = -r at tx), wrow the wavefront's word groups of the step
// (group j of column col at j * HWC + col).
template <bool HASW, int HWC, bool DEC>
__device__ __forceinline__ void band_row(const WelchArgs &a, const float4 *hrow, const uint4 *wrow, int tx, int x,
                                         int y, int dy, const float (&gp)[kNF], float (&den)[3], float (&num)[3]) {
    constexpr int F4 = band_head4<HASW>();
    const int r = a.r;
    const int w0 = band_w0(a, dy);
    const float spy = a.sp_scale * (float)(dy * dy);
    const float4 *hq = hrow + tx * F4;  // column dx = -r
    if (dy < 0) {
        const int i0 = fwd_index(r, -dy, r) - a.band_base;  // the neighbour's bit of column dx = -r
        const uint4 *wq = wrow + tx;
        for (int dx = -r; dx <= r; ++dx, hq += F4, ++wq) {
            const int idx = i0 - (dx + r);
            const uint4 wd = wq[((idx >> 5) - w0) * HWC];
            band_col<HASW, DEC>(a, hq, wd.x, wd.y, wd.z, idx & 31, dx, dy, x, y, spy, gp, den, num);
        }
    } else if (dy == 0) {
        const uint4 own = wrow[tx + r];
        const uint4 *wq = wrow + tx;
        for (int dx = -r; dx <= r; ++dx, hq += F4, ++wq) {
            if (dx == 0) continue;
            const uint4 wd = dx > 0 ? own : *wq;
            band_col<HASW, DEC>(a, hq, wd.x, wd.y, wd.z, dx > 0 ? dx - 1 : -dx - 1, dx, dy, x, y, spy, gp, den,
                                num);
        }
    } else {
        const int lo = fwd_index(-r, dy, r) - a.band_base;  // own bit of column dx = -r
        const int k0 = (lo >> 5) - w0, sh = lo & 31;
        const uint4 o0 = wrow[k0 * HWC + tx + r];
        const uint4 o1 = k0 + 1 < kBandWordSlots ? wrow[(k0 + 1) * HWC + tx + r] : make_uint4(0u, 0u, 0u, 0u);
        const uint4 o2 = k0 + 2 < kBandWordSlots ? wrow[(k0 + 2) * HWC + tx + r] : make_uint4(0u, 0u, 0u, 0u);
        // bits lo .. lo + 63 of the row: column dx = -r + j takes bit j
        const uint32_t m0[3] = {__builtin_amdgcn_alignbit(o1.x, o0.x, (uint32_t)sh),
                                __builtin_amdgcn_alignbit(o1.y, o0.y, (uint32_t)sh),
                                __builtin_amdgcn_alignbit(o1.z, o0.z, (uint32_t)sh)};
        const uint32_t m1[3] = {__builtin_amdgcn_alignbit(o2.x, o1.x, (uint32_t)sh),
                                __builtin_amdgcn_alignbit(o2.y, o1.y, (uint32_t)sh),
                                __builtin_amdgcn_alignbit(o2.z, o1.z, (uint32_t)sh)};
        const int n0 = min(2 * r + 1, 32);
        int dx = -r;
        for (int j = 0; j < n0; ++j, ++dx, hq += F4)
            band_col<HASW, DEC>(a, hq, m0[0], m0[1], m0[2], j, dx, dy, x, y, spy, gp, den, num);
        for (int j = 32; j <= 2 * r; ++j, ++dx, hq += F4)
            band_col<HASW, DEC>(a, hq, m1[0], m1[1], m1[2], j - 32, dx, dy, x, y, spy, gp, den, num);
    }
}

template <bool HASW, int HWC, int TY, bool DEC>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(
    1, lds_waves_per_simd(BandStage<HASW, HWC, TY>::LDS_BYTES, 64 * TY)))) k_welch_band(WelchArgs a) {
    using B = BandStage<HASW, HWC, TY>;
    constexpr int F4 = B::F4, NS = TY + 1;
    extern __shared__ float4 bsh[];
    float4 *ring = bsh;                                  // NS rows of HITEMS
    uint4 *wbuf = (uint4 *)(bsh + NS * B::HITEMS);       // 2 steps of WITEMS
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = (t % tiles_x) * TX, y0 = (t / tiles_x) * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < a.W && y < a.H;
    const size_t pix = (size_t)y * a.ld + x;
    const bool first = a.band_dy0 == 0, last = a.band_dy1 == a.r;

    float gp[kNF], den[3], num[3];
#pragma unroll
    for (int k = 0; k < kNF; ++k) gp[k] = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) den[c] = 1.f, num[c] = 0.f;
    if (active) {
        const float4 *h = a.head + pix * F4;
        const float4 h0 = h[0];
        if (HASW) {
            const float4 h1 = h[1], h2 = h[2];
            const float g[kNF] = {h0.w, h1.x, h1.y, h1.z, h1.w, h2.x, h2.y};
#pragma unroll
            for (int k = 0; k < kNF; ++k) gp[k] = g[k];
        }
        const float film[3] = {h0.x, h0.y, h0.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            den[c] = first ? 1.f : a.part[(size_t)c * a.plane + pix];           // the pixel itself, weight 1
            num[c] = first ? film[c] : a.part[(size_t)(3 + c) * a.plane + pix];
        }
        if (DEC && first) {
            const int win = 2 * a.r + 1;
            const size_t k0 = ((size_t)(y * a.W + x) * win * win + (size_t)a.r * win + a.r) * 3;
            a.dec[k0] = a.dec[k0 + 1] = a.dec[k0 + 2] = 1;
        }
    }

    // sweeps of consecutive window rows: [-d1, d1] for the first band,
    // [-d1, -d0] and [d0, d1] for the second
    const int nsw = first ? 1 : 2;
    for (int sw = 0; sw < nsw; ++sw) {
        const int dya = sw == 0 ? -a.band_dy1 : a.band_dy0;
        const int dyb = (first || sw == 1) ? a.band_dy1 : -a.band_dy0;
        const int steps = dyb - dya + 1;
        B S;
        // ring row j is image row y0 + dya + j; wavefront ty reads row ty + k at step k
        for (int j = 0; j < TY; ++j) {
            band_load_head<HASW, HWC, TY>(a, S, y0 + dya + j, x0);
            band_store_head<HASW, HWC, TY>(S, ring + (j % NS) * B::HITEMS);
        }
        band_load_words<HASW, HWC, TY>(a, S, dya, x0, y0);
        band_store_words<HASW, HWC, TY>(S, wbuf);
        __syncthreads();
        for (int k = 0; k < steps; ++k) {
            const int dy = dya + k;
            const bool more = k + 1 < steps;
            if (more) {
                band_load_head<HASW, HWC, TY>(a, S, y0 + dya + k + TY, x0);
                band_load_words<HASW, HWC, TY>(a, S, dy + 1, x0, y0);
            }
            if (active)
                band_row<HASW, HWC, DEC>(a, ring + ((ty + k) % NS) * B::HITEMS, wbuf + (k & 1) * B::WITEMS +
                                         ty * kBandWordSlots * HWC, tx, x, y, dy, gp, den, num);
            if (more) {
                // row k + TY replaces row k - 1, read by nobody since the last barrier
                band_store_head<HASW, HWC, TY>(S, ring + ((k + TY) % NS) * B::HITEMS);
                band_store_words<HASW, HWC, TY>(S, wbuf + ((k + 1) & 1) * B::WITEMS);
            }
            __syncthreads();
        }
    }
    if (!active) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (last) {
            *(float *)((char *)a.out + (int64_t)y * a.out_pitch + (int64_t)x * a.out_xs + c * a.out_cs) = num[c] / den[c];
        } else {
            a.part[(size_t)c * a.plane + pix] = den[c];
            a.part[(size_t)(3 + c) * a.plane + pix] = num[c];
        }
    }
}

// Both bands: decision pass, weighting pass, decision pass, weighting pass.
template <int S>
static int welch_band_launch(const WelchArgs &a, bool hasw, hipStream_t st) {
    constexpr int HWC = TX + 40, DTY = kDecideTY, TY = kTY;
    const size_t dlds = (size_t)DTY * DecideLayout<S>::FLOATS * HWC * 4;
    STATER_HIP(hipFuncSetAttribute((const void *)k_welch_decide<S, HWC, DTY, true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)dlds));
    const size_t blds = hasw ? BandStage<true, HWC, TY>::LDS_BYTES : BandStage<false, HWC, TY>::LDS_BYTES;
    const bool dec = a.dec != nullptr;
    const void *bk = hasw ? (dec ? (const void *)k_welch_band<true, HWC, TY, true> : (const void *)k_welch_band<true, HWC, TY, false>)
                          : (dec ? (const void *)k_welch_band<false, HWC, TY, true> : (const void *)k_welch_band<false, HWC, TY, false>);
    STATER_HIP(hipFuncSetAttribute(bk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)blds));
    const dim3 dgrid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + DTY - 1) / DTY)));
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY)));
    WelchArgs b = a;
    b.rec_f4 = b.rec_head = 0;
    b.bits = (uint32_t *)a.rec;  // words: the variance-filter record planes
    b.bits_words = kBandWords;
    b.part = (float *)a.bits;    // sums, then weighting records: the decision-word planes
    b.head4 = hasw ? 3 : 1;
    for (int band = 0; band < 2; ++band) {
        b.band_dy0 = band == 0 ? 0 : kBandRows0 + 1;
        b.band_dy1 = band == 0 ? kBandRows0 : a.r;
        b.band_base = band == 0 ? 0 : a.r + kBandRows0 * (2 * a.r + 1);  // fwd_index(-r, 11, r)
        b.band_count = (band == 0 ? a.r + kBandRows0 * (2 * a.r + 1) : (a.r - kBandRows0) * (2 * a.r + 1));
        b.head = band == 0 ? (float4 *)(b.part + 6 * a.plane) : nullptr;
        hipLaunchKernelGGL((k_welch_decide<S, HWC, DTY, true>), dgrid, dim3(64 * DTY), dlds, st, b);
        STATER_HIP(hipGetLastError());
        b.head = (float4 *)(b.part + 6 * a.plane);
        if (hasw && dec) hipLaunchKernelGGL((k_welch_band<true, HWC, TY, true>), grid, dim3(64 * TY), blds, st, b);
        else if (hasw) hipLaunchKernelGGL((k_welch_band<true, HWC, TY, false>), grid, dim3(64 * TY), blds, st, b);
        else if (dec) hipLaunchKernelGGL((k_welch_band<false, HWC, TY, true>), grid, dim3(64 * TY), blds, st, b);
        else hipLaunchKernelGGL((k_welch_band<false, HWC, TY, false>), grid, dim3(64 * TY), blds, st, b);
        STATER_HIP(hipGetLastError());
    }
    return STATER_OK;
}

// HWC: halo row width of the radius class, TX + 20 (r <= 10, where the
// weighting pass may take records) or TX + 40 (r <= 20, bit planes only).
template <int S, int HWC>
static int welch_sym_launch(const WelchArgs &a, bool hasw, hipStream_t st) {
    constexpr int TY = kTY, DTY = kDecideTY;
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY)));
    const dim3 dgrid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + DTY - 1) / DTY)));
    size_t lds = (size_t)DTY * DecideLayout<S>::FLOATS * HWC * 4;
    STATER_HIP(hipFuncSetAttribute((const void *)k_welch_decide<S, HWC, DTY>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (!a.dec && a.rec && a.r <= kRecBitsRadius) {  // the decision pass writes the weighting pass's records itself
        WelchArgs b = a;
        b.rec_f4 = hasw ? AccRec<true>::FLOAT4S : AccRec<false>::FLOAT4S;
        b.rec_head = hasw ? AccRec<true>::HEAD4 : AccRec<false>::HEAD4;
        hipLaunchKernelGGL((k_welch_decide<S, HWC, DTY>), dgrid, dim3(64 * DTY), lds, st, b);
        STATER_HIP(hipGetLastError());
        return hasw ? accum1_launch<true>(b, st) : accum1_launch<false>(b, st);
    }
    hipLaunchKernelGGL((k_welch_decide<S, HWC, DTY>), dgrid, dim3(64 * DTY), lds, st, a);
    STATER_HIP(hipGetLastError());
    if (hasw) {
        lds = (size_t)TY * AccumLayout<true, HWC>::FLOATS * HWC * 4;
        STATER_HIP(hipFuncSetAttribute((const void *)k_welch_accum<true, HWC, TY>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_welch_accum<true, HWC, TY>), grid, dim3(64 * TY), lds, st, a);
    } else {
        lds = (size_t)TY * AccumLayout<false, HWC>::FLOATS * HWC * 4;
        STATER_HIP(hipFuncSetAttribute((const void *)k_welch_accum<false, HWC, TY>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_welch_accum<false, HWC, TY>), grid, dim3(64 * TY), lds, st, a);
    }
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

// One tile of 64 x TY pixels per workgroup, as for the variance filter.
template <int S, bool HASW, bool PLAIN, bool MODE3, int HWC, int TY>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(
    1, lds_waves_per_simd(TY * WelchLayout<S, HASW, PLAIN>::FLOATS * HWC * 4, 64 * TY)))) k_welch(WelchArgs a) {
    extern __shared__ float ring[];
    const int tiles_x = (a.W + TX - 1) / TX;
    const int t = xcd_tile(tiles_x * ((a.H + TY - 1) / TY));
    if (t < 0) return;
    welch_tile<S, HASW, PLAIN, MODE3, HWC, TY>(a, t % tiles_x, t / tiles_x, ring);
}

template <int S, bool HASW, bool PLAIN, bool MODE3, int HWC, int TY>
static int welch_launch5(const WelchArgs &a, hipStream_t st) {
    const size_t lds = (size_t)TY * WelchLayout<S, HASW, PLAIN>::FLOATS * HWC * 4;
    STATER_HIP(hipFuncSetAttribute((const void *)k_welch<S, HASW, PLAIN, MODE3, HWC, TY>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid(xcd_grid(((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY)));
    hipLaunchKernelGGL((k_welch<S, HASW, PLAIN, MODE3, HWC, TY>), grid, dim3(64 * TY), lds, st, a);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

template <int S, bool HASW, bool PLAIN, bool MODE3>
static int welch_launch3(const WelchArgs &a, int ty, hipStream_t st) {
    if (a.r > 10) return welch_launch5<S, HASW, PLAIN, MODE3, TX + 40, kTY>(a, st);
    if (HASW && !PLAIN && !MODE3 && ty != kTY)  // tile heights 2 and 8: the tuned mode-2 filter
        return ty == 2 ? welch_launch5<S, HASW, PLAIN, MODE3, TX + 20, 2>(a, st)
                       : welch_launch5<S, HASW, PLAIN, MODE3, TX + 20, 8>(a, st);
    return welch_launch5<S, HASW, PLAIN, MODE3, TX + 20, kTY>(a, st);
}

template <int S>
static int welch_launch2(const WelchArgs &a, bool hasw, int ty, hipStream_t st) {
    if (S > 1 && a.mode3)
        return hasw ? welch_launch3<S, true, false, S != 1>(a, ty, st) : welch_launch3<S, false, false, S != 1>(a, ty, st);
    return hasw ? welch_launch3<S, true, false, false>(a, ty, st) : welch_launch3<S, false, false, false>(a, ty, st);
}

static void fill_welch_args(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g, bool plain,
                            WelchArgs &a) {
    a.n = bs->n;
    a.mean = bs->stat(0, ST_MEAN, 0);
    a.set_stride = (size_t)N_STATS * 3 * bs->plane;
    if (plain || !p->cascade) {
        a.var = bs->stat(0, ST_M2, 0);
        a.var_is_m2 = 1;
    } else {
        a.var = p->use_levc ? bs->varCD : bs->varD;
        a.var_is_m2 = 0;
    }
    a.film = bs->stat(0, ST_MEAN, 0);
    a.film_pitch = (int64_t)bs->ld * 4;
    a.film_xs = 4;
    a.film_cs = (int64_t)bs->plane * 4;
    a.out = plain ? bs->meanD : bs->filmD;
    a.out_pitch = (int64_t)bs->ld * 4;
    a.out_xs = 4;
    a.out_cs = (int64_t)bs->plane * 4;
    a.jb = bs->jb;
    a.discr = bs->discr;
    a.varW = bs->varD;
    a.f = plain ? FeatArgs{} : make_feats(g);
    a.W = bs->W; a.H = bs->H; a.ld = bs->ld; a.r = p->radius;
    a.mode3 = (!plain && p->mt_mode == 3 && bs->S > 1) ? 1 : 0;
    a.plane = bs->plane;
    a.sp_scale = plain ? 0.f : p->ds_factor * kLog2e;
    a.alpha = (double)p->alpha;
    a.tp = tcrit_for(bs->ctx, (double)p->alpha);
}

WelchArgs make_welch_args(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g) {
    WelchArgs a{};
    fill_welch_args(bs, p, g, false, a);
    return a;
}

static int welch_common(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                        const stater_plane *film, const stater_plane *film_out, bool plain, uint8_t *dec,
                        hipStream_t st, int ty) {
    WelchArgs a{};
    fill_welch_args(bs, p, g, plain, a);
    // planar planes: channel c at c * channel_stride; interleaved (the
    // reference's CV_32FC3 film, estimator.cpp:421-457): pixels 12 bytes apart
    if (film) {
        const bool il = film->layout == STATER_INTERLEAVED;
        a.film = (const float *)film->ptr;
        a.film_pitch = film->pitch_bytes;
        a.film_xs = il ? 12 : 4;
        a.film_cs = il ? 4 : film->channel_stride_bytes;
    }
    if (film_out) {
        const bool il = film_out->layout == STATER_INTERLEAVED;
        a.out = (float *)film_out->ptr;
        a.out_pitch = film_out->pitch_bytes;
        a.out_xs = il ? 12 : 4;
        a.out_cs = il ? 4 : film_out->channel_stride_bytes;
    }
    a.dec = dec;
    a.bits = bs->welch_bits;
    a.bits_words = kBitsWords;
    a.band_dy0 = 0;
    a.band_dy1 = a.r;
    a.band_base = 0;
    a.band_count = 2 * a.r * a.r + 2 * a.r;
    a.rec = (float4 *)bs->var_rec;
    a.zero = (const float4 *)bs->var_zero;
    const int S = plain ? 1 : bs->S;
    const bool hasw = a.f.nf > 0 || a.sp_scale != 0.f;
    LaunchTimer t(bs->ctx, st, plain ? "welch_plain" : "welch_filter");
    if (plain) return welch_launch3<1, false, true, false>(a, kTY, st);
    if (!a.mode3 && a.bits && a.r >= 1 && a.r <= kRecBitsRadius) {  // mode 2: each pair tested once
        switch (S) {
            case 1: return welch_sym_launch<1, TX + 2 * kRecBitsRadius>(a, hasw, st);
            case 2: return welch_sym_launch<2, TX + 2 * kRecBitsRadius>(a, hasw, st);
            case 3: return welch_sym_launch<3, TX + 2 * kRecBitsRadius>(a, hasw, st);
            case 4: return welch_sym_launch<4, TX + 2 * kRecBitsRadius>(a, hasw, st);
            case 5: return welch_sym_launch<5, TX + 2 * kRecBitsRadius>(a, hasw, st);
            default: return welch_sym_launch<6, TX + 2 * kRecBitsRadius>(a, hasw, st);
        }
    }
    if (!a.mode3 && a.bits && a.rec && a.r > kRecBitsRadius && a.r <= 20) {  // mode 2, wider windows: in two bands
        switch (S) {
            case 1: return welch_band_launch<1>(a, hasw, st);
            case 2: return welch_band_launch<2>(a, hasw, st);
            case 3: return welch_band_launch<3>(a, hasw, st);
            case 4: return welch_band_launch<4>(a, hasw, st);
            case 5: return welch_band_launch<5>(a, hasw, st);
            default: return welch_band_launch<6>(a, hasw, st);
        }
    }
    switch (S) {
        case 1: return welch_launch2<1>(a, hasw, ty, st);
        case 2: return welch_launch2<2>(a, hasw, ty, st);
        case 3: return welch_launch2<3>(a, hasw, ty, st);
        case 4: return welch_launch2<4>(a, hasw, ty, st);
        case 5: return welch_launch2<5>(a, hasw, ty, st);
        default: return welch_launch2<6>(a, hasw, ty, st);
    }
}

int launch_welch_filter(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                        const stater_plane *film, const stater_plane *film_out, bool plain, hipStream_t st, int ty) {
    return welch_common(bs, p, g, film, film_out, plain, nullptr, st, ty);
}

int launch_welch_decisions(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                           const stater_plane *film, int plain, uint8_t *dec, hipStream_t st) {
    return welch_common(bs, p, g, film, nullptr, plain != 0, dec, st, kTY);
}

// ============================================================ mean vars
__global__ void __launch_bounds__(256) k_mean_var(const int32_t *n, const float *m2, float *out, int W, int ld,
                                                  size_t plane) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const size_t pix = (size_t)y * ld + x;
    const float nf = (float)n[pix];
    for (int c = 0; c < 3; ++c) out[c * plane + pix] = m2[c * plane + pix] / ((nf - 1.f) * nf);
}

int launch_mean_vars(stater_bufset *bs, hipStream_t st) {
    LaunchTimer t(bs->ctx, st, "mean_var");
    hipLaunchKernelGGL(k_mean_var, dim3((bs->W + 255) / 256, bs->H), dim3(256), 0, st, bs->n,
                       bs->stat(0, ST_M2, 0), bs->meanVar, bs->W, bs->ld, bs->plane);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

// ==================================================== adaptive sampling
// Integer-exact water filling.  e_p (FP32, fixed operation order) is turned
// into q_p = floor(e_p 2^16); f(lam) = sum_p min(floor(lam q_p), h_p) with
// h_p = cap - n_p is an integer sum, so every reduction is exact and the
// result does not depend on the reduction order.  A search over lam in FP64
// finds lo < hi with f(lo) <= B < f(hi); targets are min(floor(lo q), h) plus
// the first R = B - f(lo) of the increments to hi, in pixel order.
// When the pixels with q > 0 cannot take the budget (sum of their h <= B,
// e.g. a flat background beside pixels near the cap), f saturates below B:
// those pixels then take their whole headroom and the rest of the budget is
// shared uniformly over the q = 0 pixels (weight 1 each), which is also the
// rule when every q is 0.
constexpr int kAsCand = 16;

struct AsCtl {
    unsigned long long sumq, sumh, sumhq;
    double lo, hi;
    double cand[kAsCand];
    unsigned long long f[kAsCand];
    unsigned long long flo;
    int phase;  // 0 bracket search, 1 refine, 2 done, 3 all pixels capped
    int uniform;
    unsigned long long B;
    unsigned long long remaining;
};

struct AsArgs {
    const int32_t *n;
    const float *varD, *mean;
    int64_t n_ld;           // row pitch of n, elements
    int64_t v_ld, v_cs;     // varD: row pitch and channel stride, elements
    int64_t m_ld, m_cs;     // mean: row pitch and channel stride, elements
    int var_is_m2;          // 1: varD holds m2 and the variance is m2/(n-1) (AS_MODE 0)
    int W, H;
    uint32_t cap;
    uint32_t *q;            // per-pixel q and h kept by the error pass, or null: recomputed
    int32_t *h;             // from n, varD and mean wherever they are needed
    int32_t *inc;           // per-pixel increments lo -> hi, or null: recomputed
    int64_t *bsum;
    AsCtl *ctl;
    int32_t *out;
    int64_t out_pitch;
};

// q_p and h_p of pixel p (row-major index over W x H), in the fixed FP32
// operation order; every pass that needs them gets the same values.
__device__ __forceinline__ void as_qh(const AsArgs &a, int p, uint32_t &q, int32_t &h) {
#pragma clang fp contract(off)
    const int y = p / a.W, x = p - y * a.W;
    const int n = a.n[(size_t)y * a.n_ld + x];
    float e = 0.f;
    if (n > 0) {
        const float nf = (float)n;
        float ec[3];
        for (int c = 0; c < 3; ++c) {
            float vd = a.varD[c * a.v_cs + (size_t)y * a.v_ld + x];
            if (a.var_is_m2) vd = n >= 2 ? vd / (float)(n - 1) : 0.f;
            vd = fmaxf(vd, 0.f);
            const float m = fmaxf(fabsf(a.mean[c * a.m_cs + (size_t)y * a.m_ld + x]), 1e-3f);
            ec[c] = sqrtf(vd / nf) / m;  // correctly rounded (IEEE) square root
        }
        e = (ec[0] + ec[1] + ec[2]) / 3.f;
    }
    const float qe = e * 65536.f;
    q = !(qe > 0.f) ? 0u : (qe >= 4294967295.f ? 4294967295u : (uint32_t)qe);
    const long long hh = (long long)a.cap - (long long)n;
    h = hh > 0 ? (int32_t)hh : 0;
}

__device__ __forceinline__ void as_get_qh(const AsArgs &a, int p, uint32_t &q, int32_t &h) {
    if (a.q) { q = a.q[p]; h = a.h[p]; }
    else as_qh(a, p, q, h);
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// Blocks of the passes that sum over all pixels (error sums, the search's
// counts): each thread walks the pixels with a grid stride and each block
// leaves its partial sums in bsum, at most kAsCand per block (bsum holds
// max(nb, 32) entries, so G kAsCand fits).  No atomics: a few hundred
// partials per sum are added up by one workgroup.
inline int as_sum_blocks(int nb) {
    const int g = nb / kAsCand;
    return g < 1 ? 1 : (g > 2048 ? 2048 : g);
}

__global__ void __launch_bounds__(256) k_as_err(AsArgs a) {
    __shared__ unsigned long long sh[4];
    const int P = a.W * a.H;
    unsigned long long sq = 0, shh = 0, shq = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        uint32_t q;
        int32_t h;
        as_qh(a, p, q, h);
        if (a.q) { a.q[p] = q; a.h[p] = h; }
        sq += q;
        shh += (unsigned long long)h;
        shq += q ? (unsigned long long)h : 0ull;
    }
    const unsigned long long bq = block_sum_u64(sq, sh);
    const unsigned long long bh = block_sum_u64(shh, sh);
    const unsigned long long bhq = block_sum_u64(shq, sh);
    if (threadIdx.x == 0) {
        a.bsum[3 * blockIdx.x] = (int64_t)bq;
        a.bsum[3 * blockIdx.x + 1] = (int64_t)bh;
        a.bsum[3 * blockIdx.x + 2] = (int64_t)bhq;
    }
}

// Sum of the G partials of sum k (stride `stride`, offset k) into thread 0.
__device__ __forceinline__ unsigned long long as_partials(const int64_t *part, int G, int stride, int k,
                                                          unsigned long long *sh) {
    unsigned long long v = 0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) v += (unsigned long long)part[b * stride + k];
    return block_sum_u64(v, sh);
}

__global__ void __launch_bounds__(256) k_as_init(AsCtl *c, const int64_t *part, int G, unsigned long long B, int P) {
    __shared__ unsigned long long sh[4];
    const unsigned long long sumq = as_partials(part, G, 3, 0, sh);
    const unsigned long long sumh = as_partials(part, G, 3, 1, sh);
    const unsigned long long sumhq = as_partials(part, G, 3, 2, sh);
    if (threadIdx.x != 0) return;
    c->sumq = sumq;
    c->sumh = sumh;
    c->sumhq = sumhq;
    c->B = B;
    for (int i = 0; i < kAsCand; ++i) c->f[i] = 0;
    if (c->sumh <= B) { c->phase = 3; return; }
    c->uniform = c->sumhq <= B;
    const double sq = c->uniform ? (double)P : (double)c->sumq;
    c->lo = 0.0;
    c->hi = (double)B / sq;
    for (int i = 0; i < kAsCand; ++i) c->cand[i] = c->hi * (double)(1ull << i);
    c->phase = 0;
}

__device__ __forceinline__ long long as_take(double lam, uint32_t q, int32_t h) {
    const double t = floor(lam * (double)q);
    return t >= (double)h ? (long long)h : (long long)t;
}

// A pixel's share at lam: in the uniform rule the q > 0 pixels are saturated
// and every q = 0 pixel has weight 1.
__device__ __forceinline__ long long as_share(double lam, bool uniform, uint32_t q, int32_t h) {
    if (uniform) return q ? (long long)h : as_take(lam, 1u, h);
    return as_take(lam, q, h);
}

__global__ void __launch_bounds__(256) k_as_count(AsArgs a) {
    __shared__ unsigned long long sh[4];
    const AsCtl *c = a.ctl;
    if (c->phase >= 2) return;
    const int P = a.W * a.H;
    const bool uniform = c->uniform;
    double cand[kAsCand];
    unsigned long long f[kAsCand];
#pragma unroll
    for (int i = 0; i < kAsCand; ++i) { cand[i] = c->cand[i]; f[i] = 0; }
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        uint32_t q;
        int32_t h;
        as_get_qh(a, p, q, h);
#pragma unroll
        for (int i = 0; i < kAsCand; ++i) f[i] += (unsigned long long)as_share(cand[i], uniform, q, h);
    }
#pragma unroll
    for (int i = 0; i < kAsCand; ++i) {
        const unsigned long long s = block_sum_u64(f[i], sh);
        if (threadIdx.x == 0) a.bsum[blockIdx.x * kAsCand + i] = (int64_t)s;
    }
}

__global__ void __launch_bounds__(256) k_as_step(AsCtl *c, const int64_t *part, int G) {
    __shared__ unsigned long long sh[4];
    if (c->phase >= 2) return;  // the same for every thread
    unsigned long long fs[kAsCand];
    for (int i = 0; i < kAsCand; ++i) fs[i] = as_partials(part, G, kAsCand, i, sh);
    if (threadIdx.x != 0) return;
    for (int i = 0; i < kAsCand; ++i) c->f[i] = fs[i];
    const unsigned long long B = c->B;
    if (c->phase == 0) {
        int j = -1;
        for (int i = 0; i < kAsCand; ++i)
            if (c->f[i] > B) { j = i; break; }
        if (j < 0) {
            c->lo = c->cand[kAsCand - 1];
            c->flo = c->f[kAsCand - 1];
            const double base = c->cand[kAsCand - 1] * 2.0;
            for (int i = 0; i < kAsCand; ++i) c->cand[i] = base * (double)(1ull << i);
        } else {
            if (j > 0) { c->lo = c->cand[j - 1]; c->flo = c->f[j - 1]; }
            c->hi = c->cand[j];
            c->phase = 1;
        }
    } else {
        double nlo = c->lo, nhi = c->hi;
        unsigned long long nflo = c->flo;
        for (int i = 0; i < kAsCand; ++i) {
            if (c->f[i] <= B) { nlo = c->cand[i]; nflo = c->f[i]; }
            else { nhi = c->cand[i]; break; }
        }
        c->lo = nlo; c->hi = nhi; c->flo = nflo;
    }
    if (c->phase == 1) {
        bool moved = false;
        for (int i = 0; i < kAsCand; ++i) {
            const double t = c->lo + (c->hi - c->lo) * (double)(i + 1) / (double)(kAsCand + 1);
            c->cand[i] = t;
            if (t > c->lo && t < c->hi) moved = true;
        }
        if (!moved) c->phase = 2;
    }
    for (int i = 0; i < kAsCand; ++i) c->f[i] = 0;
}

// Per-block sums of the increments lo -> hi (first pass of the scan).
__global__ void __launch_bounds__(256) k_as_inc(AsArgs a) {
    __shared__ unsigned long long sh[4];
    const AsCtl *c = a.ctl;
    const int P = a.W * a.H;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    long long inc = 0;
    if (p < P && c->phase != 3) {
        uint32_t q;
        int32_t h;
        as_get_qh(a, p, q, h);
        inc = as_share(c->hi, c->uniform, q, h) - as_share(c->lo, c->uniform, q, h);
        if (a.inc) a.inc[p] = (int32_t)inc;
    }
    const unsigned long long s = block_sum_u64((unsigned long long)inc, sh);
    if (threadIdx.x == 0) a.bsum[blockIdx.x] = (int64_t)s;
}

// Exclusive scan of the block sums (one workgroup, sequential chunks).
__global__ void __launch_bounds__(256) k_as_scan_blocks(int64_t *bsum, int nb) {
    __shared__ int64_t carry;
    __shared__ int64_t tmp[256];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 256) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        tmp[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int64_t t = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
            __syncthreads();
            tmp[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + tmp[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 255) carry += tmp[255];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_as_final(AsArgs a) {
    __shared__ int64_t tmp[256];
    const AsCtl *c = a.ctl;
    const int P = a.W * a.H;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t q = 0;
    int32_t h = 0;
    if (p < P) as_get_qh(a, p, q, h);
    int64_t inc = 0;
    if (p < P && c->phase != 3)
        inc = a.inc ? a.inc[p] : as_share(c->hi, c->uniform, q, h) - as_share(c->lo, c->uniform, q, h);
    tmp[threadIdx.x] = inc;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int64_t t = threadIdx.x >= o ? tmp[threadIdx.x - o] : 0;
        __syncthreads();
        tmp[threadIdx.x] += t;
        __syncthreads();
    }
    if (p >= P) return;
    const int y = p / a.W, x = p - y * a.W;
    int32_t *o = (int32_t *)((char *)a.out + (int64_t)y * a.out_pitch + (int64_t)x * 4);
    if (c->phase == 3) { *o = h; return; }
    const int64_t base = as_share(c->lo, c->uniform, q, h);
    const int64_t R = (int64_t)(c->B - c->flo);
    const int64_t E = a.bsum[blockIdx.x] + tmp[threadIdx.x] - inc;  // exclusive prefix
    int64_t add = R - E;
    add = add < 0 ? 0 : (add > inc ? inc : add);
    *o = (int32_t)(base + add);
}

// The AS map's launches.  a.bsum must hold max(nb, 32, as_sum_blocks(nb) *
// kAsCand) int64 entries (nb = 256-pixel blocks) and a.ctl one control block.
static int run_target_spp(const AsArgs &a, uint64_t budget, hipStream_t st) {
    const int P = a.W * a.H;
    const int nb = (P + 255) / 256;
    int rc = launch_zero(a.ctl, sizeof(AsCtl), st);
    if (rc) return rc;
    const int G = as_sum_blocks(nb);
    hipLaunchKernelGGL(k_as_err, dim3(G), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_as_init, dim3(1), dim3(256), 0, st, a.ctl, (const int64_t *)a.bsum, G,
                       (unsigned long long)budget * (unsigned long long)P, P);
    // Each count/step pair narrows [lo, hi] 17-fold (bracketing first); 40 pairs
    // exceed what a double can resolve, and finished searches skip the work.
    for (int it = 0; it < 40; ++it) {
        hipLaunchKernelGGL(k_as_count, dim3(G), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_as_step, dim3(1), dim3(256), 0, st, a.ctl, (const int64_t *)a.bsum, G);
    }
    hipLaunchKernelGGL(k_as_inc, dim3(nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_as_scan_blocks, dim3(1), dim3(256), 0, st, a.bsum, nb);
    hipLaunchKernelGGL(k_as_final, dim3(nb), dim3(256), 0, st, a);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

int launch_target_spp(stater_bufset *bs, uint64_t budget, uint32_t cap, int as_mode, const stater_plane *out,
                      hipStream_t st) {
    const int P = bs->W * bs->H;
    static_assert(sizeof(AsCtl) <= kAsCtlBytes, "AS control block outgrew its slot");
    if (!bs->as_scratch || bs->as_scratch_bytes < as_scratch_bytes_for((size_t)P))
        return fail(STATER_E_STATE, "stater_calc_target_spp: set has no adaptive-sampling scratch");
    char *sp = (char *)bs->as_scratch;
    AsArgs a{};
    a.n = bs->n;
    a.varD = as_mode ? bs->varD : bs->stat(0, ST_M2, 0);
    a.var_is_m2 = as_mode ? 0 : 1;
    a.mean = bs->stat(0, ST_MEAN, 0);
    a.n_ld = a.v_ld = a.m_ld = bs->ld;
    a.v_cs = a.m_cs = (int64_t)bs->plane;
    a.W = bs->W; a.H = bs->H; a.cap = cap;
    const size_t pb = ((size_t)P * 4 + 255) / 256 * 256;
    a.ctl = (AsCtl *)sp; sp += kAsCtlBytes;
    a.q = (uint32_t *)sp; sp += pb;
    a.h = (int32_t *)sp; sp += pb;
    a.inc = (int32_t *)sp; sp += pb;
    a.bsum = (int64_t *)sp;
    a.out = out ? (int32_t *)out->ptr : bs->targetspp;
    a.out_pitch = out ? out->pitch_bytes : (int64_t)bs->ld * 4;
    LaunchTimer t(bs->ctx, st, "target_spp");
    return run_target_spp(a, budget, st);
}

// ---- the AS map without allocation ----------------------------------------
// Device memory of stater_calc_target_spp_noalloc: one control block and the
// block partials of a film of at most STATER_NOALLOC_MAX_PIXELS pixels
// (65536 256-pixel blocks).  Each device holds its own copy for the life of
// the process; hipGetSymbolAddress gives its address on the current device.
constexpr size_t kNoallocBsum = STATER_NOALLOC_MAX_PIXELS / 256;
static_assert(kNoallocBsum >= (size_t)2048 * kAsCand, "block partials of the count passes must fit");
constexpr size_t kNoallocBytes = kAsCtlBytes + kNoallocBsum * sizeof(int64_t);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Called by stater_create with `device` current.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(STATER_E_UNSUPPORTED, "stater_create: device index too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return STATER_OK;
    void *ptr = nullptr;
    STATER_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(STATER_E_HIP, "hipGetSymbolAddress returned no address for the AS scratch");
    g_scratch_ptr[device] = ptr;
    return STATER_OK;
}

static int check_plane(const stater_plane *pl, int W, int H, int ch, stater_dtype dt, const char *what) {
    if (!pl || !pl->ptr || pl->width != W || pl->height != H || pl->channels != ch || pl->dtype != dt ||
        pl->layout != STATER_PLANAR || pl->pitch_bytes < (int64_t)W * 4 || pl->pitch_bytes % 4 ||
        (ch > 1 && pl->channel_stride_bytes % 4))
        return fail(STATER_E_ARG, std::string("stater_calc_target_spp_noalloc: bad ") + what + " plane");
    return STATER_OK;
}

int launch_target_spp_noalloc(stater_ctx *ctx, int W, int H, const stater_plane *n, const stater_plane *varD,
                              const stater_plane *mean, uint64_t budget, uint32_t cap, const stater_plane *out,
                              hipStream_t st) {
    if (W <= 0 || H <= 0) return fail(STATER_E_ARG, "stater_calc_target_spp_noalloc: empty film");
    if ((uint64_t)W * (uint64_t)H > STATER_NOALLOC_MAX_PIXELS)
        return fail(STATER_E_ARG, "stater_calc_target_spp_noalloc: film larger than STATER_NOALLOC_MAX_PIXELS");
    int rc;
    if ((rc = check_plane(n, W, H, 1, STATER_I32, "n")) || (rc = check_plane(varD, W, H, 3, STATER_F32, "varD")) ||
        (rc = check_plane(mean, W, H, 3, STATER_F32, "mean")) || (rc = check_plane(out, W, H, 1, STATER_I32, "out")))
        return rc;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
    }
    if (!base) return fail(STATER_E_STATE, "stater_calc_target_spp_noalloc: no scratch on the context's device");
    AsArgs a{};
    a.n = (const int32_t *)n->ptr;
    a.n_ld = n->pitch_bytes / 4;
    a.varD = (const float *)varD->ptr;
    a.v_ld = varD->pitch_bytes / 4;
    a.v_cs = varD->channel_stride_bytes / 4;
    a.mean = (const float *)mean->ptr;
    a.m_ld = mean->pitch_bytes / 4;
    a.m_cs = mean->channel_stride_bytes / 4;
    a.W = W; a.H = H; a.cap = cap;
    a.ctl = (AsCtl *)base;
    a.bsum = (int64_t *)((char *)base + kAsCtlBytes);
    a.out = (int32_t *)out->ptr;
    a.out_pitch = out->pitch_bytes;
    return run_target_spp(a, budget, st);
}

}  // namespace stater

extern "C" int stater_calc_target_spp_noalloc(stater_ctx *ctx, int width, int height, const stater_plane *n,
                                              const stater_plane *varD, const stater_plane *mean,
                                              uint64_t budget_spp, uint32_t cap, const stater_plane *out,
                                              void *stream) {
    if (!ctx) return stater::fail(STATER_E_ARG, "stater_calc_target_spp_noalloc: NULL context");
    if (budget_spp > (uint64_t)1 << 32) return stater::fail(STATER_E_ARG, "stater_calc_target_spp_noalloc: budget too large");
    return stater::launch_target_spp_noalloc(ctx, width, height, n, varD, mean, budget_spp, cap, out,
                                             stream ? (hipStream_t)stream : ctx->stream);
}

extern "C" int stater_debug_film_decisions(stater_bufset *bs, const stater_denoise_params *p,
                                           const stater_gbuffers *gbufs, const stater_plane *film, int plain,
                                           uint8_t *dec, void *stream) {
    if (!bs || !p || !dec) return stater::fail(STATER_E_ARG, "stater_debug_film_decisions: NULL argument");
    return stater::launch_welch_decisions(bs, p, gbufs, film, plain, dec, stater::pick_stream(bs, stream));
}
