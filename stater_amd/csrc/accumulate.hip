// accumulate.hip — per-sample moment accumulation on gfx950.
//
// Replaces the CPU hot loop of the reference: for every sample the renderer
// calls StatTile<Vec3>::AddTransformSampleM4C once per lambda slot
// (src/statistics/statpath.cpp:341-344 -> estimator.h:312-324 -> :278-305),
// then merges the tile into host buffers (estimator.cpp:301-343) that are
// uploaded to the GPU every iteration (estimator.cpp:346-353).  Here the
// statistics never leave device memory.
//
// Mapping (sequential mode): one thread per pixel, 64 consecutive pixels of a
// row per wavefront, so every sample-plane load of a wavefront is one
// coalesced 256-byte segment.  Each thread keeps the three channels'
// statistics of every set in registers for the whole chunk and walks the
// samples in order, so the update stays the reference's sequential
// recurrence; the guard, the count and the per-sample factors are computed
// once per pixel-sample for all channels and sets.  The FP64-merge mode runs
// one thread per (pixel, channel): three wavefronts per 64 pixels.
//
// Numerics:
//  * set 0 (untransformed) runs the reference update in its exact FP32
//    operation order (no contraction, IEEE division and square root: sqrtf,
//    not __fsqrt_rn, which lowers to the approximate v_sqrt_f32), so its
//    mean, m2, m3, m4, c2X, c2Y and m1 are bit-identical to StatTile<Vec3>
//    fed the same samples;
//  * the Box-Cox sets are accumulated relative to a per-pixel reference
//    point: with L = ln x1 of the pixel's first sample, K = (e^{lL} - 1)/l and
//    w = e^{lL}/l, every transformed value is taken as
//        y - K = w * expm1(l (ln x - L)),
//    which keeps full relative precision for every lambda (the reference's
//    (pow(x,l) - 1)/l, estimator.h:56,72-74, cancels for small l, and for
//    l = 3 on dark pixels FP32 y = -1/3 + tiny loses the variation entirely).
//    Central moments, c2X, c2Y and m1 do not depend on the shift; the stored
//    mean is K + mean(y - K); mean(y - K) and x1 persist in internal planes
//    so that a later call continues the recurrence bit for bit.  One log per
//    sample-channel is shared by all sets; expm1 is evaluated in base 2
//    (wexpm1 below) and the log is the hardware log2.  Sets with lambda 1/4,
//    1/2, 3/4 or 3 take the same value from the ratio x / x1 without the log
//    (bc_pow below).  The update is the
//    reference recurrence in factored form,
//        m2 += d^2 (1 - 1/n)
//        m3 += -3 (d/n) m2 + d^3 (1 - 1/n^2)
//        m4 += -4 (d/n) m3 - 6 (d^2/n^2) m2 + d^4 (1 - 1/n^3)
//        m1 += |d| sqrt(1 - 1/n),
//    with the per-sample factors shared by all sets; it is exact algebra of
//    estimator.h:288-304 and has no cancellation at n = 1.  The Box-Cox sets
//    are checked against an FP64 evaluation of the same recurrence.

#include "exact.hpp"
#include "kernels.hpp"

namespace stater {



// Reference update, estimator.h:226-305, evaluated exactly as written, with
// d = x - mean and dN = d / n formed by the caller for the three channels;
// for MM 5 it returns the radicand d (d - dN) of m1's increment, whose
// square root the caller takes and adds (the update's last step).
template <int MM>
__device__ __forceinline__ float meng_exact(float *st, float x, float d, float dN, float xD, float yD) {
#pragma clang fp contract(off)
    float &mean = st[0];
    if (MM == 1) {
        mean += dN;
        return 0.f;
    }
    if (MM == 2) {
        mean += dN;
        st[1] += d * (d - dN);
        return 0.f;
    }
    float &m2 = st[1];
    float &m3 = st[2];
    const float d2 = d * d;
    const float dN2 = dN * dN;
    mean += dN;
    m2 += d * (d - dN);
    m3 += -3.f * dN * m2 + d * (d2 - dN2);
    if (MM == 3) return 0.f;
    float &m4 = st[3];
    m4 += -4.f * dN * m3 - 6.f * dN2 * m2 + d * (d * d2 - dN * dN2);
    if (MM == 4) return 0.f;
    st[4] += (x - mean) * xD;
    st[5] += (x - mean) * yD;
    return d * (d - dN);
}

// Wt (e^z - 1) for z = w ln 2, evaluated in base 2: w = k + f with k the
// nearest integer (f = w - k is exact, |f| <= 1/2), 2^f - 1 = f p(f) with p a
// degree-4 least-relative-error fit to (2^f - 1)/f on [-1/2, 1/2]
// (|relative error| 4.3e-7 in exact arithmetic), and
//     Wt (2^w - 1) = (Wt 2^k)(2^f - 1) + (Wt 2^k - Wt),
// so the scale Wt rides on the exponent and costs no multiply.  The relative
// error of the result is at most 6.9e-7 (12 ulp) at every w, largest for
// |w| < 1/2, far below the sampling noise of the moments it feeds.  w is
// clamped at -160, where Wt 2^k is zero (or negligible against Wt): the limit
// x -> 0, which a sample of 0 reaches as ln-ratio -1e30.  The clamp also
// keeps f = w - k exact: without it the compiler may contract the multiply
// that made w with this subtraction, fma(dl, l log2 e, -k), which for a
// huge |w| is the product's rounding error, not 0 (a zero sample then gave
// NaN).  NaN w gives -Wt here; the callers turn a NaN sample into NaN.
__device__ __forceinline__ int cvt_i32_sat(float k) {
    int i;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(i) : "v"(k));
    return i;
}

__device__ __forceinline__ float wexpm1(float w, float Wt) {
    w = fmaxf(w, -160.f);
    const float k = __builtin_rintf(w);
    const float f = w - k;
    float p = 1.3333528768271208e-03f;
    p = __builtin_fmaf(p, f, 9.666242636740208e-03f);
    p = __builtin_fmaf(p, f, 5.550500378012657e-02f);
    p = __builtin_fmaf(p, f, 2.4022351205348969e-01f);
    p = __builtin_fmaf(p, f, 6.931471228599548e-01f);
    p *= f;  // 2^f - 1
    const float wk = __builtin_ldexpf(Wt, cvt_i32_sat(k));
    return __builtin_fmaf(wk, p, wk - Wt);
}

constexpr float kLog2eAcc = 1.4426950408889634f;

// Wt (e^z - 1) for |z| <= 1/2 as (Wt z) p(z), p a degree-5
// least-relative-error fit to (e^z - 1)/z on [-1/2, 1/2] (relative error
// 9.6e-8 in exact arithmetic, 2.3e-7 evaluated in FP32): full-rate
// instructions only, where wexpm1 needs a clamp, a rounding, a conversion
// and an ldexp.  With z = l (ln x - L) and l = 0.01 (the log-based set of the
// reference's lambda lists) it covers every ratio x / x1 in [e^-50, e^50]; a
// zero, negative, infinite or NaN sample gives |z| > 1/2 or a NaN z, and its
// lane takes the general form instead.
constexpr float kExpm1SmallZ = 0.5f;
__device__ __forceinline__ float wexpm1_small(float z, float Wt) {
    float p = 1.3888798421248794e-03f;
    p = __builtin_fmaf(p, z, 8.407683111727238e-03f);
    p = __builtin_fmaf(p, z, 4.166870191693306e-02f);
    p = __builtin_fmaf(p, z, 1.6665974259376526e-01f);
    p = __builtin_fmaf(p, z, 4.999998211860657e-01f);
    p = __builtin_fmaf(p, z, 1.0000001192092896f);
    return (Wt * z) * p;
}

// Box-Cox sets whose lambda is 1/4, 1/2, 3/4 or 3 (every transformed set of
// the reference's scene configurations but lambda = 0.01) take the shifted
// value from rho = x / x1 without a logarithm,
//     y - K = Wt (rho^l - 1),   Wt = x1^l / l,
// evaluated from u = rho - 1 = (x - x1) / x1, which has the relative
// precision of a quotient (x - x1 is exact near x1), so that every power
// minus one keeps it too:
//     rho^(1/2) - 1 = u / (rho^(1/2) + 1)                        = h2,
//                     (the root of rho = x / x1 itself, precise for small rho too)
//     rho^(1/4) - 1 = h2 / (rho^(1/4) + 1)                       = h4,
//     rho^(3/4) - 1 = h2 rho^(1/4) + h4   (= h2 + h4 + h2 h4; both terms
//                                          have the sign of u: no cancellation),
//     rho^3 - 1     = u (u (u + 3) + 3)   (Wt u taken first, so the
//                                          product overflows only where x^3 does),
// with the hardware square root and reciprocal (the square roots only in
// the denominators; one reciprocal of (rho^(1/2) + 1)(rho^(1/4) + 1) serves
// both quotients).  About twenty instructions per sample-channel for the
// four sets against twelve per set for the log-based expm1.  As pow(x, l)
// in the reference's boxCox (estimator.h:56): x = 0 gives -1/l (u = -1,
// y - K = -Wt), a negative x gives NaN for the roots and the exact cube for
// l = 3, and NaN stays NaN.  KC packs the kind of set s (1 .. T) in bits
// 3 (s - 1) .. 3 (s - 1) + 2; 0 is the log-based form.
enum : int { kBcGeneric = 0, kBcQuarter = 1, kBcHalf = 2, kBcThreeQuarter = 3, kBcCube = 4 };
__host__ __device__ constexpr int bc_kind(int KC, int s) { return (KC >> (3 * (s - 1))) & 7; }
__host__ __device__ constexpr bool bc_has(int KC, int T, int k0, int k1) {
    for (int s = 1; s <= T; ++s)
        if (bc_kind(KC, s) == k0 || bc_kind(KC, s) == k1) return true;
    return false;
}
// The one log-based set of a list whose other sets are all ratio sets (the
// reference's multi-transform lists), or 0: that set takes wexpm1_small
// where it can.
__host__ __device__ constexpr int bc_single_generic(int KC, int T) {
    if (KC == 0) return 0;
    int sg = 0;
    for (int s = 1; s <= T; ++s)
        if (bc_kind(KC, s) == kBcGeneric) {
            if (sg) return 0;
            sg = s;
        }
    return sg;
}

struct BcPow {
    float u, h2, h4, s4;  // rho - 1, rho^(1/2) - 1, rho^(1/4) - 1, rho^(1/4)
};

template <int T, int KC>
__device__ __forceinline__ BcPow bc_pow(float v, float x1, float ix1) {
    BcPow p{0.f, 0.f, 0.f, 0.f};
    if (KC == 0) return p;
    p.u = (v - x1) * ix1;
    if (bc_has(KC, T, kBcHalf, kBcThreeQuarter) || bc_has(KC, T, kBcQuarter, kBcQuarter)) {
        const float s2 = __builtin_amdgcn_sqrtf(v * ix1);  // rho itself: u + 1 loses a small rho's precision
        if (bc_has(KC, T, kBcQuarter, kBcThreeQuarter)) {
            // one reciprocal for both denominators: h4 = u / ((s2 + 1)(s4 + 1)), h2 = h4 (s4 + 1)
            p.s4 = __builtin_amdgcn_sqrtf(s2);
            const float t4 = p.s4 + 1.f;
            p.h4 = p.u * __builtin_amdgcn_rcpf(__builtin_fmaf(s2, t4, t4));
            p.h2 = __builtin_fmaf(p.h4, p.s4, p.h4);
        } else {
            p.h2 = p.u * __builtin_amdgcn_rcpf(s2 + 1.f);
        }
    }
    return p;
}

// The shifted transformed value t = y - K of set s.  The log-based form
// takes dl = ln|x| - L: a negative x then follows pow's sign rule, NaN for a
// non-integer lambda (par 0), x^l = |x|^l for an even one (y - K = Wt (r - 1),
// r = (|x| / x1)^l) and -|x|^l for an odd one (y - K = -Wt (r + 1)).  The
// reference lists' log-form set (lambda = 0.01) is known not to be an
// integer.
template <int KC>
__device__ __forceinline__ float bc_shifted(int s, float lam, float dl, const BcPow &p, float Wt, float x, int par) {
    const int kd = bc_kind(KC, s);
    if (kd == kBcGeneric) {
        const float e = wexpm1(lam * kLog2eAcc * dl, Wt);
        if (KC != 0 || !(x < 0.f)) return x >= 0.f ? e : __builtin_nanf("");  // NaN x: NaN
        return par == 0 ? __builtin_nanf("") : par == 1 ? e : -e - 2.f * Wt;
    }
    if (kd == kBcCube) return (Wt * p.u) * __builtin_fmaf(p.u, p.u + 3.f, 3.f);
    const float h = kd == kBcQuarter ? p.h4 : kd == kBcHalf ? p.h2 : __builtin_fmaf(p.h2, p.s4, p.h4);
    return Wt * h;
}

// ln|x| - L for the log-based form; x = 0 gives -1e30, where expm1 is -1 and
// y(0) = -1/l (as pow(0, l) in estimator.h:56).
__device__ __forceinline__ float log_ratio(float x, float L) {
    return fmaxf(__builtin_amdgcn_logf(fabsf(x)) * 0.69314718055994531f - L, -1e30f);
}

// Per-sample factors of the factored recurrence, shared by all Box-Cox sets.
struct NFac {
    float b, a1, a2, a3, b3, b4, b6, sa1;
};
constexpr int kFacSlots = 10;  // NFac, recip_rn(nf), nf

__device__ __forceinline__ NFac nfac(float nf) {
    NFac f;
    f.b = __builtin_amdgcn_rcpf(nf);
    const float b2 = f.b * f.b;
    f.a1 = 1.f - f.b;
    f.a2 = 1.f - b2;
    f.a3 = 1.f - b2 * f.b;
    f.b3 = -3.f * f.b;
    f.b4 = -4.f * f.b;
    f.b6 = -6.f * b2;
    f.sa1 = __builtin_amdgcn_sqrtf(f.a1);
    return f;
}

// The reference recurrence in factored form on a shifted value t = y - K
// (st[0] holds the shifted mean), grouped by powers of d:
//     m3 += d (d^2 a2 - 3 b m2),   m4 += d (-4 b m3 + d (d^2 a3 - 6 b^2 m2))
// with the updated m2 and m3, as in the reference.  The c2X, c2Y terms
// (t - mean') xD use t - mean' = d (1 - b): xa = xD (1 - b) and ya = yD (1 - b)
// are formed once per sample for all sets.
template <int MM>
__device__ __forceinline__ void meng_factored(float *st, float t, const NFac &f, float xa, float ya) {
    float &mean = st[0];
    const float d = t - mean;
    mean = __builtin_fmaf(d, f.b, mean);
    if (MM == 1) return;
    const float d2 = d * d;
    float &m2 = st[1];
    m2 = __builtin_fmaf(d2, f.a1, m2);
    if (MM == 2) return;
    float &m3 = st[2];
    m3 = __builtin_fmaf(d, __builtin_fmaf(f.b3, m2, d2 * f.a2), m3);
    if (MM == 3) return;
    float &m4 = st[3];
    const float u = __builtin_fmaf(d2, f.a3, f.b6 * m2);
    m4 = __builtin_fmaf(d, __builtin_fmaf(d, u, f.b4 * m3), m4);
    if (MM == 4) return;
    st[4] = __builtin_fmaf(d, xa, st[4]);
    st[5] = __builtin_fmaf(d, ya, st[5]);
    st[6] = __builtin_fmaf(fabsf(d), f.sa1, st[6]);
}

constexpr int nstats(int MM) { return MM >= 5 ? 7 : MM; }

// statpath.cpp:313-332: NaN channel, luminance < -1e-5 or infinite -> black.
// The reference compares the float luminance with the double -1e-5; for every
// float y that is y < (float)-1e-5 (the float nearest -1e-5 lies just above
// it, and no float lies between it and -1e-5), so no FP64 is needed.
__device__ __forceinline__ bool keep_sample(float r, float g, float b) {
#pragma clang fp contract(off)
    const bool nan = __builtin_isnan(r) || __builtin_isnan(g) || __builtin_isnan(b);
    const float y = 0.212671f * r + 0.715160f * g + 0.072169f * b;
    return !(nan || y < -1e-5f || __builtin_isinf(y));
}

// The own channel is read by one wavefront only (streamed, non-temporal);
// with the guard every wavefront also reads the other two channels, which
// then stay cacheable for the neighbouring channel wavefronts.
__device__ __forceinline__ float load_sample(const char *base, int64_t off, int64_t comp, int c, bool sanitize) {
    if (!sanitize) return __builtin_nontemporal_load((const float *)(base + off + c * comp));
    const float r = *(const float *)(base + off);
    const float g = *(const float *)(base + off + comp);
    const float b = *(const float *)(base + off + 2 * comp);
    const float v = c == 0 ? r : (c == 1 ? g : b);
    return keep_sample(r, g, b) ? v : 0.f;
}

// The Film of the render (STATER_BUFSET_FILM): pbrt creates the FilmTiles
// once per render loop (src/statistics/statpath.cpp:144) and every sample
// adds L * 1 * 1 to its pixel's contribSum and 1 to its weight sum
// (FilmTile::AddSample, src/core/film.h:152-190, box filter of radius 1/2),
// so both sums run over all the pixel's samples in order, across calls.
// film_load gives the running sum a call continues from (0 for a fresh call),
// film_store writes it back with the weight sum from the R lane.
__device__ __forceinline__ float film_load(const AccArgs &a, int c, uint32_t pix) {
    return (a.film_sum && !a.fresh) ? a.film_sum[(size_t)c * a.plane + pix] : 0.f;
}
__device__ __forceinline__ void film_store(const AccArgs &a, int c, uint32_t pix, float sum, int cnt) {
    if (!a.film_sum) return;
    a.film_sum[(size_t)c * a.plane + pix] = sum;
    if (c == 0) {
        float *w = a.film_sum + 3 * a.plane + pix;
        *w = (a.fresh ? 0.f : *w) + (float)cnt;  // exact: the weights are 1 and the sum stays below 2^24
    }
}

// A fresh call leaves what stater_bufset_reset leaves in the statistic slots
// a pixel-channel does not accumulate (all of them when it takes no sample).
__device__ __forceinline__ void clear_slots(const AccArgs &a, int S, int c, int k0, uint32_t pix) {
    for (int s = 0; s < S; ++s)
        for (int k = k0; k < N_STATS; ++k) a.stats[(((size_t)s * N_STATS + k) * 3 + c) * a.plane + pix] = 0.f;
}

// One thread per pixel (sequential mode): the three channels' statistics of
// every set live in one lane, so the sanitising guard, the sample count, the
// per-sample factors of the factored recurrence and the (u, v) loads are
// done once per pixel-sample instead of once per channel, and each sample
// costs five loads per pixel instead of fifteen.
template <int T, int MM, int KC>
__device__ __forceinline__ void accp_tile(const AccArgs a, int x, int y) {
    constexpr bool GEN = KC == 0 || bc_has(KC, T, kBcGeneric, kBcGeneric);  // some set takes the log
    constexpr int SG = bc_single_generic(KC, T);  // the log-based set taking wexpm1_small, or 0
    constexpr int NS = nstats(MM);
    if (x >= a.W) return;
    uint32_t pix = (uint32_t)y * (uint32_t)a.ld + (uint32_t)x;
    int cnt = a.spp;
    if (a.counts)
        cnt = min(cnt, *(const int32_t *)((const char *)a.counts + (int64_t)y * a.counts_pitch + x * 4) -
                           a.count_offset);
    if (cnt <= 0) {
        if (a.fresh) {
            a.n[pix] = 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                clear_slots(a, T + 1, c, 0, pix);
                film_store(a, c, pix, 0.f, 0);
            }
        }
        return;
    }
    uint32_t n = a.fresh ? 0u : (uint32_t)a.n[pix];
    const char *base = (const char *)a.data + (int64_t)y * a.pitch + (int64_t)x * 4;

    float st[3][T + 1][NS];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int s = 0; s <= T; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k)
                st[c][s][k] = a.fresh ? 0.f : a.stats[(((size_t)s * N_STATS + k) * 3 + c) * a.plane + pix];

    float L[3] = {0.f, 0.f, 0.f}, Wt[3][T + 1], x1[3] = {1.f, 1.f, 1.f}, ix1[3] = {1.f, 1.f, 1.f};
    // z = l ln2 log2(x) - l L of set SG; l ln2 kept in a VGPR (an FMA reading an SGPR issues at half rate)
    float gz0[3] = {0.f, 0.f, 0.f}, gz1 = SG ? a.lam[SG] * 0.69314718055994531f : 0.f;
    if (SG) asm volatile("" : "+v"(gz1));
    if (T > 0) {
        float v0[3] = {0.f, 0.f, 0.f};
        if (n == 0) {
            v0[0] = *(const float *)base;
            v0[1] = *(const float *)(base + a.comp);
            v0[2] = *(const float *)(base + 2 * a.comp);
            if (a.sanitize && !keep_sample(v0[0], v0[1], v0[2])) v0[0] = v0[1] = v0[2] = 0.f;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t lp = (size_t)c * a.plane + pix;
            // the reference point: the pixel-channel's first sample (1 when not positive and finite)
            if (n == 0) {
                x1[c] = (v0[c] > 0.f && __builtin_isfinite(v0[c])) ? v0[c] : 1.f;
                a.lref[lp] = x1[c];
            } else {
                x1[c] = a.lref[lp];
            }
            L[c] = __builtin_amdgcn_logf(x1[c]) * 0.69314718055994531f;
            if (SG) gz0[c] = -a.lam[SG] * L[c];
#pragma unroll
            for (int s = 1; s <= T; ++s) {
                Wt[c][s] = expf(a.lam[s] * L[c]) * a.inv_lam[s];
                st[c][s][0] = n ? a.mshift[((size_t)(s - 1) * 3 + c) * a.plane + pix] : 0.f;
            }
            if (KC != 0) ix1[c] = 1.f / x1[c];
        }
    }

    float fsum[3] = {film_load(a, 0, pix), film_load(a, 1, pix), film_load(a, 2, pix)};
    // The next sample's loads are issued before this sample's arithmetic.
    auto load = [&](int j, float (&r)[5]) {
        const char *sj = base + (int64_t)j * a.sstride;
#pragma unroll
        for (int k = 0; k < (MM >= 5 ? 5 : 3); ++k) r[k] = __builtin_nontemporal_load((const float *)(sj + k * a.comp));
    };
    float nx[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    load(0, nx);
    // Sample j of the call with the per-sample factors of count nf (f, and
    // y = recip_rn(nf) for set 0's quotients).
    auto sample = [&](int j, const NFac &f, float nf, float y) {
        float v[3] = {nx[0], nx[1], nx[2]};
        float xD = 0.f, yD = 0.f;
        if (MM >= 5) {
            xD = nx[3] - 0.5f;
            yD = nx[4] - 0.5f;
        }
        if (j + 1 < cnt) load(j + 1, nx);
        if (a.sanitize && !keep_sample(v[0], v[1], v[2])) v[0] = v[1] = v[2] = 0.f;
        // The Box-Cox sets' transforms first: they need no count factors, so
        // they cover the latency of the factors' reads.
        float dl[3] = {0.f, 0.f, 0.f};
        BcPow pw[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dl[c] = GEN && !SG ? log_ratio(v[c], L[c]) : 0.f;
            pw[c] = bc_pow<T, KC>(v[c], x1[c], ix1[c]);
            // kept here, ahead of set 0's branches (the compiler would sink them past them)
            if (T > 0 && GEN && !SG) asm volatile("" : "+v"(dl[c]));
            if (T > 0 && KC != 0) asm volatile("" : "+v"(pw[c].u), "+v"(pw[c].h2), "+v"(pw[c].h4));
            if (T > 0 && bc_has(KC, T, kBcThreeQuarter, kBcThreeQuarter)) asm volatile("" : "+v"(pw[c].s4));
        }
        // Set SG's shifted values: the small-z form, or the general one on a
        // lane with a channel outside it.
        float tg[3] = {0.f, 0.f, 0.f};
        if (SG) {
            bool wide = false;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float z = __builtin_fmaf(__builtin_amdgcn_logf(v[c]), gz1, gz0[c]);
                tg[c] = wexpm1_small(z, Wt[c][SG]);
                wide |= !(fabsf(z) <= kExpm1SmallZ);
            }
            if (__builtin_expect(wide, 0)) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    tg[c] = bc_shifted<KC>(SG, a.lam[SG], log_ratio(v[c], L[c]), pw[c], Wt[c][SG], v[c], a.lam_par[SG]);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) asm volatile("" : "+v"(tg[c]));
        }
        // Set 0 takes div_rn and sqrt_rn_large; a lane with a tiny or an
        // infinite deviation (never, in practice) redoes its quotients with
        // the IEEE division before the update and its roots with sqrtf after
        // it, so that the update itself stays one basic block.
        float d[3], dN[3], p[3], r1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            d[c] = v[c] - st[c][0][0];
            dN[c] = div_rn(d[c], y, nf);
        }
        const bool tiny = exact_tiny(d[0]) | exact_tiny(d[1]) | exact_tiny(d[2]) |
                          (fmaxf(fmaxf(fabsf(d[0]), fabsf(d[1])), fabsf(d[2])) == __builtin_inff());
        if (__builtin_expect(tiny, 0)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) dN[c] = d[c] / nf;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            fsum[c] += v[c];
            p[c] = meng_exact<MM>(st[c][0], v[c], d[c], dN[c], xD, yD);
            r1[c] = MM >= 5 ? sqrt_rn_large(p[c]) : 0.f;
        }
        if (T > 0) {
            const float xa = xD * f.a1, ya = yD * f.a1;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int s = 1; s <= T; ++s)
                    meng_factored<MM>(st[c][s],
                                      s == SG ? tg[c]
                                              : bc_shifted<KC>(s, a.lam[s], dl[c], pw[c], Wt[c][s], v[c], a.lam_par[s]),
                                      f, xa, ya);
        }
        if (MM >= 5) {
            if (__builtin_expect(tiny, 0)) {
#pragma unroll
                for (int c = 0; c < 3; ++c) r1[c] = sqrtf(p[c]);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) st[c][0][6] += r1[c];
        }
    };
    // Where every lane of the wavefront starts from the same count (a fresh
    // call, or one after uniform calls), sample j has the same count in every
    // lane, and the factors of 64 consecutive counts are computed once, one
    // count per lane, into LDS; each sample then reads its own from there (a
    // broadcast read) instead of computing them: two transcendentals and a
    // dozen other instructions per sample.  The block is one wavefront, so
    // the LDS accesses are in program order; the lanes still active fill the
    // 64 slots between them (all of them at once in a full wavefront).
    __shared__ float fac[kFacSlots][64];
    auto put = [&](int k, float nk) {
        const NFac fk = nfac(nk);
        fac[0][k] = fk.b, fac[1][k] = fk.a1, fac[2][k] = fk.a2, fac[3][k] = fk.a3;
        fac[4][k] = fk.b3, fac[5][k] = fk.b4, fac[6][k] = fk.b6, fac[7][k] = fk.sa1;
        fac[8][k] = recip_rn(nk), fac[9][k] = nk;
    };
    const bool uni = __all(n == __builtin_amdgcn_readfirstlane(n));
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < cnt; ++j) {
        int k = lane;
        if (uni) {
            if ((j & 63) == 0) {  // slots i = 0 .. 63 <- count n + 1 + i, shared by the active lanes
                const uint64_t act = __ballot(1);
                const int m = __popcll(act);
                const int r = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                for (int i = r; i < 64; i += m) put(i, (float)(n + 1u + (uint32_t)i));
            }
            k = j & 63;
        } else {
            put(lane, (float)(n + 1u));  // own slot, own count
        }
        __builtin_amdgcn_wave_barrier();
        const NFac f{fac[0][k], fac[1][k], fac[2][k], fac[3][k], fac[4][k], fac[5][k], fac[6][k], fac[7][k]};
        ++n;
        sample(j, f, fac[9][k], fac[8][k]);
        __builtin_amdgcn_wave_barrier();
    }
    asm volatile("" : "+v"(pix));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int s = 1; s <= T; ++s) {
            a.mshift[((size_t)(s - 1) * 3 + c) * a.plane + pix] = st[c][s][0];
            st[c][s][0] += expm1f(a.lam[s] * L[c]) * a.inv_lam[s];  // K: shift of the reference point
        }
#pragma unroll
        for (int s = 0; s <= T; ++s)
#pragma unroll
            for (int k = 0; k < NS; ++k) a.stats[(((size_t)s * N_STATS + k) * 3 + c) * a.plane + pix] = st[c][s][k];
        if (NS < N_STATS && a.fresh) clear_slots(a, T + 1, c, NS, pix);
        film_store(a, c, pix, fsum[c], cnt);
    }
    a.n[pix] = (int32_t)n;
}

// FP64 merge of two partial moment sets (Chan, Golub & LeVeque 1979; Pebay
// 2008): a = (na, mean, m2, m3, m4) absorbs b.
__device__ __forceinline__ void merge64(double na, double *a, double nb, const double *b, int mm) {
    const double n = na + nb;
    if (nb == 0.0) return;
    if (na == 0.0) { for (int i = 0; i < 4; ++i) a[i] = b[i]; return; }
    const double d = b[0] - a[0], d2 = d * d;
    const double mean = a[0] + d * nb / n;
    double m2 = 0, m3 = 0, m4 = 0;
    if (mm >= 2) m2 = a[1] + b[1] + d2 * na * nb / n;
    if (mm >= 3) m3 = a[2] + b[2] + d * d2 * na * nb * (na - nb) / (n * n) + 3.0 * d * (na * b[1] - nb * a[1]) / n;
    if (mm >= 4)
        m4 = a[3] + b[3] + d2 * d2 * na * nb * (na * na - na * nb + nb * nb) / (n * n * n) +
             6.0 * d2 * (na * na * b[1] + nb * nb * a[1]) / (n * n) + 4.0 * d * (na * b[2] - nb * a[2]) / n;
    a[0] = mean; a[1] = m2; a[2] = m3; a[3] = m4;
}

// STATER_ACC_MERGE64: the call's samples build FP32 partial moments from
// n = 0 (factored recurrence, every set), which are merged into the resident
// moments in FP64 and rounded once; c2X, c2Y and m1, which depend on the
// running mean of the whole sequence, continue sequentially on a running
// mean carried from the resident state.
template <int T, int MM, int KC>
__device__ __forceinline__ void accm_tile(const AccArgs a, int bx, int by) {
    constexpr bool GEN = KC == 0 || bc_has(KC, T, kBcGeneric, kBcGeneric);
    constexpr int SG = bc_single_generic(KC, T);  // the log-based set taking wexpm1_small, or 0
    constexpr int NS = nstats(MM);
    constexpr int NP = MM >= 4 ? 4 : MM;  // partial moments kept
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = bx * 64 + (threadIdx.x & 63);
    const int y = by;
    if (x >= a.W) return;
    // 32-bit pixel offset on uniform plane bases (width, height <= 16384)
    uint32_t pix = (uint32_t)y * (uint32_t)a.ld + (uint32_t)x;
    int cnt = a.spp;
    if (a.counts)
        cnt = min(cnt, *(const int32_t *)((const char *)a.counts + (int64_t)y * a.counts_pitch + x * 4) -
                           a.count_offset);
    if (cnt <= 0) {
        if (a.fresh) {
            if (c == 0) a.n[pix] = 0;
            clear_slots(a, T + 1, c, 0, pix);
            film_store(a, c, pix, 0.f, 0);
        }
        return;
    }
    const uint32_t n0 = a.fresh ? 0u : (uint32_t)a.n[pix];
    const char *base = (const char *)a.data + (int64_t)y * a.pitch + (int64_t)x * 4;

    float st[T + 1][NS];
#pragma unroll
    for (int s = 0; s <= T; ++s)
#pragma unroll
        for (int k = 0; k < NS; ++k)
            st[s][k] = a.fresh ? 0.f : a.stats[(((size_t)s * N_STATS + k) * 3 + c) * a.plane + pix];
    float L = 0.f, Wt[T + 1], x1 = 1.f, ix1 = 1.f;
    if (T > 0) {
        const size_t lp = (size_t)c * a.plane + pix;
        if (n0 == 0) {
            const float v0 = load_sample(base, 0, a.comp, c, a.sanitize);
            x1 = (v0 > 0.f && __builtin_isfinite(v0)) ? v0 : 1.f;
            a.lref[lp] = x1;
        } else {
            x1 = a.lref[lp];
        }
        L = __builtin_amdgcn_logf(x1) * 0.69314718055994531f;
#pragma unroll
        for (int s = 1; s <= T; ++s) {
            Wt[s] = expf(a.lam[s] * L) * a.inv_lam[s];
            st[s][0] = n0 ? a.mshift[((size_t)(s - 1) * 3 + c) * a.plane + pix] : 0.f;
        }
        if (KC != 0) ix1 = 1.f / x1;
    }
    float pm[T + 1][NP];
    float mg[T + 1];  // running mean of the whole sequence (c2X, c2Y, m1)
#pragma unroll
    for (int s = 0; s <= T; ++s) {
#pragma unroll
        for (int k = 0; k < NP; ++k) pm[s][k] = 0.f;
        mg[s] = st[s][0];
    }

    float fsum = film_load(a, c, pix);  // the Film's running sum of this pixel-channel's samples
    for (int j = 0; j < cnt; ++j) {
        const int64_t off = (int64_t)j * a.sstride;
        const float v = load_sample(base, off, a.comp, c, a.sanitize);
        fsum += v;
        float xD = 0.f, yD = 0.f;
        if (MM >= 5) {
            xD = *(const float *)(base + off + 3 * a.comp) - 0.5f;
            yD = *(const float *)(base + off + 4 * a.comp) - 0.5f;
        }
        const NFac fr = nfac((float)(j + 1));
        NFac fg;
        if (MM >= 5) fg = nfac((float)(n0 + j + 1));
        const float dl = T > 0 && GEN ? log_ratio(v, L) : 0.f;
        const BcPow pw = bc_pow<T, KC>(v, x1, ix1);
        float tg = 0.f;  // set SG's shifted value, as in the sequential mode
        if (SG) {
            const float z = a.lam[SG] * dl;  // dl is of |v|: a negative v takes the general form's NaN
            tg = fabsf(z) <= kExpm1SmallZ && v >= 0.f ? wexpm1_small(z, Wt[SG])
                                                      : bc_shifted<KC>(SG, a.lam[SG], dl, pw, Wt[SG], v, a.lam_par[SG]);
        }
#pragma unroll
        for (int s = 0; s <= T; ++s) {
            const float t = s == 0 ? v : s == SG ? tg : bc_shifted<KC>(s, a.lam[s], dl, pw, Wt[s], v, a.lam_par[s]);
            meng_factored<NP>(pm[s], t, fr, 0.f, 0.f);
            if (MM >= 5) {
                const float dg = t - mg[s];
                mg[s] = __builtin_fmaf(dg, fg.b, mg[s]);
                const float r = t - mg[s];
                st[s][4] = __builtin_fmaf(r, xD, st[s][4]);
                st[s][5] = __builtin_fmaf(r, yD, st[s][5]);
                st[s][6] = __builtin_fmaf(fabsf(dg), fg.sa1, st[s][6]);
            }
        }
    }
    // The statistic addresses are rebuilt after the loop rather than held in
    // 2 x (T+1) x NS registers across it.
    asm volatile("" : "+v"(pix));
#pragma unroll
    for (int s = 0; s <= T; ++s) {
        double A[4] = {st[s][0], NS > 1 ? st[s][1] : 0.0, NS > 2 ? st[s][2] : 0.0, NS > 3 ? st[s][3] : 0.0};
        const double B[4] = {pm[s][0], NP > 1 ? pm[s][1] : 0.0, NP > 2 ? pm[s][2] : 0.0, NP > 3 ? pm[s][3] : 0.0};
        merge64((double)n0, A, (double)cnt, B, MM);
#pragma unroll
        for (int k = 0; k < NP; ++k) st[s][k] = (float)A[k];
    }
#pragma unroll
    for (int s = 1; s <= T; ++s) {
        a.mshift[((size_t)(s - 1) * 3 + c) * a.plane + pix] = st[s][0];
        st[s][0] += expm1f(a.lam[s] * L) * a.inv_lam[s];  // K: shift of the reference point
    }
#pragma unroll
    for (int s = 0; s <= T; ++s)
#pragma unroll
        for (int k = 0; k < NS; ++k) a.stats[(((size_t)s * N_STATS + k) * 3 + c) * a.plane + pix] = st[s][k];
    if (NS < N_STATS && a.fresh) clear_slots(a, T + 1, c, NS, pix);
    if (c == 0) a.n[pix] = (int32_t)(n0 + cnt);
    film_store(a, c, pix, fsum, cnt);
}

// At least three waves per SIMD (at most 168 VGPRs); the six-set body with
// the store addresses rebuilt after the sample loop needs about 100.
#define STATER_ACC_OCC __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(3)))

template <int T, int MM, int KC>
__global__ void __launch_bounds__(64) k_accumulate_px(AccArgs a) {
    accp_tile<T, MM, KC>(a, blockIdx.x * 64 + threadIdx.x, blockIdx.y);
}

template <int T, int MM, int KC>
__global__ void STATER_ACC_OCC k_accumulate_merge(AccArgs a) { accm_tile<T, MM, KC>(a, blockIdx.x, blockIdx.y); }

__global__ void k_zero(uint32_t *p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// The Film image after an iteration: Film::Clear and one MergeFilmTile per
// pixel (src/core/film.cpp:127-149, called at statpath.cpp:176,388) leave
// xyz = 0 + RGBToXYZ(contribSum) (src/core/spectrum.h:72-76) and the weight
// sum; Film::UpdateImage (film.cpp:188-220, no splats) then takes XYZToRGB
// (spectrum.h:66-70), times 1/weight and clamped at 0 (std::max(0, v): +0
// for -0 and NaN) when the weight is not 0, times scale.  Every operation is
// evaluated as written, uncontracted.
__global__ void __launch_bounds__(256) k_film_resolve(const float *sum, int W, int ld, size_t plane, float scale,
                                                      float *out, int64_t out_pitch, int64_t out_cs, int out_il) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const size_t pix = (size_t)y * ld + x;
    const float r = sum[pix], g = sum[plane + pix], b = sum[2 * plane + pix], w = sum[3 * plane + pix];
    const float X = 0.f + (0.412453f * r + 0.357580f * g + 0.180423f * b);
    const float Y = 0.f + (0.212671f * r + 0.715160f * g + 0.072169f * b);
    const float Z = 0.f + (0.019334f * r + 0.119193f * g + 0.950227f * b);
    float rgb[3];
    rgb[0] = 3.240479f * X - 1.537150f * Y - 0.498535f * Z;
    rgb[1] = -0.969256f * X + 1.875991f * Y + 0.041556f * Z;
    rgb[2] = 0.055648f * X - 0.204043f * Y + 1.057311f * Z;
    if (w != 0.f) {
        const float inv = 1.f / w;
        for (int c = 0; c < 3; ++c) {
            const float v = rgb[c] * inv;
            rgb[c] = 0.f < v ? v : 0.f;
        }
    }
    char *row = (char *)out + (int64_t)y * out_pitch;
    for (int c = 0; c < 3; ++c) {
        float *dst = out_il ? (float *)(row + ((int64_t)x * 3 + c) * 4) : (float *)(row + (int64_t)x * 4 + c * out_cs);
        *dst = rgb[c] * scale;
    }
}

int launch_film_resolve(stater_bufset *bs, float scale, const stater_plane *out, hipStream_t st) {
    LaunchTimer t(bs->ctx, st, "film_resolve");
    hipLaunchKernelGGL(k_film_resolve, dim3((bs->W + 255) / 256, bs->H), dim3(256), 0, st, bs->film_sum, bs->W,
                       bs->ld, bs->plane, scale, (float *)out->ptr, out->pitch_bytes, out->channel_stride_bytes,
                       out->layout == STATER_INTERLEAVED ? 1 : 0);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

int launch_zero(void *p, size_t bytes, hipStream_t st) {
    const size_t words = bytes / 4;
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 65536);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, st, (uint32_t *)p, words);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

template <int T, int MM, int KC = 0>
static void launch_tm(const AccArgs &a, bool merge, dim3 grid, hipStream_t st) {
    if (merge) hipLaunchKernelGGL((k_accumulate_merge<T, MM, KC>), grid, dim3(192), 0, st, a);
    else hipLaunchKernelGGL((k_accumulate_px<T, MM, KC>), grid, dim3(64), 0, st, a);
}

// Kind codes of the reference's multi-transform scene configurations
// (scenes/render-denoise-mt{1..5}.pbrt: lambdas [.5], [.5 .75], [.25 1],
// [.25 .5 .75 1], [.01 .25 .5 .75 1 3]; a set per lambda other than 1, in
// list order), built with max_moment 5; every other list runs the generic form.
constexpr int kKcMt1 = kBcHalf;
constexpr int kKcMt2 = kBcHalf | kBcThreeQuarter << 3;
constexpr int kKcMt3 = kBcQuarter;
constexpr int kKcMt4 = kBcQuarter | kBcHalf << 3 | kBcThreeQuarter << 6;
constexpr int kKcMt5 = kBcGeneric | kBcQuarter << 3 | kBcHalf << 6 | kBcThreeQuarter << 9 | kBcCube << 12;

template <int T>
static void launch_t(int mm, int kc, const AccArgs &a, bool merge, dim3 grid, hipStream_t st) {
    switch (mm) {
        case 1: launch_tm<T, 1>(a, merge, grid, st); break;
        case 2: launch_tm<T, 2>(a, merge, grid, st); break;
        case 3: launch_tm<T, 3>(a, merge, grid, st); break;
        case 4: launch_tm<T, 4>(a, merge, grid, st); break;
        default:
            if (T == 1 && kc == kKcMt1) launch_tm<T, 5, (T == 1 ? kKcMt1 : 0)>(a, merge, grid, st);
            else if (T == 1 && kc == kKcMt3) launch_tm<T, 5, (T == 1 ? kKcMt3 : 0)>(a, merge, grid, st);
            else if (T == 2 && kc == kKcMt2) launch_tm<T, 5, (T == 2 ? kKcMt2 : 0)>(a, merge, grid, st);
            else if (T == 3 && kc == kKcMt4) launch_tm<T, 5, (T == 3 ? kKcMt4 : 0)>(a, merge, grid, st);
            else if (T == 5 && kc == kKcMt5) launch_tm<T, 5, (T == 5 ? kKcMt5 : 0)>(a, merge, grid, st);
            else launch_tm<T, 5>(a, merge, grid, st);
            break;
    }
}

// The kind code of a set's lambdas (sets 1 ..).
static int bc_code(const stater_bufset *bs) {
    int kc = 0;
    for (int s = 1; s < bs->S; ++s) {
        const float l = bs->set_lambda[s];
        const int k = l == 0.25f ? kBcQuarter : l == 0.5f ? kBcHalf : l == 0.75f ? kBcThreeQuarter
                                                                               : l == 3.f ? kBcCube : kBcGeneric;
        kc |= k << (3 * (s - 1));
    }
    return kc;
}

AccArgs make_acc_args(stater_bufset *bs, const stater_samples *s) {
    AccArgs a{};
    a.data = s->data;
    a.pitch = s->pitch_bytes;
    a.comp = s->comp_stride_bytes;
    a.sstride = s->sample_stride_bytes;
    a.spp = s->spp;
    a.sanitize = s->sanitize;
    a.counts = s->counts;
    a.counts_pitch = s->counts_pitch_bytes;
    a.count_offset = s->count_offset;
    a.n = bs->n;
    a.stats = bs->stats;
    a.lref = bs->lref;
    a.mshift = bs->mshift;
    a.ld = bs->ld;
    a.plane = bs->plane;
    a.W = bs->W;
    a.H = bs->H;
    for (int k = 0; k < kMaxSets; ++k) {
        a.lam[k] = k < bs->S ? bs->set_lambda[k] : 1.f;
        a.inv_lam[k] = 1.f / a.lam[k];
        const float l = a.lam[k];
        a.lam_par[k] = l != std::floor(l) || std::fabs(l) > 16777216.f ? 0 : (std::fmod(l, 2.f) == 0.f ? 1 : 2);
    }
    a.film_sum = bs->film_sum;
    return a;
}

int launch_accumulate(stater_bufset *bs, const stater_samples *s, int mode, hipStream_t st) {
    if (s->pitch_bytes % 4 || s->comp_stride_bytes % 4 || s->sample_stride_bytes % 4 ||
        s->pitch_bytes < (int64_t)bs->W * 4)
        return fail(STATER_E_ARG, "stater_accumulate: strides must be float multiples covering a row");
    AccArgs a = make_acc_args(bs, s);
    a.fresh = (mode & STATER_ACC_FRESH) ? 1 : 0;
    dim3 grid((bs->W + 63) / 64, bs->H);
    const bool merge = (mode & ~STATER_ACC_FRESH) == STATER_ACC_MERGE64;
    LaunchTimer t(bs->ctx, st, merge ? "accumulate_merge64" : "accumulate");
    const int kc = bc_code(bs);
    switch (bs->S - 1) {
        case 0: launch_t<0>(bs->mm, kc, a, merge, grid, st); break;
        case 1: launch_t<1>(bs->mm, kc, a, merge, grid, st); break;
        case 2: launch_t<2>(bs->mm, kc, a, merge, grid, st); break;
        case 3: launch_t<3>(bs->mm, kc, a, merge, grid, st); break;
        case 4: launch_t<4>(bs->mm, kc, a, merge, grid, st); break;
        default: launch_t<5>(bs->mm, kc, a, merge, grid, st); break;
    }
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

}  // namespace stater

// ---------------------------------------------------------------- testing
namespace stater {

// Checks of set 0's fast exact forms against the IEEE ones, one float bit
// pattern per thread over all 2^32: counts[0] counts the non-negative x
// outside (0, 2^-104), -0 and NaN included, where sqrt_rn_large(x) and
// sqrtf(x) differ (NaN equal to NaN; a radicand d (d - d/n) is never
// negative: the hardware root of a negative subnormal is -0, not NaN),
// counts[1] the finite d without exact_tiny(d) where div_rn, with recip_rn of
// a pseudo-random n in [1, 2^24], differs from d / n, and counts[2] the
// floats nf in [1, 2^32] (the first 2^28 + 1 threads) where recip_rn(nf)
// differs from 1.f / nf.
__global__ void __launch_bounds__(256) k_exact_forms(uint64_t base, unsigned long long *counts) {
    const uint64_t g = base + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = (uint32_t)g;
    const float x = __uint_as_float(i);
    auto same = [](float a, float b) { return __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b); };
    const bool in_range = x == 0.f || x >= 0x1p-104f || x != x;
    if (in_range && !same(sqrt_rn_large(x), sqrtf(x))) atomicAdd(&counts[0], 1ull);
    uint32_t h = i * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    const float nf = (float)(1u + (h & 0xFFFFFFu));
    if (!exact_tiny(x) && fabsf(x) != __builtin_inff() && !same(div_rn(x, recip_rn(nf), nf), x / nf))
        atomicAdd(&counts[1], 1ull);
    if (g < ((uint64_t)1 << 28)) {
        const float m = __uint_as_float(0x3f800000u + i);  // 1 <= m < 2^32
        if (!same(recip_rn(m), 1.f / m)) atomicAdd(&counts[2], 1ull);
        if (g == 0 && !same(recip_rn(0x1p32f), 1.f / 0x1p32f)) atomicAdd(&counts[2], 1ull);
    }
}

// div_rn(d, recip_rn(m), m) against the IEEE d / m for every significand
// d in [1, 2) (2^23 values, thread t takes t, t + 256, ...) and every
// integer m in [m0, m0 + gridDim.x): counts[0] counts the differences.
__global__ void __launch_bounds__(256) k_quotient_sweep(uint32_t m0, unsigned long long *counts) {
    const float m = (float)(m0 + blockIdx.x);
    const float y = recip_rn(m);
    uint32_t bad = 0;
    for (uint32_t j = threadIdx.x; j < (1u << 23); j += 256) {
        const float d = __uint_as_float(0x3f800000u | j);
        bad += __float_as_uint(div_rn(d, y, m)) != __float_as_uint(d / m);
    }
    if (bad) atomicAdd(&counts[0], (unsigned long long)bad);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

}  // namespace stater

extern "C" int stater_debug_exact_forms(stater_ctx *ctx, unsigned long long *counts, void *stream) {
    using namespace stater;
    if (!ctx || !counts) return fail(STATER_E_ARG, "stater_debug_exact_forms: NULL argument");
    STATER_HIP(hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    k_zero_u64<<<1, 64, 0, st>>>(counts, 3);
    const uint64_t per = (uint64_t)1 << 28;  // 2^20 blocks of 256 threads per launch, 16 launches
    for (uint64_t b = 0; b < ((uint64_t)1 << 32); b += per) k_exact_forms<<<(unsigned)(per / 256), 256, 0, st>>>(b, counts);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

extern "C" int stater_debug_quotient_sweep(stater_ctx *ctx, uint32_t m_lo, uint32_t m_hi, unsigned long long *counts,
                                           void *stream) {
    using namespace stater;
    if (!ctx || !counts) return fail(STATER_E_ARG, "stater_debug_quotient_sweep: NULL argument");
    if (m_lo < 1 || m_hi <= m_lo || m_hi > (1u << 24) + 1)
        return fail(STATER_E_ARG, "stater_debug_quotient_sweep: need 1 <= m_lo < m_hi <= 2^24 + 1");
    STATER_HIP(hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    k_zero_u64<<<1, 64, 0, st>>>(counts, 1);
    const uint32_t per = 4096;  // 2^35 pairs per launch
    for (uint32_t m = m_lo; m < m_hi; m += per) k_quotient_sweep<<<min(per, m_hi - m), 256, 0, st>>>(m, counts);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

namespace stater {

// VALU issue probe: 16 independent FP32 FMA chains per lane, a <- a a + a
// (one v_fmac with register operands only: an SGPR or literal operand, a
// compare or a select issue at a lower rate), so a launch of many waves
// measures the rate at which the SIMDs take non-packed FP32 instructions.
// Starting in (-1, 0) every chain decays towards 0 (no inf or NaN).
__global__ void __launch_bounds__(256) k_valu_rate(float *out, int iters, float c1, float c2) {
    float a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = c1 + c2 * (float)i - 1e-6f * (float)threadIdx.x;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) a[i] = __builtin_fmaf(a[i], a[i], a[i]);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += a[i];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace stater

extern "C" int stater_debug_valu_rate(stater_ctx *ctx, int blocks, int iters, float *out, void *stream) {
    using namespace stater;
    if (!ctx || !out || blocks <= 0 || iters <= 0) return fail(STATER_E_ARG, "stater_debug_valu_rate: bad argument");
    STATER_HIP(hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    k_valu_rate<<<blocks, 256, 0, st>>>(out, iters, -0.5f, 1e-3f);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}
