// exact.hpp — correctly rounded FP32 quotients and square roots in fewer
// instructions than the IEEE expansions, on the input ranges where they are
// exact (the accumulation's reference update, the Welch filter's discr).
#pragma once

#include <hip/hip_runtime.h>

namespace stater {

// 1/nf correctly rounded, for every float nf in [1, 2^32]: one Newton step
// y0 + y0 (1 - nf y0) from the hardware reciprocal (checked against the IEEE
// 1.f / nf on all 2^28 + 1 such floats by stater_debug_exact_forms).
__device__ __forceinline__ float recip_rn(float nf) {
    const float y0 = __builtin_amdgcn_rcpf(nf);
    return __builtin_fmaf(__builtin_fmaf(-nf, y0, 1.f), y0, y0);
}

// d / n correctly rounded, from y = recip_rn(nf), by Markstein's correction
// step: q = d y is close to the quotient, the residual n q - d is exact in
// the FMA, and q - (n q - d) y rounds to the IEEE quotient.  Three FP32
// instructions where a correctly rounded FP64 product took three at half
// rate (two conversions and a multiply).  Exact for every d with
// 2^-40 <= |d| < inf (and d = 0 or NaN) and every float nf in [1, 2^32]:
// the steps scale exactly with powers of two of d and of nf on that range
// (no subnormal intermediate; recip_rn(2^k m) = 2^-k recip_rn(m), both being
// correctly rounded), so the pairs of a significand of d in [1, 2) and an
// integer m in [2^23, 2^24] stand for all of them, and
// stater_debug_quotient_sweep checks those 2^46 pairs against the IEEE
// division.  d = +-inf would give NaN, and a tiny d can have a subnormal
// quotient that falls on a midpoint exactly (d = 3 x 2^-149, n = 2): callers
// send both to the IEEE division.
__device__ __forceinline__ float div_rn(float d, float y, float nf) {
    const float q = d * y;
    // q - (n q - d) y: the residual taken as n q - d keeps the sign of a zero
    // d (d - n q would turn -0 / n into +0)
    return __builtin_fmaf(-__builtin_fmaf(q, nf, -d), y, q);
}

// The correctly rounded square root of x >= 2^-104 (and of +-0, inf, NaN): the
// hardware root s is within an ulp, and the residuals x - s'' s of its two
// neighbours s'' decide which of the three is the rounded root (the
// expansion the compiler emits for sqrtf, without its scaling of tiny
// inputs and its 0/inf select, neither needed on this range).  Checked
// against sqrtf on all 2^32 floats: among the non-negative ones it differs
// only for 0 < x < 2^-104 (stater_debug_exact_forms).
__device__ __forceinline__ float sqrt_rn_large(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = s;
    if (__builtin_fmaf(-sd, s, x) <= 0.f) r = sd;
    if (__builtin_fmaf(-su, s, x) > 0.f) r = su;
    return r;
}

// A deviation d = x - mean with 0 < |d| < 2^-40 (absurdly small against any
// radiance) takes the IEEE division and sqrtf: everywhere else the quotient
// d / n is a normal float or zero, and the radicand d (d - d/n) of m1 is 0
// (n = 1) or at least 2^-81, where div_rn and sqrt_rn_large are exact.
__device__ __forceinline__ bool exact_tiny(float d) { return d != 0.f && fabsf(d) < 0x1p-40f; }

// x / n correctly rounded for any x and an integer-valued float n >= 1 with
// y = recip_rn(nf): div_rn, except where the quotient could be subnormal
// (0 < |x| < 2^-100) or x is infinite, which take the IEEE division.
__device__ __forceinline__ float div_exact_by_n(float x, float y, float nf) {
    const float ax = fabsf(x);
    if (__builtin_expect((x != 0.f && ax < 0x1p-100f) || ax == __builtin_inff(), 0)) return x / nf;
    return div_rn(x, y, nf);
}

}  // namespace stater
