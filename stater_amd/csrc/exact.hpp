// exact.hpp — correctly rounded FP32 quotients and square roots in fewer
// instructions than the IEEE expansions, on the input ranges where they are
// exact (the accumulation's reference update, the Welch filter's discr).
#pragma once

#include <hip/hip_runtime.h>

namespace stater {

// The quotient d / n, correctly rounded, from an FP64 reciprocal rn of n
// (within 2^-52 relative): RN32(d * rn) is RN32(d / n) wherever the quotient
// is a normal float, because d / n lies at least 2^-49 (relative) away from
// every midpoint of two adjacent floats there: d carries 24 significant bits
// and n is an integer of at most 24 (times a power of two), so d - M n, for
// a midpoint M with 25 significant bits, is a nonzero multiple of half of
// M's last bit.  A subnormal quotient can fall on a midpoint exactly
// (d = 3 x 2^-149, n = 2); it needs |d| < n 2^-126, which exact_tiny covers.
__device__ __forceinline__ float div_by_n(float d, double rn) { return (float)((double)d * rn); }

// 1/nf in FP64 from the FP32 reciprocal b = rcp(nf) (within a few ulp): two
// Newton steps take its relative error e to e^2 and then below 2^-52.
__device__ __forceinline__ double recip64(float nf, float b) {
    const double n = (double)nf, y0 = (double)b;
    const double y1 = __builtin_fma(y0, __builtin_fma(-n, y0, 1.0), y0);
    return __builtin_fma(y1, __builtin_fma(-n, y1, 1.0), y1);
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
// (n = 1) or at least 2^-81, where div_by_n and sqrt_rn_large are exact.
__device__ __forceinline__ bool exact_tiny(float d) { return d != 0.f && fabsf(d) < 0x1p-40f; }

// x / n correctly rounded for any x and an integer-valued float n >= 1 with
// FP64 reciprocal rn: div_by_n, except where the quotient could be subnormal
// (0 < |x| < 2^-100), which takes the IEEE division.
__device__ __forceinline__ float div_exact_by_n(float x, double rn, float nf) {
    if (__builtin_expect(x != 0.f && fabsf(x) < 0x1p-100f, 0)) return x / nf;
    return div_by_n(x, rn);
}

}  // namespace stater
