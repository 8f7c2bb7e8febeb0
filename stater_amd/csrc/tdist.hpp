// tdist.hpp — Student-t tail used by the Welch test, host and device.
//
// The two-sided tail of t^2 with nu degrees of freedom is the regularised
// incomplete beta I_x(nu/2, 1/2) with x = nu/(nu + t^2).  It is evaluated by
// its continued fraction (DLMF 8.17.22, modified Lentz evaluation) in FP64.
// The Welch kernel calls it only for pairs whose FP32 statistic lies in a
// narrow band around the critical value; everywhere else it compares t^2
// against a polynomial in u = 1/nu that the host fits to this function.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace stater {

// The continued fraction of DLMF 8.17.22,
//     I_x(a, b) = x^a (1 - x)^b / (a B(a, b)) / (1 + d_1/(1 + d_2/(1 + ...))),
//     d_{2m}   =  m (b - m) x / ((a + 2m - 1)(a + 2m)),
//     d_{2m+1} = -(a + m)(a + b + m) x / ((a + 2m)(a + 2m + 1)),
// returned as 1/(1 + d_1/(1 + ...)).  The denominator is evaluated front to
// back, one partial numerator d_j per step, by the modified Lentz method
// (Thompson & Barnett, J. Comput. Phys. 64 (1986) 490): with C_0 = f_0 = 1,
// D_0 = 0, C_j = 1 + d_j/C_{j-1}, D_j = 1/(1 + d_j D_{j-1}) and
// f_j = f_{j-1} C_j D_j; a vanishing C_j or D_j denominator is replaced by
// a tiny number.  It converges fast for x < (a + 1)/(a + b + 2).
__host__ __device__ inline double dlmf_beta_cf(double a, double b, double x) {
    const double tiny = 1e-300, tol = 1e-16;
    double f = 1.0, C = 1.0, D = 0.0;
    for (int j = 1; j <= 200000; ++j) {
        const double m = (double)(j >> 1);
        const double dj = (j & 1) ? -(a + m) * (a + b + m) * x / ((a + 2.0 * m) * (a + 2.0 * m + 1.0))
                                  : m * (b - m) * x / ((a + 2.0 * m - 1.0) * (a + 2.0 * m));
        D = 1.0 + dj * D;
        if (fabs(D) < tiny) D = tiny;
        D = 1.0 / D;
        C = 1.0 + dj / C;
        if (fabs(C) < tiny) C = tiny;
        const double delta = C * D;
        f *= delta;
        if (fabs(delta - 1.0) < tol) break;
    }
    return 1.0 / f;
}

// I_x(a, b) from the fraction, using the symmetry I_x(a, b) = 1 - I_{1-x}(b, a)
// where the direct fraction would converge slowly.
__host__ __device__ inline double betainc_reg(double a, double b, double x) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    const double front = exp(lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x));
    if (x < (a + 1.0) / (a + b + 2.0)) return front * dlmf_beta_cf(a, b, x) / a;
    return 1.0 - front * dlmf_beta_cf(b, a, 1.0 - x) / b;
}

// P(|T_nu| >= sqrt(t2)); nu = +inf gives the normal tail.
__host__ __device__ inline double t2_tail(double t2, double nu) {
    if (!(nu < 1e300)) return erfc(sqrt(0.5 * t2));
    return betainc_reg(0.5 * nu, 0.5, nu / (nu + t2));
}

// Degree of the critical-value fit on u <= 1/16, and of the short fit on
// u <= 1/128 that windows whose every n - 1 is at least 128 use.
constexpr int kTCritDeg = 4;
constexpr int kTCritDegShort = 2;
constexpr double kTCritUmax = 1.0 / 16.0;
constexpr double kTCritUmaxShort = 1.0 / 128.0;

// Critical t^2 of the two-sided level-alpha test as a polynomial in u = 1/nu
// on [0, umax], i.e. for nu >= 1/umax.  A fast test t2 > crit is trusted
// outside t2 in [(1 - band) crit, (1 + band) crit]; inside, and for u > umax,
// the decision takes the exact tail.
struct TCritPoly {
    float c[kTCritDeg + 1];  // monomial coefficients in u, c[0] = z^2
    float umax;
    float band;              // relative half-width of the exact-tail band
    float neg_hi, neg_lo;    // -(1 + band), -(1 - band)
    float band_abs;          // band x the largest crit on [0, umax]: t2 - crit beyond +-band_abs is decided
    float cs[kTCritDegShort + 1];  // the short fit on [0, kTCritUmaxShort]
    float band_abs_short;          // its band_abs
    double alpha;
};

// Host: fit the polynomials for a level (Chebyshev interpolation on
// [0, 1/16], i.e. nu >= 16, where smaller nu takes the exact tail, and on
// [0, 1/128]).
TCritPoly make_tcrit_poly(double alpha);
double t2_crit_host(double nu, double alpha);

}  // namespace stater
