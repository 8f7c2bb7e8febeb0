// tdist.hpp — Student-t tail used by the Welch test, host and device.
//
// The two-sided tail of t^2 with nu degrees of freedom is the regularised
// incomplete beta I_x(nu/2, 1/2) with x = nu/(nu + t^2).  It is evaluated by
// the modified Lentz continued fraction (Numerical Recipes, 3rd ed., 6.4) in
// FP64.  The Welch kernel calls it only for pairs whose FP32 statistic lies in
// a narrow band around the critical value; everywhere else it compares t^2
// against a polynomial in u = 1/nu that the host fits to this function.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace stater {

__host__ __device__ inline double betacf(double a, double b, double x) {
    const double EPS = 1e-16, FPMIN = 1e-300;
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (fabs(d) < FPMIN) d = FPMIN;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m <= 100000; ++m) {
        const double m2 = 2.0 * m;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
        c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d; h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
        c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < EPS) break;
    }
    return h;
}

__host__ __device__ inline double betainc_reg(double a, double b, double x) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    const double lbt = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x);
    if (x < (a + 1.0) / (a + b + 2.0)) return exp(lbt) * betacf(a, b, x) / a;
    return 1.0 - exp(lbt) * betacf(b, a, 1.0 - x) / b;
}

// P(|T_nu| >= sqrt(t2)); nu = +inf gives the normal tail.
__host__ __device__ inline double t2_tail(double t2, double nu) {
    if (!(nu < 1e300)) return erfc(sqrt(0.5 * t2));
    return betainc_reg(0.5 * nu, 0.5, nu / (nu + t2));
}

// Degree of the critical-value fit.
constexpr int kTCritDeg = 4;

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
    double alpha;
};

// Host: fit the polynomial for a level (Chebyshev interpolation on umax = 1/16,
// i.e. nu >= 16; smaller nu takes the exact tail).
TCritPoly make_tcrit_poly(double alpha);
double t2_crit_host(double nu, double alpha);

}  // namespace stater
