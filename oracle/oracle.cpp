// oracle.cpp — CPU oracle for the statistics-and-denoise path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (stater_amd/, include/)
// links, loads or calls this file; only tests/, the smoke check and the
// cpu_baseline leg of bench.py do, as the checker and as the timed CPU port.
//
// What it restates, with the StatER file:line each part follows:
//   * Meng's online central-moment update, in the reference's exact FP32
//     operation order:  src/statistics/estimator.h:226-305
//     (AddStatSampleM1/M2/M3/M4/M4C), the shared n++ of
//     AddSample/AddTransformSample estimator.h:306-324, and the Box-Cox
//     transform boxCox(val, l) estimator.h:43-76 (lambda == 1 is the identity).
//   * The per-sample guard of statpath.cpp:313-332 (NaN channel, luminance
//     below -1e-5 or infinite -> the sample is counted as black) with the
//     RGBSpectrum luminance weights of src/core/spectrum.h:483-486.
//   * 16x16 image tiles processed by a thread pool (statpath.cpp:82-90,178;
//     src/core/parallel.cpp:243-300) for the CPU baseline.
//   * calcMatMeanVar: meanVar = m2/((n-1)n)  (estimator.cpp:462-474).
// The denoiser kernels of the reference (stat_denoiser.cu) are not part of the
// reference checkout, so the filter, Bonett, Jarque-Bera, LEV and AS functions
// below restate the published methods the build documents (Welch 1947;
// Jarque & Bera 1980; Bonett 2006; OLS on the in-pixel position; Chan, Golub &
// LeVeque 1979 / Pebay 2008 for merging), evaluated in FP64 with exact
// incomplete-beta tails (the continued fraction of DLMF 8.17.22).
//
// Pinning (tests/test_oracle.py): Philox4x32-10 against the Random123
// known-answer vectors; the FP64 recurrence against exact two-pass moments;
// closed forms (constant input, two-sample m2/m3/m4); the incomplete beta,
// Student-t quantiles, Welch p-values and Jarque-Bera against scipy fixtures
// committed under tests/golden/; the filters against brute force and
// invariants.  Parity unpinned: c2X/c2Y/m1 beyond the reference recurrence
// itself (they are defined only by it), and every denoiser-side choice that
// the absent stat_denoiser.cu would fix.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "../include/stater.h"

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                   uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t c[4] = {c0, c1, c2, c3};
    uint32_t k[2] = {k0, k1};
    for (int i = 0; i < 10; ++i) {
        if (i) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

inline double u32_unit(uint32_t r) { return (double)r * (1.0 / 4294967296.0); }
inline double u24_unit(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }
inline double u24_open(uint32_t r) { return (double)((r >> 8) + 1u) * (1.0 / 16777216.0); }

const double kPi = 3.14159265358979323846;

// ------------------------------------------------------- synthetic scene
void scene_init(stater_scene *s, int model, uint32_t seed, uint32_t image_id, int W, int H) {
    std::memset(s, 0, sizeof(*s));
    s->width = W; s->height = H; s->seed = seed; s->image_id = image_id; s->model = model;
    s->base = 0.05;
    uint32_t r[4];
    for (int k = 0; k < 3; ++k) {
        philox(0xA0000000u + k, 0x5CE7E000u, 0, 0, seed, image_id, r);
        s->fx[k] = 0.5 + 3.5 * u32_unit(r[0]);
        s->fy[k] = 0.5 + 3.5 * u32_unit(r[1]);
        philox(0xA0000010u + k, 0x5CE7E000u, 0, 0, seed, image_id, r);
        for (int c = 0; c < 3; ++c) s->amp[k][c] = 0.6 * u32_unit(r[c]);
        philox(0xA0000020u + k, 0x5CE7E000u, 0, 0, seed, image_id, r);
        for (int c = 0; c < 3; ++c) {
            const double phi = 2.0 * kPi * u32_unit(r[c]);
            s->cos_phi[k][c] = std::cos(phi);
            s->sin_phi[k][c] = std::sin(phi);
        }
    }
    s->firefly_mult = 100.0;
    s->firefly_thresh = 4294967u;  // floor(1e-3 * 2^32)
    s->cell_size = 32;
    s->normal_jitter = 0.02;
    s->depth_min = 1.0; s->depth_max = 100.0;
    if (model == 1) {
        s->n_edges = 1;
        s->edge_vertical[0] = 1;
        s->edge_pos[0] = 0.5 * W;
        s->edge_factor[0] = 4.0;
        s->s_min = s->s_max = 0.8;
        s->albedo_modulation = 0;
    } else {
        s->n_edges = 12;
        for (int e = 0; e < 12; ++e) {
            philox(0xA0000100u + e, 0x5CE7E000u, 0, 0, seed, image_id, r);
            s->edge_vertical[e] = (int)(r[0] & 1u);
            s->edge_pos[e] = u32_unit(r[1]) * (s->edge_vertical[e] ? W : H);
            s->edge_factor[e] = std::pow(4.0, 2.0 * u32_unit(r[2]) - 1.0);
        }
        s->s_min = 0.3; s->s_max = 1.2;
        s->albedo_modulation = 1;
    }
}

// Features of the region [x0, x0+w) x [y0, y0+h) of the scene, stored
// planar over the region (P = w*h).
void features(const stater_scene *s, int x0, int y0, int w, int h, float *albedo, float *normal, float *depth) {
    const int W = s->width, P = w * h;
    const int cz = s->cell_size;
    for (int y = y0; y < y0 + h; ++y)
        for (int x = x0; x < x0 + w; ++x) {
            const int gx = x / cz, gy = y / cz;
            const double px = x + 0.5, py = y + 0.5;
            double best = 1e300; int bi = 0, bj = 0;
            uint32_t r[4];
            for (int j = gy - 1; j <= gy + 1; ++j)
                for (int i = gx - 1; i <= gx + 1; ++i) {
                    philox((uint32_t)i, (uint32_t)j, 0xCE11u, 0, s->seed, s->image_id, r);
                    const double cx = cz * (i + u32_unit(r[0]));
                    const double cy = cz * (j + u32_unit(r[1]));
                    const double d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
                    if (d2 < best) { best = d2; bi = i; bj = j; }
                }
            const int p = (y - y0) * w + (x - x0);
            philox((uint32_t)bi, (uint32_t)bj, 0xCE11u, 1, s->seed, s->image_id, r);
            for (int c = 0; c < 3; ++c) albedo[c * P + p] = (float)u32_unit(r[c]);
            depth[p] = (float)(s->depth_min + (s->depth_max - s->depth_min) * u32_unit(r[3]));
            philox((uint32_t)bi, (uint32_t)bj, 0xCE11u, 2, s->seed, s->image_id, r);
            const double cth = u32_unit(r[0]);
            const double ph = 2.0 * kPi * u32_unit(r[1]);
            const double sth = std::sqrt(1.0 - cth * cth);
            double n[3] = {sth * std::cos(ph), sth * std::sin(ph), cth};
            philox((uint32_t)(y * W + x), 0xFFFFFFFEu, 0, 0, s->seed, s->image_id, r);
            const double ra = std::sqrt(-2.0 * std::log(u24_open(r[0])));
            const double ta = 2.0 * kPi * u24_unit(r[1]);
            const double rb = std::sqrt(-2.0 * std::log(u24_open(r[2])));
            const double tb = 2.0 * kPi * u24_unit(r[3]);
            n[0] += s->normal_jitter * ra * std::cos(ta);
            n[1] += s->normal_jitter * ra * std::sin(ta);
            n[2] += s->normal_jitter * rb * std::cos(tb);
            const double inv = 1.0 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int c = 0; c < 3; ++c) normal[c * P + p] = (float)(n[c] * inv);
        }
}

inline double pixel_s(const stater_scene *s, uint32_t p) {
    if (s->s_min == s->s_max) return s->s_min;
    uint32_t r[4];
    philox(p, 0xFFFFFFFFu, 0, 0, s->seed, s->image_id, r);
    return s->s_min + (s->s_max - s->s_min) * u32_unit(r[0]);
}

// One sample of pixel (x, y): rgb + in-pixel position.
// albedo: the pixel's three albedo values (stride astride) or NULL.
inline void sample(const stater_scene *s, const float *albedo, size_t astride, int x, int y, uint32_t k,
                   double sp, float out[5]) {
    const int W = s->width;
    const uint32_t p = (uint32_t)(y * W + x);
    uint32_t r[4];
    philox(p, k, 0, 0, s->seed, s->image_id, r);
    const double u = u24_unit(r[0]), v = u24_unit(r[1]);
    const double z = std::sqrt(-2.0 * std::log(u24_open(r[2]))) * std::cos(2.0 * kPi * u24_unit(r[3]));
    const uint32_t ffw = (r[0] & 255u) | ((r[1] & 255u) << 8) | ((r[2] & 255u) << 16) | ((r[3] & 255u) << 24);
    const double F = ffw < s->firefly_thresh ? s->firefly_mult : 1.0;
    const double L = std::exp(sp * z - 0.5 * sp * sp);
    const double X = x + u, Y = y + v;
    double st[3], ct[3];
    for (int k2 = 0; k2 < 3; ++k2) {
        const double th = 2.0 * kPi * (s->fx[k2] * X / W + s->fy[k2] * Y / s->height);
        st[k2] = std::sin(th); ct[k2] = std::cos(th);
    }
    double edge = 1.0;
    for (int e = 0; e < s->n_edges; ++e) {
        const double c = s->edge_vertical[e] ? X : Y;
        if (c >= s->edge_pos[e]) edge *= s->edge_factor[e];
    }
    for (int c = 0; c < 3; ++c) {
        double mu = s->base;
        for (int k2 = 0; k2 < 3; ++k2)
            mu += s->amp[k2][c] * 0.5 * (1.0 + st[k2] * s->cos_phi[k2][c] + ct[k2] * s->sin_phi[k2][c]);
        mu *= edge;
        if (s->albedo_modulation) mu *= 0.2 + 0.8 * (double)albedo[c * astride];
        out[c] = (float)(mu * L * F);
    }
    out[3] = (float)u; out[4] = (float)v;
}

// ------------------------------------------------------------ thread pool
template <class F>
void parallel_tiles(int W, int H, int nthreads, F &&fn) {
    const int ts = 16;
    const int ntx = (W + ts - 1) / ts, nty = (H + ts - 1) / ts, nt = ntx * nty;
    if (nthreads <= 1) {
        for (int t = 0; t < nt; ++t) {
            const int tx = t % ntx, ty = t / ntx;
            fn(tx * ts, std::min(W, tx * ts + ts), ty * ts, std::min(H, ty * ts + ts));
        }
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int i = 0; i < nthreads; ++i)
        pool.emplace_back([&] {
            for (;;) {
                const int t = next.fetch_add(1);
                if (t >= nt) break;
                const int tx = t % ntx, ty = t / ntx;
                fn(tx * ts, std::min(W, tx * ts + ts), ty * ts, std::min(H, ty * ts + ts));
            }
        });
    for (auto &th : pool) th.join();
}

template <class F>
void parallel_rows(int H, int nthreads, F &&fn) {
    if (nthreads <= 1) { for (int y = 0; y < H; ++y) fn(y); return; }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int i = 0; i < nthreads; ++i)
        pool.emplace_back([&] {
            for (;;) { const int y = next.fetch_add(1); if (y >= H) break; fn(y); }
        });
    for (auto &th : pool) th.join();
}

// --------------------------------------------------------- accumulation
enum { ST_MEAN = 0, ST_M2, ST_M3, ST_M4, ST_C2X, ST_C2Y, ST_M1, N_STATS };

// boxCox(val, l) of estimator.h:44-57 in the reference's FP32 form.
inline float boxcox_ref(float x, float l) {
    if (l == 1.f) return x;
    return (std::pow(x, l) - 1.f) / l;
}
// The same transform evaluated in FP64 (the mathematical value): expm1 of
// l ln x for x >= 0, and pow for a negative x, which is real for an integer
// lambda only (NaN otherwise), as std::pow in the reference.
inline double boxcox64(double x, double l) {
    if (l == 1.0) return x;
    if (x < 0.0) return (std::pow(x, l) - 1.0) / l;
    return std::expm1(l * std::log(x)) / l;
}

// Meng's update in the operation order of estimator.h:226-305; T = float
// reproduces the reference bit for bit when the compiler neither contracts
// nor reorders (built with -ffp-contract=off, no fast-math).
template <typename T>
inline void meng(int mm, uint64_t n, T *st, size_t cs, T sample, float u, float v) {
    T &mean = st[ST_MEAN * cs];
    if (mm == 1) {
        const T d = sample - mean;
        const T dN = d / (T)n;
        mean += dN;
        return;
    }
    T &m2 = st[ST_M2 * cs];
    if (mm == 2) {
        const T d = sample - mean;
        const T dN = d / (T)n;
        mean += dN;
        m2 += d * (d - dN);
        return;
    }
    T &m3 = st[ST_M3 * cs];
    const T d = sample - mean;
    const T d2 = d * d;
    const T dN = d / (T)n;
    const T dN2 = dN * dN;
    mean += dN;
    m2 += d * (d - dN);
    m3 += -(T)3 * dN * m2 + d * (d2 - dN2);
    if (mm == 3) return;
    T &m4 = st[ST_M4 * cs];
    m4 += -(T)4 * dN * m3 - (T)6 * dN2 * m2 + d * (d * d2 - dN * dN2);
    if (mm == 4) return;
    const float xD = (u - 0.5f);
    const float yD = (v - 0.5f);
    st[ST_C2X * cs] += (sample - mean) * (T)xD;
    st[ST_C2Y * cs] += (sample - mean) * (T)yD;
    st[ST_M1 * cs] += std::sqrt(d * (d - dN));
}

// statpath.cpp:313-332 (spectrum.h:483-486 luminance).
inline void sanitize(float &r, float &g, float &b) {
    const bool nan = std::isnan(r) || std::isnan(g) || std::isnan(b);
    const float y = 0.212671f * r + 0.715160f * g + 0.072169f * b;
    if (nan || (double)y < -1e-5 || std::isinf(y)) r = g = b = 0.f;
}

template <typename T, bool REF_BOXCOX>
void accumulate(int W, int H, int S, const float *lambdas, int mm, const float *samples, int spp,
                const int32_t *counts, int san, int32_t *n, T *stats, int nthreads) {
    const size_t P = (size_t)W * H;
    const size_t cs = 3 * P;  // stride between statistics of one set
    parallel_tiles(W, H, nthreads, [&](int x0, int x1, int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                const size_t p = (size_t)y * W + x;
                const int cnt = counts ? std::min(counts[p], spp) : spp;
                uint64_t nn = (uint64_t)n[p];
                for (int j = 0; j < cnt; ++j) {
                    const float *sj = samples + (size_t)j * 5 * P + p;
                    float rgb[3] = {sj[0], sj[P], sj[2 * P]};
                    const float u = sj[3 * P], v = sj[4 * P];
                    if (san) sanitize(rgb[0], rgb[1], rgb[2]);
                    ++nn;
                    for (int s = 0; s < S; ++s) {
                        T *st = stats + (size_t)s * N_STATS * cs;
                        for (int c = 0; c < 3; ++c) {
                            T val;
                            if (REF_BOXCOX) val = (T)boxcox_ref(rgb[c], lambdas[s]);
                            else val = (T)boxcox64((double)rgb[c], (double)lambdas[s]);
                            meng<T>(mm, nn, st + c * P + p, cs, val, u, v);
                        }
                    }
                }
                n[p] = (int32_t)nn;
            }
    });
}

// ---------------------------------------------------- special functions
// Regularised incomplete beta I_x(a, b) by its continued fraction (DLMF
// 8.17.22): I_x(a, b) = x^a (1-x)^b / (a B(a, b)) / (1 + d_1/(1 + d_2/(1 + ...)))
// with d_{2m} = m(b-m)x/((a+2m-1)(a+2m)) and
// d_{2m+1} = -(a+m)(a+b+m)x/((a+2m)(a+2m+1)).  The denominator K is summed
// front to back by the modified Lentz method (Thompson & Barnett 1986),
// f_j = f_{j-1} C_j D_j, C_j = 1 + d_j/C_{j-1}, D_j = 1/(1 + d_j D_{j-1});
// for x beyond the mean (a+1)/(a+b+2) the symmetry I_x(a,b) = 1 - I_{1-x}(b,a)
// keeps the fraction in its fast-converging range.
double cf_denominator(double a, double b, double x) {
    const double tiny = 1e-300;
    double f = 1.0, C = 1.0, D = 0.0;
    for (int j = 1; j <= 200000; ++j) {
        const double m = (double)(j / 2);
        const double dj = (j % 2) ? -(a + m) * (a + b + m) * x / ((a + 2 * m) * (a + 2 * m + 1))
                                  : m * (b - m) * x / ((a + 2 * m - 1) * (a + 2 * m));
        D = 1.0 + dj * D;
        if (std::fabs(D) < tiny) D = tiny;
        D = 1.0 / D;
        C = 1.0 + dj / C;
        if (std::fabs(C) < tiny) C = tiny;
        f *= C * D;
        if (std::fabs(C * D - 1.0) < 1e-16) break;
    }
    return f;
}

double betainc(double a, double b, double x) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    const double front = std::exp(std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) +
                                  b * std::log1p(-x));
    if (x < (a + 1.0) / (a + b + 2.0)) return front / (a * cf_denominator(a, b, x));
    return 1.0 - front / (b * cf_denominator(b, a, 1.0 - x));
}

// Two-sided Student-t tail P(|T_nu| >= t) for t^2; nu = inf -> normal.
double t2_pvalue(double t2, double nu) {
    if (!(nu < 1e300)) return std::erfc(std::sqrt(0.5 * t2));
    return betainc(0.5 * nu, 0.5, nu / (nu + t2));
}

double t2_crit(double nu, double alpha) {
    if (!(nu < 1e300)) {
        double lo = 0, hi = 100;
        for (int i = 0; i < 200; ++i) {
            const double mid = 0.5 * (lo + hi);
            if (std::erfc(std::sqrt(0.5 * mid)) > alpha) lo = mid; else hi = mid;
        }
        return 0.5 * (lo + hi);
    }
    // I_x(nu/2, 1/2) increases with x = nu/(nu+t2); find I_x = alpha.
    double lo = 0.0, hi = 1.0;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (betainc(0.5 * nu, 0.5, mid) < alpha) lo = mid; else hi = mid;
    }
    const double x = 0.5 * (lo + hi);
    return nu * (1.0 - x) / x;
}

// Bracketing table of exact critical values: t2crit(nu) falls monotonically in
// nu, so for nu in [nu_i, nu_{i+1}] a statistic below t2crit(nu_{i+1}) is
// accepted and one above t2crit(nu_i) rejected without evaluating the tail;
// only statistics between the two bracket values take the exact tail.
struct CritTable {
    double alpha = -1, lnu0 = 0, dl = 0;
    std::vector<double> t2;  // t2crit at nu_i = exp(lnu0 + i*dl)
    double t2inf = 0;
};

const CritTable &crit_table(double alpha) {
    static CritTable tab;
    if (tab.alpha == alpha) return tab;
    const int N = 4001;
    tab.alpha = alpha;
    tab.lnu0 = std::log(0.5);
    tab.dl = (std::log(1e7) - tab.lnu0) / (N - 1);
    tab.t2.resize(N);
    for (int i = 0; i < N; ++i) tab.t2[i] = t2_crit(std::exp(tab.lnu0 + i * tab.dl), alpha);
    tab.t2inf = t2_crit(INFINITY, alpha);
    return tab;
}

// Welch acceptance from per-pixel mean m, v = var/n, w = v^2/(n-1).  A test
// whose operands are not all finite abstains (accepts): a set whose FP32
// statistics overflowed carries no information about the pair.
inline bool welch_accept(const CritTable &tab, double mp, double vp, double wp, double mq, double vq, double wq,
                         double alpha) {
    if (!(std::isfinite(mp) && std::isfinite(mq) && std::isfinite(vp) && std::isfinite(vq))) return true;
    const double V = vp + vq, dm = mp - mq;
    if (!(V > 0.0)) return dm == 0.0;
    const double t2 = dm * dm / V;
    const double Wd = wp + wq;
    const double nu = Wd > 0.0 ? V * V / Wd : INFINITY;
    const double fi = (std::log(nu) - tab.lnu0) / tab.dl;
    if (nu < 1e300 && fi >= 0.0 && fi < (double)tab.t2.size() - 1) {
        const size_t i = (size_t)fi;
        if (t2 < tab.t2[i + 1]) return true;   // below the smallest critical value of the bracket
        if (t2 > tab.t2[i]) return false;      // above the largest
    } else if (!(nu < 1e300) || fi >= (double)tab.t2.size() - 1) {
        if (t2 < tab.t2inf) return true;
        if (fi >= 0.0 && t2 > tab.t2.back()) return false;
    }
    return t2_pvalue(t2, nu) > alpha;
}

struct Feat {
    int n;
    const float *const *planes;
    const int *channels;
    const double *dr;
    size_t P;
};

inline double feature_logw(const Feat &f, size_t p, size_t q) {
    double acc = 0.0;
    for (int k = 0; k < f.n; ++k) {
        double d2 = 0.0;
        for (int c = 0; c < f.channels[k]; ++c) {
            const double d = (double)f.planes[k][c * f.P + p] - (double)f.planes[k][c * f.P + q];
            d2 += d * d;
        }
        acc += f.dr[k] * d2;
    }
    return acc;
}

}  // namespace

// ====================================================================== API
ORC_API void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

ORC_API int orc_scene_init(stater_scene *s, int model, uint32_t seed, uint32_t image_id, int W, int H) {
    if (model != 1 && model != 2) return -1;
    scene_init(s, model, seed, image_id, W, H);
    return 0;
}

ORC_API void orc_features(const stater_scene *s, int x0, int y0, int w, int h, float *albedo, float *normal,
                          float *depth) {
    features(s, x0, y0, w, h, albedo, normal, depth);
}

// Samples of the region [x0, x0+w) x [y0, y0+h): out[j][k][p] with
// p = (y-y0)*w + (x-x0), k in R,G,B,u,v, sample index first + j.  albedo is
// the region's planar albedo (3 x w*h) or NULL.
ORC_API void orc_samples(const stater_scene *s, const float *albedo, int x0, int y0, int w, int h, uint32_t first,
                         int spp, float *out, int nthreads) {
    const size_t P = (size_t)w * h;
    parallel_rows(h, nthreads, [&](int yy) {
        const int y = y0 + yy;
        for (int xx = 0; xx < w; ++xx) {
            const int x = x0 + xx;
            const size_t p = (size_t)yy * w + xx;
            const double sp = pixel_s(s, (uint32_t)(y * s->width + x));
            for (int j = 0; j < spp; ++j) {
                float v[5];
                sample(s, albedo ? albedo + p : nullptr, P, x, y, first + (uint32_t)j, sp, v);
                for (int k = 0; k < 5; ++k) out[((size_t)j * 5 + k) * P + p] = v[k];
            }
        }
    });
}

// Reference FP32 accumulation.  stats: [S][7][3][P] float, n: [P] int32.
// lambdas[s] is the Box-Cox lambda of set s (1 = untransformed).
ORC_API void orc_accumulate_f32(int W, int H, int S, const float *lambdas, int mm, const float *samples,
                                int spp, const int32_t *counts, int san, int32_t *n, float *stats,
                                int nthreads) {
    accumulate<float, true>(W, H, S, lambdas, mm, samples, spp, counts, san, n, stats, nthreads);
}

// ---- Film: the image denoiseFilm filters --------------------------------
// pbrt-v3's Film with the box filter of radius 1/2, as the render loop drives
// it.  The FilmTiles are created once per render loop
// (src/statistics/statpath.cpp:144), so a pixel's contribSum and weight sum
// run over all its samples, across iterations: FilmTile::AddSample
// (src/core/film.h:152-190) adds L * 1 * 1 and 1 in sample order.  Each
// iteration then runs Film::Clear (src/core/film.cpp:127-134, statpath.cpp:176)
// and merges every tile once (Film::MergeFilmTile, film.cpp:136-149,
// statpath.cpp:388): the Film pixel becomes 0 + RGBToXYZ(contribSum)
// (src/core/spectrum.h:72-76) and 0 + the weight sum.  Samples exactly on a
// pixel's left or top edge count in their own pixel only (the box filter
// would add them to the neighbour across that edge too).
// sums: [4][P] = contribSum R, G, B and the weight sum, continued by each call.
ORC_API void orc_film_add(int W, int H, const float *samples, int spp, const int32_t *counts, int san,
                          float *sums) {
    const size_t P = (size_t)W * H;
    for (size_t p = 0; p < P; ++p) {
        const int cnt = counts ? std::min(counts[p], spp) : spp;
        for (int j = 0; j < cnt; ++j) {
            const float *sj = samples + (size_t)j * 5 * P + p;
            float rgb[3] = {sj[0], sj[P], sj[2 * P]};
            if (san) sanitize(rgb[0], rgb[1], rgb[2]);
            for (int c = 0; c < 3; ++c) sums[(size_t)c * P + p] += rgb[c] * 1.f * 1.f;
            sums[3 * P + p] += 1.f;
        }
    }
}

// The Film's pixels after Film::Clear and one MergeFilmTile per pixel:
// xyzw = (0 + X, 0 + Y, 0 + Z, 0 + weight) from the running sums.
ORC_API void orc_film_xyz(int P, const float *sums, float *xyzw) {
    for (int p = 0; p < P; ++p) {
        const float r = sums[p], g = sums[(size_t)P + p], b = sums[2 * (size_t)P + p];
        const float xyz[3] = {0.412453f * r + 0.357580f * g + 0.180423f * b,
                              0.212671f * r + 0.715160f * g + 0.072169f * b,
                              0.019334f * r + 0.119193f * g + 0.950227f * b};
        for (int i = 0; i < 3; ++i) xyzw[(size_t)i * P + p] = 0.f + xyz[i];
        xyzw[3 * (size_t)P + p] = 0.f + sums[3 * (size_t)P + p];
    }
}

// Film::UpdateImage (src/core/film.cpp:188-220) without splats: XYZToRGB
// (src/core/spectrum.h:66-70), times 1/weight and std::max(0, v) when the
// weight sum is not 0, times scale.  rgb: [3][P].
ORC_API void orc_film_resolve(int P, const float *xyzw, float scale, float *rgb) {
    for (int p = 0; p < P; ++p) {
        const float X = xyzw[p], Y = xyzw[P + p], Z = xyzw[2 * (size_t)P + p], w = xyzw[3 * (size_t)P + p];
        float o[3];
        o[0] = 3.240479f * X - 1.537150f * Y - 0.498535f * Z;
        o[1] = -0.969256f * X + 1.875991f * Y + 0.041556f * Z;
        o[2] = 0.055648f * X - 0.204043f * Y + 1.057311f * Z;
        if (w != 0.f) {
            const float inv = 1.f / w;
            for (int c = 0; c < 3; ++c) o[c] = std::max(0.f, o[c] * inv);
        }
        for (int c = 0; c < 3; ++c) rgb[(size_t)c * P + p] = o[c] * scale;
    }
}

// The same recurrence in FP64 with the FP64 Box-Cox.
ORC_API void orc_accumulate_f64(int W, int H, int S, const float *lambdas, int mm, const float *samples,
                                int spp, const int32_t *counts, int san, int32_t *n, double *stats,
                                int nthreads) {
    accumulate<double, false>(W, H, S, lambdas, mm, samples, spp, counts, san, n, stats, nthreads);
}

// Exact two-pass central moments (FP64): out[S][4][3][P] = mean, m2, m3, m4
// as sums of powered deviations.
ORC_API void orc_moments_exact(int W, int H, int S, const float *lambdas, const float *samples, int spp,
                               const int32_t *counts, int san, double *out) {
    const size_t P = (size_t)W * H;
    std::vector<double> vals;
    for (size_t p = 0; p < P; ++p) {
        const int cnt = counts ? std::min(counts[p], spp) : spp;
        for (int s = 0; s < S; ++s)
            for (int c = 0; c < 3; ++c) {
                vals.assign(cnt, 0.0);
                for (int j = 0; j < cnt; ++j) {
                    const float *sj = samples + (size_t)j * 5 * P + p;
                    float rgb[3] = {sj[0], sj[P], sj[2 * P]};
                    if (san) sanitize(rgb[0], rgb[1], rgb[2]);
                    vals[j] = boxcox64((double)rgb[c], (double)lambdas[s]);
                }
                double mean = 0.0;
                for (double v : vals) mean += v;
                mean = cnt ? mean / cnt : 0.0;
                double m2 = 0, m3 = 0, m4 = 0;
                for (double v : vals) {
                    const double d = v - mean;
                    m2 += d * d; m3 += d * d * d; m4 += d * d * d * d;
                }
                double *o = out + ((size_t)s * 4 * 3 + c) * P + p;
                o[0] = mean; o[3 * P] = m2; o[6 * P] = m3; o[9 * P] = m4;
            }
    }
}

// Pebay / Chan-Golub-LeVeque merge of partial moments (FP64), one channel.
// a and b hold {n, mean, m2, m3, m4}; result into a.
ORC_API void orc_merge_moments(double *a, const double *b) {
    const double na = a[0], nb = b[0], n = na + nb;
    if (nb == 0) return;
    if (na == 0) { for (int i = 0; i < 5; ++i) a[i] = b[i]; return; }
    const double d = b[1] - a[1], d2 = d * d;
    const double mean = a[1] + d * nb / n;
    const double m2 = a[2] + b[2] + d2 * na * nb / n;
    const double m3 = a[3] + b[3] + d * d2 * na * nb * (na - nb) / (n * n) + 3.0 * d * (na * b[2] - nb * a[2]) / n;
    const double m4 = a[4] + b[4] + d2 * d2 * na * nb * (na * na - na * nb + nb * nb) / (n * n * n) +
                      6.0 * d2 * (na * na * b[2] + nb * nb * a[2]) / (n * n) + 4.0 * d * (na * b[3] - nb * a[3]) / n;
    a[0] = n; a[1] = mean; a[2] = m2; a[3] = m3; a[4] = m4;
}

// meanVar = m2 / ((n-1) n), estimator.cpp:462-474, FP32 as the reference.
ORC_API void orc_mean_var(int P, const int32_t *n, const float *m2, float *out) {
    for (int c = 0; c < 3; ++c)
        for (int p = 0; p < P; ++p) {
            const float nf = (float)n[p];
            out[(size_t)c * P + p] = m2[(size_t)c * P + p] / ((nf - 1.f) * nf);
        }
}

ORC_API double orc_betainc(double a, double b, double x) { return betainc(a, b, x); }
ORC_API double orc_t2_pvalue(double t2, double nu) { return t2_pvalue(t2, nu); }
ORC_API double orc_t2_crit(double nu, double alpha) { return t2_crit(nu, alpha); }

// Per-pixel Jarque-Bera (mean over RGB), Bonett lnS2/se2 and LEV m2C in FP64
// from FP32 statistics.  stats: [S][7][3][P] float; outputs [S][P] / [S][3][P].
// Pixels with n < 4 are degenerate (all outputs 0).  A set-channel whose m2
// or m4 is not finite (FP32 overflow, e.g. lambda = 3 on fireflies) gets
// lnS2 = se2 = jb = +inf so that its variance tests abstain.
template <typename T>
void prepass_impl(int P, int S, const int32_t *n, const T *stats, double z, double *jb, double *lnS2, double *se2,
                  double *m2C) {
    const size_t cs = 3 * (size_t)P;
    for (int s = 0; s < S; ++s) {
        const T *st = stats + (size_t)s * N_STATS * cs;
        for (int p = 0; p < P; ++p) {
            const double nn = (double)n[p];
            double jbs = 0.0;
            for (int c = 0; c < 3; ++c) {
                const size_t i = (size_t)c * P + p;
                const size_t o = (size_t)s * cs + i;
                const double m2 = st[ST_M2 * cs + i], m3 = st[ST_M3 * cs + i], m4 = st[ST_M4 * cs + i];
                const double cx = st[ST_C2X * cs + i], cy = st[ST_C2Y * cs + i];
                if (n[p] < 4) { lnS2[o] = 0; se2[o] = 0; m2C[o] = 0; continue; }
                if (!(std::isfinite(m2) && std::isfinite(m4))) {
                    lnS2[o] = INFINITY; se2[o] = INFINITY; jbs = INFINITY;
                    m2C[o] = m2;
                    continue;
                }
                const double s2 = m2 / (nn - 1.0);
                const double cb = nn / (nn - z);
                if (m2 > 0) {
                    const double mu2 = m2 / nn;
                    const double S3 = (m3 / nn) / std::pow(mu2, 1.5);
                    const double K = (m4 / nn) / (mu2 * mu2);
                    jbs += nn / 6.0 * (S3 * S3 + 0.25 * (K - 3.0) * (K - 3.0));
                    lnS2[o] = std::log(std::max(cb * s2, 1e-30));
                    se2[o] = cb * cb * (K - (nn - 3.0) / nn) / (nn - 1.0);
                    m2C[o] = std::max(0.0, m2 - 12.0 * (cx * cx + cy * cy) / nn);
                } else {
                    lnS2[o] = std::log(1e-30);
                    se2[o] = 0.0;
                    m2C[o] = 0.0;
                }
            }
            jb[(size_t)s * P + p] = n[p] < 4 ? 0.0 : jbs / 3.0;
        }
    }
}

ORC_API void orc_prepass(int P, int S, const int32_t *n, const float *stats, double z, double *jb,
                         double *lnS2, double *se2, double *m2C) {
    prepass_impl<float>(P, S, n, stats, z, jb, lnS2, se2, m2C);
}

// The same from FP64 statistics (the full-FP64 pipeline).
ORC_API void orc_prepass64(int P, int S, const int32_t *n, const double *stats, double z, double *jb,
                           double *lnS2, double *se2, double *m2C) {
    prepass_impl<double>(P, S, n, stats, z, jb, lnS2, se2, m2C);
}

// Variance filter (FP64 from FP32 inputs): per set s and channel c,
// varD = sum_q w acc s2_q / sum_q w acc, varCD likewise with m2C_q/(n_q-1);
// acc = Bonett equality test (lnS2_p - lnS2_q)^2 <= z^2 (se2_p + se2_q).
// m2, lnS2, se2, m2C: [S][3][P] float.  Outputs [S][3][P] double and, if
// given, the accepted-neighbour count [S][3][P].
template <typename T>
void denoise_vars_impl(int W, int H, int S, const int32_t *n, const T *m2, const T *lnS2, const T *se2,
                       const T *m2C, int radius, double dsF, double z, int ng, const float *const *gplanes,
                       const int *gch, const double *gdr, double *varD, double *varCD, int32_t *count,
                       int32_t *ambiguous, int nthreads) {
    const size_t P = (size_t)W * H;
    Feat f{ng, gplanes, gch, gdr, P};
    const double z2 = z * z;
    parallel_rows(H, nthreads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            const size_t p = (size_t)y * W + x;
            for (int s = 0; s < S; ++s)
                for (int c = 0; c < 3; ++c) {
                    const size_t ip = ((size_t)s * 3 + c) * P;
                    double num = 0, numC = 0, den = 0;
                    int cnt = 0, amb = 0;
                    for (int dy = -radius; dy <= radius; ++dy) {
                        const int yq = y + dy;
                        if (yq < 0 || yq >= H) continue;
                        for (int dx = -radius; dx <= radius; ++dx) {
                            const int xq = x + dx;
                            if (xq < 0 || xq >= W) continue;
                            const size_t q = (size_t)yq * W + xq;
                            bool acc;
                            if (q == p) acc = true;
                            else if (n[p] < 4 || n[q] < 4) acc = false;
                            else {
                                // NaN/inf operands abstain (accept), see orc_prepass.
                                const double dl = (double)lnS2[ip + p] - (double)lnS2[ip + q];
                                const double rhs = z2 * ((double)se2[ip + p] + (double)se2[ip + q]);
                                acc = !(dl * dl > rhs);
                                // within FP32 rounding of the threshold: the decision is not pinned
                                if (std::fabs(dl * dl - rhs) <= 1e-5 * rhs) ++amb;
                            }
                            if (!acc) continue;
                            ++cnt;
                            const double w = std::exp(dsF * (dx * dx + dy * dy) + feature_logw(f, p, q));
                            const double nq1 = n[q] >= 2 ? (double)n[q] - 1.0 : 1.0;
                            const double s2 = n[q] >= 2 ? (double)m2[ip + q] / nq1 : 0.0;
                            const double sc = n[q] >= 2 ? (double)m2C[ip + q] / nq1 : 0.0;
                            num += w * s2; numC += w * sc; den += w;
                        }
                    }
                    varD[ip + p] = den > 0 ? num / den : 0.0;
                    varCD[ip + p] = den > 0 ? numC / den : 0.0;
                    if (count) count[ip + p] = cnt;
                    if (ambiguous) ambiguous[ip + p] = amb;
                }
        }
    });
}


ORC_API void orc_denoise_vars(int W, int H, int S, const int32_t *n, const float *m2, const float *lnS2,
                              const float *se2, const float *m2C, int radius, double dsF, double z,
                              int ng, const float *const *gplanes, const int *gch, const double *gdr,
                              double *varD, double *varCD, int32_t *count, int32_t *ambiguous, int nthreads) {
    denoise_vars_impl<float>(W, H, S, n, m2, lnS2, se2, m2C, radius, dsF, z, ng, gplanes, gch, gdr, varD, varCD,
                             count, ambiguous, nthreads);
}

// The same on FP64 inputs (the full-FP64 pipeline).
ORC_API void orc_denoise_vars64(int W, int H, int S, const int32_t *n, const double *m2, const double *lnS2,
                                const double *se2, const double *m2C, int radius, double dsF, double z,
                                int ng, const float *const *gplanes, const int *gch, const double *gdr,
                                double *varD, double *varCD, int32_t *count, int32_t *ambiguous, int nthreads) {
    denoise_vars_impl<double>(W, H, S, n, m2, lnS2, se2, m2C, radius, dsF, z, ng, gplanes, gch, gdr, varD, varCD,
                              count, ambiguous, nthreads);
}

// The variance filter's accept/reject decisions alone (the same rule as
// denoise_vars_impl): dec[(((s*3 + c)*P + p)*win + dy+r)*win + dx+r], 1 =
// accepted; neighbours outside the image are 0.
template <typename T>
void var_decisions_impl(int W, int H, int S, const int32_t *n, const T *lnS2, const T *se2, int radius, double z,
                        uint8_t *dec) {
    const size_t P = (size_t)W * H;
    const int win = 2 * radius + 1;
    const double z2 = z * z;
    for (int s = 0; s < S; ++s)
        for (int c = 0; c < 3; ++c) {
            const size_t ip = ((size_t)s * 3 + c) * P;
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    const size_t p = (size_t)y * W + x;
                    uint8_t *o = dec + (ip + p) * win * win;
                    for (int dy = -radius; dy <= radius; ++dy)
                        for (int dx = -radius; dx <= radius; ++dx) {
                            const int yq = y + dy, xq = x + dx;
                            bool acc = false;
                            if (yq >= 0 && yq < H && xq >= 0 && xq < W) {
                                const size_t q = (size_t)yq * W + xq;
                                if (q == p) acc = true;
                                else if (n[p] >= 4 && n[q] >= 4) {
                                    if (std::is_same<T, float>::value) {
                                        // the FP32 path's own arithmetic: Z = z^2 se2 rounded, reject iff
                                        // fma(dl, dl, -Z_q) > Z_p (a NaN operand abstains: accept)
                                        const float z2f = (float)z * (float)z;
                                        const float Zp = z2f * (float)se2[ip + p], Zq = z2f * (float)se2[ip + q];
                                        const float dl = (float)lnS2[ip + p] - (float)lnS2[ip + q];
                                        acc = !(std::fma(dl, dl, -Zq) > Zp);
                                    } else {
                                        const double dl = (double)lnS2[ip + p] - (double)lnS2[ip + q];
                                        acc = !(dl * dl > z2 * ((double)se2[ip + p] + (double)se2[ip + q]));
                                    }
                                }
                            }
                            o[(dy + radius) * win + dx + radius] = acc ? 1 : 0;
                        }
                }
        }
}

ORC_API void orc_var_decisions(int W, int H, int S, const int32_t *n, const float *lnS2, const float *se2,
                               int radius, double z, uint8_t *dec) {
    var_decisions_impl<float>(W, H, S, n, lnS2, se2, radius, z, dec);
}

ORC_API void orc_var_decisions64(int W, int H, int S, const int32_t *n, const double *lnS2, const double *se2,
                                 int radius, double z, uint8_t *dec) {
    var_decisions_impl<double>(W, H, S, n, lnS2, se2, radius, z, dec);
}

// Welch-test mean filter (FP64 decisions with exact tails).
//   mean: [S][3][P] per-set means (set 0: the film or raw mean the test uses)
//   var : [S][3][P] variance fed to the test (v = var/n, w = v^2/(n-1))
//   jb  : [S][P] (mode 3) or NULL
//   film: [3][P] values that are averaged.
// Outputs filmD [3][P]; varW [3][P] (mean of accepted var of set 0) if given;
// decisions [P][(2r+1)^2][3] (uint8) if given; count [3][P] if given.
template <typename T>
void denoise_film_impl(int W, int H, int S, const int32_t *n, const T *mean, const T *var, const T *jb, int radius,
                       double dsF, double alpha, int mt_mode, int ng, const float *const *gplanes, const int *gch,
                       const double *gdr, const T *film, double *filmD, double *varW, uint8_t *decisions,
                       int32_t *count, int nthreads) {
    const size_t P = (size_t)W * H;
    const int win = 2 * radius + 1;
    Feat f{ng, gplanes, gch, gdr, P};
    const CritTable &tab = crit_table(alpha);
    parallel_rows(H, nthreads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            const size_t p = (size_t)y * W + x;
            int s_only = -1;
            if (mt_mode == 3 && jb) {
                s_only = 0;
                for (int s = 1; s < S; ++s)
                    if (jb[(size_t)s * P + p] < jb[(size_t)s_only * P + p]) s_only = s;
            }
            for (int c = 0; c < 3; ++c) {
                double num = 0, numV = 0, den = 0;
                int cnt = 0;
                for (int dy = -radius; dy <= radius; ++dy)
                    for (int dx = -radius; dx <= radius; ++dx) {
                        const int yq = y + dy, xq = x + dx;
                        const int k = (dy + radius) * win + (dx + radius);
                        bool acc = false;
                        if (yq >= 0 && yq < H && xq >= 0 && xq < W) {
                            const size_t q = (size_t)yq * W + xq;
                            if (q == p) acc = true;
                            else if (n[p] >= 2 && n[q] >= 2) {
                                acc = true;
                                for (int s = 0; s < S && acc; ++s) {
                                    if (s_only >= 0 && s != s_only) continue;
                                    const size_t i = ((size_t)s * 3 + c) * P;
                                    // v = var/n in the input's precision: for FP32 inputs the
                                    // per-pixel discr value.
                                    const double np = n[p], nq = n[q];
                                    const double vp = (double)(var[i + p] / (T)n[p]);
                                    const double vq = (double)(var[i + q] / (T)n[q]);
                                    acc = welch_accept(tab, mean[i + p], vp, vp * vp / (np - 1.0), mean[i + q], vq,
                                                       vq * vq / (nq - 1.0), alpha);
                                }
                            }
                            if (acc) {
                                ++cnt;
                                const double w = std::exp(dsF * (dx * dx + dy * dy) + feature_logw(f, p, q));
                                num += w * film[(size_t)c * P + q];
                                if (varW) numV += w * var[(size_t)c * P + q];
                                den += w;
                            }
                        }
                        if (decisions) decisions[(p * win * win + k) * 3 + c] = acc ? 1 : 0;
                    }
                filmD[(size_t)c * P + p] = num / den;
                if (varW) varW[(size_t)c * P + p] = numV / den;
                if (count) count[(size_t)c * P + p] = cnt;
            }
        }
    });
}

ORC_API void orc_denoise_film(int W, int H, int S, const int32_t *n, const float *mean, const float *var,
                              const float *jb, int radius, double dsF, double alpha, int mt_mode, int ng,
                              const float *const *gplanes, const int *gch, const double *gdr,
                              const float *film, double *filmD, double *varW, uint8_t *decisions,
                              int32_t *count, int nthreads) {
    denoise_film_impl<float>(W, H, S, n, mean, var, jb, radius, dsF, alpha, mt_mode, ng, gplanes, gch, gdr, film,
                             filmD, varW, decisions, count, nthreads);
}

// The same on FP64 inputs (the full-FP64 pipeline).
ORC_API void orc_denoise_film64(int W, int H, int S, const int32_t *n, const double *mean, const double *var,
                                const double *jb, int radius, double dsF, double alpha, int mt_mode, int ng,
                                const float *const *gplanes, const int *gch, const double *gdr,
                                const double *film, double *filmD, double *varW, uint8_t *decisions,
                                int32_t *count, int nthreads) {
    denoise_film_impl<double>(W, H, S, n, mean, var, jb, radius, dsF, alpha, mt_mode, ng, gplanes, gch, gdr, film,
                              filmD, varW, decisions, count, nthreads);
}

// pbrt's Film image (the same Clear / merge / UpdateImage steps as
// orc_film_add, orc_film_xyz and orc_film_resolve) evaluated in FP64, the
// matrices taken at their FP32 values: the film of the full-FP64 pipeline.
ORC_API void orc_film_image64(int W, int H, const float *samples, int spp, const int32_t *counts, int san,
                              double *rgb) {
    const size_t P = (size_t)W * H;
    const double M[3][3] = {{0.412453f, 0.357580f, 0.180423f}, {0.212671f, 0.715160f, 0.072169f},
                            {0.019334f, 0.119193f, 0.950227f}};
    const double Mi[3][3] = {{3.240479f, -1.537150f, -0.498535f}, {-0.969256f, 1.875991f, 0.041556f},
                             {0.055648f, -0.204043f, 1.057311f}};
    for (size_t p = 0; p < P; ++p) {
        const int cnt = counts ? std::min(counts[p], spp) : spp;
        double sum[3] = {0, 0, 0};
        for (int j = 0; j < cnt; ++j) {
            const float *sj = samples + (size_t)j * 5 * P + p;
            float v[3] = {sj[0], sj[P], sj[2 * P]};
            if (san) sanitize(v[0], v[1], v[2]);
            for (int c = 0; c < 3; ++c) sum[c] += v[c];
        }
        double xyz[3], o[3];
        for (int i = 0; i < 3; ++i) xyz[i] = M[i][0] * sum[0] + M[i][1] * sum[1] + M[i][2] * sum[2];
        for (int i = 0; i < 3; ++i) o[i] = Mi[i][0] * xyz[0] + Mi[i][1] * xyz[1] + Mi[i][2] * xyz[2];
        for (int c = 0; c < 3; ++c) rgb[(size_t)c * P + p] = cnt > 0 ? std::max(0.0, o[c] / cnt) : o[c];
    }
}

// Adaptive-sampling map (integer exact).  e_p = mean_c sqrt(varD_c/n)/max(|mean_c|,1e-3)
// in FP32, q_p = floor(e_p 2^16) as integer, water-filled with caps
// n + target <= cap so that the targets sum to budget*P, remainders handed
// out in pixel order.  If the pixels with q > 0 cannot take the budget
// (their headroom sums to <= budget*P), they take all of it and the rest is
// shared uniformly (weight 1) over the q = 0 pixels; with every q = 0 that
// is the uniform distribution.
ORC_API void orc_target_spp(int P, const int32_t *n, const float *varD, const float *mean, uint64_t budget_spp,
                            uint32_t cap, int32_t *target) {
    std::vector<uint64_t> q(P);
    std::vector<int64_t> h(P);
    int64_t sh = 0, shq = 0;
    for (int p = 0; p < P; ++p) {
        float e = 0.f;
        if (n[p] > 0) {
            const float nf = (float)n[p];
            float ec[3];
            for (int c = 0; c < 3; ++c) {
                const float vd = std::max(varD[(size_t)c * P + p], 0.f);
                const float m = std::max(std::fabs(mean[(size_t)c * P + p]), 1e-3f);
                ec[c] = std::sqrt(vd / nf) / m;
            }
            e = (ec[0] + ec[1] + ec[2]) / 3.f;
        }
        const float qe = e * 65536.f;
        q[p] = !(qe > 0.f) ? 0ull : (qe >= 4294967295.f ? 4294967295ull : (uint64_t)qe);
        h[p] = std::max<int64_t>(0, (int64_t)cap - (int64_t)n[p]);
        sh += h[p];
        if (q[p]) shq += h[p];
    }
    const uint64_t B = budget_spp * (uint64_t)P;
    if ((uint64_t)sh <= B) { for (int p = 0; p < P; ++p) target[p] = (int32_t)h[p]; return; }
    if (B == 0) { for (int p = 0; p < P; ++p) target[p] = 0; return; }
    const bool uniform = (uint64_t)shq <= B;
    auto share = [&](double lam, int p) -> int64_t {
        if (uniform && q[p]) return h[p];
        const double t = std::floor(lam * (double)(uniform ? 1ull : q[p]));
        return t >= (double)h[p] ? h[p] : (int64_t)t;
    };
    auto f = [&](double lam, int32_t *out) -> uint64_t {
        uint64_t tot = 0;
        for (int p = 0; p < P; ++p) {
            const int64_t ti = share(lam, p);
            if (out) out[p] = (int32_t)ti;
            tot += (uint64_t)ti;
        }
        return tot;
    };
    // f rises to sh > B (uniform) or to shq > B, so the doubling ends.
    double lo = 0.0, hi = 1.0;
    while (f(hi, nullptr) <= B) { lo = hi; hi *= 2.0; }
    for (int it = 0; it < 4000; ++it) {  // until lo and hi are adjacent doubles
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (f(mid, nullptr) <= B) lo = mid; else hi = mid;
    }
    std::vector<int32_t> base(P), up(P);
    const uint64_t fl = f(lo, base.data());
    f(hi, up.data());
    int64_t R = (int64_t)(B - fl);
    for (int p = 0; p < P; ++p) {
        const int64_t inc = up[p] - base[p];
        const int64_t add = std::min<int64_t>(std::max<int64_t>(R, 0), inc);
        target[p] = base[p] + (int32_t)add;
        R -= add;
    }
}
