// api.cpp — the C-ABI: contexts, buffer sets, named views, dispatch, timing.
//
// The buffer set replaces the reference's Estimator buffer family
// (src/statistics/estimator.cpp:83-245): instead of one host cv::Mat plus one
// device GpuMat per "t<type>-b<slot>-[film-]<name>" buffer, uploaded every
// iteration (estimator.cpp:346-353), every statistic lives only in device
// memory as FP32 planes with a 256-byte row pitch, and the untransformed
// statistics that the reference duplicates in every lambda slot are kept once.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "internal.hpp"

namespace stater {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return STATER_E_HIP;
}

static hipEvent_t take_event(stater_ctx *ctx) {
    if (!ctx->free_events.empty()) {
        hipEvent_t e = ctx->free_events.back();
        ctx->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

LaunchTimer::LaunchTimer(stater_ctx *c, hipStream_t s, const char *n) : ctx(c), st(s), name(n) {
    if (!(ctx->flags & STATER_CTX_TIMING)) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    a = take_event(ctx);
    b = take_event(ctx);
    if (a && b) (void)hipEventRecord(a, st);
}

LaunchTimer::~LaunchTimer() {
    if (!a || !b) return;
    (void)hipEventRecord(b, st);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->pending.push_back({name, a, b});
}

hipStream_t pick_stream(const stater_bufset *bs, void *stream) {
    return stream ? (hipStream_t)stream : bs->ctx->stream;
}

// ------------------------------------------------ critical-value polynomial
static double normal_t2_crit(double alpha) {
    double lo = 0, hi = 200;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (std::erfc(std::sqrt(0.5 * mid)) > alpha) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

double t2_crit_host(double nu, double alpha) {
    if (!(nu < 1e300)) return normal_t2_crit(alpha);
    double lo = 0.0, hi = 1.0;  // x = nu/(nu+t2); I_x(nu/2,1/2) rises with x
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (betainc_reg(0.5 * nu, 0.5, mid) < alpha) lo = mid; else hi = mid;
    }
    const double x = 0.5 * (lo + hi);
    return nu * (1.0 - x) / x;
}

// Fit g(u) = t2crit(1/u) on [0, umax] by Chebyshev interpolation of degree D
// (nodes in s = 2u/umax - 1), convert to monomial coefficients in u, and
// measure the FP32 Horner evaluation against the FP64 tail on a grid.
// Returns the largest relative error; cmax gets the largest crit on the range.
template <int D>
static double fit_tcrit(double alpha, double umax, float (&c)[D + 1], double &cmax) {
    double gnode[D + 1], cheb[D + 1];
    for (int k = 0; k <= D; ++k) {
        const double s = std::cos(kPi * (k + 0.5) / (D + 1));
        const double u = 0.5 * (s + 1.0) * umax;
        gnode[k] = u == 0.0 ? normal_t2_crit(alpha) : t2_crit_host(1.0 / u, alpha);
    }
    for (int j = 0; j <= D; ++j) {
        double acc = 0;
        for (int k = 0; k <= D; ++k) acc += gnode[k] * std::cos(kPi * j * (k + 0.5) / (D + 1));
        cheb[j] = acc * (j == 0 ? 1.0 : 2.0) / (D + 1);
    }
    // T_j(s) as monomials in s (T_{j+1} = 2 s T_j - T_{j-1}), then s = a u - 1.
    double Tm[D + 1][D + 1] = {};
    Tm[0][0] = 1.0;
    Tm[1][1] = 1.0;
    for (int j = 2; j <= D; ++j)
        for (int i = 0; i <= D; ++i)
            Tm[j][i] = (i ? 2.0 * Tm[j - 1][i - 1] : 0.0) - Tm[j - 2][i];
    double ms[D + 1] = {}, mu[D + 1] = {};
    for (int i = 0; i <= D; ++i)
        for (int j = 0; j <= D; ++j) ms[i] += cheb[j] * Tm[j][i];
    const double sa = 2.0 / umax;
    for (int i = 0; i <= D; ++i) {  // (sa u - 1)^i by the binomial theorem
        double binom = 1.0;
        for (int m = 0; m <= i; ++m) {
            mu[m] += ms[i] * binom * std::pow(sa, m) * (((i - m) & 1) ? -1.0 : 1.0);
            binom = binom * (i - m) / (m + 1);
        }
    }
    for (int i = 0; i <= D; ++i) c[i] = (float)mu[i];
    double maxerr = 0;
    cmax = 0;
    for (int i = 0; i <= 1000; ++i) {
        const double u = umax * i / 1000.0;
        const float uf = (float)u;
        float acc = c[D];
        for (int k = D - 1; k >= 0; --k) acc = std::fma(acc, uf, c[k]);
        const double ref = u == 0.0 ? normal_t2_crit(alpha) : t2_crit_host(1.0 / u, alpha);
        maxerr = std::max(maxerr, std::fabs((double)acc / ref - 1.0));
        cmax = std::max(cmax, std::max((double)acc, ref));
    }
    return maxerr;
}

TCritPoly make_tcrit_poly(double alpha) {
    TCritPoly tp{};
    double cmax, cmax_s;
    const double maxerr = fit_tcrit<kTCritDeg>(alpha, kTCritUmax, tp.c, cmax);
    const double maxerr_s = fit_tcrit<kTCritDegShort>(alpha, kTCritUmaxShort, tp.cs, cmax_s);
    tp.umax = (float)kTCritUmax;
    tp.alpha = alpha;
    // The exact-tail band sits well above the fit's FP32 evaluation error.
    tp.band = (float)std::max(2e-6, 8.0 * maxerr + 2e-6);
    tp.neg_hi = -(1.f + tp.band);
    tp.neg_lo = -(1.f - tp.band);
    // crit rises with u and is flat between grid points to far below 1 %
    tp.band_abs = (float)((double)tp.band * cmax * 1.01);
    const double band_s = std::max(2e-6, 8.0 * maxerr_s + 2e-6);
    tp.band_abs_short = (float)(band_s * cmax_s * 1.01);
    return tp;
}

const TCritPoly &tcrit_for(stater_ctx *ctx, double alpha) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tcrit.find(alpha);
    if (it == ctx->tcrit.end()) it = ctx->tcrit.emplace(alpha, make_tcrit_poly(alpha)).first;
    return it->second;
}

}  // namespace stater

using namespace stater;

// ================================================================ context
extern "C" {

const char *stater_last_error(void) { return g_err.c_str(); }
int stater_abi_version(void) { return STATER_ABI_VERSION; }

void stater_print_macro_configuration(void) {
    std::printf("==== stater (gfx950) configuration ====\n");
    std::printf("max sets = %d (untransformed + Box-Cox), max radius = %d\n", kMaxSets, STATER_MAX_RADIUS);
    std::printf("FILM_MULTI_TRANSFORM_MODE = runtime (mt_mode 2 | 3)\n");
    std::printf("CASCADE_MODE = runtime (cascade), USE_LEVC = runtime (use_levc)\n");
    std::printf("accumulation: untransformed set bit-exact FP32 (Meng), Box-Cox sets expm1 form\n");
}

int stater_create(int device, unsigned flags, stater_ctx **out) {
    if (!out) return fail(STATER_E_ARG, "stater_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    STATER_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(STATER_E_ARG, "stater_create: no such device");
    STATER_HIP(hipSetDevice(device));
    int rc = noalloc_scratch_init(device);  // the static scratch's address on this device
    if (rc) return rc;
    auto *ctx = new stater_ctx();
    ctx->device = device;
    ctx->flags = flags;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) { delete ctx; return hip_fail(e, "hipGetDeviceProperties"); }
    ctx->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete ctx; return hip_fail(e, "hipStreamCreate"); }
    *out = ctx;
    return STATER_OK;
}

int stater_destroy(stater_ctx *ctx) {
    if (!ctx) return STATER_OK;
    if (ctx->live_bufsets.load() > 0)
        return fail(STATER_E_STATE, "stater_destroy: destroy the context's buffer sets and servers first");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return STATER_OK;
}

int stater_ctx_stream(stater_ctx *ctx, void **stream_out) {
    if (!ctx || !stream_out) return fail(STATER_E_ARG, "stater_ctx_stream: NULL argument");
    *stream_out = (void *)ctx->stream;
    return STATER_OK;
}

static int flush_timing(stater_ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto &p : ctx->pending) {
        STATER_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        STATER_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto &acc = ctx->stats[p.name];
        acc.launches++;
        acc.ms += ms;
        ctx->free_events.push_back(p.a);
        ctx->free_events.push_back(p.b);
    }
    ctx->pending.clear();
    return STATER_OK;
}

int stater_timing_reset(stater_ctx *ctx) {
    if (!ctx) return fail(STATER_E_ARG, "stater_timing_reset: NULL context");
    int rc = flush_timing(ctx);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->stats.clear();
    return rc;
}

int stater_timing_count(stater_ctx *ctx) {
    if (!ctx) return fail(STATER_E_ARG, "stater_timing_count: NULL context");
    int rc = flush_timing(ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return (int)ctx->stats.size();
}

int stater_timing_get(stater_ctx *ctx, int index, char *name, int name_len, uint64_t *launches,
                      double *total_ms) {
    if (!ctx) return fail(STATER_E_ARG, "stater_timing_get: NULL context");
    int rc = flush_timing(ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (index < 0 || index >= (int)ctx->stats.size()) return fail(STATER_E_ARG, "stater_timing_get: index");
    auto it = ctx->stats.begin();
    std::advance(it, index);
    if (name && name_len > 0) {
        std::strncpy(name, it->first.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    return STATER_OK;
}

// ============================================================ buffer sets
struct Layout {
    int S = 1, nst = 7, ld = 0;
    size_t plane = 0, n_planes = 0;
    size_t as_bytes = 0;    // adaptive-sampling scratch after the planes (sets with derived planes)
    size_t zero_bytes = 0;  // the variance filter's zero record, last
    size_t bytes() const { return n_planes * plane * 4 + as_bytes + zero_bytes; }
    float set_lambda[kMaxSets] = {1.f};
    int slot_set[STATER_MAX_LAMBDAS] = {};
};

static int make_layout(int W, int H, int n_lambdas, const float *lambdas, int mm, unsigned flags, Layout &L) {
    if (W <= 0 || H <= 0 || W > 16384 || H > 16384) return fail(STATER_E_ARG, "bufset: width/height out of range");
    if (n_lambdas < 0 || n_lambdas > STATER_MAX_LAMBDAS || (n_lambdas && !lambdas))
        return fail(STATER_E_ARG, "bufset: bad lambda list");
    if (mm < 1 || mm > 5) return fail(STATER_E_ARG, "bufset: max_moment must be 1..5");
    L.S = 1;
    L.set_lambda[0] = 1.f;
    for (int j = 0; j < n_lambdas; ++j) {
        const float l = lambdas[j];
        if (!(l > 0.f) || !std::isfinite(l)) return fail(STATER_E_ARG, "bufset: lambdas must be positive");
        int s = -1;
        for (int k = 0; k < L.S; ++k)
            if (L.set_lambda[k] == l) s = k;
        if (s < 0) {
            if (L.S >= kMaxSets) return fail(STATER_E_UNSUPPORTED, "bufset: at most 5 distinct lambdas != 1");
            s = L.S++;
            L.set_lambda[s] = l;
        }
        L.slot_set[j] = s;
    }
    L.nst = mm >= 5 ? 7 : mm;
    L.ld = (W + kRowAlign - 1) / kRowAlign * kRowAlign;
    L.plane = (size_t)L.ld * H;
    // n + moments (all 7 slots) + Box-Cox reference point + shifted Box-Cox means
    size_t np = 1 + (size_t)L.S * N_STATS * 3 + 3 + (size_t)(L.S - 1) * 3;
    if (flags & STATER_BUFSET_FILM) np += 3 + 4;  // resolved film, RGB contribSum + weight sum
    if (!(flags & STATER_BUFSET_NO_DERIVED)) {
        np += (size_t)L.S * (1 + 6 * 3) + 9 + 1 + 3 * kBitsWords + kVarRecFloats;
        L.as_bytes = as_scratch_bytes_for((size_t)W * H);
        L.zero_bytes = kVarZeroBytes;
    }
    L.n_planes = np;
    return STATER_OK;
}

int64_t stater_bufset_bytes_for(int width, int height, int n_lambdas, const float *lambdas, int max_moment,
                                unsigned flags) {
    Layout L;
    if (make_layout(width, height, n_lambdas, lambdas, max_moment, flags, L)) return -1;
    return (int64_t)L.bytes();
}

// Carves the planes of a set out of mem (L.bytes() bytes, 256-byte aligned).
static stater_bufset *bufset_over(stater_ctx *ctx, const Layout &L, int width, int height, int n_lambdas,
                                  const float *lambdas, int max_moment, unsigned flags, char *mem, bool owns) {
    auto *bs = new stater_bufset();
    bs->ctx = ctx;
    bs->W = width; bs->H = height; bs->ld = L.ld; bs->plane = L.plane; bs->S = L.S;
    bs->n_lambdas = n_lambdas;
    for (int j = 0; j < n_lambdas; ++j) { bs->lambdas[j] = lambdas[j]; bs->slot_set[j] = L.slot_set[j]; }
    for (int s = 0; s < L.S; ++s) bs->set_lambda[s] = L.set_lambda[s];
    bs->mm = max_moment;
    bs->flags = flags;
    bs->bytes = L.bytes();
    bs->base = mem;
    bs->owns_memory = owns;
    float *p = (float *)bs->base;
    bs->n = (int32_t *)p; p += L.plane;
    bs->stats = p; p += (size_t)L.S * N_STATS * 3 * L.plane;
    bs->lref = p; p += 3 * L.plane;
    bs->mshift = p; p += (size_t)(L.S - 1) * 3 * L.plane;
    if (flags & STATER_BUFSET_FILM) {
        bs->film_rgb = p; p += 3 * L.plane;
        bs->film_sum = p; p += 4 * L.plane;
    }
    if (!(flags & STATER_BUFSET_NO_DERIVED)) {
        bs->jb = p; p += (size_t)L.S * L.plane;
        float **d[6] = {&bs->lnS2, &bs->se2, &bs->m2C, &bs->varD, &bs->varCD, &bs->discr};
        for (auto q : d) { *q = p; p += (size_t)L.S * 3 * L.plane; }
        bs->meanVar = p; p += 3 * L.plane;
        bs->meanD = p; p += 3 * L.plane;
        bs->filmD = p; p += 3 * L.plane;
        bs->targetspp = (int32_t *)p; p += L.plane;
        bs->welch_bits = (uint32_t *)p; p += 3 * kBitsWords * L.plane;
        bs->var_rec = p; p += (size_t)kVarRecFloats * L.plane;
        bs->as_scratch = p;
        bs->as_scratch_bytes = L.as_bytes;
        bs->var_zero = (float *)((char *)p + L.as_bytes);
    }
    ctx->live_bufsets.fetch_add(1);
    return bs;
}

int stater_bufset_create(stater_ctx *ctx, int width, int height, int n_lambdas, const float *lambdas,
                         int max_moment, unsigned flags, stater_bufset **out) {
    if (!ctx || !out) return fail(STATER_E_ARG, "stater_bufset_create: NULL argument");
    *out = nullptr;
    Layout L;
    int rc = make_layout(width, height, n_lambdas, lambdas, max_moment, flags, L);
    if (rc) return rc;
    STATER_HIP(hipSetDevice(ctx->device));
    char *mem = nullptr;
    hipError_t e = hipMalloc((void **)&mem, L.bytes());
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(STATER_E_NOMEM, "stater_bufset_create: device allocation failed");
    }
    // cleared on the context's stream and waited for: that stream does not
    // synchronise with the null stream, so a plain hipMemset could still be
    // running when the first kernel of the set starts
    e = hipMemsetAsync(mem, 0, L.bytes(), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { (void)hipFree(mem); return hip_fail(e, "hipMemsetAsync"); }
    *out = bufset_over(ctx, L, width, height, n_lambdas, lambdas, max_moment, flags, mem, true);
    return STATER_OK;
}

int stater_bufset_create_in(stater_ctx *ctx, int width, int height, int n_lambdas, const float *lambdas,
                            int max_moment, unsigned flags, void *mem, uint64_t mem_bytes, void *stream,
                            stater_bufset **out) {
    if (!ctx || !out || !mem) return fail(STATER_E_ARG, "stater_bufset_create_in: NULL argument");
    *out = nullptr;
    if ((uintptr_t)mem % 256) return fail(STATER_E_ARG, "stater_bufset_create_in: memory must be 256-byte aligned");
    Layout L;
    int rc = make_layout(width, height, n_lambdas, lambdas, max_moment, flags, L);
    if (rc) return rc;
    if (mem_bytes < L.bytes()) return fail(STATER_E_ARG, "stater_bufset_create_in: memory smaller than the layout");
    auto *bs = bufset_over(ctx, L, width, height, n_lambdas, lambdas, max_moment, flags, (char *)mem, false);
    rc = stater_bufset_reset(bs, stream);
    if (rc) { stater_bufset_destroy(bs); return rc; }
    *out = bs;
    return STATER_OK;
}

int stater_bufset_destroy(stater_bufset *bs) {
    if (!bs) return STATER_OK;
    if (bs->owns_memory) {
        (void)hipSetDevice(bs->ctx->device);
        (void)hipStreamSynchronize(bs->ctx->stream);
        (void)hipFree(bs->base);
    }
    bs->ctx->live_bufsets.fetch_sub(1);
    delete bs;
    return STATER_OK;
}

int64_t stater_bufset_bytes(const stater_bufset *bs) { return bs ? (int64_t)bs->bytes : -1; }

int stater_bufset_reset(stater_bufset *bs, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_bufset_reset: NULL set");
    const size_t bytes = (1 + (size_t)bs->S * N_STATS * 3) * bs->plane * 4;
    int rc = launch_zero(bs->base, bytes, pick_stream(bs, stream));
    if (!rc && bs->film_sum) rc = launch_zero(bs->film_sum, 4 * bs->plane * 4, pick_stream(bs, stream));
    return rc;
}

static void fill_plane(const stater_bufset *bs, void *ptr, int channels, stater_dtype dt, stater_plane *out) {
    out->ptr = ptr;
    out->pitch_bytes = (int64_t)bs->ld * 4;
    out->channel_stride_bytes = (int64_t)bs->plane * 4;
    out->width = bs->W;
    out->height = bs->H;
    out->channels = channels;
    out->dtype = dt;
    out->layout = STATER_PLANAR;
    out->reserved = 0;
}

int stater_bufset_set_stat_type(stater_bufset *bs, int type) {
    if (!bs) return fail(STATER_E_ARG, "stater_bufset_set_stat_type: NULL set");
    if (type < 0 || type > 7) return fail(STATER_E_ARG, "stater_bufset_set_stat_type: type must be 0..7");
    bs->stat_type = type;
    return STATER_OK;
}

// Single-channel stat types (StatTile<Float>, src/statistics/statpath.cpp:
// 119-129): the MIS win rates and the material-id and depth G-buffers.
static bool float_type(int type) { return type >= 1 && type <= 4; }

int stater_bufset_view(const stater_bufset *bs, const char *name, stater_plane *out) {
    if (!bs || !name || !out) return fail(STATER_E_ARG, "stater_bufset_view: NULL argument");
    const std::string nm(name);
    const bool derived = !(bs->flags & STATER_BUFSET_NO_DERIVED);
    auto need_derived = [&]() { return fail(STATER_E_NOTFOUND, "buffer '" + nm + "' not allocated (NO_DERIVED)"); };
    if (nm == "n") { fill_plane(bs, bs->n, 1, STATER_I32, out); return STATER_OK; }
    if (nm == "film") {
        fill_plane(bs, bs->film_rgb ? bs->film_rgb : bs->stat(0, ST_MEAN, 0), 3, STATER_F32, out);
        return STATER_OK;
    }
    if (nm == "filmD") { if (!derived) return need_derived(); fill_plane(bs, bs->filmD, 3, STATER_F32, out); return STATER_OK; }
    if (nm == "targetspp") { if (!derived) return need_derived(); fill_plane(bs, bs->targetspp, 1, STATER_I32, out); return STATER_OK; }
    // t<type>-b<slot>-[film-]<stat> of the set's stat type
    int type = -1, slot = -1, consumed = 0;
    if (std::sscanf(name, "t%d-b%d-%n", &type, &slot, &consumed) != 2 || consumed == 0 || type != bs->stat_type)
        return fail(STATER_E_NOTFOUND, "unknown buffer '" + nm + "'");
    const int ch = float_type(type) ? 1 : 3;
    const int nslots = std::max(1, bs->n_lambdas);
    if (slot < 0 || slot >= nslots) return fail(STATER_E_NOTFOUND, "no slot in '" + nm + "'");
    std::string rest = nm.substr(consumed);
    bool film = false;
    if (rest.rfind("film-", 0) == 0) { film = true; rest = rest.substr(5); }
    const int s = film ? 0 : (bs->n_lambdas ? bs->slot_set[slot] : 0);
    static const char *stat_names[N_STATS] = {"mean", "m2", "m3", "m4", "c2X", "c2Y", "m1"};
    for (int k = 0; k < N_STATS; ++k)
        if (rest == stat_names[k]) { fill_plane(bs, bs->stat(s, k, 0), ch, STATER_F32, out); return STATER_OK; }
    if (film && rest == "n") { fill_plane(bs, bs->n, 1, STATER_I32, out); return STATER_OK; }
    if (!derived) return need_derived();
    if (rest == "jb") { fill_plane(bs, bs->jb + (size_t)s * bs->plane, 1, STATER_F32, out); return STATER_OK; }
    struct { const char *n; float *b; } der[] = {{"lnS2", bs->lnS2}, {"se2", bs->se2}, {"m2C", bs->m2C},
                                                 {"varD", bs->varD}, {"varCD", bs->varCD}, {"discr", bs->discr}};
    for (auto &d : der)
        if (rest == d.n) { fill_plane(bs, bs->per_set(d.b, s, 0), ch, STATER_F32, out); return STATER_OK; }
    if (film && rest == "meanVar") { fill_plane(bs, bs->meanVar, ch, STATER_F32, out); return STATER_OK; }
    if (film && rest == "meanD") { fill_plane(bs, bs->meanD, ch, STATER_F32, out); return STATER_OK; }
    return fail(STATER_E_NOTFOUND, "unknown buffer '" + nm + "'");
}

// ============================================================== dispatch
int stater_accumulate(stater_bufset *bs, const stater_samples *samples, int mode, void *stream) {
    if (!bs || !samples || !samples->data) return fail(STATER_E_ARG, "stater_accumulate: NULL argument");
    if (samples->spp < 0) return fail(STATER_E_ARG, "stater_accumulate: spp < 0");
    const int base = mode & ~STATER_ACC_FRESH;
    if (base != STATER_ACC_SEQUENTIAL && base != STATER_ACC_MERGE64)
        return fail(STATER_E_ARG, "stater_accumulate: unknown mode");
    if (samples->spp == 0) return (mode & STATER_ACC_FRESH) ? stater_bufset_reset(bs, stream) : STATER_OK;
    return launch_accumulate(bs, samples, mode, pick_stream(bs, stream));
}

int stater_film_resolve(stater_bufset *bs, float scale, const stater_plane *out, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_film_resolve: NULL set");
    if (!bs->film_sum) return fail(STATER_E_STATE, "stater_film_resolve: set was created without STATER_BUFSET_FILM");
    stater_plane own;
    if (!out) {
        fill_plane(bs, bs->film_rgb, 3, STATER_F32, &own);
        out = &own;
    }
    if (!out->ptr || out->width != bs->W || out->height != bs->H || out->channels != 3 || out->dtype != STATER_F32)
        return fail(STATER_E_ARG, "stater_film_resolve: out must be a W x H 3-channel f32 plane");
    return launch_film_resolve(bs, scale, out, pick_stream(bs, stream));
}

// The reference's integrator defaults (CreateStatPathIntegrator,
// src/statistics/statpath.cpp:1070-1071: filterSD 10, "filterradius" 20) and
// Estimator constants (z = 2.80703 = Phi^-1(1 - 0.005/2), estimator.h:353).
void stater_default_params(stater_denoise_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->radius = 20;
    p->ds_factor = -0.5f / (10.f * 10.f);
    p->var_z = 2.80703f;
    p->alpha = 0.005f;
    p->mt_mode = 2;
    p->cascade = 1;
    p->use_levc = 1;
}

static int check_params(const stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g) {
    if (!p) return fail(STATER_E_ARG, "denoise: NULL params");
    if (p->radius < 0 || p->radius > STATER_MAX_RADIUS) return fail(STATER_E_ARG, "denoise: radius out of range");
    if (!(p->alpha > 0.f && p->alpha < 1.f)) return fail(STATER_E_ARG, "denoise: alpha must be in (0,1)");
    if (!(p->var_z > 0.f)) return fail(STATER_E_ARG, "denoise: var_z must be positive");
    if (p->mt_mode != 2 && p->mt_mode != 3) return fail(STATER_E_ARG, "denoise: mt_mode must be 2 or 3");
    if (!(p->ds_factor <= 0.f)) return fail(STATER_E_ARG, "denoise: ds_factor must be <= 0");
    if (bs->flags & STATER_BUFSET_NO_DERIVED) return fail(STATER_E_STATE, "denoise: set was created NO_DERIVED");
    if (g) {
        if (g->count < 0 || g->count > STATER_MAX_GBUFFERS) return fail(STATER_E_ARG, "denoise: gbuffer count");
        int total = 0;
        for (int k = 0; k < g->count; ++k) {
            const stater_plane &pl = g->planes[k];
            const int64_t esz = pl.dtype == STATER_BF16 ? 2 : 4;
            if (!pl.ptr || pl.width != bs->W || pl.height != bs->H || pl.channels < 1 || pl.channels > 3 ||
                (pl.dtype != STATER_F32 && pl.dtype != STATER_BF16) || pl.layout != STATER_PLANAR ||
                pl.pitch_bytes % esz || pl.channel_stride_bytes % esz)
                return fail(STATER_E_ARG, "denoise: gbuffer plane must be planar f32 or bf16 of the film's size");
            if (!(g->dr_factors[k] <= 0.f)) return fail(STATER_E_ARG, "denoise: gbuffer factor must be <= 0");
            total += pl.channels;
        }
        if (total > 7) return fail(STATER_E_UNSUPPORTED, "denoise: at most 7 feature channels");
    }
    return STATER_OK;
}

int stater_denoise_vars(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *gbufs,
                        void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_denoise_vars: NULL set");
    int rc = check_params(bs, p, gbufs);
    if (rc) return rc;
    if (bs->mm < 4) return fail(STATER_E_STATE, "stater_denoise_vars: needs max_moment >= 4");
    return launch_var_filter(bs, p, gbufs, pick_stream(bs, stream));
}

static int check_film_plane(const stater_bufset *bs, const stater_plane *f, const char *what) {
    if (!f) return STATER_OK;
    const bool il = f->layout == STATER_INTERLEAVED;
    if (!f->ptr || f->width != bs->W || f->height != bs->H || f->channels != 3 || f->dtype != STATER_F32 ||
        (f->layout != STATER_PLANAR && !il) || f->pitch_bytes % 4 || (!il && f->channel_stride_bytes % 4) ||
        f->pitch_bytes < (int64_t)bs->W * (il ? 12 : 4))
        return fail(STATER_E_ARG, std::string("stater_denoise_film: bad ") + what + " plane");
    return STATER_OK;
}

int stater_denoise_film(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *gbufs,
                        const stater_plane *film, const stater_plane *film_out, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_denoise_film: NULL set");
    int rc = check_params(bs, p, gbufs);
    if (rc) return rc;
    if (bs->mm < 2) return fail(STATER_E_STATE, "stater_denoise_film: needs max_moment >= 2");
    if ((rc = check_film_plane(bs, film, "film")) || (rc = check_film_plane(bs, film_out, "film_out"))) return rc;
    return launch_welch_filter(bs, p, gbufs, film, film_out, false, pick_stream(bs, stream));
}

int stater_denoise_plain(stater_bufset *bs, int radius, float alpha, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_denoise_plain: NULL set");
    stater_denoise_params p;
    stater_default_params(&p);
    p.radius = radius;
    p.alpha = alpha;
    p.ds_factor = 0.f;
    p.cascade = 0;
    p.use_levc = 0;
    int rc = check_params(bs, &p, nullptr);
    if (rc) return rc;
    if (bs->mm < 2) return fail(STATER_E_STATE, "stater_denoise_plain: needs max_moment >= 2");
    return launch_welch_filter(bs, &p, nullptr, nullptr, nullptr, true, pick_stream(bs, stream));
}

int stater_calc_target_spp(stater_bufset *bs, uint64_t budget_spp, uint32_t cap, const stater_plane *out,
                           void *stream) {
    return stater_calc_target_spp_ex(bs, budget_spp, cap, 1, out, stream);
}

int stater_calc_target_spp_ex(stater_bufset *bs, uint64_t budget_spp, uint32_t cap, int as_mode,
                              const stater_plane *out, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_calc_target_spp: NULL set");
    if (as_mode != 0 && as_mode != 1) return fail(STATER_E_ARG, "stater_calc_target_spp: as_mode must be 0 or 1");
    if (as_mode == 0 && bs->mm < 2) return fail(STATER_E_STATE, "stater_calc_target_spp: as_mode 0 needs m2");
    if (bs->flags & STATER_BUFSET_NO_DERIVED) return fail(STATER_E_STATE, "stater_calc_target_spp: NO_DERIVED set");
    if (out && (!out->ptr || out->width != bs->W || out->height != bs->H || out->dtype != STATER_I32 ||
                out->pitch_bytes % 4))
        return fail(STATER_E_ARG, "stater_calc_target_spp: bad output plane");
    if (budget_spp > (uint64_t)1 << 32) return fail(STATER_E_ARG, "stater_calc_target_spp: budget too large");
    return launch_target_spp(bs, budget_spp, cap, as_mode, out, pick_stream(bs, stream));
}

int stater_calc_mean_vars(stater_bufset *bs, void *stream) {
    if (!bs) return fail(STATER_E_ARG, "stater_calc_mean_vars: NULL set");
    if (bs->flags & STATER_BUFSET_NO_DERIVED) return fail(STATER_E_STATE, "stater_calc_mean_vars: NO_DERIVED set");
    if (bs->mm < 2) return fail(STATER_E_STATE, "stater_calc_mean_vars: needs max_moment >= 2");
    return launch_mean_vars(bs, pick_stream(bs, stream));
}

int stater_synchronize(void *stream) {
    STATER_HIP(hipStreamSynchronize((hipStream_t)stream));
    return STATER_OK;
}

// ============================================================== synthesis
int stater_synth_features(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo,
                          const stater_plane *normal, const stater_plane *depth, void *stream) {
    if (!ctx || !s || !albedo || !normal || !depth) return fail(STATER_E_ARG, "stater_synth_features: NULL argument");
    const stater_plane *pl[3] = {albedo, normal, depth};
    const int ch[3] = {3, 3, 1};
    for (int k = 0; k < 3; ++k)
        if (!pl[k]->ptr || pl[k]->width != s->width || pl[k]->height != s->height || pl[k]->channels != ch[k] ||
            pl[k]->dtype != STATER_F32 || pl[k]->layout != STATER_PLANAR)
            return fail(STATER_E_ARG, "stater_synth_features: planes must be planar f32 of the scene's size");
    return launch_synth_features(ctx, s, albedo, normal, depth, stream ? (hipStream_t)stream : ctx->stream);
}

int stater_synth_samples(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo, uint32_t first,
                         const stater_plane *first_plane, const stater_samples *out, void *stream) {
    if (!ctx || !s || !out || !out->data) return fail(STATER_E_ARG, "stater_synth_samples: NULL argument");
    if (s->albedo_modulation && (!albedo || !albedo->ptr))
        return fail(STATER_E_ARG, "stater_synth_samples: scene modulates by albedo; pass the albedo plane");
    if (out->spp <= 0) return fail(STATER_E_ARG, "stater_synth_samples: spp must be positive");
    if (first_plane && (!first_plane->ptr || first_plane->dtype != STATER_I32 || first_plane->width != s->width ||
                        first_plane->height != s->height))
        return fail(STATER_E_ARG, "stater_synth_samples: first must be an i32 plane of the scene's size");
    return launch_synth_samples(ctx, s, albedo, first, first_plane, out,
                                stream ? (hipStream_t)stream : ctx->stream);
}

int stater_scene_init(stater_scene *s, int model, uint32_t seed, uint32_t image_id, int width, int height);

}  // extern "C"
