// synth.hip — seeded synthetic workload: scene parameters, feature buffers and
// radiance samples.  Not the hot path: it produces the inputs the statistics
// kernels consume.  Every draw is a Philox4x32-10 block keyed by
// (seed, image_id) and addressed by (pixel, sample index), and all arithmetic
// runs in FP64 before the final rounding to FP32, so the samples do not depend
// on the launch shape and a CPU evaluation of the same formulas reproduces them.
//
// Sample model (x, y = pixel; u, v = in-pixel position; k = sample index):
//   r = Philox(p, k, 0, 0); u, v = top 24 bits of r.x, r.y
//   z = sqrt(-2 ln u1) cos(2 pi u2), u1, u2 from r.z, r.w (Box-Muller)
//   F = firefly_mult if the low bytes of r.x..r.w < firefly_thresh else 1
//   L = exp(s z - s^2/2)                       (LogNormal with mean 1)
//   mu_c(X, Y) = base + sum_k amp[k][c] (1 + sin(2 pi (fx_k X/W + fy_k Y/H) + phi_kc))/2,
//   times every step edge crossed, times (0.2 + 0.8 albedo_c) if modulated
//   sample_c = mu_c(x + u, y + v) L F

#include <cmath>
#include <cstring>

#include "kernels.hpp"

namespace stater {

static void scene_init_host(stater_scene *s, int model, uint32_t seed, uint32_t image_id, int W, int H) {
    std::memset(s, 0, sizeof(*s));
    s->width = W;
    s->height = H;
    s->seed = seed;
    s->image_id = image_id;
    s->model = model;
    s->base = 0.05;
    for (int k = 0; k < 3; ++k) {
        U4 r = philox4x32_10(0xA0000000u + k, 0x5CE7E000u, 0, 0, seed, image_id);
        s->fx[k] = 0.5 + 3.5 * unit32(r.x);
        s->fy[k] = 0.5 + 3.5 * unit32(r.y);
        r = philox4x32_10(0xA0000010u + k, 0x5CE7E000u, 0, 0, seed, image_id);
        const uint32_t ra[3] = {r.x, r.y, r.z};
        for (int c = 0; c < 3; ++c) s->amp[k][c] = 0.6 * unit32(ra[c]);
        r = philox4x32_10(0xA0000020u + k, 0x5CE7E000u, 0, 0, seed, image_id);
        const uint32_t rp[3] = {r.x, r.y, r.z};
        for (int c = 0; c < 3; ++c) {
            const double phi = 2.0 * kPi * unit32(rp[c]);
            s->cos_phi[k][c] = std::cos(phi);
            s->sin_phi[k][c] = std::sin(phi);
        }
    }
    s->firefly_mult = 100.0;
    s->firefly_thresh = 4294967u;  // floor(1e-3 * 2^32): firefly probability 1e-3
    s->cell_size = 32;
    s->normal_jitter = 0.02;
    s->depth_min = 1.0;
    s->depth_max = 100.0;
    if (model == 1) {  // 128x128 plain case: one vertical edge at W/2 scaling mu by 4
        s->n_edges = 1;
        s->edge_vertical[0] = 1;
        s->edge_pos[0] = 0.5 * W;
        s->edge_factor[0] = 4.0;
        s->s_min = s->s_max = 0.8;
        s->albedo_modulation = 0;
    } else {  // 12 random axis-aligned edges, factor log-uniform in [1/4, 4]
        s->n_edges = 12;
        for (int e = 0; e < 12; ++e) {
            const U4 r = philox4x32_10(0xA0000100u + e, 0x5CE7E000u, 0, 0, seed, image_id);
            s->edge_vertical[e] = (int)(r.x & 1u);
            s->edge_pos[e] = unit32(r.y) * (s->edge_vertical[e] ? W : H);
            s->edge_factor[e] = std::pow(4.0, 2.0 * unit32(r.z) - 1.0);
        }
        s->s_min = 0.3;
        s->s_max = 1.2;
        s->albedo_modulation = 1;
    }
}

// Jittered-grid Voronoi features (cell_size px): albedo U[0,1]^3 and depth
// U[depth_min, depth_max] per cell, a per-cell unit normal plus N(0, jitter)
// per pixel, renormalised.
__global__ void __launch_bounds__(256) k_features(stater_scene s, float *alb, int64_t alb_pitch, int64_t alb_cs,
                                                  float *nrm, int64_t nrm_pitch, int64_t nrm_cs, float *dep,
                                                  int64_t dep_pitch) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= s.width) return;
    const int cz = s.cell_size;
    const int gx = x / cz, gy = y / cz;
    const double px = x + 0.5, py = y + 0.5;
    double best = 1e300;
    int bi = 0, bj = 0;
    for (int j = gy - 1; j <= gy + 1; ++j)
        for (int i = gx - 1; i <= gx + 1; ++i) {
            const U4 r = philox4x32_10((uint32_t)i, (uint32_t)j, 0xCE11u, 0, s.seed, s.image_id);
            const double cx = cz * (i + unit32(r.x));
            const double cy = cz * (j + unit32(r.y));
            const double d2 = (px - cx) * (px - cx) + (py - cy) * (py - cy);
            if (d2 < best) { best = d2; bi = i; bj = j; }
        }
    const uint32_t p = (uint32_t)(y * s.width + x);
    U4 r = philox4x32_10((uint32_t)bi, (uint32_t)bj, 0xCE11u, 1, s.seed, s.image_id);
    char *ab = (char *)alb + (int64_t)y * alb_pitch + (int64_t)x * 4;
    *(float *)(ab) = (float)unit32(r.x);
    *(float *)(ab + alb_cs) = (float)unit32(r.y);
    *(float *)(ab + 2 * alb_cs) = (float)unit32(r.z);
    *(float *)((char *)dep + (int64_t)y * dep_pitch + (int64_t)x * 4) =
        (float)(s.depth_min + (s.depth_max - s.depth_min) * unit32(r.w));
    r = philox4x32_10((uint32_t)bi, (uint32_t)bj, 0xCE11u, 2, s.seed, s.image_id);
    const double cth = unit32(r.x);
    const double ph = 2.0 * kPi * unit32(r.y);
    const double sth = sqrt(1.0 - cth * cth);
    double n0 = sth * cos(ph), n1 = sth * sin(ph), n2 = cth;
    r = philox4x32_10(p, 0xFFFFFFFEu, 0, 0, s.seed, s.image_id);
    const double ra = sqrt(-2.0 * log(unit24_open(r.x)));
    const double ta = 2.0 * kPi * unit24(r.y);
    const double rb = sqrt(-2.0 * log(unit24_open(r.z)));
    const double tb = 2.0 * kPi * unit24(r.w);
    n0 += s.normal_jitter * ra * cos(ta);
    n1 += s.normal_jitter * ra * sin(ta);
    n2 += s.normal_jitter * rb * cos(tb);
    const double inv = 1.0 / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    char *nb = (char *)nrm + (int64_t)y * nrm_pitch + (int64_t)x * 4;
    *(float *)(nb) = (float)(n0 * inv);
    *(float *)(nb + nrm_cs) = (float)(n1 * inv);
    *(float *)(nb + 2 * nrm_cs) = (float)(n2 * inv);
}


__device__ __forceinline__ void samples_tile(const stater_scene &s, const float *alb, int64_t alb_pitch,
                                             int64_t alb_cs, uint32_t first0, const SynthOut &o, int bx, int by) {
    const int x = bx * 256 + threadIdx.x;
    const int y = by;
    if (x >= s.width) return;
    const int W = s.width, H = s.height;
    const uint32_t p = (uint32_t)(y * W + x);
    int spp = o.spp;
    if (o.counts) {
        const int c = *(const int32_t *)((const char *)o.counts + (int64_t)y * o.counts_pitch + (int64_t)x * 4);
        spp = min(max(c - o.count_offset, 0), spp);
    }
    uint32_t first = first0;
    if (o.first) first += (uint32_t)*(const int32_t *)((const char *)o.first + (int64_t)y * o.first_pitch + (int64_t)x * 4);
    float *out = o.out;
    const int64_t pitch = o.pitch, comp = o.comp, sstride = o.sstride;
    double sp = s.s_min;
    if (s.s_min != s.s_max) {
        const U4 r = philox4x32_10(p, 0xFFFFFFFFu, 0, 0, s.seed, s.image_id);
        sp = s.s_min + (s.s_max - s.s_min) * unit32(r.x);
    }
    double mod[3] = {1.0, 1.0, 1.0};
    if (s.albedo_modulation) {
        const char *ab = (const char *)alb + (int64_t)y * alb_pitch + (int64_t)x * 4;
        for (int c = 0; c < 3; ++c) mod[c] = 0.2 + 0.8 * (double)*(const float *)(ab + c * alb_cs);
    }
    char *po = (char *)out + (int64_t)y * pitch + (int64_t)x * 4;
    for (int j = 0; j < spp; ++j) {
        const uint32_t k = first + (uint32_t)j;
        const U4 r = philox4x32_10(p, k, 0, 0, s.seed, s.image_id);
        const double u = unit24(r.x), v = unit24(r.y);
        const double z = sqrt(-2.0 * log(unit24_open(r.z))) * cos(2.0 * kPi * unit24(r.w));
        const uint32_t ffw = (r.x & 255u) | ((r.y & 255u) << 8) | ((r.z & 255u) << 16) | ((r.w & 255u) << 24);
        const double F = ffw < s.firefly_thresh ? s.firefly_mult : 1.0;
        const double L = exp(sp * z - 0.5 * sp * sp);
        const double X = x + u, Y = y + v;
        double st[3], ct[3];
        for (int q = 0; q < 3; ++q) {
            const double th = 2.0 * kPi * (s.fx[q] * X / W + s.fy[q] * Y / H);
            st[q] = sin(th);
            ct[q] = cos(th);
        }
        double edge = 1.0;
        for (int e = 0; e < s.n_edges; ++e) {
            const double c = s.edge_vertical[e] ? X : Y;
            if (c >= s.edge_pos[e]) edge *= s.edge_factor[e];
        }
        char *oj = po + (int64_t)j * sstride;
        for (int c = 0; c < 3; ++c) {
            double mu = s.base;
            for (int q = 0; q < 3; ++q)
                mu += s.amp[q][c] * 0.5 * (1.0 + st[q] * s.cos_phi[q][c] + ct[q] * s.sin_phi[q][c]);
            mu *= edge;
            if (s.albedo_modulation) mu *= mod[c];
            *(float *)(oj + c * comp) = (float)(mu * L * F);
        }
        *(float *)(oj + 3 * comp) = (float)u;
        *(float *)(oj + 4 * comp) = (float)v;
    }
}

int launch_synth_features(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo,
                          const stater_plane *normal, const stater_plane *depth, hipStream_t st) {
    dim3 grid((s->width + 255) / 256, s->height);
    LaunchTimer t(ctx, st, "synth_features");
    hipLaunchKernelGGL(k_features, grid, dim3(256), 0, st, *s, (float *)albedo->ptr, albedo->pitch_bytes,
                       albedo->channel_stride_bytes, (float *)normal->ptr, normal->pitch_bytes,
                       normal->channel_stride_bytes, (float *)depth->ptr, depth->pitch_bytes);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

__global__ void __launch_bounds__(256) k_samples(stater_scene s, const float *alb, int64_t alb_pitch, int64_t alb_cs,
                                                 uint32_t first0, SynthOut o) {
    samples_tile(s, alb, alb_pitch, alb_cs, first0, o, blockIdx.x, blockIdx.y);
}



SynthOut make_synth_out(const stater_samples *out, const stater_plane *first_plane) {
    return SynthOut{(float *)out->data, out->pitch_bytes, out->comp_stride_bytes, out->sample_stride_bytes, out->spp,
                    out->counts, out->counts_pitch_bytes, out->count_offset,
                    first_plane ? (const int32_t *)first_plane->ptr : nullptr,
                    first_plane ? first_plane->pitch_bytes : 0};
}

int launch_synth_samples(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo, uint32_t first,
                         const stater_plane *first_plane, const stater_samples *out, hipStream_t st) {
    dim3 grid((s->width + 255) / 256, s->height);
    const SynthOut o = make_synth_out(out, first_plane);
    LaunchTimer t(ctx, st, "synth_samples");
    hipLaunchKernelGGL(k_samples, grid, dim3(256), 0, st, *s, albedo ? (const float *)albedo->ptr : nullptr,
                       albedo ? albedo->pitch_bytes : 0, albedo ? albedo->channel_stride_bytes : 0, first, o);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}


}  // namespace stater

extern "C" int stater_scene_init(stater_scene *s, int model, uint32_t seed, uint32_t image_id, int width,
                                 int height) {
    if (!s) return stater::fail(STATER_E_ARG, "stater_scene_init: NULL scene");
    if (model != 1 && model != 2) return stater::fail(STATER_E_ARG, "stater_scene_init: model must be 1 or 2");
    if (width <= 0 || height <= 0) return stater::fail(STATER_E_ARG, "stater_scene_init: bad size");
    stater::scene_init_host(s, model, seed, image_id, width, height);
    return STATER_OK;
}
