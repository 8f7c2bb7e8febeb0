// internal.hpp — library-internal objects shared by the C-ABI and the kernels.
#pragma once

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "tdist.hpp"

struct stater_ctx {
    int device = 0;
    unsigned flags = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    std::mutex mu;
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    struct Acc { uint64_t launches = 0; double ms = 0.0; };
    std::map<std::string, Acc> stats;
    std::vector<hipEvent_t> free_events;
    std::map<double, stater::TCritPoly> tcrit;  // fitted critical-value polynomials per level
    std::atomic<int> live_bufsets{0};           // a context outlives its buffer sets
};

struct stater_bufset {
    stater_ctx *ctx = nullptr;
    int W = 0, H = 0;
    int ld = 0;          // row pitch in elements
    size_t plane = 0;    // elements per plane (ld * H)
    int S = 1;           // statistics sets: 0 = untransformed, 1.. = Box-Cox
    float set_lambda[stater::kMaxSets] = {1.f};
    int n_lambdas = 0;
    float lambdas[STATER_MAX_LAMBDAS] = {};
    int slot_set[STATER_MAX_LAMBDAS] = {};
    int mm = 5;
    unsigned flags = 0;
    int stat_type = 0;   // the reference's StatTypeIndex the buffer names carry (t<type>-...)
    char *base = nullptr;
    size_t bytes = 0;
    // device planes
    int32_t *n = nullptr;
    float *stats = nullptr;  // [S][7][3][plane]
    float *lref = nullptr;   // [3][plane] Box-Cox reference sample x1 of each pixel-channel (sets 1..)
    float *mshift = nullptr; // [S-1][3][plane] Box-Cox means relative to the reference point
    float *film_rgb = nullptr;  // STATER_BUFSET_FILM: [3][plane] resolved film
    float *film_sum = nullptr;  // [4][plane] the Film's per-pixel RGB contribSum and weight sum,
                                // carried across calls (FilmTile pixels of the whole render)
    float *jb = nullptr;     // [S][plane]
    float *lnS2 = nullptr, *se2 = nullptr, *m2C = nullptr, *varD = nullptr, *varCD = nullptr,
          *discr = nullptr;  // [S][3][plane]
    float *meanVar = nullptr, *meanD = nullptr, *filmD = nullptr;  // [3][plane]
    int32_t *targetspp = nullptr;
    // Welch decisions of the forward half-window (radius <= kBitsMaxRadius):
    // [3 channels][kBitsWords][plane] words, bit i of word k = offset 32k + i
    uint32_t *welch_bits = nullptr;
    float *var_rec = nullptr;   // packed variance-filter records of one set group, kVarRecFloats per pixel
    float *var_zero = nullptr;  // one all-zero record (the filter's halo outside the image)
    // adaptive-sampling scratch (inside the set's memory, after the planes)
    void *as_scratch = nullptr;
    size_t as_scratch_bytes = 0;
    bool owns_memory = true;  // false: created over caller memory

    float *stat(int s, int st, int c) const {
        return stats + (((size_t)s * stater::N_STATS + st) * 3 + c) * plane;
    }
    float *per_set(float *b, int s, int c) const { return b + ((size_t)s * 3 + c) * plane; }
};

namespace stater {

// The symmetric Welch filter stores one decision bit per channel for every
// forward neighbour offset (half of the (2r+1)^2 - 1 window) of every pixel:
// 32-bit words per channel for a radius r.  The set's bit planes cover
// r <= kBitsMaxRadius, as does the weighting pass that reads them from
// per-pixel records (kRecBitsRadius).  Wider windows (r <= 20) split the
// forward rows into two bands whose words fit the variance-filter record
// planes (denoise.hip, banded form).
constexpr int bits_words(int r) { return (2 * r * r + 2 * r + 31) / 32; }
constexpr int kBitsMaxRadius = 10;
constexpr int kBitsWords = bits_words(kBitsMaxRadius);  // 7
constexpr int kRecBitsRadius = 10;
constexpr int kRecWords = bits_words(kRecBitsRadius);  // 7

// Adaptive-sampling scratch of a set: control block, then q, h and inc
// (4 B per pixel each), then one int64 per 256-pixel block.
// Variance filter input records: ok, 7 features, then (lnS2, z^2 se2, s^2,
// m2C/(n-1)) per set-channel of a group of up to 3 sets, padded to an odd
// number of 16-byte items: 44 floats per pixel.
constexpr int kVarRecFloats = 44;
constexpr size_t kVarZeroBytes = 256;

constexpr size_t kAsCtlBytes = 512;
inline size_t as_scratch_bytes_for(size_t pixels) {
    const size_t nb = (pixels + 255) / 256;
    auto r = [](size_t b) { return (b + 255) / 256 * 256; };
    return kAsCtlBytes + 3 * r(pixels * 4) + r(nb * 8);
}

// Kernel-launch timing: wraps one launch with events when the context asks.
struct LaunchTimer {
    stater_ctx *ctx;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    const char *name;
    LaunchTimer(stater_ctx *c, hipStream_t s, const char *n);
    ~LaunchTimer();
};

hipStream_t pick_stream(const stater_bufset *bs, void *stream);
const TCritPoly &tcrit_for(stater_ctx *ctx, double alpha);

// kernels (accumulate.hip, denoise.hip, synth.hip)
int launch_accumulate(stater_bufset *bs, const stater_samples *s, int mode, hipStream_t st);
// The variance filter's pack pass also runs the prepass (lnS2, se2, m2C, jb).
// ty: output rows per tile.  Variance filter: 8 (one row per filtering
// wavefront) or, for the r <= 10 feature-weighted filter, 4 (two rows per
// wavefront); r > 10 always runs 7.  Welch filter, ordered form (mode 3 and
// the plain method): 2, 4 or 8 (one row per wavefront; 2 and 8 only for the
// feature-weighted mode-2 filter at r <= 10, otherwise 4).
constexpr int kVarTileRows = 8;
int launch_var_filter(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                      hipStream_t st, int ty = kVarTileRows);
int launch_welch_filter(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                        const stater_plane *film, const stater_plane *film_out, bool plain, hipStream_t st,
                        int ty = 4);
// as_mode 1: the denoised variance varD of set 0; 0: the sample variance m2/(n-1).
int launch_target_spp(stater_bufset *bs, uint64_t budget, uint32_t cap, int as_mode, const stater_plane *out,
                      hipStream_t st);
int launch_mean_vars(stater_bufset *bs, hipStream_t st);
// The AS map over caller planes with the static scratch (stater_calc_target_spp_noalloc);
// noalloc_scratch_init takes that scratch's address on the current device.
int noalloc_scratch_init(int device);
int launch_target_spp_noalloc(stater_ctx *ctx, int W, int H, const stater_plane *n, const stater_plane *varD,
                              const stater_plane *mean, uint64_t budget, uint32_t cap, const stater_plane *out,
                              hipStream_t st);
int launch_zero(void *p, size_t bytes, hipStream_t st);
int launch_film_resolve(stater_bufset *bs, float scale, const stater_plane *out, hipStream_t st);
int launch_welch_decisions(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g,
                           const stater_plane *film, int plain, uint8_t *dec, hipStream_t st);
int launch_synth_features(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo,
                          const stater_plane *normal, const stater_plane *depth, hipStream_t st);
int launch_synth_samples(stater_ctx *ctx, const stater_scene *s, const stater_plane *albedo,
                         uint32_t first, const stater_plane *first_plane, const stater_samples *out,
                         hipStream_t st);

}  // namespace stater
