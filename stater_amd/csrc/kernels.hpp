// kernels.hpp — kernel argument blocks shared by the single-image launches
// and the batch runtime (which keeps tables of them in device memory).
#pragma once

#include "internal.hpp"

namespace stater {

constexpr int kMaxFeat = 7;

struct AccArgs {
    const float *data;
    int64_t pitch, comp, sstride;
    int spp;
    int sanitize;
    const int32_t *counts;
    int64_t counts_pitch;
    int count_offset;
    int32_t *n;
    float *stats;
    float *lref;    // [3][plane] the reference sample x1 of each pixel-channel (1 if not positive and finite)
    float *mshift;  // [S-1][3][plane] Box-Cox means relative to the reference point
    int ld;
    size_t plane;
    int W, H;
    float lam[kMaxSets];
    float inv_lam[kMaxSets];
    int lam_par[kMaxSets];  // 0: lambda not an integer, 1: even, 2: odd (the sign rule of pow for x < 0)
    float *film_sum;    // [4][plane] the Film's RGB contribSum and weight sum (continued), or null
    int fresh;          // start from an empty set: old n and statistics are neither read nor kept
};

struct FeatArgs {
    const float *ptr[kMaxFeat];
    int64_t pitch[kMaxFeat];
    float scale[kMaxFeat];  // sqrt(-dr log2 e): w = exp2(sp - sum (g'_p - g'_q)^2)
    int nf;
    unsigned bf16;          // bit k: feature k is stored as bfloat16
};

struct VarArgs {
    const int32_t *n;
    const float *m2;     // set-0 base of the m2 statistic
    const float *stats;  // statistics block (the pack pass reads m2, m3, m4, c2X, c2Y)
    float *lnS2, *se2, *m2C, *jb;  // prepass outputs, written by the pack pass
    float *varD, *varCD;
    FeatArgs f;
    int W, H, ld, r, g0;       // g0: first set of this group
    size_t plane, set_stride;  // distance between sets of the statistics block
    float z, z2, sp_scale;     // sp_scale = ds_factor * log2 e
    int lev;                   // c2X, c2Y present (max_moment 5): m2C by LEV
    float4 *rec;               // packed records of the set group (written by the pack pass)
    const float4 *zero;        // one all-zero record
};

struct WelchArgs {
    const int32_t *n;
    const float *mean;     // statistics block base (set s mean = mean + s*set_stride)
    const float *var;      // variance fed to the test, [S][3][plane] (or m2 for the raw test)
    int var_is_m2;         // 1: var = m2/(n-1) computed on the fly from the statistics
    const float *film;     // filtered values and set-0 test mean (planar, 3 ch)
    int64_t film_pitch, film_xs, film_cs;  // bytes: row, pixel (4 planar, 12 interleaved), channel
    const float *jb;
    float *out;            // filmD
    int64_t out_pitch, out_xs, out_cs;
    float *discr;          // [S][3][plane]
    float *varW;           // plain method: mean of accepted s^2 (set 0) -> varD
    uint8_t *dec;          // optional per-pair decisions [P][(2r+1)^2][3]
    uint32_t *bits;        // forward half-window decisions [3][bits_words][plane] (symmetric form)
    int bits_words;        // words per channel in bits
    int band_dy0, band_dy1;  // forward window rows of this pass pair (all of 0 .. r for r <= 10)
    int band_base, band_count;  // forward index of the band's first offset, offsets in the band
    float *part;           // banded form: [6][plane] den[3], num[3] carried from the first band to the second
    float4 *head;          // banded form: per-pixel weighting records (film, features), head4 items each
    int head4;
    float4 *rec;           // pass-2 records (the buffer set's variance-filter record planes)
    const float4 *zero;    // one all-zero record
    int rec_f4, rec_head;  // > 0: pass 1 writes the records (float4 items per record, head items)
    FeatArgs f;
    int W, H, ld, r, mode3;
    size_t plane, set_stride;
    float sp_scale;
    double alpha;
    TCritPoly tp;
};

struct SynthOut {
    float *out;
    int64_t pitch, comp, sstride;
    int spp;
    const int32_t *counts;
    int64_t counts_pitch;
    int count_offset;
    const int32_t *first;
    int64_t first_pitch;
};


// Argument builders of the launches.
AccArgs make_acc_args(stater_bufset *bs, const stater_samples *s);
FeatArgs make_feats(const stater_gbuffers *g);
VarArgs make_var_args(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g, int g0);
WelchArgs make_welch_args(stater_bufset *bs, const stater_denoise_params *p, const stater_gbuffers *g);
SynthOut make_synth_out(const stater_samples *out, const stater_plane *first_plane);

}  // namespace stater
