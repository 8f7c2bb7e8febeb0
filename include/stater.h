/*
 * stater.h — C-ABI of the MI355X (gfx950) statistics-and-denoise library.
 *
 * This is the drop-in boundary for StatER's statistics-buffer and denoise path.
 * Every entry point is plain C: device pointers, sizes and an int status.
 * No exception crosses it.  Each call enqueues its kernels on the HIP stream it
 * is given (NULL = the context's own stream) and returns without waiting,
 * like the reference's calls on a cv::cuda::Stream.
 *
 * Reference interfaces replaced (paths relative to the StatER checkout):
 *   stater_create / stater_destroy          cv::cuda::stat_denoiser::setup()
 *                                           (src/statistics/estimator.h:383)
 *   stater_print_macro_configuration        stat_denoiser::printMacroConfiguration()
 *                                           (src/statistics/estimator.cpp:519-521)
 *   stater_bufset_create / _view            Estimator::AllocateBuffers + BUF_LIST /
 *                                           FILM_BUF_LIST names
 *                                           (estimator.cpp:83-245, estimator.h:459-479,
 *                                           name scheme estimator.cpp:21-24)
 *   stater_accumulate                       StatTile<T>::Add[Transform]SampleM1..M4C
 *                                           (estimator.h:201-324) fed by
 *                                           statpath.cpp:313-349, followed by
 *                                           Estimator::Merge[Transform]Tiles
 *                                           (estimator.cpp:267-343) and Upload
 *                                           (estimator.cpp:346-353): the moments
 *                                           are built directly in device memory
 *   stater_denoise_vars                     Estimator::DenoiseVars ->
 *                                           stat_denoiser::denoiseVars<float3>
 *                                           (estimator.cpp:365-416)
 *   stater_denoise_film                     Estimator::DenoiseMeans ->
 *                                           stat_denoiser::denoiseFilm
 *                                           (estimator.cpp:419-459)
 *   stater_calc_target_spp                  Estimator::CalcTargetSPPs ->
 *                                           stat_denoiser::calcTargetSPP<float3>
 *                                           (estimator.cpp:498-513)
 *   stater_calc_mean_vars                   Estimator::CalcMeanVars / calcMatMeanVar
 *                                           (estimator.cpp:462-496)
 *   stater_synchronize                      stat_denoiser::synchronize(stream)
 *                                           (estimator.cpp:515-517)
 *
 * Error convention: 0 = ok, negative = STATER_E_*; stater_last_error() gives a
 * thread-local message for the last failing call on the calling thread.
 * Threading: calls on distinct buffer sets are thread-safe; calls on one
 * buffer set must be ordered by the caller (one stream per set).
 */
#ifndef STATER_H
#define STATER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STATER_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define STATER_OK             0
#define STATER_E_ARG         -1  /* invalid argument (size, pointer, name) */
#define STATER_E_HIP         -2  /* a HIP runtime call failed */
#define STATER_E_NOMEM       -3  /* device allocation failed */
#define STATER_E_STATE       -4  /* call not valid in the object's state */
#define STATER_E_UNSUPPORTED -5  /* configuration not supported */
#define STATER_E_NOTFOUND    -6  /* unknown buffer name */

/* ---- limits ------------------------------------------------------------ */
#define STATER_MAX_LAMBDAS   8   /* Box-Cox lambda slots per buffer set */
#define STATER_MAX_RADIUS   20   /* filter radius (reference default is 20) */
#define STATER_MAX_GBUFFERS  4   /* feature buffers per filter */
#define STATER_MAX_EDGES    16   /* synthetic-scene step edges */

typedef struct stater_ctx stater_ctx;
typedef struct stater_bufset stater_bufset;

/* STATER_BF16: bfloat16 storage (feature planes only). */
typedef enum { STATER_F32 = 0, STATER_I32 = 1, STATER_BF16 = 2 } stater_dtype;
typedef enum { STATER_PLANAR = 0, STATER_INTERLEAVED = 1 } stater_layout;

/* A 2-D image plane in device memory.  PLANAR: channel c starts at
 * ptr + c*channel_stride_bytes.  INTERLEAVED: channels are adjacent
 * (the reference's CV_32FC3 GpuMat, a PtrStepSzb with step = pitch_bytes). */
typedef struct {
    void   *ptr;
    int64_t pitch_bytes;
    int64_t channel_stride_bytes;
    int32_t width;
    int32_t height;
    int32_t channels;
    int32_t dtype;   /* stater_dtype */
    int32_t layout;  /* stater_layout */
    int32_t reserved;
} stater_plane;

/* ---- context ----------------------------------------------------------- */
#define STATER_CTX_TIMING 1u  /* record HIP events around every kernel launch */

int  stater_create(int device, unsigned flags, stater_ctx **out);
int  stater_destroy(stater_ctx *ctx);
/* The context's own HIP stream (hipStream_t), created at default priority. */
int  stater_ctx_stream(stater_ctx *ctx, void **stream_out);
const char *stater_last_error(void);
int  stater_abi_version(void);
/* Prints the denoiser-side configuration (the reference's macro dump). */
void stater_print_macro_configuration(void);

/* Per-kernel timing (contexts created with STATER_CTX_TIMING).  Reading the
 * statistics synchronises the recorded events. */
int  stater_timing_reset(stater_ctx *ctx);
int  stater_timing_count(stater_ctx *ctx);
int  stater_timing_get(stater_ctx *ctx, int index, char *name, int name_len,
                       uint64_t *launches, double *total_ms);

/* ---- buffer sets (the Estimator's statistics buffers, device resident) ---
 * One set holds, for a W x H RGB film: the sample count n (i32) and, for the
 * untransformed samples (set 0) and every distinct Box-Cox lambda != 1
 * (sets 1..T), the statistics mean, m2, m3, m4, c2X, c2Y, m1 plus the derived
 * jb, lnS2, se2, m2C, varD, varCD, discr; film-level meanVar, meanD, filmD and
 * targetspp.  All planes are PLANAR f32 with a 256-byte aligned row pitch.
 *
 * max_moment follows the reference's maxMoment (statpath.cpp:1154):
 * 1 = mean, 2 = +m2, 3 = +m3, 4 = +m4, 5 = +m4 with c2X, c2Y and m1 ("M4C").
 * lambdas are the reference's "lambdas" list (scenes/render-denoise-mt5.pbrt:26);
 * n_lambdas = 0 means no transform (the reference's transform = false). */
#define STATER_BUFSET_NO_DERIVED 1u  /* accumulate-only set: skip denoise planes */
#define STATER_BUFSET_FILM       2u  /* also keep the Film's pixel sums (film = Film::UpdateImage) */

int stater_bufset_create(stater_ctx *ctx, int width, int height,
                         int n_lambdas, const float *lambdas, int max_moment,
                         unsigned flags, stater_bufset **out);
/* The same set over caller memory (mem_bytes >= stater_bufset_bytes_for(...),
 * 256-byte aligned, owned by the caller and kept until the set is destroyed):
 * allocates nothing on the device; n and the moments are zeroed on the stream.
 * For callers that pool device memory (the serving runtime's arena). */
int stater_bufset_create_in(stater_ctx *ctx, int width, int height,
                            int n_lambdas, const float *lambdas, int max_moment,
                            unsigned flags, void *mem, uint64_t mem_bytes,
                            void *stream, stater_bufset **out);
int stater_bufset_destroy(stater_bufset *bs);
/* Zero the accumulated statistics (n and every moment) on the stream. */
int stater_bufset_reset(stater_bufset *bs, void *stream);
/* Look up a buffer by the reference's name: "n", "film", "filmD", "targetspp",
 * "t<type>-b<j>-film-<stat>" (untransformed statistics of lambda slot j) or
 * "t<type>-b<j>-<stat>" (Box-Cox statistics of slot j), type the set's stat
 * type (0 unless stater_bufset_set_stat_type changed it), stat in
 * mean m2 m3 m4 c2X c2Y m1 jb lnS2 se2 m2C varD varCD discr meanVar meanD n.
 * "film" is the resolved Film with STATER_BUFSET_FILM, else the set-0 mean. */
int stater_bufset_view(const stater_bufset *bs, const char *name, stater_plane *out);
/* The set's stat type, the reference's StatTypeIndex (src/statistics/
 * statpath.h:28-37): 0 Radiance (the default), 1/2 MIS win rates,
 * 3 material id, 4 depth, 5 normal, 6 albedo, 7 iteration radiance.  Its
 * buffers are named "t<type>-b<j>-..." (estimator.cpp:21-22): a G-buffer set
 * (no lambdas, max_moment 1, fed one feature value per sample as
 * statpath.cpp:366-379 does) of type 5 answers to "t5-b0-film-mean", the
 * normal buffer the reference's denoiser reads.  Types 1..4 are
 * single-channel (StatTile<Float>): their views hold channel 0. */
int stater_bufset_set_stat_type(stater_bufset *bs, int type);
/* Copy a plane into another of the same shape, converting the layout
 * (PLANAR <-> INTERLEAVED, e.g. a caller's CV_32FC3 film) and the storage type
 * (f32 <-> bf16, round to nearest even).  Feature planes may be bf16. */
int stater_convert_plane(stater_ctx *ctx, const stater_plane *src, const stater_plane *dst, void *stream);
/* Bytes of device memory the set holds (the admission unit for serving). */
int64_t stater_bufset_bytes(const stater_bufset *bs);
/* Bytes a set of this shape would hold, without allocating. */
int64_t stater_bufset_bytes_for(int width, int height, int n_lambdas,
                                const float *lambdas, int max_moment, unsigned flags);

/* ---- film (pbrt Film fed the same samples) ------------------------------
 * With STATER_BUFSET_FILM every stater_accumulate call also adds its samples
 * to the Film's pixel sums as the reference's render loop does: the film
 * tiles live for the whole render (statpath.cpp:144), so each pixel's RGB
 * contribSum and weight sum run over all its (sanitised) samples in order,
 * across calls (FilmTile::AddSample, box filter of radius 1/2: weight 1).
 * The reference clears the Film and merges each tile once per iteration,
 * which leaves xyz = 0 + RGBToXYZ(contribSum).  stater_film_resolve is that
 * and then Film::UpdateImage (src/core/film.cpp:188-220, no splats): RGB =
 * max(0, XYZToRGB(xyz) / weight sum) * scale, into out (NULL: the set's
 * "film" plane, which is then what denoise_film's film argument should be).
 * STATER_ACC_FRESH and stater_bufset_reset start the sums from zero. */
int stater_film_resolve(stater_bufset *bs, float scale, const stater_plane *out, void *stream);

/* ---- on-disk buffer interchange (PFM) ------------------------------------
 * The reference's file format: OpenCV imwrite/imread of each buffer as
 * <stem>-<spp>-<name>.pfm (src/statistics/buffer.cpp:44-57 writes,
 * src/statistics/statpath.cpp:479-547 reads them in --denoise mode).  "PF"
 * (RGB) or "Pf" (1 channel), rows bottom to top, little-endian floats;
 * integer planes are stored as floats and rounded to nearest when read.
 * Host arrays are top row first, channels interleaved.  Plane copies are
 * synchronous with the stream. */
int stater_pfm_info(const char *path, int *width, int *height, int *channels);
int stater_pfm_read_host(const char *path, float *data, int width, int height, int channels);
int stater_pfm_write_host(const char *path, const float *data, int width, int height, int channels);
/* Read a file into / write a device plane (f32, i32 or bf16; 1 or 3 channels). */
int stater_pfm_read(const char *path, const stater_plane *dst, void *stream);
int stater_pfm_write(const char *path, const stater_plane *src, void *stream);
/* OutputBufferSelection::Write (buffer.cpp:44-57) for a set: every buffer
 * whose reference name matches the ECMAScript regex (whole name; NULL = all)
 * goes to <stem>-<spp>-<name>.pfm.  Names: film, filmD, t0-b<j>-film-<stat>,
 * t0-b<j>-<stat> (see stater_bufset_view); "film" is the set-0 sample mean. */
int stater_bufset_write_pfm(const stater_bufset *bs, const char *stem, uint64_t spp, const char *regex,
                            void *stream, int *written);
/* The loader of StatPathIntegrator::Denoise (statpath.cpp:505-545) for a set:
 * every <stem>-<spp>-<name>.pfm whose name is one of the set's buffers is
 * read into it (in name order).  film.pfm is left to the caller (the
 * reference reads it into its Film, the input of denoise_film), as are
 * other types' buffers (t4/t5/t6 G-buffers). */
int stater_bufset_read_pfm(stater_bufset *bs, const char *stem, uint64_t spp, void *stream, int *nread);

/* ---- accumulation -----------------------------------------------------
 * A sample chunk: for sample index j (0..spp-1) and component k (R,G,B,u,v)
 * the plane starts at data + j*sample_stride_bytes + k*comp_stride_bytes, with
 * rows pitch_bytes apart.  (u,v) is the in-pixel position in [0,1)^2
 * (statpath.cpp:339).  counts (optional, i32 plane) limits pixel p to its
 * first clamp(counts[p] - count_offset, 0, spp) samples of the chunk, so a
 * per-pixel target (e.g. targetspp) can be consumed in passes of spp. */
typedef struct {
    const float   *data;
    int64_t        pitch_bytes;
    int64_t        comp_stride_bytes;
    int64_t        sample_stride_bytes;
    int32_t        spp;
    int32_t        sanitize;   /* NaN / negative / inf guard of statpath.cpp:313-332 */
    const int32_t *counts;
    int64_t        counts_pitch_bytes;
    int32_t        count_offset;
    int32_t        reserved;
} stater_samples;

/* Sequential update in the reference's exact FP32 operation order for the
 * untransformed statistics (bit-identical to estimator.h:278-305). */
#define STATER_ACC_SEQUENTIAL 0
/* Per-call partial moments merged into the resident ones in FP64
 * (Chan-Golub-LeVeque / Pebay); c2X, c2Y, m1 continue sequentially. */
#define STATER_ACC_MERGE64    1
/* OR'ed into the mode: the call starts from an empty set, with the result of
 * stater_bufset_reset on the same stream followed by the call (for every
 * pixel of the film, and the Film sums with STATER_BUFSET_FILM), without the
 * separate clearing pass or the read of the old statistics. */
#define STATER_ACC_FRESH      16

int stater_accumulate(stater_bufset *bs, const stater_samples *samples,
                      int mode, void *stream);

/* ---- denoising ---------------------------------------------------------- */
typedef struct {
    int32_t radius;       /* window radius (filterRadius), <= STATER_MAX_RADIUS */
    float   ds_factor;    /* spatial factor -0.5/filterSD^2; 0 = uniform (estimator.h:353) */
    float   var_z;        /* variance-test z (varZValue, default 2.80703) */
    float   alpha;        /* two-sided Welch level for the mean filter (0.005) */
    int32_t mt_mode;      /* 2: accept iff every lambda set accepts;
                             3: test only in the set with the smallest jb */
    int32_t cascade;      /* 1: Welch uses the denoised variance (varD/varCD) */
    int32_t use_levc;     /* 1: cascade uses the LEV-corrected varCD */
    int32_t reserved;
} stater_denoise_params;

/* Feature (G-) buffers for the range weights exp(dr * |g_p - g_q|^2),
 * dr = -0.5/sd^2 (estimator.cpp:10-18). */
typedef struct {
    int32_t      count;
    int32_t      reserved;
    stater_plane planes[STATER_MAX_GBUFFERS];
    float        dr_factors[STATER_MAX_GBUFFERS];
} stater_gbuffers;

void stater_default_params(stater_denoise_params *p);

/* Per-pixel Jarque-Bera, Bonett lnS2/se2, LEV m2C, then the variance filter
 * varD/varCD for every set.  gbufs may be NULL. */
int stater_denoise_vars(stater_bufset *bs, const stater_denoise_params *p,
                        const stater_gbuffers *gbufs, void *stream);
/* Welch-test mean filter.  film: the mean to filter and to test set 0 with
 * (NULL = the set-0 sample mean).  film_out: NULL = the set's filmD plane.
 * Both may be PLANAR or INTERLEAVED f32 (the reference's CV_32FC3 film and
 * filmD, estimator.cpp:421-457), read and written in place.
 * Writes discr (per set) and filmD. */
int stater_denoise_film(stater_bufset *bs, const stater_denoise_params *p,
                        const stater_gbuffers *gbufs, const stater_plane *film,
                        const stater_plane *film_out, void *stream);
/* The plain method: Welch test on the raw sample variances, uniform weights,
 * no features; writes meanD (the mean over accepted neighbours) and varD of
 * set 0 (the mean of their sample variances). */
int stater_denoise_plain(stater_bufset *bs, int radius, float alpha, void *stream);

/* Adaptive-sampling map: distributes budget_spp * W * H samples over the
 * pixels in proportion to the relative standard error of the mean,
 * sqrt(varD/n)/max(|mean|,1e-3) averaged over RGB, with n + target <= cap.
 * The targets sum to the budget exactly unless every pixel is capped.
 * out: NULL = the set's targetspp plane. */
int stater_calc_target_spp(stater_bufset *bs, uint64_t budget_spp, uint32_t cap,
                           const stater_plane *out, void *stream);
/* The same map with the variance it is built from chosen by as_mode, the
 * reference's AS_MODE build flag (scripts/alternate-builds/build-asn+mcl.sh:5,
 * README.md:140-154; calcTargetSPP receives both m2 and varD,
 * estimator.cpp:502-511): 1 ("ASM", the default, what stater_calc_target_spp
 * does) = the denoised variance varD of set 0; 0 ("ASN") = the pixel's own
 * sample variance m2/(n-1). */
int stater_calc_target_spp_ex(stater_bufset *bs, uint64_t budget_spp, uint32_t cap, int as_mode,
                              const stater_plane *out, void *stream);
/* The adaptive-sampling map of stater_calc_target_spp on caller planes,
 * allocating nothing.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  n (i32, 1 channel), varD and mean
 * (f32, 3 planar channels) and out (i32) are W x H device planes; the result
 * equals stater_calc_target_spp on the same values.  Its scratch is a static
 * device array per GPU (one control block and 256-pixel block partials,
 * 512 KiB + 512 B, held until the process ends), whose address is taken when
 * the context is created: films above STATER_NOALLOC_MAX_PIXELS are rejected
 * with STATER_E_ARG.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (different streams or threads) race: the caller orders
 * them. */
#define STATER_NOALLOC_MAX_PIXELS (4096 * 4096)
int stater_calc_target_spp_noalloc(stater_ctx *ctx, int width, int height, const stater_plane *n,
                                   const stater_plane *varD, const stater_plane *mean,
                                   uint64_t budget_spp, uint32_t cap, const stater_plane *out, void *stream);
/* meanVar = m2 / ((n-1) n) of set 0 (estimator.cpp:462-474). */
int stater_calc_mean_vars(stater_bufset *bs, void *stream);
int stater_synchronize(void *stream);

/* ---- synthetic scenes (workload generator, not the hot path) -------------
 * Seeded, counter-based (Philox4x32-10) sample model: every draw is keyed by
 * (seed, image_id) and addressed by (pixel, sample index), so any launch
 * shape produces the same samples. */
typedef struct {
    int32_t  width, height;
    uint32_t seed, image_id;
    int32_t  model;          /* 1 = plain (128x128 case), 2 = features + edges */
    int32_t  n_edges;
    double   base;
    double   amp[3][3];      /* [sinusoid][channel] */
    double   fx[3], fy[3];
    double   cos_phi[3][3], sin_phi[3][3];
    int32_t  edge_vertical[STATER_MAX_EDGES];
    double   edge_pos[STATER_MAX_EDGES];
    double   edge_factor[STATER_MAX_EDGES];
    double   s_min, s_max;
    double   firefly_mult;
    uint32_t firefly_thresh; /* firefly iff a 32-bit draw < thresh */
    int32_t  albedo_modulation;
    int32_t  cell_size;
    int32_t  pad0;
    double   normal_jitter;
    double   depth_min, depth_max;
} stater_scene;

int stater_scene_init(stater_scene *s, int model, uint32_t seed, uint32_t image_id,
                      int width, int height);
/* Albedo (3 ch), normal (3 ch) and depth (1 ch) planes, PLANAR f32. */
int stater_synth_features(stater_ctx *ctx, const stater_scene *s,
                          const stater_plane *albedo, const stater_plane *normal,
                          const stater_plane *depth, void *stream);
/* Writes samples first_sample + first[p] + j (j = 0..count-1) of every pixel
 * into a chunk laid out as stater_samples describes; count is spp, or the
 * clamped per-pixel count of out->counts.  first (optional i32 plane, e.g. a
 * set's "n" view) continues each pixel's sample sequence.  albedo may be NULL
 * unless the scene modulates by albedo. */
int stater_synth_samples(stater_ctx *ctx, const stater_scene *s,
                         const stater_plane *albedo, uint32_t first_sample,
                         const stater_plane *first, const stater_samples *out,
                         void *stream);

/* ---- batches --------------------------------------------------------------
 * Several films of uneven size (and lambda count) processed together: each
 * film's chain (reset, passes of chunk_spp generated samples + accumulation,
 * prepass, variance filter, Welch filter) is spread with the others over a
 * few HIP streams (longest first, least loaded stream), so small films'
 * launches fill the device beside large films' tails.  use_graph captures the
 * whole fan-out/fan-in once into a hipGraph and replays it.  The buffers each
 * film ends with are bit-identical to running it on its own
 * (stater_accumulate + stater_denoise_vars + stater_denoise_film).  A set
 * created with STATER_BUFSET_FILM runs the renderer's loop: guarded samples
 * (sanitize = 1), stater_film_resolve(scale 1) after the last pass, and the
 * resolved Film as the Welch filter's film (statpath.cpp:313-336, 411, 428-432).
 * Replaces nothing in the reference (it renders one film per process). */
typedef struct stater_batch stater_batch;
typedef struct {
    stater_bufset         *set;     /* full set (max_moment >= 4) of the film's size */
    const stater_scene    *scene;   /* seeded sample source (copied), used when samples is NULL */
    const stater_plane    *albedo;  /* the scene's albedo plane when it modulates */
    const stater_gbuffers *gbufs;   /* range-weight features or NULL (copied) */
    int32_t                spp;     /* samples per pixel of this image */
    int32_t                reserved;
    const stater_samples  *samples; /* the caller's samples (a renderer's chunk, copied; its device
                                       memory must stay valid while the batch runs): the first spp
                                       of them are taken in passes of chunk_spp, counts and
                                       count_offset as for stater_accumulate.  NULL: generated
                                       from scene */
} stater_batch_image;

int stater_batch_create(stater_ctx *ctx, int count, const stater_batch_image *images,
                        const stater_denoise_params *p, int chunk_spp, stater_batch **out);
/* Times the filters' tile heights (2, 4, 8 rows) on every film and keeps the
 * fastest per film; drops a captured graph. */
int stater_batch_tune(stater_batch *b, void *stream);
int stater_batch_run(stater_batch *b, int use_graph, void *stream);
/* image < 0: returns the number of films; else the film's set count, stream
 * (branch) and chosen tile heights. */
int stater_batch_info(const stater_batch *b, int image, int *sets, int *branch,
                      int *tile_rows_var, int *tile_rows_welch);
int stater_batch_destroy(stater_batch *b);

/* ---- serving ------------------------------------------------------------
 * A request loop over one context: jobs (a film of any size and lambda count,
 * one round or several adaptive-sampling rounds) are submitted from any
 * thread; each is admitted when its device memory fits the server's budget
 * (FIFO) and a slot is free.  Every slot is a worker with its own HIP stream,
 * so admitted jobs run concurrently and their kernels overlap on the device;
 * a job's statistics stay resident across its rounds, and a kept job stays
 * resident across requests: a continuation (continue_id) adds samples to it
 * and filters again.  Job memory comes from one arena allocated at creation:
 * admission never calls hipMalloc/hipFree.  A job's buffers are the same
 * whether it runs alone or among others.
 * Samples come from the job's seeded scene or from the caller (a renderer's
 * chunk, device memory that must stay valid until the request is waited
 * for).  The caller's adaptive loop: submit a kept job with its first chunk
 * and next_spp, wait, read "targetspp" (stater_server_view), render that many
 * samples per pixel, submit a continuation whose chunk takes the job's
 * "targetspp" plane as its counts, and so on; release the job at the end.
 * Replaces nothing in the reference (it renders one film per process); it is
 * the serving form of the statistics-and-denoise path. */
typedef struct stater_server stater_server;
typedef struct {
    uint64_t max_bytes;       /* device memory for job state (the admission budget) */
    int32_t  slots;           /* jobs in flight, one worker thread and HIP stream each */
    int32_t  chunk_spp;       /* samples generated per pass */
    int32_t  max_width, max_height;  /* largest film (sizes the per-slot sample chunks) */
    int32_t  acc_mode;        /* STATER_ACC_SEQUENTIAL or STATER_ACC_MERGE64 */
    float    feature_sds[3];  /* range sds of albedo, normal, depth (<= 0: feature unused) */
    stater_denoise_params params;
    int32_t  reserved[4];
} stater_server_config;

typedef struct {
    int32_t  width, height;
    int32_t  n_lambdas;
    float    lambdas[STATER_MAX_LAMBDAS];
    int32_t  spp;             /* samples per pixel of round 1, per-pixel budget of later rounds */
    int32_t  rounds;          /* 1: one shot; > 1: adaptive-sampling rounds (calc_target_spp) */
    uint32_t cap_spp;         /* per-pixel cap of the adaptive rounds */
    int32_t  scene_model;     /* synthetic scene (stater_scene_init model) */
    uint32_t seed, image_id;
    int32_t  feature_dtype;   /* STATER_F32 or STATER_BF16 storage of the filters' features */
    int32_t  keep;            /* 1: keep the buffers after completion (stater_server_view) */
    int32_t  film;            /* 1: the renderer's loop: guarded samples, the set's pbrt Film
                                 (STATER_BUFSET_FILM) resolved and used as the Welch
                                 filter's film; 0: the radiance mean is the film */
    int32_t  next_spp;        /* > 0: after the filters, the AS map of a next round of
                                 next_spp x pixels samples (cap cap_spp) into "targetspp" */
    uint64_t continue_id;     /* 0: a new job.  Else the id of a finished kept job whose
                                 resident state this request continues: its size, lambdas,
                                 film and features are the kept job's; with samples the
                                 chunk is accumulated, without them `rounds` adaptive rounds
                                 of `spp` x pixels samples are generated from its scene
                                 (synthetic jobs); then the filters run again */
    const stater_samples  *samples;  /* the caller's samples of this request (copied; one
                                        accumulate call over the chunk, counts honoured), or
                                        NULL: generated from the seeded scene */
    const stater_gbuffers *gbufs;    /* a new caller-sample job's features (copied; planes of
                                        the film's size kept valid while the job is kept), or
                                        NULL: no range weights */
    int32_t  reserved[2];
} stater_job;

typedef struct {
    uint64_t id;
    int32_t  status;          /* STATER_OK or the status of the failing call */
    int32_t  slot;
    double   t_submit, t_admit, t_done;  /* seconds on the server clock */
    uint64_t samples;         /* samples accumulated over all rounds */
    uint64_t bytes;           /* admitted device memory */
} stater_job_result;

void stater_server_default_config(stater_server_config *cfg);
int  stater_server_create(stater_ctx *ctx, const stater_server_config *cfg, stater_server **out);
/* Device memory a job is admitted with (E_ARG / E_UNSUPPORTED as negative). */
int64_t stater_server_job_bytes(const stater_server *s, const stater_job *job);
int  stater_server_submit(stater_server *s, const stater_job *job, uint64_t *id);
/* Blocks until the job has finished; fills res. */
int  stater_server_wait(stater_server *s, uint64_t id, stater_job_result *res);
/* A finished job's buffer by name (jobs submitted with keep = 1; a
 * continuation's results are its kept job's buffers). */
int  stater_server_view(stater_server *s, uint64_t id, const char *name, stater_plane *out);
/* Frees a kept job's buffers (STATER_E_STATE while a continuation of it is
 * pending or running). */
int  stater_server_release(stater_server *s, uint64_t id);
/* Blocks until every submitted job has finished. */
int  stater_server_drain(stater_server *s);
double stater_server_clock(const stater_server *s);
int  stater_server_destroy(stater_server *s);

/* ---- testing ------------------------------------------------------------
 * Runs the Welch mean filter (plain != 0: the plain method) and also writes
 * every accept/reject decision to dec (device, uint8):
 * dec[((y*W + x)*(2r+1)^2 + (dy+r)*(2r+1) + (dx+r))*3 + c]. */
int stater_debug_film_decisions(stater_bufset *bs, const stater_denoise_params *p,
                                const stater_gbuffers *gbufs, const stater_plane *film,
                                int plain, uint8_t *dec, void *stream);
/* Checks the fast exact forms of the accumulation's untransformed update
 * against the IEEE ones over all 2^32 float bit patterns x (device counts[3],
 * written by the call): counts[0] = non-negative x (-0 and NaN included)
 * outside (0, 2^-104) whose neighbour-checked square root differs from
 * sqrtf(x); counts[1] = finite x with |x| >= 2^-40 or x = 0 whose quotient
 * by a pseudo-random n in [1, 2^24] through the correctly rounded FP32
 * reciprocal differs from x / n; counts[2] = floats m in [1, 2^32] whose
 * reciprocal (one Newton step from the hardware one) differs from 1.f / m.
 * All three are 0 on a working build; no reference counterpart (it checks
 * this build's arithmetic). */
int stater_debug_exact_forms(stater_ctx *ctx, unsigned long long *counts, void *stream);
/* Checks the accumulation's quotient form (a product with the correctly
 * rounded reciprocal and one residual correction) against the IEEE x / m for
 * every significand x in [1, 2) and every integer m in [m_lo, m_hi)
 * (1 <= m_lo < m_hi <= 2^24 + 1): device counts[0] receives the number of
 * differences, 0 on a working build.  m_lo = 2^23, m_hi = 2^24 + 1 covers
 * every pair of significands (2^46 quotients, some seconds); no reference
 * counterpart. */
int stater_debug_quotient_sweep(stater_ctx *ctx, uint32_t m_lo, uint32_t m_hi, unsigned long long *counts,
                                void *stream);
/* VALU issue probe: blocks x 256 threads, each running 16 independent FP32
 * FMA chains (register operands only) for iters steps (16 x iters FMAs per
 * thread); out (device,
 * blocks x 256 floats) receives the chains' sums.  Timing it measures the
 * non-packed FP32 instruction rate the kernels' rooflines are priced
 * against; no reference counterpart. */
int stater_debug_valu_rate(stater_ctx *ctx, int blocks, int iters, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* STATER_H */
