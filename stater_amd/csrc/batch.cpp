// batch.cpp — films of uneven size processed together.
//
// Every film's chain (passes of chunk_spp samples -- generated from the film's
// seeded scene, or taken from the caller's sample chunk -- each followed by
// accumulation, the first from an empty set; prepass and variance filter;
// Welch filter) runs as the single-film launches, and the chains are spread
// over a few HIP streams (longest chain first, onto the least loaded
// stream), so the launches of small films fill the device beside the large
// films' tails and nothing waits for the host.  With use_graph the fan-out
// and fan-in are captured once into a hipGraph and replayed.
// stater_batch_tune times the filters' tile heights (2, 4 and 8 rows) on
// every film and keeps the fastest for it.
//
// Each film runs the same kernels in the same order as on its own, so a batch
// leaves bit-identical buffers.  A set created with STATER_BUFSET_FILM runs
// the renderer's loop (src/statistics/statpath.cpp:313-336, 388, 411,
// 428-432): samples pass the NaN / negative / infinity guard, the pbrt Film is
// built beside the statistics, resolved after the last pass, and is the Welch
// filter's film.

#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.hpp"

using namespace stater;

namespace {

constexpr int kBranches = 4;
constexpr int kVarRowOptions[2] = {4, 8};        // variance filter: output rows per tile
constexpr int kWelchRowOptions[3] = {2, 4, 8};  // Welch filter

}  // namespace

struct stater_batch {
    stater_ctx *ctx = nullptr;
    stater_denoise_params p{};
    int chunk_spp = 32;
    struct Img {
        stater_bufset *bs;
        stater_scene sc;
        stater_samples samples;  // the caller's chunk (has_samples) or unused
        bool has_samples;
        stater_plane albedo;
        bool has_alb;
        stater_gbuffers g;
        bool has_g;
        stater_plane film;  // resolved pbrt Film (STATER_BUFSET_FILM sets)
        bool has_film;
        int spp;
        int branch;
        int ty_var, ty_welch;
        double cost;
    };
    std::vector<Img> imgs;
    std::vector<std::vector<int>> order;  // films of each branch, in launch order
    std::vector<hipStream_t> streams;
    std::vector<char *> chunks;           // one sample chunk per branch
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

stater_samples chunk_desc(const stater_batch *b, const stater_batch::Img &I, int spp) {
    stater_samples s{};
    s.data = (const float *)b->chunks[I.branch];
    s.pitch_bytes = (int64_t)I.bs->ld * 4;
    s.comp_stride_bytes = (int64_t)I.bs->plane * 4;
    s.sample_stride_bytes = 5 * s.comp_stride_bytes;
    s.spp = spp;
    s.sanitize = I.has_film ? 1 : 0;
    return s;
}

const stater_gbuffers *gb(const stater_batch::Img &I) { return I.has_g ? &I.g : nullptr; }
const stater_plane *film(const stater_batch::Img &I) { return I.has_film ? &I.film : nullptr; }

int welch(const stater_batch *b, const stater_batch::Img &I, hipStream_t st, int ty) {
    return launch_welch_filter(I.bs, &b->p, gb(I), film(I), nullptr, false, st, ty);
}

// Pass [off, off + spp) of the caller's chunk: its data from sample off on,
// and its per-pixel counts (if any) shifted by off.
stater_samples caller_pass(const stater_batch::Img &I, int off, int spp) {
    stater_samples s = I.samples;
    s.data = (const float *)((const char *)I.samples.data + (int64_t)off * I.samples.sample_stride_bytes);
    s.spp = spp;
    s.count_offset = I.samples.count_offset + off;
    if (I.has_film) s.sanitize = 1;
    return s;
}

// One film's chain on stream st.
int enqueue_film(const stater_batch *b, const stater_batch::Img &I, hipStream_t st) {
    int rc = STATER_OK;
    for (int off = 0; off < I.spp && !rc; off += b->chunk_spp) {
        const int k = std::min(b->chunk_spp, I.spp - off);
        const stater_samples s = I.has_samples ? caller_pass(I, off, k) : chunk_desc(b, I, k);
        if (!I.has_samples)
            rc = launch_synth_samples(b->ctx, &I.sc, I.has_alb ? &I.albedo : nullptr, (uint32_t)off, nullptr, &s, st);
        // the first pass starts the film from an empty set (no clearing pass)
        if (!rc) rc = launch_accumulate(I.bs, &s, STATER_ACC_SEQUENTIAL | (off == 0 ? STATER_ACC_FRESH : 0), st);
    }
    if (!rc) rc = launch_var_filter(I.bs, &b->p, gb(I), st, I.ty_var);
    if (!rc && I.has_film) rc = launch_film_resolve(I.bs, 1.f, &I.film, st);
    if (!rc) rc = welch(b, I, st, I.ty_welch);
    return rc;
}

int enqueue(stater_batch *b, hipStream_t st) {
    STATER_HIP(hipEventRecord(b->fork, st));
    for (size_t k = 0; k < b->streams.size(); ++k) STATER_HIP(hipStreamWaitEvent(b->streams[k], b->fork, 0));
    for (size_t k = 0; k < b->streams.size(); ++k)
        for (int i : b->order[k]) {
            const int rc = enqueue_film(b, b->imgs[i], b->streams[k]);
            if (rc) return rc;
        }
    for (size_t k = 0; k < b->streams.size(); ++k) {
        STATER_HIP(hipEventRecord(b->joins[k], b->streams[k]));
        STATER_HIP(hipStreamWaitEvent(st, b->joins[k], 0));
    }
    return STATER_OK;
}

void drop_graph(stater_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void destroy(stater_batch *b) {
    drop_graph(b);
    for (auto s : b->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (auto c : b->chunks) (void)hipFree(c);
    for (auto e : b->joins) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
}

}  // namespace

extern "C" {

int stater_batch_create(stater_ctx *ctx, int count, const stater_batch_image *images, const stater_denoise_params *p,
                        int chunk_spp, stater_batch **out) {
    if (!ctx || !images || !p || !out || count <= 0) return fail(STATER_E_ARG, "stater_batch_create: bad arguments");
    *out = nullptr;
    if (p->radius < 0 || p->radius > STATER_MAX_RADIUS || !(p->alpha > 0.f && p->alpha < 1.f) || !(p->var_z > 0.f) ||
        !(p->ds_factor <= 0.f) || (p->mt_mode != 2 && p->mt_mode != 3))
        return fail(STATER_E_ARG, "stater_batch_create: bad denoise parameters");
    if (chunk_spp <= 0 || chunk_spp > 1024) return fail(STATER_E_ARG, "stater_batch_create: chunk_spp in 1..1024");
    STATER_HIP(hipSetDevice(ctx->device));
    auto *b = new stater_batch();
    b->ctx = ctx;
    b->p = *p;
    b->chunk_spp = chunk_spp;
    for (int i = 0; i < count; ++i) {
        const stater_batch_image &im = images[i];
        const bool set_ok = im.set && im.spp > 0 && !(im.set->flags & STATER_BUFSET_NO_DERIVED) &&
                            im.set->ctx == ctx && im.set->mm >= 4;
        const bool src_ok =
            im.samples ? (im.samples->data && im.spp <= im.samples->spp && im.samples->pitch_bytes % 4 == 0 &&
                          im.samples->comp_stride_bytes % 4 == 0 && im.samples->sample_stride_bytes % 4 == 0 &&
                          im.samples->pitch_bytes >= (int64_t)im.set->W * 4)
                       : (im.scene && im.scene->width == im.set->W && im.scene->height == im.set->H &&
                          (!im.scene->albedo_modulation || (im.albedo && im.albedo->ptr)));
        if (!set_ok || !src_ok) {
            delete b;
            return fail(STATER_E_ARG, "stater_batch_create: image " + std::to_string(i) +
                                          " needs a full set (max_moment >= 4) and a scene of its size or a "
                                          "sample chunk of at least spp samples");
        }
        stater_batch::Img I{};
        I.bs = im.set;
        I.has_samples = im.samples != nullptr;
        if (I.has_samples) I.samples = *im.samples;
        else I.sc = *im.scene;
        I.has_alb = im.albedo != nullptr;
        if (I.has_alb) I.albedo = *im.albedo;
        I.has_g = im.gbufs != nullptr;
        if (I.has_g) I.g = *im.gbufs;
        I.spp = im.spp;
        I.has_film = im.set->film_sum != nullptr;
        if (I.has_film) (void)stater_bufset_view(im.set, "film", &I.film);
        I.ty_var = kVarTileRows;
        I.ty_welch = 4;
        // work estimate: samples x sets (generation + accumulation) + window tests
        const double P = (double)im.set->W * im.set->H, S = im.set->S;
        const double win = (2.0 * p->radius + 1) * (2.0 * p->radius + 1);
        I.cost = P * ((double)im.spp * (2.0 + S) + 0.05 * win * S * 3.0);
        b->imgs.push_back(I);
    }
    // longest chain first, each onto the least loaded branch
    const int nb = std::min(kBranches, count);
    std::vector<int> idx(count);
    for (int i = 0; i < count; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return b->imgs[a].cost > b->imgs[c].cost; });
    std::vector<double> load(nb, 0.0);
    b->order.assign(nb, {});
    for (int i : idx) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->imgs[i].branch = k;
        load[k] += b->imgs[i].cost;
        b->order[k].push_back(i);
    }
    auto fail_out = [&](int rc, const std::string &what) {
        destroy(b);
        (void)hipGetLastError();
        return fail(rc, "stater_batch_create: " + what);
    };
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return fail_out(STATER_E_HIP, "event creation failed");
    for (int k = 0; k < nb; ++k) {
        size_t need = 0;
        for (int i : b->order[k])
            if (!b->imgs[i].has_samples) need = std::max(need, (size_t)chunk_spp * 5 * b->imgs[i].bs->plane * 4);
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        char *c = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return fail_out(STATER_E_HIP, "stream");
        b->streams.push_back(s);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail_out(STATER_E_HIP, "event");
        b->joins.push_back(e);
        if (hipMalloc((void **)&c, std::max<size_t>(need, 256)) != hipSuccess)
            return fail_out(STATER_E_NOMEM, "sample chunks do not fit");
        b->chunks.push_back(c);
    }
    *out = b;
    return STATER_OK;
}

int stater_batch_destroy(stater_batch *b) {
    if (!b) return STATER_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    destroy(b);
    return STATER_OK;
}

int stater_batch_run(stater_batch *b, int use_graph, void *stream) {
    if (!b) return fail(STATER_E_ARG, "stater_batch_run: NULL batch");
    hipStream_t st = stream ? (hipStream_t)stream : b->ctx->stream;
    if (!use_graph) return enqueue(b, st);
    if (!b->exec) {
        STATER_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(b, st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        STATER_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    STATER_HIP(hipGraphLaunch(b->exec, st));
    return STATER_OK;
}

int stater_batch_tune(stater_batch *b, void *stream) {
    if (!b) return fail(STATER_E_ARG, "stater_batch_tune: NULL batch");
    hipStream_t st = stream ? (hipStream_t)stream : b->ctx->stream;
    int rc = enqueue(b, st);  // statistics to filter
    if (rc) return rc;
    hipEvent_t e0, e1;
    STATER_HIP(hipEventCreate(&e0));
    STATER_HIP(hipEventCreate(&e1));
    auto time_one = [&](auto &&launch, float *ms) {
        int r = 0;
        for (int rep = 0; rep < 2 && !r; ++rep) {  // the first launch warms up
            (void)hipEventRecord(e0, st);
            r = launch();
            (void)hipEventRecord(e1, st);
            if (r) break;
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(ms, e0, e1);
        }
        return r;
    };
    for (auto &I : b->imgs) {
        const bool hasw = I.has_g || b->p.ds_factor != 0.f;
        if (!hasw || b->p.radius > 10) continue;  // only 4-row tiles are built there
        float best_v = 1e30f, best_w = 1e30f;
        for (int ty : kVarRowOptions) {
            float ms = 0.f;
            rc = time_one([&] { return launch_var_filter(I.bs, &b->p, gb(I), st, ty); }, &ms);
            if (rc) break;
            if (ms < best_v) { best_v = ms; I.ty_var = ty; }
        }
        for (int ty : kWelchRowOptions) {
            if (rc || b->p.mt_mode != 2) break;
            float ms = 0.f;
            rc = time_one([&] { return welch(b, I, st, ty); }, &ms);
            if (rc) break;
            if (ms < best_w) { best_w = ms; I.ty_welch = ty; }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(b);  // re-capture with the chosen shapes
    return rc;
}

int stater_batch_info(const stater_batch *b, int image, int *sets, int *branch, int *tile_rows_var,
                      int *tile_rows_welch) {
    if (!b) return fail(STATER_E_ARG, "stater_batch_info: NULL batch");
    if (image < 0) return (int)b->imgs.size();
    if (image >= (int)b->imgs.size()) return fail(STATER_E_ARG, "stater_batch_info: no such image");
    const auto &I = b->imgs[image];
    if (sets) *sets = I.bs->S;
    if (branch) *branch = I.branch;
    if (tile_rows_var) *tile_rows_var = I.ty_var;
    if (tile_rows_welch) *tile_rows_welch = I.ty_welch;
    return STATER_OK;
}

}  // extern "C"
