// batch.cpp — uneven images through persistent launches, captured in a hipGraph.
//
// A batch owns, per image, a device chunk of chunk_spp samples and, per
// (sets, max moment, feature weights) group, tables of kernel argument
// blocks in device memory.  One run is: reset -> for every pass of chunk_spp
// samples: generate (all images, one persistent launch) + accumulate (one
// persistent launch per group) -> prepass, variance filter and Welch filter
// (persistent launches per group).  Every persistent launch walks a work list
// of tiles over all images of its group with the XCD-aware static schedule
// of common.hpp, so small and large images share the chip without tails.
// The filters' tile height (wavefronts per tile) is chosen per group by
// stater_batch_tune, which times the built candidates on the batch itself.
// With use_graph the whole run is captured once and replayed.
//
// Each pixel's arithmetic is the single-image path's (same tile bodies, same
// per-pixel loop order), so a batch produces bit-identical buffers.

#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.hpp"

using namespace stater;

namespace {

// Host staging of all descriptor tables; uploaded once into one allocation.
struct Arena {
    std::vector<char> host;
    size_t add(const void *p, size_t bytes) {
        const size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + bytes);
        std::memcpy(host.data() + off, p, bytes);
        return off;
    }
};

constexpr int kTyOptions[3] = {2, 4, 8};

}  // namespace

struct stater_batch {
    stater_ctx *ctx = nullptr;
    stater_denoise_params p{};
    int chunk_spp = 32;
    int passes = 0;
    struct Img {
        stater_bufset *bs;
        stater_scene sc;
        stater_plane albedo;
        bool has_alb;
        stater_gbuffers g;
        bool has_g;
        int spp;
        size_t chunk_off;
    };
    std::vector<Img> imgs;
    char *chunk_mem = nullptr;
    char *dev = nullptr;  // descriptor tables
    // generator, per pass
    std::vector<size_t> synth_jobs, synth_prefix;
    std::vector<int> synth_grid;
    int nimg = 0;
    struct Group {
        int S, mm;
        bool hasw;
        std::vector<int> members;
        std::vector<size_t> acc_args, acc_prefix;  // per pass
        std::vector<int> acc_grid;
        size_t pre_args = 0, pre_prefix = 0;
        int pre_grid = 0;
        std::vector<int> g0s, Gs;
        std::vector<size_t> var_args;              // per g0
        size_t welch_args = 0;
        size_t tile_prefix[3] = {0, 0, 0};         // per tile-height option
        int var_grid[3][3] = {};                   // [g0 index][ty option]
        int welch_grid[3] = {};
        int ty_var = 1, ty_welch = 1;              // index into kTyOptions (4 rows)
    };
    std::vector<Group> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <class T>
const T *devp(const stater_batch *b, size_t off) { return (const T *)(b->dev + off); }

int grid_for(stater_ctx *ctx, const void *k, int block, size_t lds, int ntiles) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, block, lds) != hipSuccess || occ < 1) occ = 1;
    int g = std::min(occ * ctx->num_cu, std::max(ntiles, 1));
    g = (g + kXcds - 1) / kXcds * kXcds;
    return std::max(g, kXcds);
}

bool ty_built(bool hasw, int ti) { return hasw || kTyOptions[ti] == 4; }

int enqueue(stater_batch *b, hipStream_t st) {
    for (auto &im : b->imgs) {
        int rc = launch_zero(im.bs->base, (1 + (size_t)im.bs->S * N_STATS * 3) * im.bs->plane * 4, st);
        if (rc) return rc;
    }
    for (int k = 0; k < b->passes; ++k) {
        int rc = launch_samples_batch(devp<SampleJob>(b, b->synth_jobs[k]), devp<int>(b, b->synth_prefix[k]), b->nimg,
                                      b->synth_grid[k], st);
        if (rc) return rc;
        for (auto &g : b->groups) {
            if (!g.acc_grid[k]) continue;
            rc = launch_accumulate_batch(g.S, g.mm, false, devp<AccArgs>(b, g.acc_args[k]), devp<int>(b, g.acc_prefix[k]),
                                         (int)g.members.size(), g.acc_grid[k], st);
            if (rc) return rc;
        }
    }
    for (auto &g : b->groups) {
        const int m = (int)g.members.size();
        int rc = launch_prepass_batch(devp<PreArgs>(b, g.pre_args), devp<int>(b, g.pre_prefix), m, g.pre_grid, st);
        if (rc) return rc;
        for (size_t i = 0; i < g.g0s.size(); ++i) {
            rc = launch_var_batch(g.Gs[i], g.hasw, b->p.radius, kTyOptions[g.ty_var], devp<VarArgs>(b, g.var_args[i]),
                                  devp<int>(b, g.tile_prefix[g.ty_var]), m, g.var_grid[i][g.ty_var], st);
            if (rc) return rc;
        }
        rc = launch_welch_batch(g.S, g.hasw, b->p.radius, kTyOptions[g.ty_welch], devp<WelchArgs>(b, g.welch_args),
                                devp<int>(b, g.tile_prefix[g.ty_welch]), m, g.welch_grid[g.ty_welch], st);
        if (rc) return rc;
    }
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}

void drop_graph(stater_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

int stater_batch_create(stater_ctx *ctx, int count, const stater_batch_image *images, const stater_denoise_params *p,
                        int chunk_spp, stater_batch **out) {
    if (!ctx || !images || !p || !out || count <= 0) return fail(STATER_E_ARG, "stater_batch_create: bad arguments");
    *out = nullptr;
    if (p->radius < 0 || p->radius > 10)
        return fail(STATER_E_UNSUPPORTED, "stater_batch_create: batched filters are built for radius <= 10");
    if (p->mt_mode != 2) return fail(STATER_E_UNSUPPORTED, "stater_batch_create: batches run mode 2");
    if (!(p->alpha > 0.f && p->alpha < 1.f) || !(p->var_z > 0.f) || !(p->ds_factor <= 0.f))
        return fail(STATER_E_ARG, "stater_batch_create: bad denoise parameters");
    if (chunk_spp <= 0 || chunk_spp > 1024) return fail(STATER_E_ARG, "stater_batch_create: chunk_spp in 1..1024");
    STATER_HIP(hipSetDevice(ctx->device));
    auto *b = new stater_batch();
    b->ctx = ctx;
    b->p = *p;
    b->chunk_spp = chunk_spp;
    b->nimg = count;
    size_t chunk_total = 0;
    int max_spp = 0;
    for (int i = 0; i < count; ++i) {
        const stater_batch_image &im = images[i];
        if (!im.set || !im.scene || im.spp <= 0 || (im.set->flags & STATER_BUFSET_NO_DERIVED) || im.set->ctx != ctx ||
            im.set->mm < 4 || im.scene->width != im.set->W || im.scene->height != im.set->H ||
            (im.scene->albedo_modulation && (!im.albedo || !im.albedo->ptr))) {
            delete b;
            return fail(STATER_E_ARG, "stater_batch_create: image " + std::to_string(i) +
                                          " needs a full set (max_moment >= 4) of its scene's size and a scene");
        }
        stater_batch::Img I{};
        I.bs = im.set;
        I.sc = *im.scene;
        I.has_alb = im.albedo != nullptr;
        if (I.has_alb) I.albedo = *im.albedo;
        I.has_g = im.gbufs != nullptr;
        if (I.has_g) I.g = *im.gbufs;
        I.spp = im.spp;
        I.chunk_off = chunk_total;
        chunk_total += (size_t)chunk_spp * 5 * im.set->plane * 4;
        max_spp = std::max(max_spp, im.spp);
        b->imgs.push_back(I);
    }
    b->passes = (max_spp + chunk_spp - 1) / chunk_spp;
    if (hipMalloc((void **)&b->chunk_mem, chunk_total) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return fail(STATER_E_NOMEM, "stater_batch_create: sample chunks do not fit");
    }

    Arena ar;
    auto samples_of = [&](const stater_batch::Img &I, int k) {
        stater_samples s{};
        s.data = (const float *)(b->chunk_mem + I.chunk_off);
        s.pitch_bytes = (int64_t)I.bs->ld * 4;
        s.comp_stride_bytes = (int64_t)I.bs->plane * 4;
        s.sample_stride_bytes = 5 * s.comp_stride_bytes;
        s.spp = std::max(0, std::min(chunk_spp, I.spp - k * chunk_spp));
        return s;
    };
    // generator tables
    for (int k = 0; k < b->passes; ++k) {
        std::vector<SampleJob> jobs(count);
        std::vector<int> prefix(count + 1, 0);
        for (int i = 0; i < count; ++i) {
            const auto &I = b->imgs[i];
            stater_samples s = samples_of(I, k);
            SampleJob &j = jobs[i];
            j.s = I.sc;
            j.alb = I.has_alb ? (const float *)I.albedo.ptr : nullptr;
            j.alb_pitch = I.has_alb ? I.albedo.pitch_bytes : 0;
            j.alb_cs = I.has_alb ? I.albedo.channel_stride_bytes : 0;
            j.first0 = (uint32_t)(k * chunk_spp);
            j.o = make_synth_out(&s, nullptr);
            prefix[i + 1] = prefix[i] + (s.spp > 0 ? ((I.bs->W + 255) / 256) * I.bs->H : 0);
        }
        b->synth_jobs.push_back(ar.add(jobs.data(), jobs.size() * sizeof(SampleJob)));
        b->synth_prefix.push_back(ar.add(prefix.data(), prefix.size() * sizeof(int)));
        b->synth_grid.push_back(std::max(kXcds, std::min(4 * ctx->num_cu, (prefix[count] + 7) / 8 * 8)));
    }
    // groups by (sets, max moment, weights)
    for (int i = 0; i < count; ++i) {
        const auto &I = b->imgs[i];
        const FeatArgs f = make_feats(I.has_g ? &I.g : nullptr);
        const bool hasw = f.nf > 0 || p->ds_factor != 0.f;
        auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const stater_batch::Group &g) {
            return g.S == I.bs->S && g.mm == I.bs->mm && g.hasw == hasw;
        });
        if (it == b->groups.end()) {
            stater_batch::Group g;
            g.S = I.bs->S;
            g.mm = I.bs->mm;
            g.hasw = hasw;
            b->groups.push_back(g);
            it = b->groups.end() - 1;
        }
        it->members.push_back(i);
    }
    for (auto &g : b->groups) {
        const int m = (int)g.members.size();
        for (int k = 0; k < b->passes; ++k) {
            std::vector<AccArgs> args(m);
            std::vector<int> prefix(m + 1, 0);
            for (int j = 0; j < m; ++j) {
                const auto &I = b->imgs[g.members[j]];
                stater_samples s = samples_of(I, k);
                args[j] = make_acc_args(I.bs, &s);
                prefix[j + 1] = prefix[j] + (s.spp > 0 ? ((I.bs->W + 63) / 64) * I.bs->H : 0);
            }
            g.acc_args.push_back(ar.add(args.data(), args.size() * sizeof(AccArgs)));
            g.acc_prefix.push_back(ar.add(prefix.data(), prefix.size() * sizeof(int)));
            g.acc_grid.push_back(prefix[m] ? grid_for(ctx, accumulate_batch_kernel(g.S, g.mm, false), 192, 0, prefix[m])
                                           : 0);
        }
        std::vector<PreArgs> pre(m);
        std::vector<int> pprefix(m + 1, 0);
        for (int j = 0; j < m; ++j) {
            const auto &I = b->imgs[g.members[j]];
            pre[j] = make_pre_args(I.bs, p->var_z);
            pprefix[j + 1] = pprefix[j] + ((I.bs->W + 255) / 256) * I.bs->H;
        }
        g.pre_args = ar.add(pre.data(), pre.size() * sizeof(PreArgs));
        g.pre_prefix = ar.add(pprefix.data(), pprefix.size() * sizeof(int));
        g.pre_grid = std::max(kXcds, std::min(4 * ctx->num_cu, (pprefix[m] + 7) / 8 * 8));
        for (int g0 = 0; g0 < g.S;) {
            const int G = std::min(3, g.S - g0);
            std::vector<VarArgs> va(m);
            for (int j = 0; j < m; ++j) {
                const auto &I = b->imgs[g.members[j]];
                va[j] = make_var_args(I.bs, p, I.has_g ? &I.g : nullptr, g0);
            }
            g.g0s.push_back(g0);
            g.Gs.push_back(G);
            g.var_args.push_back(ar.add(va.data(), va.size() * sizeof(VarArgs)));
            g0 += G;
        }
        std::vector<WelchArgs> wa(m);
        for (int j = 0; j < m; ++j) {
            const auto &I = b->imgs[g.members[j]];
            wa[j] = make_welch_args(I.bs, p, I.has_g ? &I.g : nullptr);
        }
        g.welch_args = ar.add(wa.data(), wa.size() * sizeof(WelchArgs));
        for (int ti = 0; ti < 3; ++ti) {
            const int ty = kTyOptions[ti];
            std::vector<int> tprefix(m + 1, 0);
            for (int j = 0; j < m; ++j) {
                const auto &I = b->imgs[g.members[j]];
                tprefix[j + 1] = tprefix[j] + ((I.bs->W + 63) / 64) * ((I.bs->H + ty - 1) / ty);
            }
            g.tile_prefix[ti] = ar.add(tprefix.data(), tprefix.size() * sizeof(int));
            if (!ty_built(g.hasw, ti)) continue;
            for (size_t gi = 0; gi < g.Gs.size(); ++gi) {
                size_t lds = 0;
                const void *k = var_batch_kernel(g.Gs[gi], g.hasw, p->radius, ty, &lds, 0);
                (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                g.var_grid[gi][ti] = grid_for(ctx, k, 64 * ty, lds, tprefix[m]);
            }
            size_t lds = 0;
            const void *k = welch_batch_kernel(g.S, g.hasw, p->radius, ty, &lds);
            (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            g.welch_grid[ti] = grid_for(ctx, k, 64 * ty, lds, tprefix[m]);
        }
    }
    if (hipMalloc((void **)&b->dev, std::max<size_t>(ar.host.size(), 256)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(b->chunk_mem);
        delete b;
        return fail(STATER_E_NOMEM, "stater_batch_create: descriptor tables do not fit");
    }
    hipError_t e = hipMemcpy(b->dev, ar.host.data(), ar.host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(b->dev);
        (void)hipFree(b->chunk_mem);
        delete b;
        return hip_fail(e, "hipMemcpy");
    }
    *out = b;
    return STATER_OK;
}

int stater_batch_destroy(stater_batch *b) {
    if (!b) return STATER_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    (void)hipFree(b->dev);
    (void)hipFree(b->chunk_mem);
    delete b;
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
    for (auto &g : b->groups) {
        const int m = (int)g.members.size();
        float best_v = 1e30f, best_w = 1e30f;
        for (int ti = 0; ti < 3; ++ti) {
            if (!ty_built(g.hasw, ti)) continue;
            const int ty = kTyOptions[ti];
            for (int rep = 0; rep < 2; ++rep) {  // first launch warms up
                (void)hipEventRecord(e0, st);
                for (size_t i = 0; i < g.g0s.size() && !rc; ++i)
                    rc = launch_var_batch(g.Gs[i], g.hasw, b->p.radius, ty, devp<VarArgs>(b, g.var_args[i]),
                                          devp<int>(b, g.tile_prefix[ti]), m, g.var_grid[i][ti], st);
                (void)hipEventRecord(e1, st);
                if (rc) break;
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep && ms < best_v) { best_v = ms; g.ty_var = ti; }
            }
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                (void)hipEventRecord(e0, st);
                rc = launch_welch_batch(g.S, g.hasw, b->p.radius, ty, devp<WelchArgs>(b, g.welch_args),
                                        devp<int>(b, g.tile_prefix[ti]), m, g.welch_grid[ti], st);
                (void)hipEventRecord(e1, st);
                if (rc) break;
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep && ms < best_w) { best_w = ms; g.ty_welch = ti; }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(b);  // re-capture with the chosen shapes
    return rc;
}

int stater_batch_info(const stater_batch *b, int group, int *sets, int *images, int *tile_rows_var,
                      int *tile_rows_welch) {
    if (!b) return fail(STATER_E_ARG, "stater_batch_info: NULL batch");
    if (group < 0) return (int)b->groups.size();
    if (group >= (int)b->groups.size()) return fail(STATER_E_ARG, "stater_batch_info: no such group");
    const auto &g = b->groups[group];
    if (sets) *sets = g.S;
    if (images) *images = (int)g.members.size();
    if (tile_rows_var) *tile_rows_var = kTyOptions[g.ty_var];
    if (tile_rows_welch) *tile_rows_welch = kTyOptions[g.ty_welch];
    return STATER_OK;
}

}  // extern "C"
