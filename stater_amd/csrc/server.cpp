// server.cpp — the serving runtime of the statistics-and-denoise path.
//
// Jobs are admitted in arrival order when their memory fits one device arena
// (allocated once: admission never calls hipMalloc or hipFree, which would
// synchronise the device) and a slot is free.  Each slot is a worker thread
// with its own HIP stream and sample chunk; a job runs its rounds there:
// samples from its seeded scene (synthetic features) or from the caller's
// chunk (caller features), accumulation into its resident statistics, then
// the variance and Welch filters, with adaptive-sampling rounds in between
// when asked for (the reference's round driver, statpath.cpp:152-259 and
// :424-441).  A kept job stays resident between requests: a continuation
// accumulates more samples into it and filters again.  Jobs on different slots run
// concurrently and their kernels overlap on the device.  Every job owns its
// buffers and every kernel is deterministic, so a job's results do not depend
// on what runs beside it.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace stater;

namespace {

constexpr size_t kBlock = size_t(2) << 20;  // arena granularity: 2 MiB
size_t round_up(size_t b, size_t a) { return (b + a - 1) / a * a; }

// First-fit allocator over [0, size) with coalescing frees.
class Arena {
   public:
    void init(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = round_up(bytes, kBlock);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, o = it->first;
            free_.erase(it);
            if (rest) free_[o + bytes] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round_up(bytes, kBlock);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<size_t, size_t> free_;  // offset -> length
};

struct JobRec {
    uint64_t id = 0;
    stater_job spec{};
    int status = STATER_OK;
    int slot = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t samples = 0;
    size_t bytes = 0, off = 0;  // arena block while held
    size_t admitted = 0;        // bytes the job was admitted with
    bool done = false;
    stater_bufset *bs = nullptr;
    // the request's samples: the caller's chunk, or the scene's
    stater_samples chunk{};
    bool has_chunk = false;
    // a continuation: the kept job whose state it continues (and does not own)
    JobRec *base = nullptr;
    // state a kept job carries between requests
    bool busy = false;          // a request is running on this job's state
    int conts = 0;              // continuations submitted and not finished
    bool synthetic = true;      // samples from the seeded scene
    stater_scene sc{};
    stater_plane alb{};
    stater_gbuffers g{};        // the filters' features (f32 or bf16 planes)
};

// Memory of one job inside the arena: the buffer set, f32 features (albedo 3,
// normal 3, depth 1 planes) of a synthetic job and, for bf16 storage, the
// bf16 copies the filters read (7 channels).
struct JobLayout {
    size_t set_bytes = 0, feat_off = 0, bf16_off = 0, total = 0;
    int ld = 0;
};

int job_layout(const stater_job &j, JobLayout &L) {
    const int64_t sb =
        stater_bufset_bytes_for(j.width, j.height, j.n_lambdas, j.lambdas, 5, j.film ? STATER_BUFSET_FILM : 0u);
    if (sb < 0) return (int)STATER_E_ARG;
    L.ld = (j.width + 63) / 64 * 64;
    const size_t plane = (size_t)L.ld * j.height;
    L.set_bytes = (size_t)sb;
    L.feat_off = round_up(L.set_bytes, 256);
    L.bf16_off = round_up(L.feat_off + (j.samples ? 0 : 7 * plane * 4), 256);
    L.total = L.bf16_off + (j.feature_dtype == STATER_BF16 ? round_up(7 * plane * 2, 256) : 0);
    return STATER_OK;
}

int check_samples(const stater_samples *d, int W, const char *who) {
    if (!d->data || d->spp < 0 || d->pitch_bytes % 4 || d->comp_stride_bytes % 4 || d->sample_stride_bytes % 4 ||
        d->pitch_bytes < (int64_t)W * 4 || (d->counts && d->counts_pitch_bytes < (int64_t)W * 4))
        return fail(STATER_E_ARG, std::string(who) + ": bad sample chunk");
    return STATER_OK;
}

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct stater_server {
    stater_ctx *ctx = nullptr;
    stater_server_config cfg{};
    char *mem = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<JobRec *> pending;
    std::map<uint64_t, std::unique_ptr<JobRec>> jobs;
    uint64_t next_id = 1;
    int inflight = 0;
    bool stop = false;
    struct Slot {
        hipStream_t st = nullptr;
        char *chunk = nullptr;
        int32_t *host_tgt = nullptr;
        std::thread th;
    };
    std::vector<Slot> slots;
    size_t chunk_bytes = 0;
    std::chrono::steady_clock::time_point t0;
};

namespace {

stater_plane planar(void *ptr, int W, int H, int ch, int ld, int dtype) {
    const int64_t esz = dtype == STATER_BF16 ? 2 : 4;
    stater_plane p{};
    p.ptr = ptr;
    p.pitch_bytes = (int64_t)ld * esz;
    p.channel_stride_bytes = (int64_t)ld * H * esz;
    p.width = W;
    p.height = H;
    p.channels = ch;
    p.dtype = dtype;
    p.layout = STATER_PLANAR;
    return p;
}

#define JOB_CALL(x)               \
    do {                          \
        const int rc_ = (x);      \
        if (rc_) return rc_;      \
    } while (0)

// A new job's state in its arena block: the buffer set and its features
// (generated from the scene, or the caller's planes), converted to bf16 when
// the job stores its features so.
int setup_job(stater_server *s, JobRec *j, hipStream_t st) {
    const stater_job &J = j->spec;
    const int W = J.width, H = J.height;
    JobLayout L;
    JOB_CALL(job_layout(J, L));
    char *m = s->mem + j->off;
    JOB_CALL(stater_bufset_create_in(s->ctx, W, H, J.n_lambdas, J.lambdas, 5, J.film ? STATER_BUFSET_FILM : 0u, m,
                                     L.set_bytes, st, &j->bs));
    const size_t plane = (size_t)L.ld * H;
    stater_gbuffers src{};
    if (j->synthetic) {
        float *feat = (float *)(m + L.feat_off);
        stater_plane alb = planar(feat, W, H, 3, L.ld, STATER_F32);
        stater_plane nrm = planar(feat + 3 * plane, W, H, 3, L.ld, STATER_F32);
        stater_plane dep = planar(feat + 6 * plane, W, H, 1, L.ld, STATER_F32);
        JOB_CALL(stater_scene_init(&j->sc, J.scene_model, J.seed, J.image_id, W, H));
        JOB_CALL(stater_synth_features(s->ctx, &j->sc, &alb, &nrm, &dep, st));
        j->alb = alb;
        const stater_plane *fp[3] = {&alb, &nrm, &dep};
        for (int k = 0; k < 3; ++k) {
            if (!(s->cfg.feature_sds[k] > 0.f)) continue;
            src.planes[src.count] = *fp[k];
            src.dr_factors[src.count] = -0.5f / (s->cfg.feature_sds[k] * s->cfg.feature_sds[k]);
            src.count++;
        }
    } else {
        src = j->g;  // the caller's planes, copied at submission
    }
    j->g = src;
    if (J.feature_dtype == STATER_BF16) {
        uint16_t *b16 = (uint16_t *)(m + L.bf16_off);
        size_t used = 0;
        for (int k = 0; k < src.count; ++k) {
            if (src.planes[k].dtype == STATER_BF16) continue;
            stater_plane fp = planar(b16 + used, W, H, src.planes[k].channels, L.ld, STATER_BF16);
            used += (size_t)src.planes[k].channels * plane;
            JOB_CALL(stater_convert_plane(s->ctx, &src.planes[k], &fp, st));
            j->g.planes[k] = fp;
        }
    }
    return STATER_OK;
}

// Runs one request on the slot's stream: a new job's rounds, or a
// continuation on its kept job's state; returns its status.
int run_job(stater_server *s, JobRec *j, stater_server::Slot &slot) {
    JobRec *o = j->base ? j->base : j;  // the job that owns the state
    const stater_job &J = j->spec, &O = o->spec;
    const int W = O.width, H = O.height;
    hipStream_t st = slot.st;
    if (!j->base) JOB_CALL(setup_job(s, j, st));
    stater_bufset *bs = o->bs;
    const stater_gbuffers *gp = o->g.count ? &o->g : nullptr;
    stater_plane n_view, tgt_view;
    JOB_CALL(stater_bufset_view(bs, "n", &n_view));
    JOB_CALL(stater_bufset_view(bs, "targetspp", &tgt_view));
    const stater_denoise_params &p = s->cfg.params;
    if (j->has_chunk) {
        stater_samples d = j->chunk;
        if (O.film) d.sanitize = 1;
        JOB_CALL(stater_accumulate(bs, &d, s->cfg.acc_mode, st));
        if (!d.counts) {
            j->samples += ((uint64_t)d.spp * W * H);
        } else {  // the samples the counts let in
            STATER_HIP(hipMemcpy2DAsync(slot.host_tgt, (size_t)W * 4, d.counts, (size_t)d.counts_pitch_bytes,
                                        (size_t)W * 4, (size_t)H, hipMemcpyDeviceToHost, st));
            STATER_HIP(hipStreamSynchronize(st));
            uint64_t sum = 0;
            for (size_t i = 0; i < (size_t)W * H; ++i)
                sum += (uint64_t)std::min(std::max(slot.host_tgt[i] - d.count_offset, 0), d.spp);
            j->samples += (sum);
        }
    } else {
        stater_samples d{};
        const size_t plane = (size_t)bs->ld * H;
        d.data = (const float *)slot.chunk;
        d.pitch_bytes = (int64_t)bs->ld * 4;
        d.comp_stride_bytes = (int64_t)plane * 4;
        d.sample_stride_bytes = 5 * d.comp_stride_bytes;
        d.sanitize = O.film ? 1 : 0;
        const int chunk = s->cfg.chunk_spp;
        const uint32_t cap = J.cap_spp ? J.cap_spp : O.cap_spp;
        for (int r = 0; r < J.rounds; ++r) {
            int total = J.spp;
            bool counts = false;
            if (r > 0 || j->base) {  // adaptive round: per-pixel targets from the current estimate
                JOB_CALL(stater_denoise_vars(bs, &p, gp, st));
                JOB_CALL(stater_calc_target_spp(bs, (uint64_t)J.spp, cap, nullptr, st));
                STATER_HIP(hipMemcpy2DAsync(slot.host_tgt, (size_t)W * 4, tgt_view.ptr, (size_t)tgt_view.pitch_bytes,
                                            (size_t)W * 4, (size_t)H, hipMemcpyDeviceToHost, st));
                STATER_HIP(hipStreamSynchronize(st));
                int mx = 0;
                uint64_t sum = 0;
                for (size_t i = 0; i < (size_t)W * H; ++i) {
                    mx = std::max(mx, slot.host_tgt[i]);
                    sum += (uint64_t)std::max(slot.host_tgt[i], 0);
                }
                total = mx;
                counts = true;
                j->samples += (sum);
            } else {
                j->samples += ((uint64_t)J.spp * W * H);
            }
            for (int off = 0; off < total; off += chunk) {
                d.spp = std::min(chunk, total - off);
                d.counts = counts ? (const int32_t *)tgt_view.ptr : nullptr;
                d.counts_pitch_bytes = counts ? tgt_view.pitch_bytes : 0;
                d.count_offset = counts ? off : 0;
                // first = n: every pixel continues its own sample sequence
                JOB_CALL(stater_synth_samples(s->ctx, &o->sc, &o->alb, 0, &n_view, &d, st));
                JOB_CALL(stater_accumulate(bs, &d, s->cfg.acc_mode, st));
            }
        }
    }
    JOB_CALL(stater_denoise_vars(bs, &p, gp, st));
    stater_plane film;
    if (O.film) {
        JOB_CALL(stater_film_resolve(bs, 1.f, nullptr, st));
        JOB_CALL(stater_bufset_view(bs, "film", &film));
    }
    JOB_CALL(stater_denoise_film(bs, &p, gp, O.film ? &film : nullptr, nullptr, st));
    if (J.next_spp > 0)
        JOB_CALL(stater_calc_target_spp(bs, (uint64_t)J.next_spp, J.cap_spp ? J.cap_spp : O.cap_spp, nullptr, st));
    STATER_HIP(hipStreamSynchronize(st));
    return STATER_OK;
}

void finish_locked(stater_server *s, JobRec *j) {
    if (j->base) {  // a continuation: the state stays its kept job's
        j->base->busy = false;
        j->base->conts--;
        j->done = true;
        return;
    }
    j->busy = false;
    if (!j->spec.keep || j->status != STATER_OK) {
        if (j->bs) stater_bufset_destroy(j->bs);
        j->bs = nullptr;
        if (j->bytes) s->arena.release(j->off, j->bytes);
        j->bytes = 0;
    }
    j->done = true;
}

void worker(stater_server *s, int k) {
    (void)hipSetDevice(s->ctx->device);
    auto &slot = s->slots[k];
    for (;;) {
        JobRec *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                if (!s->pending.empty()) {
                    JobRec *h = s->pending.front();
                    size_t off = 0;
                    // a continuation needs no memory, only its kept job's state to itself
                    const bool ok = h->base ? !h->base->busy : s->arena.alloc(h->bytes, &off);
                    if (ok) {
                        if (h->base) h->base->busy = true;
                        else { h->off = off; h->busy = true; }
                        h->admitted = h->bytes;
                        h->slot = k;
                        h->t_admit = now_s(s->t0);
                        s->pending.pop_front();
                        s->inflight++;
                        j = h;
                        break;
                    }
                }
                s->cv_work.wait(lk);
            }
        }
        const int rc = run_job(s, j, slot);
        if (rc) (void)hipStreamSynchronize(slot.st);  // nothing of the job may still run on its memory
        {
            std::lock_guard<std::mutex> lk(s->mu);
            j->status = rc;
            j->t_done = now_s(s->t0);
            finish_locked(s, j);
            s->inflight--;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // memory came free
    }
}

}  // namespace

extern "C" {

void stater_server_default_config(stater_server_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_bytes = (uint64_t)64 << 30;
    c->slots = 4;
    c->chunk_spp = 32;
    c->max_width = 3840;
    c->max_height = 2160;
    c->acc_mode = STATER_ACC_SEQUENTIAL;
    c->feature_sds[0] = 0.02f;
    c->feature_sds[1] = 0.1f;
    c->feature_sds[2] = 0.5f;
    stater_default_params(&c->params);
}

int stater_server_create(stater_ctx *ctx, const stater_server_config *cfg, stater_server **out) {
    if (!ctx || !cfg || !out) return fail(STATER_E_ARG, "stater_server_create: NULL argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 64 || cfg->chunk_spp < 1 || cfg->chunk_spp > 1024 || cfg->max_width < 1 ||
        cfg->max_height < 1 || cfg->max_width > 16384 || cfg->max_height > 16384 || cfg->max_bytes < kBlock ||
        (cfg->acc_mode != STATER_ACC_SEQUENTIAL && cfg->acc_mode != STATER_ACC_MERGE64))
        return fail(STATER_E_ARG, "stater_server_create: bad configuration");
    STATER_HIP(hipSetDevice(ctx->device));
    auto srv = std::make_unique<stater_server>();
    srv->ctx = ctx;
    srv->cfg = *cfg;
    srv->cfg.max_bytes = cfg->max_bytes / kBlock * kBlock;
    const size_t ld = (size_t)(cfg->max_width + 63) / 64 * 64;
    srv->chunk_bytes = (size_t)cfg->chunk_spp * 5 * ld * cfg->max_height * 4;
    auto cleanup = [&](int rc, const std::string &what) {
        for (auto &sl : srv->slots) {
            if (sl.st) (void)hipStreamDestroy(sl.st);
            if (sl.chunk) (void)hipFree(sl.chunk);
            if (sl.host_tgt) (void)hipHostFree(sl.host_tgt);
        }
        if (srv->mem) (void)hipFree(srv->mem);
        (void)hipGetLastError();
        return fail(rc, "stater_server_create: " + what);
    };
    if (hipMalloc((void **)&srv->mem, srv->cfg.max_bytes) != hipSuccess)
        return cleanup(STATER_E_NOMEM, "job arena does not fit");
    srv->arena.init(srv->cfg.max_bytes);
    srv->slots.resize(cfg->slots);
    for (auto &sl : srv->slots) {
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess)
            return cleanup(STATER_E_HIP, "stream creation failed");
        if (hipMalloc((void **)&sl.chunk, srv->chunk_bytes) != hipSuccess)
            return cleanup(STATER_E_NOMEM, "sample chunks do not fit");
        if (hipHostMalloc((void **)&sl.host_tgt, (size_t)cfg->max_width * cfg->max_height * 4) != hipSuccess)
            return cleanup(STATER_E_NOMEM, "pinned host buffers do not fit");
    }
    srv->t0 = std::chrono::steady_clock::now();
    stater_server *s = srv.release();
    for (int k = 0; k < (int)s->slots.size(); ++k) s->slots[k].th = std::thread(worker, s, k);
    ctx->live_bufsets.fetch_add(1);  // the context outlives the server
    *out = s;
    return STATER_OK;
}

int64_t stater_server_job_bytes(const stater_server *s, const stater_job *job) {
    if (!s || !job) return fail(STATER_E_ARG, "stater_server_job_bytes: NULL argument");
    if (job->continue_id) return 0;  // runs on its kept job's memory
    if (job->width > s->cfg.max_width || job->height > s->cfg.max_height)
        return fail(STATER_E_UNSUPPORTED, "stater_server: film larger than the server's maximum");
    if (job->spp < 1 || job->rounds < 1 || job->rounds > 64 ||
        (job->feature_dtype != STATER_F32 && job->feature_dtype != STATER_BF16) || (job->film & ~1))
        return fail(STATER_E_ARG, "stater_server: job needs spp >= 1, 1..64 rounds, f32/bf16 features, film 0/1");
    if (job->samples && job->rounds != 1)
        return fail(STATER_E_ARG, "stater_server: a job with the caller's samples runs one round per request");
    JobLayout L;
    if (job_layout(*job, L)) return STATER_E_ARG;
    return (int64_t)round_up(L.total, kBlock);
}

int stater_server_submit(stater_server *s, const stater_job *job, uint64_t *id) {
    if (!s || !job || !id) return fail(STATER_E_ARG, "stater_server_submit: NULL argument");
    const int64_t bytes = stater_server_job_bytes(s, job);
    if (bytes < 0) return (int)bytes;
    if ((uint64_t)bytes > s->cfg.max_bytes)
        return fail(STATER_E_UNSUPPORTED, "stater_server_submit: job needs more memory than the server's budget");
    auto rec = std::make_unique<JobRec>();
    rec->spec = *job;
    rec->spec.samples = nullptr;  // the copies below are the request's own
    rec->spec.gbufs = nullptr;
    rec->bytes = (size_t)bytes;
    if (job->samples) {
        int rc = check_samples(job->samples, job->width, "stater_server_submit");
        if (rc) return rc;
        rec->chunk = *job->samples;
        rec->has_chunk = true;
    }
    if (!job->continue_id) {
        rec->synthetic = job->samples == nullptr;
        if (job->gbufs) {
            if (rec->synthetic)
                return fail(STATER_E_ARG, "stater_server_submit: a synthetic job's features come from its scene");
            const stater_gbuffers &g = *job->gbufs;
            int total = 0;
            bool ok = g.count >= 0 && g.count <= STATER_MAX_GBUFFERS;
            for (int k = 0; ok && k < g.count; ++k) {
                const stater_plane &pl = g.planes[k];
                ok = pl.ptr && pl.width == job->width && pl.height == job->height && pl.channels >= 1 &&
                     pl.channels <= 3 && (pl.dtype == STATER_F32 || pl.dtype == STATER_BF16) &&
                     pl.layout == STATER_PLANAR && g.dr_factors[k] <= 0.f;
                total += pl.channels;
            }
            if (!ok || total > 7) return fail(STATER_E_ARG, "stater_server_submit: bad feature planes");
            rec->g = g;
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return fail(STATER_E_STATE, "stater_server_submit: server is stopping");
        if (job->continue_id) {
            auto it = s->jobs.find(job->continue_id);
            JobRec *b = it == s->jobs.end() ? nullptr : it->second.get();
            if (!b || b->base || !b->done || !b->bs || b->status != STATER_OK)
                return fail(STATER_E_STATE, "stater_server_submit: continue_id names no finished kept job");
            if (!rec->has_chunk && !b->synthetic)
                return fail(STATER_E_ARG, "stater_server_submit: continuing a caller-sample job needs samples");
            if (rec->has_chunk && rec->chunk.pitch_bytes < (int64_t)b->spec.width * 4)
                return fail(STATER_E_ARG, "stater_server_submit: sample chunk narrower than the kept job's film");
            rec->base = b;
            b->conts++;
        }
        rec->id = s->next_id++;
        rec->t_submit = now_s(s->t0);
        *id = rec->id;
        s->pending.push_back(rec.get());
        s->jobs[rec->id] = std::move(rec);
    }
    s->cv_work.notify_all();
    return STATER_OK;
}

int stater_server_wait(stater_server *s, uint64_t id, stater_job_result *res) {
    if (!s) return fail(STATER_E_ARG, "stater_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(STATER_E_NOTFOUND, "stater_server_wait: unknown job");
    JobRec *j = it->second.get();
    s->cv_done.wait(lk, [&] { return j->done; });
    if (res) {
        res->id = j->id;
        res->status = j->status;
        res->slot = j->slot;
        res->t_submit = j->t_submit;
        res->t_admit = j->t_admit;
        res->t_done = j->t_done;
        res->samples = j->samples;
        res->bytes = j->admitted;
    }
    const int status = j->status;
    if (!j->bs) s->jobs.erase(it);  // nothing kept: the record goes with its result
    if (status) return fail(status, "stater_server: job " + std::to_string(id) + " failed");
    return STATER_OK;
}

int stater_server_view(stater_server *s, uint64_t id, const char *name, stater_plane *out) {
    if (!s || !name || !out) return fail(STATER_E_ARG, "stater_server_view: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || !it->second->done || !it->second->bs || it->second->busy)
        return fail(STATER_E_STATE, "stater_server_view: job not finished, not kept, or being continued");
    return stater_bufset_view(it->second->bs, name, out);
}

int stater_server_release(stater_server *s, uint64_t id) {
    if (!s) return fail(STATER_E_ARG, "stater_server_release: NULL server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end() || !it->second->done) return fail(STATER_E_STATE, "stater_server_release: job not finished");
        JobRec *j = it->second.get();
        if (j->conts || j->busy)
            return fail(STATER_E_STATE, "stater_server_release: a continuation of the job is pending or running");
        if (j->bs) stater_bufset_destroy(j->bs);
        j->bs = nullptr;
        if (j->bytes) s->arena.release(j->off, j->bytes);
        s->jobs.erase(it);
    }
    s->cv_work.notify_all();
    return STATER_OK;
}

int stater_server_drain(stater_server *s) {
    if (!s) return fail(STATER_E_ARG, "stater_server_drain: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
    return STATER_OK;
}

double stater_server_clock(const stater_server *s) { return s ? now_s(s->t0) : -1.0; }

int stater_server_destroy(stater_server *s) {
    if (!s) return STATER_OK;
    stater_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &sl : s->slots)
        if (sl.th.joinable()) sl.th.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto &kv : s->jobs)
        if (kv.second->bs) stater_bufset_destroy(kv.second->bs);
    for (auto &sl : s->slots) {
        (void)hipStreamSynchronize(sl.st);
        (void)hipStreamDestroy(sl.st);
        (void)hipFree(sl.chunk);
        (void)hipHostFree(sl.host_tgt);
    }
    (void)hipFree(s->mem);
    s->ctx->live_bufsets.fetch_sub(1);
    delete s;
    return STATER_OK;
}

}  // extern "C"
