// pfm.cpp — on-disk buffer interchange in the reference's format.
//
// The reference writes every selected buffer with OpenCV's imwrite as
// <stem>-<spp>-<name>.pfm (src/statistics/buffer.cpp:44-57; three-channel
// buffers are converted RGB -> BGR first, which OpenCV's PFM encoder stores
// as RGB again) and its --denoise mode reads them back by name
// (src/statistics/statpath.cpp:479-547, file name parsing at :515-545).  A
// PFM file (Portable Float Map) is a text header "PF" (3 channels, RGB) or
// "Pf" (1 channel), then width and height, then a scale whose sign gives the
// byte order (negative: little-endian), each followed by white space (one
// single white-space character after the scale), then float32 samples row by
// row from the bottom row up, channels interleaved.  Integer buffers (the
// sample count n) are stored as floats and rounded to nearest on reading,
// as OpenCV's convertTo does.

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <regex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace stater;

namespace {

struct PfmHeader {
    int w = 0, h = 0, ch = 0;
    bool little = true;
};

// Reads one white-space separated token; `last` is the character after it.
bool read_token(FILE *f, std::string &tok, int &last) {
    tok.clear();
    int c = std::fgetc(f);
    while (c == ' ' || c == '\t' || c == '\n' || c == '\r') c = std::fgetc(f);
    while (c != EOF && c != ' ' && c != '\t' && c != '\n' && c != '\r') {
        tok.push_back((char)c);
        if (tok.size() > 64) return false;
        c = std::fgetc(f);
    }
    last = c;
    return !tok.empty();
}

int read_header(FILE *f, const std::string &path, PfmHeader &hd) {
    std::string t;
    int last = 0;
    if (!read_token(f, t, last) || (t != "PF" && t != "Pf")) return fail(STATER_E_ARG, path + ": not a PFM file");
    hd.ch = t == "PF" ? 3 : 1;
    if (!read_token(f, t, last)) return fail(STATER_E_ARG, path + ": truncated PFM header");
    hd.w = std::atoi(t.c_str());
    if (!read_token(f, t, last)) return fail(STATER_E_ARG, path + ": truncated PFM header");
    hd.h = std::atoi(t.c_str());
    if (!read_token(f, t, last) || last == EOF) return fail(STATER_E_ARG, path + ": truncated PFM header");
    const double scale = std::atof(t.c_str());
    if (hd.w <= 0 || hd.h <= 0 || hd.w > 65536 || hd.h > 65536 || scale == 0.0 || !std::isfinite(scale))
        return fail(STATER_E_ARG, path + ": bad PFM header");
    hd.little = scale < 0.0;
    return STATER_OK;
}

bool host_little_endian() {
    const uint32_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

void swap_bytes(float *v, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        uint32_t u;
        std::memcpy(&u, &v[i], 4);
        u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
        std::memcpy(&v[i], &u, 4);
    }
}

// float -> bfloat16, round to nearest even (NaN stays NaN).
uint16_t to_bf16(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float from_bf16(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

int check_plane(const stater_plane *p, const char *what) {
    if (!p || !p->ptr) return fail(STATER_E_ARG, std::string(what) + ": NULL plane");
    if (p->width <= 0 || p->height <= 0 || (p->channels != 1 && p->channels != 3))
        return fail(STATER_E_ARG, std::string(what) + ": PFM holds 1 or 3 channels");
    if (p->dtype != STATER_F32 && p->dtype != STATER_I32 && p->dtype != STATER_BF16)
        return fail(STATER_E_ARG, std::string(what) + ": unknown dtype");
    return STATER_OK;
}

size_t elem_bytes(int dtype) { return dtype == STATER_BF16 ? 2 : 4; }

// Device plane -> host floats, top row first, channels interleaved.
int download(const stater_plane *p, std::vector<float> &out, hipStream_t st) {
    const int W = p->width, H = p->height, C = p->channels;
    const size_t eb = elem_bytes(p->dtype);
    std::vector<unsigned char> raw((size_t)W * H * C * eb);
    if (p->layout == STATER_INTERLEAVED) {
        STATER_HIP(hipMemcpy2DAsync(raw.data(), (size_t)W * C * eb, p->ptr, (size_t)p->pitch_bytes,
                                    (size_t)W * C * eb, H, hipMemcpyDeviceToHost, st));
    } else {
        for (int c = 0; c < C; ++c)
            STATER_HIP(hipMemcpy2DAsync(raw.data() + (size_t)c * W * H * eb, (size_t)W * eb,
                                        (const char *)p->ptr + c * p->channel_stride_bytes, (size_t)p->pitch_bytes,
                                        (size_t)W * eb, H, hipMemcpyDeviceToHost, st));
    }
    STATER_HIP(hipStreamSynchronize(st));
    out.resize((size_t)W * H * C);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < C; ++c) {
                const size_t src = p->layout == STATER_INTERLEAVED ? ((size_t)y * W + x) * C + c
                                                                  : ((size_t)c * H + y) * W + x;
                float v;
                if (p->dtype == STATER_F32) {
                    std::memcpy(&v, raw.data() + src * 4, 4);
                } else if (p->dtype == STATER_I32) {
                    int32_t i;
                    std::memcpy(&i, raw.data() + src * 4, 4);
                    v = (float)i;
                } else {
                    uint16_t b;
                    std::memcpy(&b, raw.data() + src * 2, 2);
                    v = from_bf16(b);
                }
                out[((size_t)y * W + x) * C + c] = v;
            }
    return STATER_OK;
}

// Host floats (top row first, interleaved) -> device plane.
int upload(const std::vector<float> &in, const stater_plane *p, hipStream_t st) {
    const int W = p->width, H = p->height, C = p->channels;
    const size_t eb = elem_bytes(p->dtype);
    std::vector<unsigned char> raw((size_t)W * H * C * eb);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < C; ++c) {
                const float v = in[((size_t)y * W + x) * C + c];
                const size_t dst = p->layout == STATER_INTERLEAVED ? ((size_t)y * W + x) * C + c
                                                                  : ((size_t)c * H + y) * W + x;
                if (p->dtype == STATER_F32) {
                    std::memcpy(raw.data() + dst * 4, &v, 4);
                } else if (p->dtype == STATER_I32) {
                    const int32_t i = std::isfinite(v) ? (int32_t)std::lrint(v) : 0;
                    std::memcpy(raw.data() + dst * 4, &i, 4);
                } else {
                    const uint16_t b = to_bf16(v);
                    std::memcpy(raw.data() + dst * 2, &b, 2);
                }
            }
    if (p->layout == STATER_INTERLEAVED) {
        STATER_HIP(hipMemcpy2DAsync(p->ptr, (size_t)p->pitch_bytes, raw.data(), (size_t)W * C * eb,
                                    (size_t)W * C * eb, H, hipMemcpyHostToDevice, st));
    } else {
        for (int c = 0; c < C; ++c)
            STATER_HIP(hipMemcpy2DAsync((char *)p->ptr + c * p->channel_stride_bytes, (size_t)p->pitch_bytes,
                                        raw.data() + (size_t)c * W * H * eb, (size_t)W * eb, (size_t)W * eb, H,
                                        hipMemcpyHostToDevice, st));
    }
    STATER_HIP(hipStreamSynchronize(st));
    return STATER_OK;
}

hipStream_t stream_of(void *stream) { return (hipStream_t)stream; }

// The reference's buffer names a set holds (src/statistics/estimator.cpp:21-24,
// 101-134; estimator.h:459-479): the film and its denoised form, then for
// every lambda slot j the untransformed buffers t<type>-bj-film-<name> and,
// with lambdas, the transformed buffers t<type>-bj-<name> (type: the set's
// stat type, 0 for radiance, 3..6 for the G-buffers).
std::vector<std::string> set_names(const stater_bufset *bs) {
    static const char *base[] = {"mean", "m2", "m3", "m4", "c2X", "c2Y", "m1"};
    static const char *derived[] = {"jb", "lnS2", "se2", "m2C", "varD", "varCD", "discr"};
    static const char *film_only[] = {"meanVar", "meanD"};
    const bool der = !(bs->flags & STATER_BUFSET_NO_DERIVED);
    std::vector<std::string> out{"film"};
    if (der) out.push_back("filmD");
    const int nslots = std::max(1, bs->n_lambdas);
    for (int j = 0; j < nslots; ++j) {
        const std::string pre = "t" + std::to_string(bs->stat_type) + "-b" + std::to_string(j) + "-";
        for (auto n : base) out.push_back(pre + "film-" + n);
        if (der)
            for (auto n : derived) out.push_back(pre + "film-" + n);
        out.push_back(pre + "film-n");
        if (der)
            for (auto n : film_only) out.push_back(pre + "film-" + n);
        if (bs->n_lambdas > 0) {
            for (auto n : base) out.push_back(pre + n);
            if (der)
                for (auto n : derived) out.push_back(pre + n);
        }
    }
    return out;
}

}  // namespace

extern "C" {

int stater_pfm_info(const char *path, int *width, int *height, int *channels) {
    if (!path) return fail(STATER_E_ARG, "stater_pfm_info: NULL path");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(STATER_E_NOTFOUND, std::string(path) + ": " + std::strerror(errno));
    PfmHeader hd;
    const int rc = read_header(f, path, hd);
    std::fclose(f);
    if (rc) return rc;
    if (width) *width = hd.w;
    if (height) *height = hd.h;
    if (channels) *channels = hd.ch;
    return STATER_OK;
}

int stater_pfm_read_host(const char *path, float *data, int width, int height, int channels) {
    if (!path || !data) return fail(STATER_E_ARG, "stater_pfm_read_host: NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(STATER_E_NOTFOUND, std::string(path) + ": " + std::strerror(errno));
    PfmHeader hd;
    int rc = read_header(f, path, hd);
    if (!rc && (hd.w != width || hd.h != height || hd.ch != channels))
        rc = fail(STATER_E_ARG, std::string(path) + ": PFM is " + std::to_string(hd.w) + "x" + std::to_string(hd.h) +
                                    "x" + std::to_string(hd.ch) + ", expected " + std::to_string(width) + "x" +
                                    std::to_string(height) + "x" + std::to_string(channels));
    const size_t row = (size_t)width * channels;
    for (int r = 0; !rc && r < height; ++r)  // file rows run bottom to top
        if (std::fread(data + (size_t)(height - 1 - r) * row, 4, row, f) != row)
            rc = fail(STATER_E_ARG, std::string(path) + ": truncated PFM data");
    std::fclose(f);
    if (rc) return rc;
    if (hd.little != host_little_endian()) swap_bytes(data, row * height);
    return STATER_OK;
}

int stater_pfm_write_host(const char *path, const float *data, int width, int height, int channels) {
    if (!path || !data) return fail(STATER_E_ARG, "stater_pfm_write_host: NULL argument");
    if (width <= 0 || height <= 0 || (channels != 1 && channels != 3))
        return fail(STATER_E_ARG, "stater_pfm_write_host: PFM holds 1 or 3 channels");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(STATER_E_ARG, std::string(path) + ": " + std::strerror(errno));
    const bool le = host_little_endian();
    std::fprintf(f, "%s\n%d %d\n%s\n", channels == 3 ? "PF" : "Pf", width, height, le ? "-1.000000" : "1.000000");
    const size_t row = (size_t)width * channels;
    bool ok = true;
    for (int r = 0; ok && r < height; ++r) ok = std::fwrite(data + (size_t)(height - 1 - r) * row, 4, row, f) == row;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? STATER_OK : fail(STATER_E_ARG, std::string(path) + ": write failed");
}

int stater_pfm_read(const char *path, const stater_plane *dst, void *stream) {
    int rc = check_plane(dst, "stater_pfm_read");
    if (rc) return rc;
    std::vector<float> host((size_t)dst->width * dst->height * dst->channels);
    rc = stater_pfm_read_host(path, host.data(), dst->width, dst->height, dst->channels);
    if (rc) return rc;
    return upload(host, dst, stream_of(stream));
}

int stater_pfm_write(const char *path, const stater_plane *src, void *stream) {
    int rc = check_plane(src, "stater_pfm_write");
    if (rc) return rc;
    std::vector<float> host;
    rc = download(src, host, stream_of(stream));
    if (rc) return rc;
    return stater_pfm_write_host(path, host.data(), src->width, src->height, src->channels);
}

int stater_bufset_write_pfm(const stater_bufset *bs, const char *stem, uint64_t spp, const char *regex, void *stream,
                            int *written) {
    if (!bs || !stem) return fail(STATER_E_ARG, "stater_bufset_write_pfm: NULL argument");
    if (written) *written = 0;
    std::regex re;
    try {
        re = std::regex(regex ? regex : ".*");
    } catch (const std::regex_error &e) {
        return fail(STATER_E_ARG, std::string("stater_bufset_write_pfm: bad regex: ") + e.what());
    }
    hipStream_t st = stream ? (hipStream_t)stream : bs->ctx->stream;
    for (const std::string &name : set_names(bs)) {
        if (!std::regex_match(name, re)) continue;
        stater_plane p;
        int rc = stater_bufset_view(bs, name.c_str(), &p);
        if (rc) return rc;
        const std::string path = std::string(stem) + "-" + std::to_string(spp) + "-" + name + ".pfm";
        rc = stater_pfm_write(path.c_str(), &p, st);
        if (rc) return rc;
        if (written) ++*written;
    }
    return STATER_OK;
}

int stater_bufset_read_pfm(stater_bufset *bs, const char *stem, uint64_t spp, void *stream, int *nread) {
    namespace fs = std::filesystem;
    if (!bs || !stem) return fail(STATER_E_ARG, "stater_bufset_read_pfm: NULL argument");
    if (nread) *nread = 0;
    const fs::path sp(stem);
    const fs::path dir = sp.has_parent_path() ? sp.parent_path() : fs::path(".");
    const std::string prefix = sp.filename().string() + "-" + std::to_string(spp) + "-";
    const std::vector<std::string> names = set_names(bs);
    hipStream_t st = stream ? (hipStream_t)stream : bs->ctx->stream;
    std::error_code ec;
    std::vector<std::pair<std::string, std::string>> files;  // (name, path), sorted for a fixed order
    for (const auto &e : fs::directory_iterator(dir, ec)) {
        const std::string fn = e.path().filename().string();
        if (fn.size() <= prefix.size() + 4 || fn.compare(0, prefix.size(), prefix) != 0 ||
            fn.compare(fn.size() - 4, 4, ".pfm") != 0)
            continue;
        const std::string id = fn.substr(prefix.size(), fn.size() - prefix.size() - 4);
        // the film is the caller's plane (the reference reads it into its Film), other
        // other types (the G-buffers of other sets) and unknown names are not statistics of this set
        if (id == "film" || std::find(names.begin(), names.end(), id) == names.end()) continue;
        files.emplace_back(id, e.path().string());
    }
    if (ec) return fail(STATER_E_NOTFOUND, dir.string() + ": " + ec.message());
    std::sort(files.begin(), files.end());
    for (const auto &np : files) {
        stater_plane p;
        int rc = stater_bufset_view(bs, np.first.c_str(), &p);
        if (rc) return rc;
        rc = stater_pfm_read(np.second.c_str(), &p, st);
        if (rc) return rc;
        if (nread) ++*nread;
    }
    return STATER_OK;
}

}  // extern "C"
