// convert.hip — plane conversion at the boundary: layout (PLANAR <->
// INTERLEAVED, the reference's CV_32FC3 GpuMats) and storage type (f32 <->
// bf16, round to nearest even).  One thread per pixel, all channels.

#include "kernels.hpp"

namespace stater {

__device__ __forceinline__ uint16_t f32_to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays a quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct PlaneRef {
    char *ptr;
    int64_t pitch, cs;
    int dtype, interleaved;
};

__device__ __forceinline__ char *elem(const PlaneRef &p, int x, int y, int c, int C) {
    const int64_t esz = p.dtype == STATER_BF16 ? 2 : 4;
    return p.ptr + (int64_t)y * p.pitch + (p.interleaved ? ((int64_t)x * C + c) * esz : c * p.cs + (int64_t)x * esz);
}

__global__ void __launch_bounds__(256) k_convert(PlaneRef src, PlaneRef dst, int W, int C) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    for (int c = 0; c < C; ++c) {
        const char *s = elem(src, x, y, c, C);
        const float v = src.dtype == STATER_BF16 ? __uint_as_float((uint32_t)*(const uint16_t *)s << 16)
                                                 : *(const float *)s;
        char *d = elem(dst, x, y, c, C);
        if (dst.dtype == STATER_BF16) *(uint16_t *)d = f32_to_bf16(v);
        else *(float *)d = v;
    }
}

static bool plane_ok(const stater_plane *p) {
    if (!p || !p->ptr || p->width <= 0 || p->height <= 0 || p->channels < 1 || p->channels > 4) return false;
    if (p->dtype != STATER_F32 && p->dtype != STATER_BF16) return false;
    const int64_t esz = p->dtype == STATER_BF16 ? 2 : 4;
    const int64_t row = (int64_t)p->width * esz * (p->layout == STATER_INTERLEAVED ? p->channels : 1);
    return p->pitch_bytes >= row && p->pitch_bytes % esz == 0 &&
           (p->layout == STATER_INTERLEAVED || (p->channel_stride_bytes % esz == 0 &&
                                                 p->channel_stride_bytes >= p->pitch_bytes * p->height));
}

}  // namespace stater

using namespace stater;

extern "C" int stater_convert_plane(stater_ctx *ctx, const stater_plane *src, const stater_plane *dst, void *stream) {
    if (!ctx) return fail(STATER_E_ARG, "stater_convert_plane: NULL context");
    if (!plane_ok(src) || !plane_ok(dst))
        return fail(STATER_E_ARG, "stater_convert_plane: planes must be f32 or bf16, 1..4 channels, pitches covering a row");
    if (src->width != dst->width || src->height != dst->height || src->channels != dst->channels)
        return fail(STATER_E_ARG, "stater_convert_plane: source and destination shapes differ");
    auto ref = [](const stater_plane *p) {
        return PlaneRef{(char *)p->ptr, p->pitch_bytes, p->channel_stride_bytes, p->dtype,
                        p->layout == STATER_INTERLEAVED ? 1 : 0};
    };
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(k_convert, dim3((src->width + 255) / 256, src->height), dim3(256), 0, st, ref(src), ref(dst),
                       src->width, src->channels);
    STATER_HIP(hipGetLastError());
    return STATER_OK;
}
