// common.hpp — shared device/host definitions of the gfx950 statistics library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/stater.h"

namespace stater {

// ------------------------------------------------------------ error state
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define STATER_HIP(call)                                            \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) return ::stater::hip_fail(e_, #call); \
    } while (0)

// ----------------------------------------------------------------- layout
// Order of the accumulated statistics of one set (BUF_LIST order of
// src/statistics/estimator.h:459-466).
enum Stat { ST_MEAN = 0, ST_M2, ST_M3, ST_M4, ST_C2X, ST_C2Y, ST_M1, N_STATS };

constexpr int kMaxSets = 6;    // untransformed + up to 5 distinct Box-Cox lambdas
constexpr int kRowAlign = 64;  // floats: 256-byte row pitch

// -------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the counter-based
// generator of the synthetic workload.  Host and device share this code.
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (i) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    }
    return U4{c0, c1, c2, c3};
}

__host__ __device__ inline double unit32(uint32_t r) { return (double)r * (1.0 / 4294967296.0); }
__host__ __device__ inline double unit24(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }
__host__ __device__ inline double unit24_open(uint32_t r) {
    return (double)((r >> 8) + 1u) * (1.0 / 16777216.0);
}

constexpr double kPi = 3.14159265358979323846;

constexpr int kXcds = 8;  // accelerator complex dies: workgroup i runs on XCD i % 8

// Static XCD-aware placement of a one-tile-per-workgroup grid of
// xcd_grid(ntiles) workgroups: workgroup i runs on XCD i % 8 and takes the
// (i / 8)-th tile of that XCD's contiguous chunk, so the tiles that share halo
// rows are read through one L2.  Returns -1 for the few padding workgroups.
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + kXcds - 1) / kXcds;
    const int t = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    return t < ntiles ? t : -1;
}
inline unsigned xcd_grid(int ntiles) { return (unsigned)((ntiles + kXcds - 1) / kXcds * kXcds); }

}  // namespace stater
