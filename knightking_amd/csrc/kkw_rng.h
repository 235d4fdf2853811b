// kkw_rng.h -- the random stream of every walk, shared by the kernels of
// libkkw and by walk programs compiled at run time (hipRTC, where the standard
// headers are absent).
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), keyed by the job seed,
// counter (walker id, step, trial, tag).  One call yields four 32-bit words; a
// trial uses words 0-1 for the candidate index, 2 for the alias coin and 3 for
// the rejection dart.  This replaces KnightKing's per-thread std::mt19937
// (include/util.hpp:45-70) with a stream any CPU replay can reproduce.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace kkw {

enum : uint32_t { TAG_TRIAL = 0, TAG_EXT = 1, TAG_INIT = 2, TAG_OUTLIER = 3 };
enum : uint32_t { TAG_RMAT = 0x100, TAG_WEIGHT = 0x101, TAG_VTYPE = 0x102 };

struct u32x4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    return u32x4{c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ u32x4 draw(uint64_t seed, uint32_t walker, uint32_t step,
                                               uint32_t trial, uint32_t tag)
{
    return philox(walker, step, trial, tag, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// [0,1) on the 2^-24 grid; exact in float and never 1.
__host__ __device__ __forceinline__ float u01(uint32_t u)
{
    return (float)(u >> 8) * 5.9604644775390625e-08f;
}

// floor(r * n / 2^64) for r = hi:lo.
__host__ __device__ __forceinline__ uint32_t bounded(uint32_t lo, uint32_t hi, uint32_t n)
{
    uint64_t a = (uint64_t)hi * n;
    uint64_t b = (uint64_t)lo * n;
    return (uint32_t)((a + (b >> 32)) >> 32);
}

}  // namespace kkw
