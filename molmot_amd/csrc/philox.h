// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): 10 rounds of
// two 32x32->64 multiplications with the published multipliers and Weyl
// key increments.  Stateless: (counter, key) -> 4 x uint32.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

struct mm_u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void mm_mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ __forceinline__ mm_u32x4 mm_philox4x32_10(mm_u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mm_mulhilo32(M0, c.x, hi0, lo0);
        mm_mulhilo32(M1, c.z, hi1, lo1);
        mm_u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// uint32 -> float in (0, 1]: (u + 1) * 2^-32 rounded to float
__device__ __forceinline__ float mm_u01_open0(uint32_t u) {
    return ((float)(u >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// One standard normal from two uint32 (Box-Muller, FP32: the variate only
// scales a random kick, its rounding is far below its own spread).
// Both Box-Muller normals of one (a, b) pair: r cos(2 pi u2), r sin(2 pi u2).
__device__ __forceinline__ void mm_normal2_from(uint32_t a, uint32_t b, float& n1, float& n2) {
    const float u1 = mm_u01_open0(a);
    const float u2 = (float)(b >> 8) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * __logf(u1));
    const float ph = 6.2831853071795865f * u2;
    n1 = r * __cosf(ph);
    n2 = r * __sinf(ph);
}

__device__ __forceinline__ float mm_normal_from(uint32_t a, uint32_t b) {
    const float u1 = mm_u01_open0(a);
    const float u2 = (float)(b >> 8) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * __logf(u1));
    return r * __cosf(6.2831853071795865f * u2);
}
