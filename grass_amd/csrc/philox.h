// Counter-based RNG shared by host and device code.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11).  Every random choice on the sampling path is keyed by the 64-bit seed and a
// 128-bit counter {instance, walker, step, draw}, so a draw depends on what is being sampled
// and never on which thread, wave or launch samples it.  This replaces the reference's
// per-thread cuRAND XORWOW states (LE/src/rand_num_gen.cuh:117-122, LE/src/grass.cu:1006-1011).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grass {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// Counter tags for draws that are not a numbered step.
constexpr uint32_t kTagStart = 0xFFFFFFFFu;     // start vertex of a walk / frontier entry
constexpr uint32_t kTagMdrwStep = 0xFFFFFFFEu;  // MDRW step: frontier pick + neighbour pick

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                     uint32_t c3, uint64_t seed) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return U4{c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Integer-only map of 64 random bits to [0, bound): floor(r * bound / 2^64).  The bias is
// below bound / 2^64 (< 2^-33 for any vertex degree), there is no float anywhere, and the
// same bits give the same index on every device.
__host__ __device__ __forceinline__ uint64_t uniform_index(uint32_t lo, uint32_t hi,
                                                           uint64_t bound) {
    return mulhi64(((uint64_t)hi << 32) | lo, bound);
}

__host__ __device__ __forceinline__ uint64_t uniform(uint64_t seed, uint32_t c0, uint32_t c1,
                                                     uint32_t c2, uint32_t c3, uint64_t bound) {
    U4 r = philox4x32_10(c0, c1, c2, c3, seed);
    return uniform_index(r.x, r.y, bound);
}

}  // namespace grass
