// Reduce-then-scan exclusive prefix sum over a generated sequence of uint64 values.
//
// Launches with no inter-workgroup waiting: per-tile sums, then each tile re-generates its values
// and writes prefixes.  A sequence of at most kScanFoldMax tiles has each tile add up the tile sums
// before it itself (a few thousand cached values); a longer one scans the tile sums in one
// workgroup in between (a third launch).  The length can be
// a host value or a device-resident count (so a sequence whose size a previous kernel
// produced needs no host round trip); the grid is sized for the host-side upper bound.
// Used for the record-length scan at load (replacing cub::DeviceScan::ExclusiveSum,
// LE/src/utils.hpp:337-342), CSR row pointers and sampler output offsets.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devutil.h"

namespace grass {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr uint64_t kScanTile = (uint64_t)kScanBlock * kScanItems;
constexpr uint64_t kScanFoldMax = 2048;  // tiles whose start each tile sums itself

struct Count {
    uint64_t imm;
    const uint64_t* dev;  // when non-null the count is *dev
    __device__ __forceinline__ uint64_t get() const { return dev ? *dev : imm; }
};

template <class Gen>
__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(Gen gen, Count cnt,
                                                                 uint64_t* partials) {
    __shared__ uint64_t lds[kScanBlock / kWave];
    const uint64_t n = cnt.get();
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t s = 0;
    if (base < n) {
#pragma unroll 4
        for (int j = 0; j < kScanItems; ++j) {
            uint64_t i = base + (uint64_t)j * kScanBlock + threadIdx.x;
            if (i < n) s += gen(i);
        }
    }
    s = block_sum<uint64_t, kScanBlock>(s, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// One workgroup: partials[0..nb) -> exclusive prefixes in place; *total = sum.
// (A template so that every translation unit including this header may instantiate it.)
template <typename T>
__global__ __launch_bounds__(1024) void scan_partials_kernel(T* partials, uint64_t nb, T* total) {
    __shared__ T lds[1024 / kWave];
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t lo = threadIdx.x * per;
    T s = 0;
    for (uint64_t i = lo; i < lo + per && i < nb; ++i) s += partials[i];
    T tot;
    T pre = block_exclusive_sum<T, 1024>(s, tot, lds);
    for (uint64_t i = lo; i < lo + per && i < nb; ++i) {
        T t = partials[i];
        partials[i] = pre;
        pre += t;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// 64-bit inclusive sum over the wave with DPP moves of the two halves (no LDS).
__device__ __forceinline__ uint64_t wave_inclusive_sum_u64(uint64_t x) {
#define GRASS_SCAN_STEP(CTRL, RM)                                                                  \
    x += ((uint64_t)dpp_move<CTRL, RM>((uint32_t)(x >> 32)) << 32) | dpp_move<CTRL, RM>((uint32_t)x)
    GRASS_SCAN_STEP(0x111, 0xf);
    GRASS_SCAN_STEP(0x112, 0xf);
    GRASS_SCAN_STEP(0x114, 0xf);
    GRASS_SCAN_STEP(0x118, 0xf);
    GRASS_SCAN_STEP(0x142, 0xa);
    GRASS_SCAN_STEP(0x143, 0xc);
#undef GRASS_SCAN_STEP
    return x;
}

// out[i] = sum of gen(0..i-1) for i < n, and out[n] = total.  Wave w of a tile takes
// kWave * kScanItems consecutive items in kScanItems rounds of kWave consecutive ones (coalesced
// loads and stores): its lanes' sums and one exchange of the wave sums give each wave its start,
// and each round is one DPP wave scan carried into the next (no workgroup scan per round).
// FOLD: partials[] are the tile sums themselves (at most kScanFoldMax of them): each tile adds up
// those before it, and the tile holding item n - 1 also writes *total; otherwise partials[] are
// the tiles' exclusive prefixes and *total was written with them.
template <class Gen, bool FOLD>
__global__ __launch_bounds__(kScanBlock) void scan_write_kernel(Gen gen, Count cnt,
                                                                const uint64_t* partials,
                                                                uint64_t* total,
                                                                uint64_t* out) {
    __shared__ uint64_t wsum[kScanBlock / kWave];
    __shared__ uint64_t fold_lds[kScanBlock / kWave];
    const uint64_t n = cnt.get();
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    if (base >= n) {
        if (base == 0 && threadIdx.x == 0) {  // empty sequence
            out[0] = 0;
            if (FOLD && total) *total = 0;
        }
        return;
    }
    uint64_t start;
    if (FOLD) {
        uint64_t p = 0;
        for (uint64_t k = threadIdx.x; k < blockIdx.x; k += kScanBlock) p += partials[k];
        start = block_sum<uint64_t, kScanBlock>(p, fold_lds);
    } else {
        start = partials[blockIdx.x];
    }
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    const uint64_t wb = base + (uint64_t)wid * kWave * kScanItems + lane;
    uint64_t x[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t i = wb + (uint64_t)j * kWave;
        x[j] = i < n ? gen(i) : 0;
        s += x[j];
    }
    const uint64_t st = wave_inclusive_sum_u64(s);
    if (lane == kWave - 1) wsum[wid] = st;
    __syncthreads();
    uint64_t carry = start;
    for (int w = 0; w < wid; ++w) carry += wsum[w];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t i = wb + (uint64_t)j * kWave;
        const uint64_t inc = wave_inclusive_sum_u64(x[j]);
        if (i < n) out[i] = carry + inc - x[j];
        carry += readlane_u64(inc, kWave - 1);
    }
    if (threadIdx.x == 0 && base + kScanTile >= n) {
        if (FOLD) {
            uint64_t t = start;
            for (int w = 0; w < kScanBlock / kWave; ++w) t += wsum[w];
            out[n] = t;
            if (total) *total = t;
        } else {
            out[n] = *total;
        }
    }
}

inline uint64_t scan_blocks(uint64_t n_max) { return n_max ? (n_max + kScanTile - 1) / kScanTile : 1; }

// Scratch: partials[scan_blocks(n_max)] and total[1].
template <class Gen>
hipError_t exclusive_scan(Gen gen, Count cnt, uint64_t n_max, uint64_t* out, uint64_t* partials,
                          uint64_t* total, hipStream_t stream) {
    const uint64_t nb = scan_blocks(n_max);
    hipLaunchKernelGGL(scan_reduce_kernel<Gen>, dim3((unsigned)nb), dim3(kScanBlock), 0, stream,
                       gen, cnt, partials);
    if (nb <= kScanFoldMax) {
        hipLaunchKernelGGL((scan_write_kernel<Gen, true>), dim3((unsigned)nb), dim3(kScanBlock), 0, stream,
                           gen, cnt, (const uint64_t*)partials, total, out);
    } else {
        hipLaunchKernelGGL(scan_partials_kernel<uint64_t>, dim3(1), dim3(1024), 0, stream, partials, nb, total);
        hipLaunchKernelGGL((scan_write_kernel<Gen, false>), dim3((unsigned)nb), dim3(kScanBlock), 0, stream,
                           gen, cnt, (const uint64_t*)partials, total, out);
    }
    return hipGetLastError();
}

}  // namespace grass
