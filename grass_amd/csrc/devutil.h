// Wave64 / workgroup primitives used by the kernels (CDNA4: 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grass {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// 64-bit value of the first active lane / of lane `src` (wave-uniform).  The builtins return a
// signed int: each half is taken back to uint32_t before widening, or a low half with bit 31 set
// would be sign-extended over the high half.
__device__ __forceinline__ uint64_t readfirstlane_u64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Lane `src` of a per-lane value read (wave-uniform) / replaced by a uniform value; the other lanes
// keep theirs.  A readlane per 32-bit half; the write is a compare with the lane id and a select
// (clang has no writelane builtin, and a v_writelane in inline asm would need M0, which the
// compiler reserves).
__device__ __forceinline__ uint32_t lane_get(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}
__device__ __forceinline__ uint64_t lane_get(uint64_t v, int src) { return readlane_u64(v, src); }
__device__ __forceinline__ uint32_t lane_set(uint32_t v, int src, uint32_t x) {
    return lane_id() == src ? x : v;
}
__device__ __forceinline__ uint64_t lane_set(uint64_t v, int src, uint64_t x) {
    return lane_id() == src ? x : v;
}

// Lanes below this one among the set bits of a 64-bit lane mask.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long lanes) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(lanes >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lanes, 0u));
}

// Orders one wave's LDS accesses across lanes (a wave-scope fence around a wave barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        T y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

// A lane's value moved by a DPP control (0 where the control has no source lane or the row/bank
// mask leaves the lane out).
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_move(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, BANK_MASK, true);
}

// 32-bit inclusive sum with DPP moves (VALU, no LDS round trips): a scan within each 16-lane row
// (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then
// rows 0-1's total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    x += dpp_move<0x111>(x);
    x += dpp_move<0x112>(x);
    x += dpp_move<0x114>(x);
    x += dpp_move<0x118>(x);
    x += dpp_move<0x142, 0xa>(x);
    x += dpp_move<0x143, 0xc>(x);
    return x;
}

// Inclusive max over lanes 0..lane, the same DPP tree (a masked-off or out-of-range source reads
// as 0, which leaves an unsigned max unchanged).
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t x) {
    x = max(x, dpp_move<0x111>(x));
    x = max(x, dpp_move<0x112>(x));
    x = max(x, dpp_move<0x114>(x));
    x = max(x, dpp_move<0x118>(x));
    x = max(x, dpp_move<0x142, 0xa>(x));
    x = max(x, dpp_move<0x143, 0xc>(x));
    return x;
}

// OR of a 64-bit value over the wave, returned uniform: the same DPP tree as the scan above, per
// 32-bit half, leaves the OR of lanes 0..63 in lane 63.
__device__ __forceinline__ unsigned long long wave_or_u64(unsigned long long v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo |= dpp_move<0x111>(lo);
    hi |= dpp_move<0x111>(hi);
    lo |= dpp_move<0x112>(lo);
    hi |= dpp_move<0x112>(hi);
    lo |= dpp_move<0x114>(lo);
    hi |= dpp_move<0x114>(hi);
    lo |= dpp_move<0x118>(lo);
    hi |= dpp_move<0x118>(hi);
    lo |= dpp_move<0x142, 0xa>(lo);
    hi |= dpp_move<0x142, 0xa>(hi);
    lo |= dpp_move<0x143, 0xc>(lo);
    hi |= dpp_move<0x143, 0xc>(hi);
    return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)hi, kWave - 1) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)lo, kWave - 1);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_max(T x) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        T y = __shfl_xor(x, off, kWave);
        x = y > x ? y : x;
    }
    return x;
}

// Segmented inclusive sum over the wave: a lane with head != 0 starts a new segment.
__device__ __forceinline__ uint32_t wave_segmented_sum(uint32_t x, int head) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t px = __shfl_up(x, off, kWave);
        int ph = __shfl_up(head, off, kWave);
        if (lane >= off && !head) {
            x += px;
            head = ph;
        }
    }
    return x;
}

// Segmented inclusive sum over runs of at most `span` lanes (a lane with head != 0 starts a
// run): log2(span) shuffle steps instead of six.
__device__ __forceinline__ uint32_t short_segmented_sum(uint32_t x, int head, uint32_t span) {
    const int lane = lane_id();
    if (span == 4) {
        // the default checkpoint interval: runs of at most four lanes, so a lane adds the values
        // of the (up to three) lanes before it back to its run's head, moved by wave shifts of
        // one lane (DPP, no LDS); lane 0 always starts a run here
        const uint32_t hf = head ? 1u : 0u;
        const uint32_t v1 = dpp_move<0x138>(x), h1 = dpp_move<0x138>(hf);
        const uint32_t v2 = dpp_move<0x138>(v1), h2 = dpp_move<0x138>(h1);
        const uint32_t v3 = dpp_move<0x138>(v2);
        const bool a1 = !hf, a2 = a1 && !h1, a3 = a2 && !h2;
        return x + (a1 ? v1 : 0u) + (a2 ? v2 : 0u) + (a3 ? v3 : 0u);
    }
    for (uint32_t off = 1; off < span && off < (uint32_t)kWave; off <<= 1) {
        const uint32_t px = __shfl_up(x, off, kWave);
        const int ph = __shfl_up(head, off, kWave);
        if (lane >= (int)off && !head) {
            x += px;
            head = ph;
        }
    }
    return x;
}

// Exclusive sum over a workgroup of BLOCK threads; `total` receives the workgroup sum.
// `lds` holds BLOCK / 64 elements.  Every thread of the workgroup must call it.
template <typename T, int BLOCK>
__device__ __forceinline__ T block_exclusive_sum(T x, T& total, T* lds) {
    static_assert(BLOCK % kWave == 0, "block must be whole waves");
    T inc = wave_inclusive_sum(x);
    const int wid = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) lds[wid] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) {
        T t = lds[i];
        pre += (i < wid) ? t : T(0);
        tot += t;
    }
    __syncthreads();
    total = tot;
    return pre + inc - x;
}

template <typename T, int BLOCK>
__device__ __forceinline__ T block_sum(T x, T* lds) {
    T s = wave_sum(x);
    const int wid = threadIdx.x / kWave;
    if (lane_id() == 0) lds[wid] = s;
    __syncthreads();
    T tot = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) tot += lds[i];
    __syncthreads();
    return tot;
}

// Bit-stream passes (MSB-first bits in big-endian-bit 32-bit words): a wave assembles the fields
// of up to 64 positions in an LDS buffer whose first word is stream word w0, then writes it out.

// ORs the `wdt`-bit field `val` (wdt <= 32, val < 2^wdt) at stream bit `at` into the buffer.
__device__ __forceinline__ void put_field(uint32_t* B, uint64_t w0, uint64_t at, uint32_t wdt, uint32_t val) {
    if (!wdt) return;
    const uint32_t rel = (uint32_t)((at >> 5) - w0), s = (uint32_t)(at & 31);
    if (s + wdt <= 32) {
        atomicOr(&B[rel], val << (32 - s - wdt));
    } else {
        atomicOr(&B[rel], val >> (s + wdt - 32));
        atomicOr(&B[rel + 1], val << (64 - s - wdt));
    }
}

// p[0..count) = value.  Used instead of hipMemsetAsync for anything a caller might capture in a
// HIP graph.
template <typename T>
__global__ void fill_kernel(T* p, uint64_t count, T value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

template <typename T, int BLOCK>
__device__ __forceinline__ T block_max(T x, T* lds) {
    T s = wave_max(x);
    const int wid = threadIdx.x / kWave;
    if (lane_id() == 0) lds[wid] = s;
    __syncthreads();
    T m = lds[0];
#pragma unroll
    for (int i = 1; i < BLOCK / kWave; ++i) m = lds[i] > m ? lds[i] : m;
    __syncthreads();
    return m;
}

}  // namespace grass
