// Multi-dimensional random walk (frontier sampling) in batches of instances.
//
// The reference's multi-RW app (LE/src/apps/multiRW/multiRW.cu:22-45,74-97) keeps 80 roots per
// sample, moves a uniformly chosen root per step and materialises nothing.  Here an instance
// keeps a frontier of m vertices, moves an entry chosen proportionally to its degree (the
// intent stated at multiRW.cu:77) and returns its traversed edges, its sorted vertex set and the
// induced edge list of that set.
//
// 1. mdrw_walk_kernel: persistent, one wave per instance taken from a device queue.  The
//    frontier ids and degrees sit in LDS.  Lane L owns the contiguous block [L*B, L*B + B) of
//    entries (B = ceil(m / 64)) and keeps its degree sum in a register, so the degree-
//    proportional pick is two wave scans: one over the 64 block sums to find the block, one over
//    that block's degrees to find the entry.  A step is then one header load, one neighbour
//    decode and one degree load, all wave-uniform.
// 2. mdrw_verts_kernel: one workgroup per instance sorts frontier + step destinations in LDS
//    (bitonic) and keeps the first of each run of equal ids.
// 3. Induced edges, per slice of up to 512 instances: a bit mask row per distinct vertex of the
//    slice marks the instances whose sample holds it.  The lists of those vertices are cut into
//    chunks of kIndChunk entries and each chunk is decoded once by one wave (LE: O(1) random
//    access per lane; GT: absolute values plus gaps turned into ids with a segmented wave sum);
//    entry y of x's list joins the induced list of every instance in mask(x) & mask(y).  A count
//    pass and a scan give every (instance, vertex, chunk) its output offset, so the write pass
//    emits each instance's edges in (vertex, position) order without any sorting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

#include "devutil.h"
#include "philox.h"
#include "runtime.h"
#include "scan.h"

namespace grass {

constexpr uint32_t kIndChunk = 2048;  // adjacency entries per chunk (32 wave iterations)
constexpr uint32_t kMdrwHdrLdsMax = 4096;  // largest frontier whose headers the walk keeps in LDS
constexpr int kSortBlock = 1024;

struct MdDesc {
    uint64_t step_base;  // bound layout: instance i's steps start here (target - m slots)
    uint64_t init_base;  // initial frontier
    uint64_t vert_base;  // bound layout of the vertex set (m + steps slots)
    uint32_t inst, m, steps, pad;
};

// ---------------------------------------------------------------------------- 1. walks

// HDR: the frontier entries' header-cache words sit in LDS next to their ids and degrees (20 bytes
// per entry instead of 8), so a step's chain is two dependent loads (the picked entry's neighbour,
// then that neighbour's header) instead of three (header, neighbour, degree).
template <int FMT, bool HDR>
__global__ __launch_bounds__(kWave) void mdrw_walk_kernel(DevGraph g, const MdDesc* __restrict__ desc,
                                                          uint32_t I, unsigned long long* queue,
                                                          uint64_t seed, uint32_t* __restrict__ init,
                                                          uint32_t* __restrict__ ssrc,
                                                          uint32_t* __restrict__ sdst,
                                                          uint32_t* __restrict__ scnt,
                                                          uint32_t mmax, uint32_t* __restrict__ err) {
    extern __shared__ uint32_t lds[];
    uint32_t* F = lds;         // frontier ids
    uint32_t* D = lds + mmax;  // their degrees
    uint32_t* Hx = lds + 2 * mmax;  // HDR: their header words x, y, w (z is the degree)
    uint32_t* Hy = lds + 3 * mmax;
    uint32_t* Hw = lds + 4 * mmax;
    const int lane = lane_id();
    for (;;) {
        unsigned long long gi = 0;
        if (lane == 0) gi = atomicAdd(queue, 1ull);
        gi = __shfl(gi, 0, kWave);
        if (gi >= I) break;  // every wave exits once the queue is drained
        const MdDesc q = desc[gi];
        const uint32_t m = q.m;
        for (uint32_t j = lane; j < m; j += kWave) {
            const uint32_t v = (uint32_t)uniform(seed, q.inst, j, kTagStart, 0, g.n);
            F[j] = v;
            if (HDR) {
                const uint4 e = g.hdr[v];
                D[j] = e.z;
                Hx[j] = e.x;
                Hy[j] = e.y;
                Hw[j] = e.w;
            } else {
                D[j] = g.deg[v];
            }
            init[q.init_base + j] = v;
        }
        __syncthreads();
        const uint32_t B = (m + kWave - 1) / kWave;
        uint64_t bs = 0;
        for (uint32_t i = 0; i < B; ++i) {
            const uint32_t j = lane * B + i;
            if (j < m) bs += D[j];
        }
        uint64_t W = wave_sum(bs);
        // lane l: degree sum of blocks 0..l, kept up to date by each step's one-entry change
        uint64_t incl = wave_inclusive_sum(bs);
        uint32_t t = 0;
        bool bad = false;
        // The draws of a step do not depend on the walk: lane l computes those of step t0 + l once
        // per 64 steps, and each step reads its own from that lane.
        U4 ahead{0, 0, 0, 0};
        for (; t < q.steps && W > 0; ++t) {
            const int tl = (int)(t & (kWave - 1));
            if (tl == 0) ahead = philox4x32_10(q.inst, t + lane, kTagMdrwStep, 0, seed);
            U4 r;
            r.x = (uint32_t)__builtin_amdgcn_readlane(ahead.x, tl);
            r.y = (uint32_t)__builtin_amdgcn_readlane(ahead.y, tl);
            r.z = (uint32_t)__builtin_amdgcn_readlane(ahead.z, tl);
            r.w = (uint32_t)__builtin_amdgcn_readlane(ahead.w, tl);
            const uint64_t x = mulhi64(((uint64_t)r.y << 32) | r.x, W);
            // block of the pick: first lane whose inclusive block sum exceeds x (exists: total W > x)
            const unsigned long long mk = __ballot(incl > x);
            const int L = __ffsll(mk) - 1;
            const uint64_t xr = x - readlane_u64(incl - bs, L);
            // entry of the pick inside block L: a 32-bit scan of its degrees when they sum below 2^32
            uint32_t j = 0;
            if (readlane_u64(bs, L) <= 0xFFFFFFFFull) {
                const uint32_t xr32 = (uint32_t)xr;
                uint32_t acc = 0;
                for (uint32_t c0 = 0; c0 < B; c0 += kWave) {
                    const uint32_t o = c0 + lane, idx = L * B + o;
                    const uint32_t w = (o < B && idx < m) ? D[idx] : 0u;
                    const uint32_t inc2 = wave_inclusive_sum(w) + acc;
                    const unsigned long long mk2 = __ballot(inc2 > xr32);
                    if (mk2) {
                        j = L * B + c0 + (__ffsll(mk2) - 1);
                        break;
                    }
                    acc = __shfl(inc2, kWave - 1, kWave);
                }
            } else {
                uint64_t acc = 0;
                for (uint32_t c0 = 0; c0 < B; c0 += kWave) {
                    const uint32_t o = c0 + lane, idx = L * B + o;
                    const uint64_t w = (o < B && idx < m) ? D[idx] : 0;
                    const uint64_t inc2 = wave_inclusive_sum(w) + acc;
                    const unsigned long long mk2 = __ballot(inc2 > xr);
                    if (mk2) {
                        j = L * B + c0 + (__ffsll(mk2) - 1);
                        break;
                    }
                    acc = __shfl(inc2, kWave - 1, kWave);
                }
            }
            const uint32_t u = F[j], du = D[j];
            Hdr h;
            if (HDR) unpack_hdr<FMT>(make_uint4(Hx[j], Hy[j], du, Hw[j]), h);
            else load_hdr<FMT>(g, u, h);
            const uint32_t k = (uint32_t)mulhi64(((uint64_t)r.w << 32) | r.z, du);
            const uint32_t v = get_neighbor<FMT>(g, h, u, k, nullptr);
            if (v >= g.n) {  // the graph decodes to garbage: stop this instance
                if (lane == 0) atomicCAS(err, kErrNone, kErrBadNeighbor);
                bad = true;
                break;
            }
            uint4 ev = make_uint4(0, 0, 0, 0);
            if (HDR) ev = g.hdr[v];
            const uint32_t dv = HDR ? ev.z : g.deg[v];
            __syncthreads();  // every lane has read F[j], D[j] (and the header words)
            if (lane == 0) {
                ssrc[q.step_base + t] = u;
                sdst[q.step_base + t] = v;
                F[j] = v;
                D[j] = dv;
                if (HDR) {
                    Hx[j] = ev.x;
                    Hy[j] = ev.y;
                    Hw[j] = ev.w;
                }
            }
            if (lane == L) bs = bs - du + dv;
            if (lane >= L) incl = incl - du + dv;
            W = W - du + dv;
            __syncthreads();
        }
        if (lane == 0) scnt[gi] = bad ? 0 : t;
        __syncthreads();  // F and D are rewritten by the next instance
    }
}

// ---------------------------------------------------------------------------- 2. vertex sets

// keys: init frontier then step destinations, padded to `cap` (a power of two) with ~0.
__global__ __launch_bounds__(kSortBlock) void mdrw_verts_kernel(const MdDesc* __restrict__ desc,
                                                                const uint32_t* __restrict__ init,
                                                                const uint32_t* __restrict__ sdst,
                                                                const uint32_t* __restrict__ scnt,
                                                                uint32_t cap, uint32_t* __restrict__ vbound,
                                                                uint32_t* __restrict__ vcnt) {
    extern __shared__ uint32_t keys[];
    __shared__ uint32_t part[kSortBlock / kWave];
    const MdDesc q = desc[blockIdx.x];
    const uint32_t m = q.m, ns = scnt[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < cap; i += kSortBlock) {
        uint32_t k = 0xFFFFFFFFu;
        if (i < m) k = init[q.init_base + i];
        else if (i < m + ns) k = sdst[q.step_base + (i - m)];
        keys[i] = k;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= cap; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < cap; i += kSortBlock) {
                const uint32_t o = i ^ stride;
                if (o > i) {
                    const uint32_t a = keys[i], b = keys[o];
                    const bool up = (i & size) == 0;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    // keep the first of each run; thread t owns the consecutive keys [t*per, (t+1)*per)
    const uint32_t per = cap / kSortBlock ? cap / kSortBlock : 1;
    const uint32_t lo = threadIdx.x * per;
    uint32_t c = 0;
    for (uint32_t i = lo; i < lo + per && i < cap; ++i) {
        const uint32_t k = keys[i];
        c += k != 0xFFFFFFFFu && (i == 0 || keys[i - 1] != k);
    }
    uint32_t tot;
    uint32_t pos = block_exclusive_sum<uint32_t, kSortBlock>(c, tot, part);
    for (uint32_t i = lo; i < lo + per && i < cap; ++i) {
        const uint32_t k = keys[i];
        if (k != 0xFFFFFFFFu && (i == 0 || keys[i - 1] != k)) vbound[q.vert_base + pos++] = k;
    }
    if (threadIdx.x == 0) vcnt[blockIdx.x] = tot;
}

struct U32Gen {
    const uint32_t* a;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return a[i]; }
};

struct U64Gen {
    const uint64_t* a;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return a[i]; }
};

// Copies instance blockIdx.x's `cnt` entries from its bound-layout region to the compact one.
__global__ __launch_bounds__(256) void mdrw_compact_kernel(const uint32_t* __restrict__ src,
                                                           const MdDesc* __restrict__ desc, int which,
                                                           const uint32_t* __restrict__ cnt,
                                                           const uint64_t* __restrict__ off,
                                                           uint32_t* __restrict__ dst) {
    const MdDesc q = desc[blockIdx.x];
    const uint64_t base = which == 0 ? q.step_base : q.vert_base;
    const uint64_t o = off[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < cnt[blockIdx.x]; i += 256) dst[o + i] = src[base + i];
}

// ---------------------------------------------------------------------------- 3. induced edges

struct ChunkGen {  // chunks of vertex gv
    const uint32_t* verts;
    const uint32_t* deg;
    __device__ __forceinline__ uint64_t operator()(uint64_t gv) const {
        return (deg[verts[gv]] + kIndChunk - 1) / kIndChunk;
    }
};

// Membership masks.  The samples of one slice of up to kSlice instances share one pass over the
// lists of their distinct vertices: the mask row of x holds one bit per instance of the slice whose
// sample contains x, so an adjacency entry (x, y) belongs to the induced lists of exactly the
// instances in mask(x) & mask(y).  Every list is decoded once per slice, however many samples
// hold its vertex.  Rows are indexed by a dense id (dmap[v] = the id, or kNoRow for a vertex no
// sample of the slice holds; dvert inverts it), so the rows that the induced pass reads at random
// (W words per sampled vertex, 162 MB for 512 instances at scale 24) fit the Infinity Cache where
// rows for all n vertices would not.
constexpr uint32_t kNoRow = 0xFFFFFFFFu;
constexpr uint32_t kSlice = 512;
constexpr uint32_t kSliceWords = kSlice / 64;
constexpr int kIndBlock = 256;
constexpr int kIndWaves = kIndBlock / kWave;
constexpr uint32_t kRowPitch = kSliceWords + 1;  // u64 words per staged row in LDS (odd: fewer bank conflicts)

// Instance of pair p: the largest i in [i0, i1) with voff[i] <= p.
__device__ __forceinline__ uint32_t inst_of(const uint64_t* __restrict__ voff, uint32_t i0, uint32_t i1,
                                            uint64_t p) {
    uint32_t lo = i0, hi = i1 - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (voff[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// dmap[x] = 1 for every vertex x that a sample of slice [i0, i1) holds (dmap starts zeroed).
__global__ __launch_bounds__(256) void mdrw_flag_kernel(const uint32_t* __restrict__ verts,
                                                        const uint64_t* __restrict__ voff, uint32_t i0,
                                                        uint32_t i1, uint32_t* __restrict__ dmap) {
    const uint64_t p0 = voff[i0], p1 = voff[i1];
    for (uint64_t p = p0 + (uint64_t)blockIdx.x * 256 + threadIdx.x; p < p1; p += (uint64_t)gridDim.x * 256)
        dmap[verts[p]] = 1u;
}

// From the flags and their exclusive scan `did`: dmap[x] = dense id (or kNoRow), dvert[id] = x.
__global__ __launch_bounds__(256) void mdrw_dense_kernel(const uint64_t* __restrict__ did, uint32_t n,
                                                         uint32_t* __restrict__ dmap, uint32_t* __restrict__ dvert) {
    for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (uint64_t)gridDim.x * 256) {
        if (dmap[x]) {
            const uint32_t d = (uint32_t)did[x];
            dmap[x] = d;
            dvert[d] = (uint32_t)x;
        } else {
            dmap[x] = kNoRow;
        }
    }
}

// The dense ids again as a bit vector with counts: word w holds one bit per vertex 64w..64w+63
// that the slice holds (x) and the dense id of the first of them (y), so that the dense id of y is
// y's word's count plus the set bits below y in the word.  The induced pass looks up every
// adjacency entry here: 16 bytes per 64 vertices (2.2 MB at scale 24) against dmap's 4 bytes per
// vertex (35 MB).
__global__ __launch_bounds__(256) void mdrw_dbits_kernel(const uint32_t* __restrict__ dmap,
                                                         const uint64_t* __restrict__ did, uint32_t n,
                                                         ulonglong2* __restrict__ dbits) {
    const uint64_t nwords = ((uint64_t)n + kWave - 1) / kWave;
    const int lane = lane_id();
    for (uint64_t w = ((uint64_t)blockIdx.x * 256 + threadIdx.x) / kWave; w < nwords;
         w += (uint64_t)gridDim.x * 256 / kWave) {
        const uint64_t x = w * kWave + lane;
        const unsigned long long bits = __ballot(x < n && dmap[x] != kNoRow);
        if (lane == 0) dbits[w] = make_ulonglong2(bits, did[w * kWave]);
    }
}

// rows[0, *nd * W) = 0
__global__ __launch_bounds__(256) void mdrw_zero_rows_kernel(unsigned long long* __restrict__ rows,
                                                             const uint64_t* __restrict__ nd, uint32_t W) {
    const uint64_t count = *nd * W;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256)
        rows[i] = 0ull;
}

// Row of x, word l / 64, gets bit l % 64 for every pair (instance i0 + l, vertex x) of the slice.
__global__ __launch_bounds__(256) void mdrw_mask_kernel(const uint32_t* __restrict__ verts,
                                                        const uint64_t* __restrict__ voff, uint32_t i0,
                                                        uint32_t i1, uint32_t W, const uint32_t* __restrict__ dmap,
                                                        unsigned long long* __restrict__ rows) {
    const uint64_t p0 = voff[i0], p1 = voff[i1];
    for (uint64_t p = p0 + (uint64_t)blockIdx.x * 256 + threadIdx.x; p < p1; p += (uint64_t)gridDim.x * 256) {
        const uint32_t l = inst_of(voff, i0, i1, p) - i0;
        atomicOr(&rows[(uint64_t)dmap[verts[p]] * W + (l >> 6)], 1ull << (l & 63));
    }
}

struct PopGen {  // instances of the slice holding dense vertex d
    const unsigned long long* rows;
    uint32_t W;
    __device__ __forceinline__ uint64_t operator()(uint64_t d) const {
        uint64_t c = 0;
        for (uint32_t w = 0; w < W; ++w) c += __popcll(rows[d * W + w]);
        return c;
    }
};

struct UChunkGen {  // chunks of the list of dense vertex d
    const uint32_t* dvert;
    const uint32_t* deg;
    __device__ __forceinline__ uint64_t operator()(uint64_t d) const {
        return (deg[dvert[d]] + kIndChunk - 1) / kIndChunk;
    }
};

// xpair[pbase[d] + r] = coff[p] for pair p = (instance i, vertex x), d = dmap[x], r = rank of i
// among the slice's instances holding x: from a row's bits to its pairs' first chunk slots (and so
// to their output positions).
__global__ __launch_bounds__(256) void mdrw_xpair_kernel(const uint32_t* __restrict__ verts,
                                                         const uint64_t* __restrict__ voff, uint32_t i0,
                                                         uint32_t i1, uint32_t W, const uint32_t* __restrict__ dmap,
                                                         const unsigned long long* __restrict__ rows,
                                                         const uint64_t* __restrict__ pbase,
                                                         const uint64_t* __restrict__ coff,
                                                         uint64_t* __restrict__ xpair) {
    const uint64_t p0 = voff[i0], p1 = voff[i1];
    for (uint64_t p = p0 + (uint64_t)blockIdx.x * 256 + threadIdx.x; p < p1; p += (uint64_t)gridDim.x * 256) {
        const uint32_t l = inst_of(voff, i0, i1, p) - i0;
        const uint64_t d = dmap[verts[p]];
        uint64_t r = 0;
        for (uint32_t w = 0; w < (l >> 6); ++w) r += __popcll(rows[d * W + w]);
        r += __popcll(rows[d * W + (l >> 6)] & ((1ull << (l & 63)) - 1));
        xpair[pbase[d] + r] = coff[p];
    }
}

// chunk c of the slice's distinct vertices: ux[c] = dense id, uk[c] = chunk index in its list
__global__ __launch_bounds__(256) void mdrw_uexpand_kernel(const uint64_t* __restrict__ uoff,
                                                           const uint64_t* __restrict__ nd,
                                                           uint32_t* __restrict__ ux, uint32_t* __restrict__ uk) {
    const uint64_t ND = *nd;
    for (uint64_t d = (uint64_t)blockIdx.x * 256 + threadIdx.x; d < ND; d += (uint64_t)gridDim.x * 256) {
        const uint64_t b = uoff[d], e = uoff[d + 1];
        for (uint64_t c = b; c < e; ++c) {
            ux[c] = (uint32_t)d;
            uk[c] = (uint32_t)(c - b);
        }
    }
}

// One wave per chunk of a distinct vertex x.  The wave decodes the chunk once; for every entry y
// and every instance i of the slice in mask(x) & mask(y) it counts (COUNT) or writes (WRITE) the
// edge (x, y) into instance i's induced list (its destination y; mdrw_src_kernel fills the
// sources, x for the whole range of pair (i, x)).  Output slot of (instance i, x, chunk k) is
// coff[pair (i, x)] + k in the instance-major chunk enumeration, so a scan of the counts gives
// every (i, x, k) its position and the edges land ordered by instance, vertex and list position.
// COUNT: each lane adds its own entry's matches to per-wave LDS counters (the order of the adds
// does not matter).  WRITE: lane l keeps the next output slot of instance 64 w + l of each mask
// word w in registers, and per matched instance one ballot ranks the wave's hits in list order.
// MODE 0 counts; MODE 1 writes with 32-bit output slots (fewer than 2^32 induced edges), MODE 2
// with 64-bit ones.
template <int FMT, int MODE>
__global__ __launch_bounds__(kIndBlock) void mdrw_induced_kernel(DevGraph g, const unsigned long long* __restrict__ rows,
                                                                 uint32_t W, const ulonglong2* __restrict__ dbits,
                                                                 const uint32_t* __restrict__ dvert,
                                                                 const uint64_t* __restrict__ nchunks,
                                                                 const uint32_t* __restrict__ ux,
                                                                 const uint32_t* __restrict__ uk,
                                                                 const uint64_t* __restrict__ pbase,
                                                                 const uint64_t* __restrict__ xpair,
                                                                 uint64_t* __restrict__ ccnt,
                                                                 const uint64_t* __restrict__ cpos,
                                                                 uint32_t* __restrict__ idst) {
    constexpr bool WRITE = MODE != 0;
    using Slot = typename std::conditional<MODE == 2, uint64_t, uint32_t>::type;
    __shared__ uint32_t counts[WRITE ? 1 : kIndWaves][kSlice];  // COUNT: per wave and instance
    __shared__ unsigned long long staged[kIndWaves][kWave * kRowPitch];  // the window's rows
    const int lane = lane_id();
    const unsigned long long below = (1ull << lane) - 1;
    const uint64_t NU = *nchunks;
    const uint64_t waves = (uint64_t)gridDim.x * kIndWaves;
    uint32_t* cnt = counts[WRITE ? 0 : threadIdx.x / kWave];
    unsigned long long* R = staged[threadIdx.x / kWave];
    for (uint64_t c = (uint64_t)blockIdx.x * kIndWaves + threadIdx.x / kWave; c < NU; c += waves) {
        const uint32_t dx = __builtin_amdgcn_readfirstlane(ux[c]);
        const uint32_t kc = __builtin_amdgcn_readfirstlane(uk[c]);
        const uint32_t x = __builtin_amdgcn_readfirstlane(dvert[dx]);
        const uint64_t pb = pbase[dx];
        unsigned long long mx[kSliceWords];
        Slot acc[kSliceWords];  // WRITE, lane l: next output slot of instance 64 w + l
        uint32_t rbase = 0;
#pragma unroll
        for (uint32_t w = 0; w < kSliceWords; ++w) {
            const unsigned long long v = w < W ? rows[(uint64_t)dx * W + w] : 0ull;
            mx[w] = readfirstlane_u64(v);
            acc[w] = 0;
            if (WRITE && ((mx[w] >> lane) & 1ull))
                acc[w] = (Slot)cpos[xpair[pb + rbase + __popcll(mx[w] & below)] + kc];
            if (!WRITE) cnt[w * 64 + lane] = 0;
            rbase += __popcll(mx[w]);
        }
        if (!WRITE) wave_sync();
        Hdr h;
        load_hdr<FMT>(g, x, h);
        UDiv d1, d0;
        d1.init(h.q + 1);
        d0.init(h.q);
        const uint32_t kb = kc * kIndChunk;
        const uint32_t ke = min(h.d, kb + kIndChunk);
        uint32_t start = kb;
        if (FMT == kGT && kb) {  // back up to the block head or checkpoint before kb
            uint32_t blk, i;
            gt_locate(h, kb, blk, i);
            start = kb - (i & ((1u << g.cp_shift) - 1));
        }
        // Window pipeline: the next window's entries are decoded and their rows requested before
        // the current window is counted or written, so the row loads (and, in WRITE, the stores
        // of the current window) are in flight together rather than one after the other.
        uint32_t carry = 0;
        // entry base + lane of the chunk -> its value y and its row dy (kNoRow when no sample of the
        // slice holds y, or the entry is not this chunk's)
        auto decode_window = [&](uint32_t base, uint32_t& y, uint32_t& dy) {
            const uint32_t k = base + lane;
            const bool act = k < ke;
            int head;
            y = element_uniform<FMT>(g, h, x, act ? k : ke - 1, d1, d0, head);
            if (FMT == kGT) {
                if (lane == 0 && !head) {
                    y += carry;
                    head = 1;
                }
                y = short_segmented_sum(y, head, 1u << g.cp_shift);  // runs of at most P entries
                carry = __shfl(y, kWave - 1, kWave);
            }
            dy = kNoRow;
            if (act && k >= kb && y < g.n) {
                const ulonglong2 bw = dbits[y >> 6];
                const unsigned long long bit = 1ull << (y & 63);
                if (bw.x & bit) dy = (uint32_t)bw.y + (uint32_t)__popcll(bw.x & (bit - 1));
            }
        };
        // The 64 rows of a window.  With W >= 2 the W / 2 lanes of a row each load 16 bytes of it,
        // so a row is one memory access instead of one per mask word; the loads land in registers
        // (fetch_rows) and go to LDS in the per-lane row layout later (stage_rows).
        const uint32_t cpr = W >> 1, per = W >= 2 ? kWave / cpr : kWave;  // chunks per row, rows per round
        ulonglong2 rv[kSliceWords / 2];
        auto fetch_rows = [&](uint32_t dy) {
            if (W == 1) {
                rv[0].x = dy != kNoRow ? rows[dy] : 0ull;
                return;
            }
#pragma unroll
            for (uint32_t r = 0; r < kSliceWords / 2; ++r) {
                if (r >= cpr) break;
                const uint32_t e = r * per + lane / cpr, q = lane % cpr;
                const uint32_t de = __shfl(dy, (int)e, kWave);
                rv[r] = make_ulonglong2(0ull, 0ull);
                if (de != kNoRow) rv[r] = *(const ulonglong2*)(rows + (uint64_t)de * W + 2 * q);
            }
        };
        auto stage_rows = [&]() {
            if (W == 1) {
                R[lane * kRowPitch] = rv[0].x;
                return;
            }
#pragma unroll
            for (uint32_t r = 0; r < kSliceWords / 2; ++r) {
                if (r >= cpr) break;
                const uint32_t e = r * per + lane / cpr, q = lane % cpr;
                R[e * kRowPitch + 2 * q] = rv[r].x;
                R[e * kRowPitch + 2 * q + 1] = rv[r].y;
            }
        };
        uint32_t y, dy;
        decode_window(start, y, dy);
        fetch_rows(dy);
        stage_rows();
        for (uint32_t base = start; base < ke; base += kWave) {
            const bool more = base + kWave < ke;  // uniform
            uint32_t yn = 0;
            if (more) decode_window(base + kWave, yn, dy);
            wave_sync();
#pragma unroll
            for (uint32_t w = 0; w < kSliceWords; ++w) {
                if (mx[w] == 0ull) continue;  // uniform
                const unsigned long long m = mx[w] & R[lane * kRowPitch + w];
                if (!WRITE) {
                    for (unsigned long long b2 = m; b2; b2 &= b2 - 1)
                        atomicAdd(&cnt[w * 64 + __builtin_ctzll(b2)], 1u);
                    continue;
                }
                // Uniform loop over the instances with at least one hit in the window (a third of
                // x's instances in the window have none), by 32-bit halves of the word: per
                // instance one ballot, one masked store of its run (the sources are filled per pair
                // by mdrw_src_kernel) and one writelane that moves its slot on.
                const unsigned long long bits = wave_or_u64(m);
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    uint32_t hb = (uint32_t)(bits >> (32 * half));
                    const uint32_t mh = (uint32_t)(m >> (32 * half));
                    while (hb) {
                        const int b = __builtin_ctz(hb);
                        hb &= hb - 1;
                        const int src = b + 32 * half;
                        const bool hit = (mh >> b) & 1u;
                        const unsigned long long lanes = __ballot(hit);
                        const Slot slot = lane_get(acc[w], src);
                        if (hit) idst[(uint64_t)slot + lanes_below(lanes)] = y;
                        acc[w] = lane_set(acc[w], src, (Slot)(slot + (Slot)__popcll(lanes)));
                    }
                }
            }
            wave_sync();  // every lane is done with the rows: the next window's replace them
            if (more) {
                fetch_rows(dy);
                stage_rows();
            }
            y = yn;
        }
        if (!WRITE) {
            wave_sync();
            rbase = 0;
#pragma unroll
            for (uint32_t w = 0; w < kSliceWords; ++w) {
                if ((mx[w] >> lane) & 1ull)
                    ccnt[xpair[pb + rbase + __popcll(mx[w] & below)] + kc] = cnt[w * 64 + lane];
                rbase += __popcll(mx[w]);
            }
            wave_sync();  // the counters are cleared for the next chunk
        }
    }
}

// The induced edges of pair p = (instance, x) fill [cpos[coff[p]], cpos[coff[p + 1]]) and all
// have source x.  A workgroup takes a tile of 256 consecutive pairs, loads their bounds and ids
// with one coalesced pass (instead of a dependent pair of random loads per pair), then its waves
// fill the tile's pairs one after another with coalesced stores.
__global__ __launch_bounds__(256) void mdrw_src_kernel(const uint32_t* __restrict__ verts,
                                                       const uint64_t* __restrict__ coff,
                                                       const uint64_t* __restrict__ cpos, uint64_t NV,
                                                       uint32_t* __restrict__ isrc) {
    __shared__ uint64_t bnd[257];
    __shared__ uint32_t xs[256];
    const int lane = lane_id(), wid = threadIdx.x / kWave;
    for (uint64_t t0 = (uint64_t)blockIdx.x * 256; t0 < NV; t0 += (uint64_t)gridDim.x * 256) {
        const uint32_t np = (uint32_t)min<uint64_t>(256, NV - t0);
        if (threadIdx.x < np) bnd[threadIdx.x] = cpos[coff[t0 + threadIdx.x]];
        if (threadIdx.x == 0) bnd[np] = cpos[coff[t0 + np]];
        if (threadIdx.x < np) xs[threadIdx.x] = verts[t0 + threadIdx.x];
        __syncthreads();
        for (uint32_t j = wid; j < np; j += 256 / kWave) {
            const uint32_t x = xs[j];
            const uint64_t e = bnd[j + 1];
            for (uint64_t i = bnd[j] + lane; i < e; i += kWave) isrc[i] = x;
        }
        __syncthreads();  // the next tile's bounds replace these
    }
}

// ind_off[i] = cpos[coff[voff[i]]]
__global__ void mdrw_ind_off_kernel(const uint64_t* __restrict__ voff, const uint64_t* __restrict__ coff,
                                    const uint64_t* __restrict__ cpos, uint32_t I, uint64_t* __restrict__ ioff) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= I) ioff[i] = cpos[coff[voff[i]]];
}

// ---------------------------------------------------------------------------- host

static uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Raises a kernel's dynamic-LDS limit when a launch needs more than 64 KB.  hipFuncSetAttribute
// records a failure as the thread's last error, which later hipGetLastError checks would pick up,
// so the error state is cleared here and the launch itself reports any real problem.
static void allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= 65536) return;
    (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    (void)hipGetLastError();
}

static int sync_read(hipStream_t s, void* host, const void* dev, size_t bytes) {
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return status_of(e);
}

// Working memory, sizes and outputs of one MDRW batch.  The eager entry point sizes everything from
// totals it reads back between the phases and frees the working memory on return.  A captured
// batch (grass_plan_create_mdrw) keeps the state of the eager run that created it: the requests and
// the seed are fixed, so every replay has the same sizes, and the replay neither allocates nor
// reads anything back.
struct MdrwState {
    std::vector<MdDesc> desc;
    uint32_t I = 0, target = 0;
    uint64_t seed = 0;
    uint64_t tot_steps = 0, tot_init = 0, tot_vb = 0;
    uint32_t mmax = 1, vmax = 1;
    uint64_t NS = 0, NV = 0, NC = 0, NI = 0;  // steps, sampled vertices, list chunks, induced edges
    DevMem tmp, tmp2, tmp3;
    grass_mdrw_result* res = nullptr;
    ~MdrwState() {
        if (res) res->release();
    }
};

// Checks the requests and fills the descriptors and bounds.
static int mdrw_prepare(const grass_mdrw_request* reqs, uint32_t I, uint32_t target, uint64_t seed,
                        MdrwState& st) {
    if ((I && !reqs) || target > GRASS_MDRW_MAX_TARGET) return GRASS_EINVAL;
    st.desc.resize(I);
    st.I = I;
    st.target = target;
    st.seed = seed;
    for (uint32_t i = 0; i < I; ++i) {
        const uint32_t m = reqs[i].frontier;
        if (m == 0 || m > GRASS_MDRW_MAX_FRONTIER) return GRASS_EINVAL;
        const uint32_t stp = target > m ? target - m : 0;
        st.desc[i] = MdDesc{st.tot_steps, st.tot_init, st.tot_vb, reqs[i].instance, m, stp, 0};
        st.tot_steps += stp;
        st.tot_init += m;
        st.tot_vb += (uint64_t)m + stp;
        st.mmax = std::max(st.mmax, m);
        st.vmax = std::max(st.vmax, m + stp);
    }
    st.res = new (std::nothrow) grass_mdrw_result();
    return st.res ? GRASS_OK : GRASS_ENOMEM;
}

// Enqueues the batch on s.  replay = false: reads the totals back as they become known and
// allocates the outputs and working memory to fit (then waits for the batch); replay = true:
// takes every size and buffer from a previous eager run on the same state and only enqueues.
static int mdrw_run(grass_graph* g, MdrwState& st, hipStream_t s, bool replay) {
    const uint32_t I = st.I;
    grass_mdrw_result* res = st.res;
    res->device = g->device;
    res->out.count = I;
    if (!replay && res->offs.alloc(3 * (uint64_t)(I + 1) * 8) != hipSuccess) return GRASS_ENOMEM;
    uint64_t* step_off = res->offs.as<uint64_t>();
    uint64_t* vert_off = step_off + (I + 1);
    uint64_t* ind_off = vert_off + (I + 1);
    res->out.step_off_dev = step_off;
    res->out.vert_off_dev = vert_off;
    res->out.ind_off_dev = ind_off;
    if (I == 0) {
        hipLaunchKernelGGL(fill_kernel<uint64_t>, dim3(1), dim3(64), 0, s, step_off, (uint64_t)3, (uint64_t)0);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && !replay) e = hipStreamSynchronize(s);
        return status_of(e);
    }
    const DevGraph dg = g->dev();
    const bool le = g->fmt == kLE;
    const uint64_t seed = st.seed;

    // phase 1: walks and vertex sets in bound layout
    const uint64_t nbI = scan_blocks(I);
    auto plan1 = [&](Carve& c) {
        c.take<MdDesc>(I);
        c.take<unsigned long long>(1);
        c.take<uint32_t>(st.tot_init);
        c.take<uint32_t>(st.tot_steps);
        c.take<uint32_t>(st.tot_steps);
        c.take<uint32_t>(I);
        c.take<uint32_t>(st.tot_vb);
        c.take<uint32_t>(I);
        c.take<uint64_t>(nbI + 2);
        c.take<uint64_t>(1);
    };
    if (!replay) {
        Carve need(nullptr);
        plan1(need);
        if (st.tmp.alloc(need.off + 256) != hipSuccess) return GRASS_ENOMEM;
    }
    Carve c(st.tmp.p);
    MdDesc* d_desc = c.take<MdDesc>(I);
    unsigned long long* d_queue = c.take<unsigned long long>(1);
    uint32_t* d_init = c.take<uint32_t>(st.tot_init);
    uint32_t* d_ssrc = c.take<uint32_t>(st.tot_steps);
    uint32_t* d_sdst = c.take<uint32_t>(st.tot_steps);
    uint32_t* d_scnt = c.take<uint32_t>(I);
    uint32_t* d_vb = c.take<uint32_t>(st.tot_vb);
    uint32_t* d_vcnt = c.take<uint32_t>(I);
    uint64_t* d_part = c.take<uint64_t>(nbI + 2);
    uint64_t* d_tot = c.take<uint64_t>(1);
    // the descriptors are uploaded once (a replay reads the copy the eager run left)
    hipError_t e = replay ? hipSuccess
                          : hipMemcpyAsync(d_desc, st.desc.data(), I * sizeof(MdDesc), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_queue, (uint64_t)1, 0ull);
    {
        // headers in LDS while two instances still fit a CU (20 bytes per frontier entry)
        const bool hdr = st.mmax <= kMdrwHdrLdsMax;
        const size_t lds = (hdr ? 5ull : 2ull) * st.mmax * 4;
        const void* kfn = le ? (hdr ? (const void*)mdrw_walk_kernel<kLE, true> : (const void*)mdrw_walk_kernel<kLE, false>)
                             : (hdr ? (const void*)mdrw_walk_kernel<kGT, true> : (const void*)mdrw_walk_kernel<kGT, false>);
        allow_lds(kfn, lds);
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kfn, kWave, lds);
        const unsigned grid = (unsigned)std::min<uint64_t>(I, (uint64_t)std::max(cus, 1) * std::max(per, 1));
#define GRASS_MDRW_WALK(F, H)                                                                                   \
    hipLaunchKernelGGL((mdrw_walk_kernel<F, H>), dim3(grid), dim3(kWave), lds, s, dg, d_desc, I, d_queue, seed, \
                       d_init, d_ssrc, d_sdst, d_scnt, st.mmax, g->d_err)
        if (le && hdr) GRASS_MDRW_WALK(kLE, true);
        else if (le) GRASS_MDRW_WALK(kLE, false);
        else if (hdr) GRASS_MDRW_WALK(kGT, true);
        else GRASS_MDRW_WALK(kGT, false);
#undef GRASS_MDRW_WALK
    }
    const uint32_t cap = next_pow2(st.vmax);
    hipLaunchKernelGGL(mdrw_verts_kernel, dim3(I), dim3(kSortBlock), (size_t)cap * 4, s, d_desc, d_init,
                       d_sdst, d_scnt, cap, d_vb, d_vcnt);
    e = exclusive_scan(U32Gen{d_scnt}, Count{I, nullptr}, I, step_off, d_part, d_tot, s);
    if (e == hipSuccess) e = exclusive_scan(U32Gen{d_vcnt}, Count{I, nullptr}, I, vert_off, d_part, d_tot, s);
    if (e != hipSuccess) return status_of(e);
    if (!replay) {
        uint64_t totals[2];
        if (hipMemcpyAsync(&totals[0], step_off + I, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return GRASS_EDEVICE;
        int rc = sync_read(s, &totals[1], vert_off + I, 8);
        if (rc) return rc;
        uint32_t err[2] = {0, 0};
        if (hipMemcpy(err, g->d_err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess) return GRASS_EDEVICE;
        if (err[0] == kErrBadNeighbor) {
            (void)hipMemset(g->d_err, 0, sizeof(err));
            return GRASS_EFORMAT;
        }
        st.NS = totals[0];
        st.NV = totals[1];
        if (res->steps.alloc(st.NS * 8) != hipSuccess || res->verts.alloc(st.NV * 4) != hipSuccess)
            return GRASS_ENOMEM;
    }
    const uint64_t NS = st.NS, NV = st.NV;
    res->out.steps = NS;
    res->out.verts = NV;
    uint32_t* o_ssrc = res->steps.as<uint32_t>();
    uint32_t* o_sdst = o_ssrc + NS;
    uint32_t* o_verts = res->verts.as<uint32_t>();
    res->out.step_src_dev = o_ssrc;
    res->out.step_dst_dev = o_sdst;
    res->out.verts_dev = o_verts;
    hipLaunchKernelGGL(mdrw_compact_kernel, dim3(I), dim3(256), 0, s, d_ssrc, d_desc, 0, d_scnt, step_off, o_ssrc);
    hipLaunchKernelGGL(mdrw_compact_kernel, dim3(I), dim3(256), 0, s, d_sdst, d_desc, 0, d_scnt, step_off, o_sdst);
    hipLaunchKernelGGL(mdrw_compact_kernel, dim3(I), dim3(256), 0, s, d_vb, d_desc, 1, d_vcnt, vert_off, o_verts);

    // phase 2: chunks of the sampled vertices' lists
    const uint64_t nbV = scan_blocks(NV);
    if (!replay && st.tmp2.alloc((NV + 1) * 8 + (nbV + 2) * 8 + 512) != hipSuccess) return GRASS_ENOMEM;
    Carve c2(st.tmp2.p);
    uint64_t* d_coff = c2.take<uint64_t>(NV + 1);
    uint64_t* d_part2 = c2.take<uint64_t>(nbV + 2);
    e = exclusive_scan(ChunkGen{o_verts, dg.deg}, Count{NV, nullptr}, NV, d_coff, d_part2, d_part2 + nbV + 1, s);
    if (e != hipSuccess) return status_of(e);
    if (!replay) {
        int rc = sync_read(s, &st.NC, d_coff + NV, 8);
        if (rc) return rc;
    }
    const uint64_t NC = st.NC;

    // phase 3: per slice of up to kSlice instances, membership rows for the slice's distinct
    // vertices; every distinct vertex has its list decoded once per pass (count, then write)
    const uint64_t n = g->n;
    // mask words per row, a power of two (16-byte aligned row chunks for the staged loads)
    uint32_t W = 1;
    while (W * 64 < std::min<uint32_t>(I, kSlice)) W <<= 1;
    const uint64_t nbN = scan_blocks(n), nbC = scan_blocks(NC);
    unsigned long long* d_rows = nullptr;
    uint64_t *d_did = nullptr, *d_pbase = nullptr, *d_uoff = nullptr, *d_nu = nullptr, *d_partN = nullptr,
             *d_ccnt = nullptr, *d_cpos = nullptr, *d_part3 = nullptr;
    uint32_t *d_dmap = nullptr, *d_dvert = nullptr, *d_ux = nullptr, *d_uk = nullptr;
    uint64_t* d_xpair = nullptr;
    ulonglong2* d_dbits = nullptr;
    auto plan3 = [&](Carve& c) {
        d_rows = c.take<unsigned long long>(n * W);  // the slice's distinct vertices use a prefix
        d_did = c.take<uint64_t>(n + 1);
        d_pbase = c.take<uint64_t>(n + 1);
        d_uoff = c.take<uint64_t>(n + 1);
        d_nu = c.take<uint64_t>(1);
        d_partN = c.take<uint64_t>(nbN + 2);
        d_dmap = c.take<uint32_t>(n);
        d_dbits = c.take<ulonglong2>((n + kWave - 1) / kWave);
        d_dvert = c.take<uint32_t>(n);
        d_xpair = c.take<uint64_t>(NV);
        d_ux = c.take<uint32_t>(NC);
        d_uk = c.take<uint32_t>(NC);
        d_ccnt = c.take<uint64_t>(NC + 1);
        d_cpos = c.take<uint64_t>(NC + 1);
        d_part3 = c.take<uint64_t>(nbC + 2);
    };
    if (!replay) {
        Carve need3(nullptr);
        plan3(need3);
        if (st.tmp3.alloc(need3.off + 256) != hipSuccess) return GRASS_ENOMEM;
    }
    Carve c3(st.tmp3.p);
    plan3(c3);
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned sgrid = (unsigned)std::max(cus, 1) * 8;  // grid-stride launches
    const uint64_t* d_nd = d_did + n;                        // distinct vertices of the slice
    // dense ids, rows, their scans, pair ranks and chunk list of slice [i0, i1)
    auto prep = [&](uint32_t i0, uint32_t i1) -> hipError_t {
        const uint32_t* verts = o_verts;
        const uint64_t* voff = vert_off;
        hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(sgrid), dim3(256), 0, s, d_dmap, n, 0u);
        hipLaunchKernelGGL(mdrw_flag_kernel, dim3(sgrid), dim3(256), 0, s, verts, voff, i0, i1, d_dmap);
        hipError_t e = exclusive_scan(U32Gen{d_dmap}, Count{n, nullptr}, n, d_did, d_partN, d_partN + nbN + 1, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(mdrw_dense_kernel, dim3(sgrid), dim3(256), 0, s, (const uint64_t*)d_did, (uint32_t)n,
                           d_dmap, d_dvert);
        hipLaunchKernelGGL(mdrw_dbits_kernel, dim3(sgrid), dim3(256), 0, s, (const uint32_t*)d_dmap,
                           (const uint64_t*)d_did, (uint32_t)n, d_dbits);
        hipLaunchKernelGGL(mdrw_zero_rows_kernel, dim3(sgrid), dim3(256), 0, s, d_rows, d_nd, W);
        hipLaunchKernelGGL(mdrw_mask_kernel, dim3(sgrid), dim3(256), 0, s, verts, voff, i0, i1, W,
                           (const uint32_t*)d_dmap, d_rows);
        e = exclusive_scan(PopGen{d_rows, W}, Count{0, d_nd}, n, d_pbase, d_partN, d_partN + nbN + 1, s);
        if (e == hipSuccess)
            e = exclusive_scan(UChunkGen{d_dvert, dg.deg}, Count{0, d_nd}, n, d_uoff, d_partN, d_nu, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(mdrw_xpair_kernel, dim3(sgrid), dim3(256), 0, s, verts, voff, i0, i1, W,
                           (const uint32_t*)d_dmap, (const unsigned long long*)d_rows, (const uint64_t*)d_pbase,
                           (const uint64_t*)d_coff, d_xpair);
        hipLaunchKernelGGL(mdrw_uexpand_kernel, dim3(sgrid), dim3(256), 0, s, (const uint64_t*)d_uoff, d_nd,
                           d_ux, d_uk);
        return hipGetLastError();
    };
    // mode 0 counts, 1 writes with 32-bit output slots, 2 with 64-bit ones
    auto induced = [&](int mode, uint32_t* idst) {
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(sgrid), dim3(kIndBlock), 0, s, dg, (const unsigned long long*)d_rows, W,
                               (const ulonglong2*)d_dbits, (const uint32_t*)d_dvert, (const uint64_t*)d_nu,
                               (const uint32_t*)d_ux, (const uint32_t*)d_uk, (const uint64_t*)d_pbase,
                               (const uint64_t*)d_xpair, d_ccnt, (const uint64_t*)d_cpos,
                               idst);
        };
        if (le)
            go(mode == 0 ? mdrw_induced_kernel<kLE, 0> : mode == 1 ? mdrw_induced_kernel<kLE, 1> : mdrw_induced_kernel<kLE, 2>);
        else
            go(mode == 0 ? mdrw_induced_kernel<kGT, 0> : mode == 1 ? mdrw_induced_kernel<kGT, 1> : mdrw_induced_kernel<kGT, 2>);
        return hipGetLastError();
    };
    for (uint32_t i0 = 0; i0 < I; i0 += kSlice) {
        e = prep(i0, std::min<uint32_t>(I, i0 + kSlice));
        if (e == hipSuccess) e = induced(0, nullptr);
        if (e != hipSuccess) return status_of(e);
    }
    e = exclusive_scan(U64Gen{d_ccnt}, Count{NC, nullptr}, NC, d_cpos, d_part3, d_part3 + nbC + 1, s);
    if (e != hipSuccess) return status_of(e);
    if (!replay) {
        int rc = sync_read(s, &st.NI, d_cpos + NC, 8);
        if (rc) return rc;
        if (res->ind.alloc(st.NI * 8) != hipSuccess) return GRASS_ENOMEM;
    }
    const uint64_t NI = st.NI;
    res->out.induced = NI;
    // The write pass keeps 32-bit output slots below 2^32 induced edges; GRASS_MDRW_SLOTS64=1
    // takes the 64-bit variant whatever the count (the tests cover it on small batches that way).
    const char* s64 = getenv("GRASS_MDRW_SLOTS64");
    const bool slots64 = s64 && s64[0] == '1';
    uint32_t* o_isrc = res->ind.as<uint32_t>();
    uint32_t* o_idst = o_isrc + NI;
    res->out.ind_src_dev = o_isrc;
    res->out.ind_dst_dev = o_idst;
    for (uint32_t i0 = 0; i0 < I; i0 += kSlice) {
        // a single slice keeps its masks and ranks from the count pass
        e = I > kSlice ? prep(i0, std::min<uint32_t>(I, i0 + kSlice)) : hipSuccess;
        if (e == hipSuccess) e = induced(NI >> 32 || slots64 ? 2 : 1, o_idst);
        if (e != hipSuccess) return status_of(e);
    }
    hipLaunchKernelGGL(mdrw_src_kernel, dim3(sgrid), dim3(256), 0, s, (const uint32_t*)o_verts,
                       (const uint64_t*)d_coff, (const uint64_t*)d_cpos, NV, o_isrc);
    hipLaunchKernelGGL(mdrw_ind_off_kernel, dim3((I + 256) / 256), dim3(256), 0, s, vert_off, d_coff, d_cpos, I,
                       ind_off);
    e = hipGetLastError();
    if (e == hipSuccess && !replay) e = hipStreamSynchronize(s);  // the working memory goes with the state
    return status_of(e);
}

int mdrw(grass_graph* g, const grass_mdrw_request* reqs, uint32_t I, uint32_t target, uint64_t seed,
         hipStream_t s, grass_mdrw_result** out) {
    if (!out) return GRASS_EINVAL;
    *out = nullptr;
    MdrwState st;
    int rc = mdrw_prepare(reqs, I, target, seed, st);
    if (!rc) rc = mdrw_run(g, st, s, false);
    if (rc) return rc;
    *out = st.res;
    st.res = nullptr;
    return GRASS_OK;
}

// ---------------------------------------------------------------------------- captured batches

struct MdrwCapture {
    MdrwState st;
};

int mdrw_capture_prepare(grass_graph* g, const grass_mdrw_request* reqs, uint32_t I, uint32_t target,
                         uint64_t seed, hipStream_t s, MdrwCapture** out) {
    std::unique_ptr<MdrwCapture> c(new (std::nothrow) MdrwCapture());
    if (!c) return GRASS_ENOMEM;
    int rc = mdrw_prepare(reqs, I, target, seed, c->st);
    if (!rc) rc = mdrw_run(g, c->st, s, false);
    if (rc) return rc;
    *out = c.release();
    return GRASS_OK;
}

int mdrw_capture_enqueue(grass_graph* g, MdrwCapture* c, hipStream_t s) { return mdrw_run(g, c->st, s, true); }

const grass_mdrw_result* mdrw_capture_result(const MdrwCapture* c) { return c->st.res; }

void mdrw_capture_free(MdrwCapture* c) { delete c; }

}  // namespace grass

extern "C" {

int grass_mdrw(grass_graph* g, const grass_mdrw_request* reqs, uint32_t count, uint32_t target,
               uint64_t seed, void* stream, grass_mdrw_result** out) {
    if (!g) return GRASS_EINVAL;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) return GRASS_EDEVICE;
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    int rc = grass::mdrw(g, reqs, count, target, seed, s, out);
    (void)hipSetDevice(prev);
    return rc;
}

int grass_mdrw_output_get(const grass_mdrw_result* res, grass_mdrw_output* out) {
    if (!res || !out) return GRASS_EINVAL;
    *out = res->out;
    return GRASS_OK;
}

int grass_mdrw_fetch(const grass_mdrw_result* res, uint64_t* step_off, uint32_t* step_src,
                     uint32_t* step_dst, uint64_t* vert_off, uint32_t* verts, uint64_t* ind_off,
                     uint32_t* ind_src, uint32_t* ind_dst) {
    if (!res) return GRASS_EINVAL;
    const grass_mdrw_output& o = res->out;
    const uint64_t no = (uint64_t)o.count + 1;
    struct Part {
        void* host;
        const void* dev;
        uint64_t bytes;
    } parts[] = {{step_off, o.step_off_dev, no * 8}, {step_src, o.step_src_dev, o.steps * 4},
                 {step_dst, o.step_dst_dev, o.steps * 4}, {vert_off, o.vert_off_dev, no * 8},
                 {verts, o.verts_dev, o.verts * 4},      {ind_off, o.ind_off_dev, no * 8},
                 {ind_src, o.ind_src_dev, o.induced * 4}, {ind_dst, o.ind_dst_dev, o.induced * 4}};
    for (const Part& p : parts) {
        if (!p.host || !p.bytes) continue;
        if (hipMemcpy(p.host, p.dev, p.bytes, hipMemcpyDeviceToHost) != hipSuccess) return GRASS_EDEVICE;
    }
    return GRASS_OK;
}

void grass_mdrw_free(grass_mdrw_result* res) {
    if (res) res->release();
}

}  // extern "C"
