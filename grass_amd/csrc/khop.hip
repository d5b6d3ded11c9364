// Batched k-hop neighbourhood sampling without replacement, per-hop dedup.
//
// The reference's k-hop app draws `stepSize(k)` neighbours with replacement per transit and
// keeps no frontier (LE/src/apps/khop/khop.cu:1-24, driven by the transit-parallel engine of
// LE/src/grass.cu:1299-1900).  Here every request of a batch runs in the same launches, with all
// counts kept on the device, so a batch of any size costs a fixed number of launches per hop and
// no host round trip:
//   frontier 0  = the request's seeds, duplicates dropped (first appearance kept);
//   hop h       = for each frontier vertex f (in order): c = min(fanout[h], deg f) distinct
//                 neighbour positions, drawn by Floyd's algorithm with draw k taken from
//                 Philox(seed, {instance, f, h, k}); edges (f, N_f[pos]) in ascending pos;
//   frontier h+1 = destinations of hop h, duplicates dropped (first appearance kept).
// Per hop: one fused count+scan (exclusive scan over items of min(fanout, deg)), a sampling
// kernel (the c draws of a frontier vertex are made in parallel by c lanes -- in groups of a
// fixed width, or packed side by side when the hop's vertices often have fewer neighbours than
// the fanout; Floyd's "taken?" test is resolved in draw order with wave shuffles; ranks give the
// sorted order), and a dedup of (request, vertex) keys in a hash table where the smallest edge
// position wins, followed by a flag scan and a compaction into the next frontier.
// Output is hop-major: edges of (hop h, request r) are [off[h][r], off[h][r+1]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "devutil.h"
#include "philox.h"
#include "runtime.h"
#include "scan.h"

namespace grass {

constexpr int kKhBlock = 256;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

struct KhReq {
    const uint32_t* seeds;
    uint64_t seed_begin;
    uint32_t inst;
    uint32_t hops;
    uint32_t fan[GRASS_KHOP_MAX_HOPS];
};

// Largest r in [0, R-1] with off[r] <= p.
__device__ __forceinline__ uint32_t seg_of(const uint64_t* __restrict__ off, uint32_t R, uint64_t p) {
    uint32_t lo = 0, hi = R - 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo + 1) / 2;
        if (off[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ hash table (dedup)
// Slot h = {key, val} in one aligned 16-byte pair, so that a probe, its insert and its atomicMin
// touch one memory line.  key: (gen << 48) | (request << 31) | vertex; a slot whose generation is
// not the current one is free.  val: ((0xFFFF - gen) << 32) | position, so atomicMin keeps the
// smallest position of the current generation over any stale value.
struct HashTab {
    unsigned long long* slot;  // 2 * (mask + 1) words
    uint64_t mask;
    uint32_t shift;
    __device__ __forceinline__ unsigned long long* key(uint64_t h) const { return slot + 2 * h; }
    __device__ __forceinline__ unsigned long long* val(uint64_t h) const { return slot + 2 * h + 1; }
};

__device__ __forceinline__ uint64_t ht_home(const HashTab& t, uint64_t key48) {
    return ((key48 * 0x9E3779B97F4A7C15ull) >> t.shift) & t.mask;
}

__global__ void ht_clear_kernel(HashTab t) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= t.mask;
         i += (uint64_t)gridDim.x * blockDim.x) {
        *t.key(i) = 0ull;
        *t.val(i) = ~0ull;
    }
}

// Value of the p-th entry of a dedup input, kInvalid when absent or past the output capacity.
struct DedupIn {
    const uint32_t* arr;
    const uint64_t* base;  // device offset added to p (hop base); nullptr = 0
    uint64_t cap;          // entries of arr that exist
    __device__ __forceinline__ uint32_t at(uint64_t p) const {
        uint64_t q = (base ? *base : 0) + p;
        return q < cap ? arr[q] : kInvalid;
    }
};

// The 16-byte generation-tagged table behind the interface the dedup kernels use.
struct GenTab {
    HashTab t;
    uint32_t gen;
    __device__ __forceinline__ void insert(uint32_t r, uint32_t v, uint64_t p, uint32_t* err) const {
        const uint64_t key48 = ((uint64_t)r << 31) | v;
        const unsigned long long full = ((unsigned long long)gen << 48) | key48;
        uint64_t h = ht_home(t, key48);
        uint64_t probes = 0;
        for (;;) {
            unsigned long long cur = __atomic_load_n(t.key(h), __ATOMIC_RELAXED);
            if (cur == full) break;
            if ((uint32_t)(cur >> 48) != gen) {
                unsigned long long prev = atomicCAS(t.key(h), cur, full);
                if (prev == cur || prev == full) break;
                continue;  // the slot was taken meanwhile: look at it again
            }
            h = (h + 1) & t.mask;
            if (++probes > t.mask) {  // cannot happen while slots > entries; never spin
                atomicCAS(err, kErrNone, kErrBadNeighbor);
                return;
            }
        }
        // min of (generation, position): read first, since most repeats come after the first
        // occurrence and then need no atomic (an atomic costs a write to memory)
        const unsigned long long mine = ((unsigned long long)(0xFFFFu - gen) << 32) | (uint32_t)p;
        if (__atomic_load_n(t.val(h), __ATOMIC_RELAXED) > mine) atomicMin(t.val(h), mine);
    }
    __device__ __forceinline__ bool first_is(uint32_t r, uint32_t v, uint64_t p) const {
        const uint64_t key48 = ((uint64_t)r << 31) | v;
        const unsigned long long full = ((unsigned long long)gen << 48) | key48;
        uint64_t h = ht_home(t, key48);
        for (uint64_t probes = 0; probes <= t.mask; ++probes) {
            const ulonglong2 kv = *(const ulonglong2*)t.key(h);  // key and val in one load
            if (kv.x == full) return (uint32_t)kv.y == (uint32_t)p;
            h = (h + 1) & t.mask;
        }
        return false;
    }
};

// 8 bytes per slot when the request, the vertex and the position fit one word:
// ((request << vb | vertex) << pb) | position, all ones = empty (positions < 2^pb - 1).  A CAS
// claims an empty slot with the key and the position at once; atomicMin on a word of the same
// key lowers only its position bits.  Half the footprint of the 16-byte table, one atomic per new
// key, and a pass clears only the slots its own entries need (no generations).
struct PackTab {
    unsigned long long* slot;
    uint64_t mask;
    uint32_t shift, vb, pb;
    __device__ __forceinline__ uint64_t home(unsigned long long k) const {
        return (((k + 1) * 0x9E3779B97F4A7C15ull) >> shift) & mask;
    }
    __device__ __forceinline__ void insert(uint32_t r, uint32_t v, uint64_t p, uint32_t* err) const {
        const unsigned long long k = ((unsigned long long)r << vb) | v;
        const unsigned long long w = (k << pb) | p;
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            // the CAS claims an empty slot, or returns what the slot holds
            const unsigned long long cur = atomicCAS(slot + h, ~0ull, w);
            if (cur == ~0ull) return;
            if ((cur >> pb) == k) {
                if (cur > w) atomicMin(slot + h, w);
                return;
            }
            h = (h + 1) & mask;
        }
        atomicCAS(err, kErrNone, kErrBadNeighbor);  // cannot happen: slots > entries
    }
    __device__ __forceinline__ bool first_is(uint32_t r, uint32_t v, uint64_t p) const {
        const unsigned long long k = ((unsigned long long)r << vb) | v;
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            const unsigned long long cur = slot[h];
            if (cur == ~0ull) break;
            if ((cur >> pb) == k) return (cur & ((1ull << pb) - 1)) == p;
            h = (h + 1) & mask;
        }
        return false;
    }
};

template <class Tab>
__global__ __launch_bounds__(kKhBlock) void dedup_insert_kernel(DedupIn in, Count E,
                                                                const uint64_t* __restrict__ eoff_r,
                                                                uint32_t R, Tab t, uint32_t* err) {
    const uint64_t n = E.get();
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = in.at(p);
        if (v == kInvalid) continue;
        t.insert(seg_of(eoff_r, R, p), v, p, err);
    }
}

// flag[p] = 1 when entry p is the first occurrence of its (request, vertex) key.
template <class Tab>
__global__ __launch_bounds__(kKhBlock) void dedup_mark_kernel(DedupIn in, Count E,
                                                              const uint64_t* __restrict__ eoff_r,
                                                              uint32_t R, Tab t,
                                                              uint32_t* __restrict__ flag) {
    const uint64_t n = E.get();
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = in.at(p);
        flag[p] = v != kInvalid && t.first_is(seg_of(eoff_r, R, p), v, p);
    }
}

struct FlagGen {
    const uint32_t* flag;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return flag[i]; }
};

__global__ __launch_bounds__(kKhBlock) void compact_kernel(DedupIn in, Count E,
                                                           const uint64_t* __restrict__ eoff_r,
                                                           uint32_t R,
                                                           const uint32_t* __restrict__ flag,
                                                           const uint64_t* __restrict__ fpos,
                                                           const uint32_t* __restrict__ deg,
                                                           uint32_t* __restrict__ fv,
                                                           uint32_t* __restrict__ fr,
                                                           uint32_t* __restrict__ fd) {
    const uint64_t n = E.get();
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (uint64_t)gridDim.x * blockDim.x) {
        if (!flag[p]) continue;
        uint64_t q = fpos[p];
        const uint32_t v = in.at(p);
        fv[q] = v;
        fr[q] = seg_of(eoff_r, R, p);
        fd[q] = deg[v];  // read here once; the count scan and the sampler read it in order
    }
}

// foff[r] = first frontier item of request r; icount = frontier size.
__global__ void front_offsets_kernel(const uint64_t* __restrict__ eoff_r, uint32_t R,
                                     const uint64_t* __restrict__ fpos, uint64_t* __restrict__ foff,
                                     uint64_t* icount) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= R; r += gridDim.x * blockDim.x) {
        uint64_t x = fpos[eoff_r[r]];
        foff[r] = x;
        if (r == R) *icount = x;
    }
}

// ------------------------------------------------------------------ seeds -> virtual edges
__global__ __launch_bounds__(kKhBlock) void seed_gather_kernel(const KhReq* __restrict__ reqs,
                                                               const uint64_t* __restrict__ sbeg,
                                                               uint32_t R, uint64_t S, uint32_t n,
                                                               uint32_t* __restrict__ out,
                                                               uint32_t* err) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < S;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t r = seg_of(sbeg, R, p);
        uint32_t v = reqs[r].seeds[p - sbeg[r]];
        if (v >= n) {
            atomicCAS(err, kErrNone, kErrBadStart);
            v = kInvalid;
        }
        out[p] = v;
    }
}

// ------------------------------------------------------------------ per hop
struct CountGen {  // draws of frontier item i at hop h: min(fanout, degree)
    const uint32_t* fd;
    const uint32_t* fr;
    const KhReq* reqs;
    uint32_t h;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const {
        const KhReq& q = reqs[fr[i]];
        uint32_t f = h < q.hops ? q.fan[h] : 0u;
        uint32_t d = fd[i];
        return d < f ? d : f;
    }
};

// eoff_r[r] = first edge (within hop h) of request r; out_off[h][r] = hop base + eoff_r[r];
// hb[h+1] = hb[h] + edges of hop h.
__global__ void hop_offsets_kernel(const uint64_t* __restrict__ foff, uint32_t R,
                                   const uint64_t* __restrict__ eoff, uint64_t* __restrict__ eoff_r,
                                   uint64_t* hb, uint32_t h, uint64_t* __restrict__ out_off) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= R; r += gridDim.x * blockDim.x) {
        uint64_t e = eoff[foff[r]];
        eoff_r[r] = e;
        out_off[(uint64_t)h * (R + 1) + r] = hb[h] + e;
        if (r == R) hb[h + 1] = hb[h] + e;
    }
}

// Lane layouts of the sampling kernel.
enum SampleLanes : int { kLanesPacked = 0, kLanesAuto = 1, kLanesGroups = 2 };

#define GRASS_SAMPLE_PARAMS                                                                              \
    DevGraph g, const uint32_t* __restrict__ fv, const uint32_t* __restrict__ fr,                      \
        const uint32_t* __restrict__ fd, uint64_t I, const KhReq* __restrict__ reqs, uint32_t h,       \
        uint64_t seed, const uint64_t* __restrict__ eoff, uint64_t hbase, uint32_t* __restrict__ osrc, \
        uint32_t* __restrict__ odst, uint64_t cap, uint32_t* err
#define GRASS_SAMPLE_ARGS g, fv, fr, fd, I, reqs, h, seed, eoff, hbase, osrc, odst, cap, err

// Groups of G lanes, one frontier item per group: kWave / G items per wave and pass; when G does
// not divide 64 the last lanes are idle.
template <int FMT, int G>
__device__ __forceinline__ void sample_groups(GRASS_SAMPLE_PARAMS) {
    constexpr int GPW = kWave / G;
    const int lane = lane_id();
    const int grp = lane / G;
    const int gl = lane - grp * G;
    const int gbase = grp * G;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kWave * GPW;
    for (uint64_t i = wave * GPW + grp;; i += groups) {
        const bool has = grp < GPW && i < I;
        if (!__any(has)) break;  // wave-uniform exit
        uint32_t v = 0, c = 0, d = 0, inst = 0;
        uint64_t ebase = 0;
        uint4 hw = make_uint4(0, 0, 0, 0);
        if (has) {
            v = fv[i];
            hw = g.hdr[v];  // issued before the draws, which do not need it
            const KhReq& q = reqs[fr[i]];
            uint32_t f = h < q.hops ? q.fan[h] : 0u;
            d = fd[i];
            c = d < f ? d : f;
            inst = q.inst;
            ebase = hbase + eoff[i];
        }
        const bool all = c == d;  // every neighbour: positions 0..d-1, no draws
        const uint32_t j = d - c + gl;
        uint32_t t = 0;
        if (gl < (int)c && !all) t = (uint32_t)uniform(seed, inst, v, h, gl, (uint64_t)j + 1);
        // Floyd: x_k = t_k unless t_k equals an earlier x, then j_k; resolved in draw order
        bool dup = false;
#pragma unroll 4
        for (int jj = 0; jj < G; ++jj) {
            uint32_t xj = __shfl(dup ? j : t, gbase + jj, kWave);
            if (gl > jj && jj < (int)c && t == xj) dup = true;
        }
        const uint32_t x = all ? (uint32_t)gl : (dup ? j : t);
        uint32_t rank = all ? (uint32_t)gl : 0u;
        if (!all) {
#pragma unroll 4
            for (int jj = 0; jj < G; ++jj) {
                uint32_t xj = __shfl(x, gbase + jj, kWave);
                if (jj < (int)c && xj < x) ++rank;
            }
        }
        if (gl < (int)c) {
            Hdr hd;
            unpack_hdr<FMT>(hw, hd);
            uint32_t nb = get_neighbor<FMT>(g, hd, v, x);
            if (nb >= g.n) {  // never let a bad id reach the next frontier
                atomicCAS(err, kErrNone, kErrBadNeighbor);
                nb = kInvalid;
            }
            const uint64_t p = ebase + rank;
            if (p < cap) {
                osrc[p] = v;
                odst[p] = nb;
            }
        }
    }
}

// Packed lanes: a wave takes a chunk of K consecutive frontier items (one per lane: their vertex,
// request, degree, edge offset and header, loaded once) and lays their c draws side by side in
// rounds of 64 lanes, an item never split between two rounds.  Lane L of a round is draw B + L of
// the chunk (edges numbered from the chunk's first), owned by the item whose [st, st + c) holds
// it: each item of the round writes its lane into an LDS slot at st - B, and an inclusive max over
// the slots hands every lane its owner.  Floyd's test and the ranks run within the owner's lanes
// as in the groups, so the edges are the same; an item of c = 4 takes 4 lanes rather than a
// group, and a round makes no dependent loads besides the neighbour's own.  K is 64, or less when
// the frontier would leave waves idle.  fmax bounds c.
template <int FMT>
__device__ __forceinline__ void sample_packed(GRASS_SAMPLE_PARAMS, int fmax) {
    __shared__ uint32_t owner_lds[kKhBlock];
    uint32_t* own = owner_lds + (threadIdx.x & ~(kWave - 1));
    const int lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / kWave;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    uint32_t K = kWave;
    while (K > 8 && K * nw > I) K >>= 1;
    for (uint64_t base = wave * K; base < I; base += nw * K) {  // wave-uniform
        const uint64_t i = base + lane;
        uint32_t v = 0, c = 0, d = 0, inst = 0;
        uint64_t e = 0;
        uint4 hw = make_uint4(0, 0, 0, 0);
        if (lane < (int)K && i < I) {
            v = fv[i];
            const KhReq& q = reqs[fr[i]];
            const uint32_t f = h < q.hops ? q.fan[h] : 0u;
            d = fd[i];
            c = d < f ? d : f;
            inst = q.inst;
            e = eoff[i];
            if (c) hw = g.hdr[v];
        }
        const uint64_t e0 = readlane_u64(e, 0);  // lane 0 holds item `base` < I
        const uint32_t st = (uint32_t)(e - e0);  // chunk-local first draw (< 64 * 64)
        unsigned long long left = __ballot(c > 0);
        while (left) {  // one round
            const uint32_t B = lane_get(st, __builtin_ctzll(left));
            const bool in = c > 0 && st >= B && st + c <= B + kWave;
            left &= ~__ballot(in);
            own[lane] = 0;
            wave_sync();
            if (in) own[st - B] = lane + 1;
            wave_sync();
            const int j = (int)wave_inclusive_max(own[lane]) - 1;  // slot 0 holds the round's first item
            wave_sync();
            const uint32_t jv = __shfl(v, j, kWave), jc = __shfl(c, j, kWave), jd = __shfl(d, j, kWave);
            const uint32_t js = __shfl(st, j, kWave), ji = __shfl(inst, j, kWave);
            const uint32_t gl = B + lane - js;  // draw of the owner this lane makes
            const bool act = gl < jc;
            const bool all = jc == jd;  // every neighbour: positions 0..d-1, no draws
            const int seg = (int)(js - B);
            const uint32_t jj = jd - jc + gl;
            uint32_t x = gl, rank = gl;
            if (__any(act && !all)) {  // wave-uniform: every lane takes part in the shuffles
                uint32_t t = 0;
                if (act && !all) t = (uint32_t)uniform(seed, ji, jv, h, gl, (uint64_t)jj + 1);
                // Floyd: x_k = t_k unless t_k equals an earlier x, then j_k; resolved in draw order
                bool dup = false;
                for (int k = 0; k < fmax; ++k) {
                    const uint32_t xk = __shfl(dup ? jj : t, seg + k, kWave);
                    if ((int)gl > k && k < (int)jc && t == xk) dup = true;
                }
                const uint32_t xd = dup ? jj : t;
                uint32_t rk = 0;
                for (int k = 0; k < fmax; ++k) {
                    const uint32_t xk = __shfl(xd, seg + k, kWave);
                    if (k < (int)jc && xk < xd) ++rk;
                }
                if (!all) {
                    x = xd;
                    rank = rk;
                }
            }
            uint4 jh;
            jh.x = __shfl(hw.x, j, kWave);
            jh.y = __shfl(hw.y, j, kWave);
            jh.z = __shfl(hw.z, j, kWave);
            jh.w = __shfl(hw.w, j, kWave);
            if (act) {
                Hdr hd;
                unpack_hdr<FMT>(jh, hd);
                uint32_t nb = get_neighbor<FMT>(g, hd, jv, x);
                if (nb >= g.n) {  // never let a bad id reach the next frontier
                    atomicCAS(err, kErrNone, kErrBadNeighbor);
                    nb = kInvalid;
                }
                const uint64_t p = hbase + e0 + js + rank;
                if (p < cap) {
                    osrc[p] = jv;
                    odst[p] = nb;
                }
            }
        }
    }
}

// Whether a hop takes packed lanes: always, never, or (kLanesAuto) when its draws would fill less
// than 70 % of the lanes of groups of G -- seeds, say, which are often short lists under a large
// fanout.  Measured on R-MAT s24 2-hop [25, 10]: hop 0 fills 29 % of 32-lane groups and runs
// twice as fast packed, hop 1 fills 98 % of 10-lane groups and runs 12 % faster in groups.  The
// frontier size and the hop's edge count are on the device by the time the hop runs.
__device__ __forceinline__ bool packed_hop(int lanes, uint64_t edges, uint64_t items, int G) {
    return lanes == kLanesPacked || (lanes == kLanesAuto && edges * 10 < items * G * 7);
}

// One hop's draws in groups of G lanes; returns at once when the hop takes packed lanes.  (The two
// layouts are separate kernels: in one kernel with both, the groups ran 18 % slower.)
template <int FMT, int G>
__global__ __launch_bounds__(kKhBlock) void sample_kernel(DevGraph g, const uint32_t* __restrict__ fv,
                                                          const uint32_t* __restrict__ fr,
                                                          const uint32_t* __restrict__ fd,
                                                          const uint64_t* __restrict__ icount,
                                                          const KhReq* __restrict__ reqs, uint32_t h,
                                                          uint64_t seed,
                                                          const uint64_t* __restrict__ eoff,
                                                          const uint64_t* __restrict__ hb, int lanes,
                                                          uint32_t* __restrict__ osrc,
                                                          uint32_t* __restrict__ odst, uint64_t cap,
                                                          uint32_t* err) {
    const uint64_t I = *icount;
    const uint64_t hbase = hb[h];
    if (packed_hop(lanes, hb[h + 1] - hbase, I, G)) return;
    sample_groups<FMT, G>(GRASS_SAMPLE_ARGS);
}

// One hop's draws in packed lanes; returns at once when the hop takes groups of G.
template <int FMT>
__global__ __launch_bounds__(kKhBlock) void sample_packed_kernel(DevGraph g, const uint32_t* __restrict__ fv,
                                                                 const uint32_t* __restrict__ fr,
                                                                 const uint32_t* __restrict__ fd,
                                                                 const uint64_t* __restrict__ icount,
                                                                 const KhReq* __restrict__ reqs, uint32_t h,
                                                                 uint64_t seed,
                                                                 const uint64_t* __restrict__ eoff,
                                                                 const uint64_t* __restrict__ hb, int fmax,
                                                                 int lanes, int G, uint32_t* __restrict__ osrc,
                                                                 uint32_t* __restrict__ odst, uint64_t cap,
                                                                 uint32_t* err) {
    const uint64_t I = *icount;
    const uint64_t hbase = hb[h];
    if (!packed_hop(lanes, hb[h + 1] - hbase, I, G)) return;
    sample_packed<FMT>(GRASS_SAMPLE_ARGS, fmax);
}
#undef GRASS_SAMPLE_PARAMS
#undef GRASS_SAMPLE_ARGS

int khop_bound(const grass_graph* g, const grass_khop_request* reqs, uint32_t count,
               uint64_t* edges, uint64_t* items, uint64_t* dedup, uint32_t* max_hops) {
    uint64_t E = 0, I = 0, D = 0;
    uint32_t H = 0;
    std::vector<uint64_t> Ih(GRASS_KHOP_MAX_HOPS, 0), Eh(GRASS_KHOP_MAX_HOPS, 0);
    for (uint32_t r = 0; r < count; ++r) {
        const grass_khop_request& q = reqs[r];
        if (q.hops > GRASS_KHOP_MAX_HOPS || (q.num_seeds && !q.seeds_dev)) return GRASS_EINVAL;
        H = std::max(H, q.hops);
        for (uint32_t h = 0; h < q.hops; ++h)
            if (q.fanout[h] > 64) return GRASS_EINVAL;
    }
    // per request: b = frontier bound, e = edge bound of each hop; a request whose hops are
    // done still holds the frontier its last hop produced (with fanout 0 from then on)
    for (uint32_t r = 0; r < count; ++r) {
        const grass_khop_request& q = reqs[r];
        uint64_t b = std::min<uint64_t>(q.num_seeds, g->n);
        for (uint32_t h = 0; h < H; ++h) {
            Ih[h] += b;
            uint64_t e = h < q.hops ? b * std::min<uint64_t>(q.fanout[h], g->max_degree) : 0;
            Eh[h] += e;
            b = std::min<uint64_t>(e, g->n);
        }
    }
    uint64_t S = 0;
    for (uint32_t r = 0; r < count; ++r) S += reqs[r].num_seeds;
    D = S;
    for (uint32_t h = 0; h < H; ++h) {
        E += Eh[h];
        I = std::max(I, Ih[h]);
        if (h + 1 < H) D = std::max(D, Eh[h]);
    }
    I = std::max(I, S);
    if (edges) *edges = E;
    if (items) *items = I;
    if (dedup) *dedup = D;
    if (max_hops) *max_hops = H;
    return GRASS_OK;
}

// Persistent grid of one sample-kernel instantiation (computed once per instantiation).
template <int FMT, int G>
static unsigned sample_grid() {
    static unsigned grid = 0;
    if (!grid) grid = persistent_grid((const void*)sample_kernel<FMT, G>, kKhBlock);
    return grid;
}

// Hop h in `lanes` layout (SampleLanes) with groups of G lanes: the smallest of 4, 6, 8, 10, 12,
// 16, 21, 32, 64 that covers the hop's largest fanout (twice that with the tuned scale 2), so that
// a fanout-10 hop runs six vertices per wave instead of four in 16-lane groups.  kLanesAuto
// launches both kernels and the one the hop does not take returns at once.
template <int FMT>
static void launch_sample(grass_graph* g, uint32_t G, uint32_t fmax, int lanes, hipStream_t s,
                          const uint32_t* fv, const uint32_t* fr, const uint32_t* fd, const uint64_t* ic,
                          const KhReq* reqs, uint32_t h, uint64_t seed, const uint64_t* eoff, const uint64_t* hb,
                          uint32_t* osrc, uint32_t* odst, uint64_t cap, uint32_t* err) {
    if (lanes != kLanesGroups) {
        static unsigned grid = 0;
        if (!grid) grid = persistent_grid((const void*)sample_packed_kernel<FMT>, kKhBlock);
        hipLaunchKernelGGL(sample_packed_kernel<FMT>, dim3(grid), dim3(kKhBlock), 0, s, g->dev(), fv, fr, fd, ic,
                           reqs, h, seed, eoff, hb, (int)fmax, lanes, (int)G, osrc, odst, cap, err);
        if (lanes == kLanesPacked) return;
    }
#define GRASS_SAMPLE(GG)                                                                                   \
    hipLaunchKernelGGL((sample_kernel<FMT, GG>), dim3(sample_grid<FMT, GG>()), dim3(kKhBlock), 0, s, g->dev(), \
                       fv, fr, fd, ic, reqs, h, seed, eoff, hb, lanes, osrc, odst, cap, err)
    if (G <= 4) GRASS_SAMPLE(4);
    else if (G <= 6) GRASS_SAMPLE(6);
    else if (G <= 8) GRASS_SAMPLE(8);
    else if (G <= 10) GRASS_SAMPLE(10);
    else if (G <= 12) GRASS_SAMPLE(12);
    else if (G <= 16) GRASS_SAMPLE(16);
    else if (G <= 21) GRASS_SAMPLE(21);
    else if (G <= 32) GRASS_SAMPLE(32);
    else GRASS_SAMPLE(64);
#undef GRASS_SAMPLE
}

static uint32_t bit_width(uint64_t x) {
    uint32_t k = 0;
    while (k < 64 && (x >> k) != 0) ++k;
    return k;
}

// Slots of the packed dedup table for `entries` positions (load <= 2/3).
static uint64_t pack_slots(uint64_t entries) {
    uint64_t slots = 1024;
    while (slots < entries + entries / 2 + 2) slots <<= 1;
    return slots;
}

// The batch's dedup table is the 8-byte packed one when (request, vertex, position) fit one word;
// GRASS_KHOP_SLOT=16 forces the 16-byte generation-tagged table.
static bool kh_packed_table(const grass_graph* g, uint32_t R, uint64_t S, uint64_t Dmax) {
    const uint32_t vb = std::max<uint32_t>(bit_width(g->n ? g->n - 1 : 0), 1u);
    const uint32_t pb = std::max<uint32_t>(bit_width(std::max(S, Dmax)), 1u);
    const char* force16 = getenv("GRASS_KHOP_SLOT");
    return bit_width(R - 1) + vb + pb <= 64 && !(force16 && strcmp(force16, "16") == 0);
}

// Device scratch of a batch of R requests, S seeds, frontier bound Imax and dedup bound Dmax.
static size_t kh_scratch_need(uint64_t R, uint64_t S, uint64_t Imax, uint64_t Dmax) {
    const uint64_t nb = std::max(scan_blocks(Imax), scan_blocks(Dmax));
    return Carve::need(R * sizeof(KhReq)) + Carve::need((R + 1) * 8) * 4 + Carve::need(S * 4 + 4) +
           Carve::need(Imax * 4) * 6 + Carve::need((Imax + 1) * 8) + Carve::need(Dmax * 4) +
           Carve::need((Dmax + 1) * 8) + Carve::need(nb * 8) + Carve::need(64 * 8);
}

static hipError_t ensure_table(Workspace& ws, uint64_t entries, HashTab& t) {
    uint64_t slots = 1024;
    while (slots < 2 * entries + 2) slots <<= 1;
    if (slots > ws.ht_slots) {
        hipError_t e = ws.ht.reserve(slots * 16);
        if (e != hipSuccess) return e;
        ws.ht_slots = slots;
        ws.ht_gen = 0xFFFF;  // force a clear
    }
    t.slot = (unsigned long long*)ws.ht.ptr;
    t.mask = ws.ht_slots - 1;
    uint32_t lg = 0;
    while ((1ull << lg) < ws.ht_slots) ++lg;
    t.shift = 64 - lg;
    return hipSuccess;
}

static uint32_t next_gen(Workspace& ws, const HashTab& t, hipStream_t s) {
    if (ws.ht_gen >= 0xFFFF) {
        hipLaunchKernelGGL(ht_clear_kernel, dim3(4096), dim3(256), 0, s, t);
        ws.ht_gen = 0;
    }
    return ++ws.ht_gen;
}

int khop_batch(grass_graph* g, const grass_khop_request* reqs, uint32_t R, uint64_t seed,
               grass_khop_output* out, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g->mu);
    int rc = khop_enqueue(g, reqs, R, seed, out, s, g->ws);
    if (rc || out->max_hops == 0) return rc;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    out->total = *g->ws.result_slot();
    return out->total > out->capacity ? GRASS_ECAPACITY : GRASS_OK;
}

int khop_reserve(const grass_graph* g, Workspace& ws, uint32_t R, uint64_t S, uint64_t I, uint64_t D) {
    if (!R) return GRASS_OK;
    const uint64_t Imax = std::max<uint64_t>(I, 1), Dmax = std::max<uint64_t>(D, 1);
    if (ws.reserve_stage(R * sizeof(KhReq) + (R + 1) * 8 + 64) != hipSuccess ||
        ws.scratch.reserve(kh_scratch_need(R, S, Imax, Dmax)) != hipSuccess)
        return GRASS_ENOMEM;
    if (kh_packed_table(g, R, S, Dmax)) {
        if (ws.ht.reserve(pack_slots(std::max(S, Dmax)) * 8) != hipSuccess) return GRASS_ENOMEM;
        ws.ht_gen = 0xFFFF;
    } else {
        HashTab t{};
        if (ensure_table(ws, Dmax, t) != hipSuccess) return GRASS_ENOMEM;
    }
    return GRASS_OK;
}

int khop_enqueue(grass_graph* g, const grass_khop_request* reqs, uint32_t R, uint64_t seed,
                 grass_khop_output* out, hipStream_t s, Workspace& ws) {
    // the edge total goes to ws.result_slot(), taken after reserve_stage may have moved it
    out->total = 0;
    out->max_hops = 0;
    if (R == 0) return GRASS_OK;
    if (R > (1u << 17)) return GRASS_EINVAL;
    uint64_t Ebound, Ibound, Dbound;
    uint32_t H;
    int rc = khop_bound(g, reqs, R, &Ebound, &Ibound, &Dbound, &H);
    if (rc) return rc;
    out->max_hops = H;
    if (H == 0) return GRASS_OK;
    if (!out->offsets_dev || (out->capacity && (!out->src_dev || !out->dst_dev))) return GRASS_EINVAL;
    if (Dbound >= 0xFFFFFFFFull) return GRASS_EINVAL;  // positions are 32-bit in the dedup table

    // descriptors staged in the workspace's pinned memory (the copy below is truly async)
    if (ws.reserve_stage(R * sizeof(KhReq) + (R + 1) * 8 + 64) != hipSuccess) return GRASS_ENOMEM;
    KhReq* hreq = (KhReq*)ws.stage();
    uint64_t* sbeg = (uint64_t*)(hreq + R);
    sbeg[0] = 0;
    for (uint32_t r = 0; r < R; ++r) {
        KhReq q{};
        q.seeds = reqs[r].seeds_dev;
        q.seed_begin = sbeg[r];
        q.inst = reqs[r].instance;
        q.hops = reqs[r].hops;
        for (int h = 0; h < GRASS_KHOP_MAX_HOPS; ++h) q.fan[h] = h < (int)reqs[r].hops ? reqs[r].fanout[h] : 0;
        hreq[r] = q;
        sbeg[r + 1] = sbeg[r] + reqs[r].num_seeds;
    }
    const uint64_t S = sbeg[R];
    const uint64_t Imax = std::max<uint64_t>(Ibound, 1), Dmax = std::max<uint64_t>(Dbound, 1);
    const uint64_t nbI = scan_blocks(Imax), nbD = scan_blocks(Dmax);
    hipError_t e = ws.scratch.reserve(kh_scratch_need(R, S, Imax, Dmax));
    if (e != hipSuccess) return GRASS_ENOMEM;
    Carve cv(ws.scratch.ptr);
    KhReq* d_req = cv.take<KhReq>(R);
    uint64_t* d_sbeg = cv.take<uint64_t>(R + 1);
    uint64_t* d_foff[2] = {cv.take<uint64_t>(R + 1), cv.take<uint64_t>(R + 1)};
    uint64_t* d_eoff_r = cv.take<uint64_t>(R + 1);
    uint32_t* d_seedv = cv.take<uint32_t>(S + 1);
    uint32_t* d_fv[2] = {cv.take<uint32_t>(Imax), cv.take<uint32_t>(Imax)};
    uint32_t* d_fr[2] = {cv.take<uint32_t>(Imax), cv.take<uint32_t>(Imax)};
    uint32_t* d_fd[2] = {cv.take<uint32_t>(Imax), cv.take<uint32_t>(Imax)};  // frontier degrees
    uint64_t* d_eoff = cv.take<uint64_t>(Imax + 1);
    uint32_t* d_flag = cv.take<uint32_t>(Dmax);
    uint64_t* d_fpos = cv.take<uint64_t>(Dmax + 1);
    uint64_t* d_part = cv.take<uint64_t>(std::max(nbI, nbD));
    uint64_t* d_sc = cv.take<uint64_t>(64);  // [0..1] item counts, [2] edges total, [3] flags total,
                                             // [8..8+H] hop bases
    uint64_t* d_icount[2] = {d_sc + 0, d_sc + 1};
    uint64_t* d_etot = d_sc + 2;
    uint64_t* d_ftot = d_sc + 3;
    uint64_t* d_hb = d_sc + 8;

    // Dedup table: 8-byte packed slots when (request, vertex, position) fit one word, else the
    // 16-byte generation-tagged table.  GRASS_KHOP_SLOT=16 forces the latter.
    const uint32_t vb = std::max<uint32_t>(bit_width(g->n ? g->n - 1 : 0), 1u);
    const uint32_t pb = std::max<uint32_t>(bit_width(std::max(S, Dmax)), 1u);  // positions < 2^pb - 1
    const bool packed = kh_packed_table(g, R, S, Dmax);
    HashTab ht{};
    if (packed) {
        if (ws.ht.reserve(pack_slots(std::max(S, Dmax)) * 8) != hipSuccess) return GRASS_ENOMEM;
        ws.ht_gen = 0xFFFF;  // the 16-byte table's next use starts from a clear
    } else {
        e = ensure_table(ws, Dmax, ht);
        if (e != hipSuccess) return GRASS_ENOMEM;
    }
    // the table for `entries` positions: packed (cleared now, only the slots it uses) or the next
    // generation of the 16-byte one
    auto pack_tab = [&](uint64_t entries) {
        const uint64_t slots = pack_slots(entries);
        const unsigned grid = (unsigned)std::min<uint64_t>((slots + 255) / 256, 4096);
        hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(grid), dim3(256), 0, s,
                           (unsigned long long*)ws.ht.ptr, slots, ~0ull);
        return PackTab{(unsigned long long*)ws.ht.ptr, slots - 1, (uint32_t)(64 - __builtin_ctzll(slots)), vb, pb};
    };
    auto dedup = [&](const DedupIn& in, Count E, const uint64_t* off_r, uint64_t entries) {
        if (packed) {
            const PackTab t = pack_tab(entries);
            hipLaunchKernelGGL(dedup_insert_kernel<PackTab>, dim3(2048), dim3(kKhBlock), 0, s, in, E, off_r, R, t,
                               err_word(g, ws));
            hipLaunchKernelGGL(dedup_mark_kernel<PackTab>, dim3(2048), dim3(kKhBlock), 0, s, in, E, off_r, R, t,
                               d_flag);
        } else {
            const GenTab t{ht, next_gen(ws, ht, s)};
            hipLaunchKernelGGL(dedup_insert_kernel<GenTab>, dim3(2048), dim3(kKhBlock), 0, s, in, E, off_r, R, t,
                               err_word(g, ws));
            hipLaunchKernelGGL(dedup_mark_kernel<GenTab>, dim3(2048), dim3(kKhBlock), 0, s, in, E, off_r, R, t,
                               d_flag);
        }
    };

    e = hipMemcpyAsync(d_req, hreq, R * sizeof(KhReq), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sbeg, sbeg, (R + 1) * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(fill_kernel<uint64_t>, dim3(1), dim3(64), 0, s, d_sc, (uint64_t)64, (uint64_t)0);

    const unsigned gridE = 2048;
    uint32_t* err = err_word(g, ws);
    // frontier 0: seeds, deduplicated
    const uint32_t n32 = (uint32_t)g->n;
    if (S) hipLaunchKernelGGL(seed_gather_kernel, dim3(gridE), dim3(kKhBlock), 0, s, d_req, d_sbeg, R, S, n32,
                              d_seedv, err);
    {
        DedupIn in{d_seedv, nullptr, S};
        dedup(in, Count{S, nullptr}, d_sbeg, S);
        e = exclusive_scan(FlagGen{d_flag}, Count{S, nullptr}, Dmax, d_fpos, d_part, d_ftot, s);
        if (e != hipSuccess) return status_of(e);
        hipLaunchKernelGGL(compact_kernel, dim3(gridE), dim3(kKhBlock), 0, s, in, Count{S, nullptr}, d_sbeg, R,
                           d_flag, d_fpos, (const uint32_t*)g->d_deg, d_fv[0], d_fr[0], d_fd[0]);
        hipLaunchKernelGGL(front_offsets_kernel, dim3((R + 256) / 256), dim3(256), 0, s, d_sbeg, R, d_fpos,
                           d_foff[0], d_icount[0]);
    }
    // group width of each hop and the hop's largest fanout; the tuned scale picks the layout:
    // 0 packed lanes, 1 packed or groups as the hop's draws fill the groups, 2 groups of twice
    // the width
    uint32_t lanes[GRASS_KHOP_MAX_HOPS], fmax[GRASS_KHOP_MAX_HOPS];
    const uint32_t scale = g->tune.khop_lane_scale.load(std::memory_order_relaxed);
    const int layout = scale == 0 ? kLanesPacked : scale == 1 ? kLanesAuto : kLanesGroups;
    for (uint32_t h = 0; h < H; ++h) {
        uint32_t f = 0;
        for (uint32_t r = 0; r < R; ++r)
            if (h < reqs[r].hops) f = std::max(f, reqs[r].fanout[h]);
        fmax[h] = f;
        const uint32_t need = f * std::max(scale, 1u);
        uint32_t G = 64;
        for (uint32_t c : {4u, 6u, 8u, 10u, 12u, 16u, 21u, 32u})
            if (c >= need) {
                G = c;
                break;
            }
        lanes[h] = G;
    }
    for (uint32_t h = 0; h < H; ++h) {
        const int cur = h & 1, nxt = cur ^ 1;
        CountGen cg{d_fd[cur], d_fr[cur], d_req, h};
        e = exclusive_scan(cg, Count{0, d_icount[cur]}, Imax, d_eoff, d_part, d_etot, s);
        if (e != hipSuccess) return status_of(e);
        hipLaunchKernelGGL(hop_offsets_kernel, dim3((R + 256) / 256), dim3(256), 0, s, d_foff[cur], R, d_eoff,
                           d_eoff_r, d_hb, h, out->offsets_dev);
        if (g->fmt == kLE)
            launch_sample<kLE>(g, lanes[h], fmax[h], layout, s, d_fv[cur], d_fr[cur], d_fd[cur], d_icount[cur],
                               d_req, h, seed, d_eoff, d_hb, out->src_dev, out->dst_dev, out->capacity, err);
        else
            launch_sample<kGT>(g, lanes[h], fmax[h], layout, s, d_fv[cur], d_fr[cur], d_fd[cur], d_icount[cur],
                               d_req, h, seed, d_eoff, d_hb, out->src_dev, out->dst_dev, out->capacity, err);
        if (h + 1 == H) break;
        DedupIn in{out->dst_dev, d_hb + h, out->capacity};
        Count E{0, d_etot};
        dedup(in, E, d_eoff_r, Dmax);
        e = exclusive_scan(FlagGen{d_flag}, E, Dmax, d_fpos, d_part, d_ftot, s);
        if (e != hipSuccess) return status_of(e);
        hipLaunchKernelGGL(compact_kernel, dim3(gridE), dim3(kKhBlock), 0, s, in, E, d_eoff_r, R, d_flag, d_fpos,
                           (const uint32_t*)g->d_deg, d_fv[nxt], d_fr[nxt], d_fd[nxt]);
        hipLaunchKernelGGL(front_offsets_kernel, dim3((R + 256) / 256), dim3(256), 0, s, d_eoff_r, R, d_fpos,
                           d_foff[nxt], d_icount[nxt]);
    }
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(ws.result_slot(), d_hb + H, 8, hipMemcpyDeviceToHost, s);
    return status_of(e);
}

}  // namespace grass
