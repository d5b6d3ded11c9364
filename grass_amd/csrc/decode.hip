// Full decode of the resident compressed graph to CSR (rowptr u64[n+1], col u32[m]).
//
// Edge-parallel: each wave64 owns a contiguous range of output slots (a tuning choice, 2048 by
// default), so writes are coalesced and work is balanced whatever the degree skew.  When 64
// consecutive slots lie in one list (most slots: the long lists hold most entries) the wave reads
// that vertex's header once and divides through shared multiply-high reciprocals; otherwise each
// lane finds its vertex from the 64 row ends after the previous owner (one coalesced load and a
// register search), parses that vertex's header and decodes its element:
//   LE: O(1) random access (LE/src/bcg.cuh:196-261);
//   GT: block heads and checkpoints are absolute and every other body is a gap, so the wave
//       turns (absolute | gap+1) values into ids with a segmented prefix sum, carrying the
//       running id from lane 63 into the next 64 slots (the cooperative decode of
//       GT/src/bcg.cuh:171-345 and GT/src/grass.cu:525-553, without its per-warp cache
//       limits).  A wave whose range starts between two absolute positions first backs up to
//       the earlier one.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "devutil.h"
#include "runtime.h"
#include "scan.h"

namespace grass {

constexpr int kDecBlock = 256;

struct DegGen {
    const uint32_t* deg;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return deg[i]; }
};

// Largest v in [lo, hi] with rowptr[v] <= e.
__device__ __forceinline__ uint32_t owner_of(const uint64_t* __restrict__ rowptr, uint32_t lo,
                                             uint32_t hi, uint64_t e) {
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo + 1) / 2;
        if (rowptr[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// wstart[w] = the vertex that holds slot w * per_wave, the first slot of decode wave w: each
// vertex writes the waves that start inside its list.
__global__ __launch_bounds__(256) void wave_start_kernel(const uint64_t* __restrict__ rowptr, uint32_t n,
                                                         uint32_t per_wave, uint64_t waves,
                                                         uint32_t* __restrict__ wstart) {
    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (uint64_t)gridDim.x * 256) {
        const uint64_t b = rowptr[v], e = rowptr[v + 1];
        for (uint64_t w = (b + per_wave - 1) / per_wave; w < waves && w * per_wave < e; ++w)
            wstart[w] = (uint32_t)v;
    }
}

// A GT list in the device stream with its wave-uniform constants: block geometry, field
// widths, the bits of a (q + 1)- and of a q-position block, and shared divisors, so that an
// element costs two multiply-high divisions and one 32-bit window.
struct GtList {
    uint64_t pay;
    uint32_t v, f1, f2, fp, q, R, big, F1, F0, sh;
    UDiv d1, d0;
    __device__ __forceinline__ void init(const Hdr& h, uint32_t vv, uint32_t shift) {
        pay = h.pay;
        v = vv;
        f1 = h.f1;
        f2 = h.f2;
        fp = h.fp;
        q = h.q;
        R = h.R;
        sh = shift;
        big = (q + 1) * R;
        F1 = (uint32_t)gtx_field(h, q + 1, shift);
        F0 = (uint32_t)gtx_field(h, q, shift);
        d1.init(q + 1);
        d0.init(q);
    }
    // Element k: the absolute id at a block head or checkpoint (head = 1), else the gap + 1.  The
    // field's bit, width and kind are chosen first, so every lane issues the same one load.
    template <class Src>
    __device__ __forceinline__ uint32_t at(const Src& w, uint32_t k, int& head) const {
        uint32_t blk, i;
        uint64_t bo;
        if (k < big) {
            blk = d1.div(k);
            i = k - blk * (q + 1);
            bo = pay + (uint64_t)blk * F1;
        } else {
            const uint32_t kk = k - big, b2 = d0.div(kk);
            i = kk - b2 * q;
            blk = b2 + R;
            bo = pay + (uint64_t)R * F1 + (uint64_t)b2 * F0;
        }
        const uint32_t j = i - 1;
        const bool hd = i == 0, abs = !hd && (i & ((1u << sh) - 1)) == 0;
        const uint32_t off = hd ? 0u : f1 + j * f2 + (j >> sh) * (32 - f2);
        const uint32_t raw = w.read(bo + off, hd ? f1 : abs ? 32u : f2);
        head = hd || abs;
        return hd ? (blk >= fp ? v + raw + 1 : v - raw - 1) : abs ? raw : raw + 1;
    }
};

// Where the GT payload words come from: the stream in global memory, or a wave's LDS copy of the
// span of the stream its range of slots covers (L[0] is stream word W0).
struct GlobalWords {
    const uint32_t* __restrict__ w;
    __device__ __forceinline__ uint32_t read(uint64_t b, uint32_t nbits) const { return read_bits(w, b, nbits); }
};
struct LdsWords {
    const uint32_t* L;
    uint64_t W0;
    __device__ __forceinline__ uint32_t read(uint64_t b, uint32_t nbits) const {
        const uint32_t o = (uint32_t)((b >> 5) - W0);
        const uint64_t x = ((uint64_t)L[o] << 32) | L[o + 1];
        return (uint32_t)(x >> (32 - (uint32_t)(b & 31))) >> (32 - nbits);
    }
};

// gt_element (codec.h) on a word source.
template <class Src>
__device__ __forceinline__ uint32_t gt_element_src(const Src& w, const Hdr& h, uint32_t v, uint32_t blk, uint32_t i,
                                                   uint32_t sh, int& head) {
    const uint64_t bo = gtx_block_bit(h, blk, sh);
    const bool hd = i == 0, abs = !hd && (i & ((1u << sh) - 1)) == 0;
    const uint32_t raw = w.read(bo + gtx_field(h, i, sh), hd ? h.f1 : abs ? 32u : h.f2);
    head = hd || abs;
    return hd ? (blk >= h.fp ? v + raw + 1 : v - raw - 1) : abs ? raw : raw + 1;
}

// The slots [e0, e1) of one wave: vlo owns e0, vhi bounds the owner of e1 - 1, decoding starts
// at eb <= e0 (a run's absolute value); GT payload words come from `src`.
template <int FMT, class Src>
__device__ __forceinline__ void decode_range(const DevGraph& g, const Src& src, const uint64_t* __restrict__ rowptr,
                                             uint32_t* __restrict__ col, uint64_t e0, uint64_t e1, uint64_t eb,
                                             uint32_t vlo, uint32_t vhi, uint32_t span) {
    const int lane = lane_id();
    uint32_t carry = 0;
    // owner of the iteration's first slot (a lower bound for every lane's) and its row bounds,
    // wave-uniform, reloaded only when it changes
    uint32_t cur = __builtin_amdgcn_readfirstlane(vlo);
    uint64_t rc = rowptr[cur], rn = rowptr[cur + 1];
    uint32_t hv = 0xFFFFFFFFu;  // vertex whose list constants are cached below
    Hdr uh;
    UDiv d1, d0;
    GtList gl;
    for (uint64_t base = eb; base < e1; base += kWave) {
        const uint64_t e = base + lane;
        const bool act = e < e1;
        const uint64_t ee = act ? e : e1 - 1;
        uint32_t x;
        int head = 1;
        if (base >= rc && base + (kWave - 1) < rn) {
            // the whole wave is inside one list: one header, shared constants, no owner search
            if (cur != hv) {
                hv = cur;
                load_hdr<FMT>(g, hv, uh);
                if (FMT == kGT) {
                    gl.init(uh, hv, g.cp_shift);
                } else {
                    d1.init(uh.q + 1);
                    d0.init(uh.q);
                }
            }
            if (FMT == kGT) x = gl.at(src, (uint32_t)(ee - rc), head);
            else x = element_uniform<FMT>(g, uh, hv, (uint32_t)(ee - rc), d1, d0, head);
        } else {
            // Owner of slot ee: one coalesced load of the 64 row ends after `cur` (and, beside
            // it, of those 64 vertices' headers), then a 6-step search over them in registers;
            // slots more than 64 vertices past `cur` (only possible across vertices without
            // edges) fall back to a binary search of rowptr.  The owner's header then comes from
            // the lane that loaded it, and the next window's row bounds from the row ends.
            // (row ends as 32-bit offsets from the start of `cur`'s list, which is at or before
            // every slot of the iteration; one list is < 2^31 entries, larger offsets saturate)
            const uint64_t bj = (uint64_t)cur + 1 + lane <= g.n ? rowptr[cur + 1 + lane] : ~0ull;
            uint4 hj = make_uint4(0u, 0u, 0u, 0u);
            if (cur + (uint32_t)lane < g.n) hj = g.hdr[cur + lane];
            const uint32_t rj = bj - rc > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)(bj - rc);
            const uint32_t er = (uint32_t)(ee - rc);
            uint32_t pos = 0;
#pragma unroll
            for (uint32_t st = 32; st >= 1; st >>= 1) {
                const uint32_t t = __shfl(rj, (int)(pos + st - 1), kWave);
                if (t <= er) pos += st;
            }
            const uint32_t last = __shfl(rj, kWave - 1, kWave);
            if (pos == kWave - 1 && last <= er) pos = kWave;
            const uint32_t prev = __shfl(rj, (int)(pos > 0 ? pos - 1 : 0), kWave);
            const int src_lane = (int)(pos < (uint32_t)kWave ? pos : 0u);
            uint4 hv4 = make_uint4(__shfl(hj.x, src_lane, kWave), __shfl(hj.y, src_lane, kWave),
                                   __shfl(hj.z, src_lane, kWave), __shfl(hj.w, src_lane, kWave));
            uint32_t v;
            uint64_t vstart;
            if (pos < (uint32_t)kWave) {
                v = cur + pos;
                vstart = rc + (pos ? prev : 0u);
            } else {
                v = owner_of(rowptr, cur + kWave, vhi, ee);
                vstart = rowptr[v];
                hv4 = g.hdr[v];
            }
            const uint32_t p63 = (uint32_t)__builtin_amdgcn_readlane((int)pos, kWave - 1);
            if (p63 < (uint32_t)kWave) {
                if (p63) {
                    cur += p63;
                    rc = readlane_u64(bj, (int)p63 - 1);
                    rn = readlane_u64(bj, (int)p63);
                }
            } else {
                cur = __builtin_amdgcn_readlane((int)v, kWave - 1);
                rc = rowptr[cur];
                rn = rowptr[cur + 1];
            }
            const uint32_t k = (uint32_t)(ee - vstart);
            Hdr h;
            unpack_hdr<FMT>(hv4, h);
            if (FMT == kLE) {
                x = le_get(g.w, h, v, k);
            } else {
                uint32_t blk, i;
                gt_locate(h, k, blk, i);
                x = gt_element_src(src, h, v, blk, i, g.cp_shift, head);
            }
        }
        if (FMT == kGT) {
            if (lane == 0 && !head) {
                x += carry;
                head = 1;
            }
            x = short_segmented_sum(x, head, span);
            carry = (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
        }
        if (act && e >= e0) col[e] = x;
    }
}


// ---------------------------------------------------------------- GT runs (P <= 4)
// A run is a block head or a checkpoint and the gap fields after it up to the next one: at most P
// positions, the first absolute, so a run decodes without any other.  The run kernel gives each
// lane one run: its slots and stream bit from the run's index in its list (no per-position
// division), its fields from one 16-byte load of the stream (a fifth word when they reach past
// it), its ids by adding the gaps in registers (no cross-lane scan).  The 64 runs of a wave's pass
// fill consecutive output slots, which are staged in LDS and stored with coalesced writes.  Runs
// are numbered over the graph (roff = exclusive scan of runs per vertex) and every wave takes a
// fixed number of them, with its first vertex from a table like the slot-based one.

constexpr uint32_t kRunMaxP = 4;  // the run kernel takes P <= kRunMaxP (the default P is 4)

// Runs of a GT vertex with d entries: R blocks of q + 1 positions and C - R of q, each
// ceil(len / P) runs.
__host__ __device__ __forceinline__ uint64_t gt_runs(uint32_t d, uint32_t C, uint32_t q, uint32_t sh) {
    if (!d) return 0;
    const uint32_t R = d - C * q, P1 = (1u << sh) - 1;
    return (uint64_t)R * ((q + 1 + P1) >> sh) + (uint64_t)(C - R) * ((q + P1) >> sh);
}

struct RunGen {  // runs of vertex i, from its header
    const uint4* hdr;
    uint32_t sh;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const {
        const uint4 e = hdr[i];
        return gt_runs(e.z, e.w & 0xFFFFu, e.y >> 16, sh);
    }
};

// One run's place: first output slot, stream bit of its absolute field, positions, whether that
// field is the block head (else a checkpoint), and its block.
struct RunAt {
    uint64_t slot, bit;
    uint32_t cnt, blk;
    bool head;
};

// Run rr (< the vertex's runs) of a vertex: R blocks of nr1 runs, then blocks of nr0.  Div divides
// by nr1 / nr0 (shared multiply-high reciprocals when the wave is in one list, else per lane).
template <class Div1, class Div0>
__device__ __forceinline__ RunAt run_at(const Hdr& h, uint64_t vstart, uint32_t rr, uint32_t sh, const Div1& div1,
                                        const Div0& div0) {
    const uint32_t P1 = (1u << sh) - 1;
    const uint32_t nr1 = (h.q + 1 + P1) >> sh, nr0 = (h.q + P1) >> sh;
    const uint32_t bigr = h.R * nr1;
    uint32_t blk, j, len;
    uint64_t bpos, bbit;
    if (rr < bigr) {
        blk = div1(rr, nr1);
        j = rr - blk * nr1;
        len = h.q + 1;
        bpos = (uint64_t)blk * (h.q + 1);
        bbit = h.pay + (uint64_t)blk * gtx_field(h, h.q + 1, sh);
    } else {
        const uint32_t t = rr - bigr, b2 = div0(t, nr0);
        j = t - b2 * nr0;
        blk = h.R + b2;
        len = h.q;
        bpos = (uint64_t)h.R * (h.q + 1) + (uint64_t)b2 * h.q;
        bbit = h.pay + (uint64_t)h.R * gtx_field(h, h.q + 1, sh) + (uint64_t)b2 * gtx_field(h, h.q, sh);
    }
    const uint32_t i0 = j << sh;
    RunAt a;
    a.cnt = min(1u << sh, len - i0);
    a.slot = vstart + bpos + i0;
    a.bit = bbit + gtx_field(h, i0, sh);
    a.head = i0 == 0;
    a.blk = blk;
    return a;
}

// The ids of one run into out[0 .. cnt): the absolute field, then each gap + 1 added on.  One
// 16-byte load holds the field's offset in its word (< 32), a 32-bit field and three gaps of up to
// 21 bits; a fifth word is read when they need more (at most 31 + 32 + 3 * 31 < 160 bits).
__device__ __forceinline__ void run_ids(const uint32_t* __restrict__ w, const Hdr& h, uint32_t v, const RunAt& a,
                                        uint32_t out[kRunMaxP]) {
    const uint32_t* p = w + (a.bit >> 5);
    const Words4 q = *(const Words4*)p;
    const uint32_t s = (uint32_t)(a.bit & 31);
    const uint32_t aw = a.head ? h.f1 : 32u, f2 = h.f2;
    const uint32_t w4 = s + aw + (a.cnt - 1) * f2 > 128 ? p[4] : 0u;
    uint64_t buf = (((uint64_t)q.x << 32) | q.y) << s;
    uint32_t avail = 64 - s, n0 = q.z, n1 = q.w, n2 = w4;  // buffered bits, then the words after them
    auto take = [&](uint32_t wdt) {  // 1 <= wdt <= 32
        if (avail < wdt) {
            buf |= (uint64_t)n0 << (32 - avail);
            avail += 32;
            n0 = n1;
            n1 = n2;
        }
        const uint32_t r = (uint32_t)(buf >> (64 - wdt));
        buf <<= wdt;
        avail -= wdt;
        return r;
    };
    const uint32_t a0 = take(aw);  // f1 >= 1 for a vertex with entries
    uint32_t x = a.head ? (a.blk >= h.fp ? v + a0 + 1 : v - a0 - 1) : a0;
    out[0] = x;
#pragma unroll
    for (uint32_t k = 1; k < kRunMaxP; ++k) {
        if (k < a.cnt) {
            x += (f2 ? take(f2) : 0u) + 1;  // (0-bit gap fields read as 0)
            out[k] = x;
        }
    }
}

struct UDivBy {  // per-lane division by a list constant (float reciprocal, corrected)
    __device__ __forceinline__ uint32_t operator()(uint32_t n, uint32_t d) const {
        return d == 1 ? n : udiv_small(n, d);  // short lists: one run per block
    }
};
struct UDivShared {  // division by the one divisor it was made for
    UDiv u;
    __device__ __forceinline__ uint32_t operator()(uint32_t n, uint32_t) const { return u.div(n); }
};

// Wave w decodes runs [w * per_wave, ...) of the graph; rstart[w] is the vertex holding its first.
__global__ __launch_bounds__(kDecBlock) void decode_runs_kernel(DevGraph g, const uint64_t* __restrict__ rowptr,
                                                                const uint64_t* __restrict__ roff,
                                                                const uint32_t* __restrict__ rstart,
                                                                uint32_t* __restrict__ col, uint64_t nruns,
                                                                uint32_t per_wave) {
    __shared__ uint32_t stage_lds[(kDecBlock / kWave) * kWave * kRunMaxP];
    __shared__ uint32_t owner_lds[kDecBlock];
    uint32_t* L = stage_lds + (threadIdx.x / kWave) * (kWave * kRunMaxP);
    uint32_t* own = owner_lds + (threadIdx.x & ~(kWave - 1));
    const uint64_t wave = ((uint64_t)blockIdx.x * kDecBlock + threadIdx.x) / kWave;
    const int lane = lane_id();
    const uint64_t r0 = wave * per_wave;
    if (r0 >= nruns) return;  // uniform per wave
    const uint64_t r1 = r0 + per_wave < nruns ? r0 + per_wave : nruns;
    const uint32_t sh = g.cp_shift;
    // the vertex of the pass's first run and its run bounds, wave-uniform
    uint32_t cur = __builtin_amdgcn_readfirstlane(rstart[wave]);
    uint64_t rc = roff[cur], rn = roff[cur + 1];
    uint32_t hv = 0xFFFFFFFFu;  // vertex whose list constants are cached below
    Hdr uh;
    uint64_t uvs = 0;
    UDivShared u1, u0;
    for (uint64_t base = r0; base < r1; base += kWave) {
        const uint64_t rho = base + lane;
        const bool act = rho < r1;
        const uint64_t rr = act ? rho : r1 - 1;
        RunAt a;
        Hdr h;
        uint32_t v;
        if (base >= rc && base + (kWave - 1) < rn) {
            // the whole pass is in one list: one header, shared divisors
            if (cur != hv) {
                hv = cur;
                load_hdr<kGT>(g, hv, uh);
                uvs = rowptr[hv];
                const uint32_t P1 = (1u << sh) - 1;
                u1.u.init((uh.q + 1 + P1) >> sh);
                u0.u.init((uh.q + P1) >> sh);
            }
            h = uh;
            v = hv;
            a = run_at(h, uvs, (uint32_t)(rr - rc), sh, u1, u0);
        } else {
            // The 64 vertices from `cur`: their run ends (relative to rc), headers and first slots
            // in one coalesced load each.  Vertex cur + j covers the pass's lanes from its first
            // run on; each vertex that starts a lane of the pass (or holds its first run) writes
            // j + 1 into that lane's LDS slot, and an inclusive max over the slots gives every
            // lane its vertex (no search).  When those vertices end before the pass does (only
            // across vertices without entries), each lane searches the run offsets instead.
            const uint64_t bj = (uint64_t)cur + 1 + lane <= g.n ? roff[cur + 1 + lane] : ~0ull;
            uint4 hj = make_uint4(0u, 0u, 0u, 0u);
            uint64_t sj = 0;
            if (cur + (uint32_t)lane < g.n) {
                hj = g.hdr[cur + lane];
                sj = rowptr[cur + lane];
            }
            const uint32_t rj = bj - rc > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)(bj - rc);
            const uint32_t pj = dpp_move<0x138>(rj);  // start of vertex cur + lane (lane 0: 0)
            const uint32_t e0 = (uint32_t)(base - rc);
            const uint32_t lastl = (uint32_t)(r1 - base < (uint64_t)kWave ? r1 - base - 1 : kWave - 1);
            const uint32_t elast = e0 + lastl;
            int jo;
            uint64_t rstart_v;
            if ((uint32_t)__builtin_amdgcn_readlane((int)rj, kWave - 1) > elast) {
                const bool cand = rj > pj && rj > e0 && pj <= elast;
                own[lane] = 0;
                wave_sync();
                if (cand) own[pj > e0 ? pj - e0 : 0u] = lane + 1;
                wave_sync();
                jo = (int)wave_inclusive_max(own[lane]) - 1;  // lane 0 always has a vertex
                wave_sync();
                jo = act ? jo : __builtin_amdgcn_readlane(jo, (int)lastl);
                v = cur + jo;
                rstart_v = rc + __shfl(pj, jo, kWave);
            } else {
                v = owner_of(roff, cur, g.n - 1, rr);
                jo = (int)(v - cur < (uint32_t)kWave ? v - cur : 0u);
                rstart_v = roff[v];
            }
            uint4 hv4 = make_uint4(__shfl(hj.x, jo, kWave), __shfl(hj.y, jo, kWave), __shfl(hj.z, jo, kWave),
                                   __shfl(hj.w, jo, kWave));
            uint64_t vs = ((uint64_t)__shfl((uint32_t)(sj >> 32), jo, kWave) << 32) | __shfl((uint32_t)sj, jo, kWave);
            if (v - cur >= (uint32_t)kWave) {
                hv4 = g.hdr[v];
                vs = rowptr[v];
            }
            // the next pass starts at or after the last active lane's vertex
            const uint32_t vl = (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lastl);
            if (vl - cur < (uint32_t)kWave) {
                const int jl = (int)(vl - cur);
                if (jl) {
                    cur = vl;
                    rc = readlane_u64(bj, jl - 1);
                    rn = readlane_u64(bj, jl);
                }
            } else {
                cur = vl;
                rc = roff[cur];
                rn = roff[cur + 1];
            }
            unpack_hdr<kGT>(hv4, h);
            a = run_at(h, vs, (uint32_t)(rr - rstart_v), sh, UDivBy{}, UDivBy{});
        }
        uint32_t ids[kRunMaxP];
        run_ids(g.w, h, v, a, ids);
        // stage the pass's slots [sb, se) and store them coalesced
        const uint64_t sb = readlane_u64(a.slot, 0);
        const uint32_t lastl = (uint32_t)(r1 - base < (uint64_t)kWave ? r1 - base - 1 : kWave - 1);
        const uint32_t count = (uint32_t)(readlane_u64(a.slot, (int)lastl) - sb) +
                               (uint32_t)__builtin_amdgcn_readlane((int)a.cnt, (int)lastl);
        if (act) {
            const uint32_t o = (uint32_t)(a.slot - sb);
#pragma unroll
            for (uint32_t k = 0; k < kRunMaxP; ++k)
                if (k < a.cnt) L[o + k] = ids[k];
        }
        wave_sync();
        for (uint32_t k = lane; k < count; k += kWave) col[sb + k] = L[k];
        wave_sync();  // the stage is read before the next pass writes it
    }
}

// LDS words a wave may copy its span into (the rest of a larger span is read from global memory).
constexpr uint32_t kDecStageMax = 2048;
__host__ __device__ constexpr uint32_t dec_stage_words(uint32_t per_wave) {
    return (per_wave < kDecStageMax ? per_wave : kDecStageMax) + 64;
}

// GT with stage != 0: the wave first copies the span of the payload stream its slots use (every
// field from its first run's absolute value to slot e1) into LDS with coalesced loads, all in
// flight together, when it fits `stage` words, and then decodes from LDS: no dependent global
// load per 64 slots.
template <int FMT>
__global__ __launch_bounds__(kDecBlock) void decode_csr_kernel(DevGraph g,
                                                               const uint64_t* __restrict__ rowptr,
                                                               const uint32_t* __restrict__ wstart,
                                                               uint32_t* __restrict__ col,
                                                               uint64_t m, uint32_t per_wave, uint32_t stage) {
    extern __shared__ uint32_t dec_lds[];
    const uint64_t wave = ((uint64_t)blockIdx.x * kDecBlock + threadIdx.x) / kWave;
    const int lane = lane_id();
    const uint64_t e0 = wave * per_wave;
    if (e0 >= m) return;  // uniform per wave
    const uint64_t e1 = e0 + per_wave < m ? e0 + per_wave : m;
    // owner of e0, and a bound on the owner of e1 - 1 (the next wave's first owner)
    const uint32_t vlo = wstart[wave];
    const uint32_t vhi = e1 < m ? wstart[wave + 1] : g.n - 1;
    // GT: an absolute value (block head or checkpoint) starts every run of at most P positions
    const uint32_t span = 1u << g.cp_shift;
    uint64_t eb = e0;
    if (FMT == kGT) {
        const uint32_t sh = g.cp_shift;
        Hdr h;
        load_hdr<kGT>(g, vlo, h);
        uint32_t blk, i;
        gt_locate(h, (uint32_t)(e0 - rowptr[vlo]), blk, i);
        const uint32_t ib = i & ~(span - 1);
        eb = e0 - (i - ib);  // back to the block head or checkpoint
        if (stage) {
            // bits [B0, B1): from eb's field to slot e1's (the stream's end for the last wave);
            // a field read touches its word and the next, so words W0 ..= ((B1 - 1) >> 5) + 1
            const uint64_t B0 = gtx_block_bit(h, blk, sh) + gtx_field(h, ib, sh);
            Hdr hu;
            load_hdr<kGT>(g, vhi, hu);
            uint64_t B1;
            if (e1 < m) {
                uint32_t b1, i1;
                gt_locate(hu, (uint32_t)(e1 - rowptr[vhi]), b1, i1);
                B1 = gtx_block_bit(hu, b1, sh) + gtx_field(hu, i1, sh);
            } else {
                B1 = hu.pay + gtx_payload_bits(hu, sh);
            }
            const uint64_t W0 = B0 >> 5;
            const uint64_t NW = B1 > B0 ? ((B1 - 1) >> 5) + 2 - W0 : 0;
            if (NW && NW <= stage) {
                uint32_t* L = dec_lds + (threadIdx.x / kWave) * stage;
                const uint32_t nw = (uint32_t)NW;
                for (uint32_t c = 0; c < nw; c += 8 * kWave) {
                    uint32_t t[8];
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) {
                        const uint32_t j = c + u * kWave + lane;
                        t[u] = j < nw ? g.w[W0 + j] : 0u;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 8; ++u) {
                        const uint32_t j = c + u * kWave + lane;
                        if (j < nw) L[j] = t[u];
                    }
                }
                wave_sync();
                decode_range<FMT>(g, LdsWords{L, W0}, rowptr, col, e0, e1, eb, vlo, vhi, span);
                return;
            }
        }
    }
    decode_range<FMT>(g, GlobalWords{g.w}, rowptr, col, e0, e1, eb, vlo, vhi, span);
}

int decode_csr(grass_graph* g, uint64_t* rowptr, uint32_t* col, hipStream_t s) {
    const uint64_t n = g->n, m = g->m;
    const uint64_t nb = scan_blocks(n);
    {
        // The scan's partials have a buffer of their own, ordered between calls on the device:
        // each call's stream waits for the previous call's scan (on whatever stream) before its
        // own scan reuses the buffer, so no call waits on the host.
        std::lock_guard<std::mutex> lock(g->mu);
        if (!g->dec_ev && hipEventCreateWithFlags(&g->dec_ev, hipEventDisableTiming) != hipSuccess)
            return GRASS_ENOMEM;
        hipError_t e = hipStreamWaitEvent(s, g->dec_ev, 0);
        if (e != hipSuccess) return status_of(e);
        if (g->dec_part.reserve((nb + 4) * 8) != hipSuccess) return GRASS_ENOMEM;
        uint64_t* part = (uint64_t*)g->dec_part.ptr;
        e = exclusive_scan(DegGen{g->d_deg}, Count{n, nullptr}, n, rowptr, part, part + nb, s);
        if (e == hipSuccess) e = hipEventRecord(g->dec_ev, s);
        if (e != hipSuccess) return status_of(e);
    }
    if (m == 0 || !col) return GRASS_OK;
    const uint32_t per_wave = g->tune.decode_edges_per_wave.load(std::memory_order_relaxed);
    uint32_t slot = 0;
    while ((512u << slot) < per_wave && slot + 1 < kWaveStartSlots) ++slot;
    if ((512u << slot) != per_wave) return GRASS_EINVAL;
    // GT with P <= 4 decodes by runs: about per_wave slots per wave, per_wave / P runs.  The run
    // offsets (a scan over the headers) and the total are built once per graph.
    const bool runs = g->fmt == kGT && (1u << g->cp_shift) <= kRunMaxP;
    const uint32_t unit = runs ? per_wave >> g->cp_shift : per_wave;
    uint64_t units = m;
    if (runs) {
        std::lock_guard<std::mutex> lock(g->mu);
        if (!g->d_roff) {
            uint64_t* r = nullptr;
            uint64_t* part = nullptr;
            if (hipMalloc(&r, (n + 1) * 8) != hipSuccess) return GRASS_ENOMEM;
            if (hipMalloc(&part, (nb + 4) * 8) != hipSuccess) {
                (void)hipFree(r);
                return GRASS_ENOMEM;
            }
            uint64_t total = 0;
            hipError_t e = exclusive_scan(RunGen{g->d_hdr, g->cp_shift}, Count{n, nullptr}, n, r, part, part + nb, s);
            if (e == hipSuccess) e = hipMemcpyAsync(&total, r + n, 8, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            (void)hipFree(part);
            if (e != hipSuccess) {
                (void)hipFree(r);
                return status_of(e);
            }
            g->d_roff = r;
            g->dec_runs = total;
        }
        units = g->dec_runs;
    }
    const uint64_t* uoff = runs ? (const uint64_t*)g->d_roff : (const uint64_t*)rowptr;
    const uint64_t waves = (units + unit - 1) / unit;
    const uint64_t blocks = (waves + kDecBlock / kWave - 1) / (kDecBlock / kWave);
    if (blocks > 0x7fffffffull) return GRASS_EINVAL;
    // The waves' first vertices depend only on the graph and per_wave: built once per per_wave
    // value from the row pointers (or the run offsets), complete before it is published, never
    // changed after.
    const uint32_t* wstart;
    {
        std::lock_guard<std::mutex> lock(g->mu);
        if (!g->d_wstart[slot]) {
            uint32_t* t = nullptr;
            if (hipMalloc(&t, (waves + 1) * 4) != hipSuccess) return GRASS_ENOMEM;
            const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 16384);
            hipLaunchKernelGGL(wave_start_kernel, dim3(grid), dim3(256), 0, s, uoff, (uint32_t)n, unit, waves, t);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                (void)hipFree(t);
                return status_of(e);
            }
            g->d_wstart[slot] = t;
        }
        wstart = g->d_wstart[slot];
    }
    if (runs) {
        hipLaunchKernelGGL(decode_runs_kernel, dim3((unsigned)blocks), dim3(kDecBlock), 0, s, g->dev(),
                           (const uint64_t*)rowptr, uoff, wstart, col, units, unit);
    } else if (g->fmt == kLE) {
        hipLaunchKernelGGL(decode_csr_kernel<kLE>, dim3((unsigned)blocks), dim3(kDecBlock), 0, s,
                           g->dev(), (const uint64_t*)rowptr, wstart, col, m, per_wave, 0u);
    } else {
        const uint32_t stage = dec_stage_words(per_wave);
        hipLaunchKernelGGL(decode_csr_kernel<kGT>, dim3((unsigned)blocks), dim3(kDecBlock),
                           (size_t)stage * 4 * (kDecBlock / kWave), s, g->dev(), (const uint64_t*)rowptr, wstart,
                           col, m, per_wave, stage);
    }
    return status_of(hipGetLastError());
}

}  // namespace grass
