// Sample export: k-hop mini-batch blocks (per-request local numbering), local ids of MDRW
// samples, and DLPack hand-over of MDRW results.
//
// k-hop blocks.  Request r's sequence is its seeds, then its hop-0 destinations, its hop-1
// destinations, ...; the sequences of all requests are laid end to end (request-major, starts
// Q[r]) so that one position space covers the batch.
//   1. rel_len_kernel + scan: Q.
//   2. rel_insert_kernel: every position p inserts the key (r, vertex) into a hash table and
//      atomicMins p into the key's slot -- the slot ends up holding the first appearance.
//   3. rel_mark_kernel: flag[p] = p is its key's first appearance; a scan of the flags gives F,
//      and the local id of the vertex at p is F[first(p)] - F[Q[r]].
//   4. rel_emit_kernel: writes each first appearance to nodes[F[p]] and the local ids of every
//      batch edge (its destination is a position, its source a key looked up in the table).
// Every step is a data-parallel pass over the positions; the result does not depend on timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>

#include "../../include/grass_export.h"
#include "devutil.h"
#include "runtime.h"
#include "scan.h"

namespace grass {
namespace {

constexpr int kExBlock = 256;
constexpr unsigned kExGrid = 4096;
constexpr uint32_t kNoId = 0xFFFFFFFFu;  // local id of a vertex that is not in the sample

struct RelReq {
    const uint32_t* seeds;
    uint64_t num;
    uint32_t hops;
};

// Largest r in [0, R) with Q[r] <= p (Q ascending).
__device__ __forceinline__ uint32_t req_of(const uint64_t* __restrict__ Q, uint32_t R, uint64_t p) {
    uint32_t lo = 0, hi = R - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (Q[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Open-addressing tables from (request r, vertex v) to the first position of v in r's sequence.
// Both have the same interface; the kernels below take either.  Linear probing from a
// multiplicative hash; a slot, once it holds a key, keeps it, and its position only decreases, so
// every slot ends at the minimum position whatever the order of the inserts.

// 16 bytes per slot: {key, first position}, key = ((r << 32) | v) + 1, key 0 = empty.  Takes any
// request count, vertex id and position; two atomics per new key (claim, then position).
struct RelTab16 {
    unsigned long long* slot;  // 2 * (mask + 1) words
    uint64_t mask;
    uint32_t shift;
    __device__ __forceinline__ static unsigned long long key(uint32_t r, uint32_t v) {
        return (((unsigned long long)r << 32) | v) + 1;
    }
    __device__ __forceinline__ unsigned long long* kp(uint64_t h) const { return slot + 2 * h; }
    __device__ __forceinline__ uint64_t home(unsigned long long k) const {
        return ((k * 0x9E3779B97F4A7C15ull) >> shift) & mask;
    }
    __device__ __forceinline__ void clear(uint64_t i) const {
        reinterpret_cast<ulonglong2*>(slot)[i] = make_ulonglong2(0ull, ~0ull);
    }
    __device__ __forceinline__ void insert(uint32_t r, uint32_t v, uint64_t p, uint32_t* err) const {
        const unsigned long long k = key(r, v);
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            unsigned long long cur = __atomic_load_n(kp(h), __ATOMIC_RELAXED);
            if (cur == 0) {
                cur = atomicCAS(kp(h), 0ull, k);
                if (cur == 0) cur = k;
            }
            if (cur == k) {
                // read first: a repeat after the first occurrence needs no atomic
                if (__atomic_load_n(kp(h) + 1, __ATOMIC_RELAXED) > p) atomicMin(kp(h) + 1, (unsigned long long)p);
                return;
            }
            h = (h + 1) & mask;
        }
        atomicOr(err, 1u);  // cannot happen: slots > positions
    }
    // First position of v in request r; ~0 when absent.
    __device__ __forceinline__ uint64_t first(uint32_t r, uint32_t v) const {
        const unsigned long long k = key(r, v);
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            const ulonglong2 kv = *(const ulonglong2*)kp(h);
            if (kv.x == k) return kv.y;
            if (kv.x == 0) break;
            h = (h + 1) & mask;
        }
        return ~0ull;
    }
};

// 8 bytes per slot when the request, the vertex and the position fit in one word:
// word = (((r << vb) | v) << pb) | position, all ones = empty (positions < 2^pb - 1 keep the low
// bits from being all ones).  A CAS claims an empty slot with the key and the position at once, and
// atomicMin on a word of the same key lowers only its position bits: one atomic per new key.
struct RelTab8 {
    unsigned long long* slot;  // mask + 1 words
    uint64_t mask;
    uint32_t shift, vb, pb;
    __device__ __forceinline__ uint64_t home(unsigned long long k) const {
        return (((k + 1) * 0x9E3779B97F4A7C15ull) >> shift) & mask;
    }
    __device__ __forceinline__ void clear(uint64_t i) const { slot[i] = ~0ull; }
    __device__ __forceinline__ void insert(uint32_t r, uint32_t v, uint64_t p, uint32_t* err) const {
        if (vb < 32 && (v >> vb) != 0) {  // not a vertex of the graph
            atomicOr(err, 1u);
            return;
        }
        const unsigned long long k = ((unsigned long long)r << vb) | v;
        const unsigned long long w = (k << pb) | p;
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            unsigned long long cur = __atomic_load_n(slot + h, __ATOMIC_RELAXED);
            if (cur == ~0ull) {
                cur = atomicCAS(slot + h, ~0ull, w);
                if (cur == ~0ull) return;
            }
            if ((cur >> pb) == k) {
                if (cur > w) atomicMin(slot + h, w);
                return;
            }
            h = (h + 1) & mask;
        }
        atomicOr(err, 1u);  // cannot happen: slots > positions
    }
    __device__ __forceinline__ uint64_t first(uint32_t r, uint32_t v) const {
        if (vb < 32 && (v >> vb) != 0) return ~0ull;
        const unsigned long long k = ((unsigned long long)r << vb) | v;
        uint64_t h = home(k);
        for (uint64_t probes = 0; probes <= mask; ++probes) {
            const unsigned long long cur = slot[h];
            if (cur == ~0ull) break;
            if ((cur >> pb) == k) return cur & ((1ull << pb) - 1);
            h = (h + 1) & mask;
        }
        return ~0ull;
    }
};

template <class Tab>
__global__ void rel_clear_kernel(Tab t) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= t.mask;
         i += (uint64_t)gridDim.x * blockDim.x)
        t.clear(i);
}

// The batch as one position space.  Every offset row has been checked (rel_check_kernel) to be
// ascending and to end at or below the batch total before any pass reads through it, so every
// edge index e below is < total.
struct RelSeq {
    const RelReq* req;
    const uint64_t* off;  // batch offsets, hop-major [H][R + 1]
    const uint32_t* src;
    const uint32_t* dst;
    const uint64_t* Q;  // R + 1 request starts
    uint32_t R, H;
    uint32_t n;  // vertices of the graph: an id >= n is not a vertex
    // Vertex at position p < Q[R]; r = its request, e = its batch edge (~0 for a seed).
    __device__ __forceinline__ uint32_t at(uint64_t p, uint32_t& r, uint64_t& e) const {
        r = req_of(Q, R, p);
        uint64_t j = p - Q[r];
        const RelReq q = req[r];
        e = ~0ull;
        if (j < q.num) return q.seeds[j];
        j -= q.num;
        for (uint32_t h = 0; h < H; ++h) {
            const uint64_t e0 = off[(uint64_t)h * (R + 1) + r];
            const uint64_t c = off[(uint64_t)h * (R + 1) + r + 1] - e0;
            if (j < c) {
                e = e0 + j;
                return dst[e];
            }
            j -= c;
        }
        return n;  // past the request's edges (the checks reject such offsets first)
    }
};

// Checks the batch offsets against the requests before any pass reads through them: every row
// ascending and ending at or below `total`, no edges of request r at hops >= its hop count, and
// the sequence lengths summing to P.  Any failure sets *invalid, and the passes then touch
// nothing.
__global__ void rel_check_kernel(const uint64_t* __restrict__ off, const RelReq* __restrict__ req,
                                 uint32_t R, uint32_t H, uint64_t total, uint32_t* invalid) {
    const uint64_t cells = (uint64_t)H * (R + 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t h = (uint32_t)(i / (R + 1)), r = (uint32_t)(i % (R + 1));
        const uint64_t a = off[i];
        bool ok = a <= total;
        if (r < R) {
            const uint64_t b = off[i + 1];
            ok = ok && a <= b && (h < req[r].hops || a == b);
        }
        if (!ok) atomicOr(invalid, 1u);
    }
}

// Sequence lengths: seeds plus the request's edges at every hop (only after rel_check_kernel).
__global__ void rel_len_kernel(const RelReq* __restrict__ req, const uint64_t* __restrict__ off, uint32_t R,
                               uint32_t H, const uint32_t* __restrict__ invalid, uint64_t* __restrict__ len) {
    const bool bad = *invalid != 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x) {
        uint64_t L = req[r].num;
        if (!bad)
            for (uint32_t h = 0; h < H; ++h) L += off[(uint64_t)h * (R + 1) + r + 1] - off[(uint64_t)h * (R + 1) + r];
        len[r] = L;
    }
}

// The sequence total must be the seeds plus the batch's edges: rows of different hops that
// overlap would make it larger.
__global__ void rel_total_kernel(const uint64_t* __restrict__ Q, uint32_t R, uint64_t P, uint32_t* invalid) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && Q[R] != P) atomicOr(invalid, 1u);
}

struct LenGen {
    const uint64_t* len;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return len[i]; }
};

template <class Tab>
__global__ __launch_bounds__(kExBlock) void rel_insert_kernel(RelSeq q, uint64_t P, Tab t,
                                                              const uint32_t* __restrict__ invalid, uint32_t* err) {
    if (*invalid) return;
    for (uint64_t p = (uint64_t)blockIdx.x * kExBlock + threadIdx.x; p < P; p += (uint64_t)gridDim.x * kExBlock) {
        uint32_t r;
        uint64_t e;
        const uint32_t v = q.at(p, r, e);
        if (v >= q.n) {  // a seed or destination that is not a vertex of the graph
            atomicOr(err, 1u);
            continue;
        }
        t.insert(r, v, p, err);
    }
}

template <class Tab>
__global__ __launch_bounds__(kExBlock) void rel_mark_kernel(RelSeq q, uint64_t P, Tab t,
                                                            const uint32_t* __restrict__ invalid,
                                                            uint32_t* __restrict__ flag) {
    const bool bad = *invalid != 0;
    for (uint64_t p = (uint64_t)blockIdx.x * kExBlock + threadIdx.x; p < P; p += (uint64_t)gridDim.x * kExBlock) {
        uint32_t r;
        uint64_t e;
        if (bad) {
            flag[p] = 0;
            continue;
        }
        const uint32_t v = q.at(p, r, e);
        flag[p] = v < q.n && t.first(r, v) == p;
    }
}

struct FlagGen {
    const uint32_t* flag;
    __device__ __forceinline__ uint64_t operator()(uint64_t i) const { return flag[i]; }
};

template <class Tab>
__global__ __launch_bounds__(kExBlock) void rel_emit_kernel(RelSeq q, uint64_t P, Tab t,
                                                            const uint32_t* __restrict__ flag,
                                                            const uint64_t* __restrict__ F,
                                                            uint32_t* __restrict__ nodes, uint64_t cap,
                                                            uint32_t* __restrict__ lsrc,
                                                            uint32_t* __restrict__ ldst,
                                                            const uint32_t* __restrict__ invalid, uint32_t* err) {
    if (*invalid) return;
    for (uint64_t p = (uint64_t)blockIdx.x * kExBlock + threadIdx.x; p < P; p += (uint64_t)gridDim.x * kExBlock) {
        uint32_t r;
        uint64_t e;
        const uint32_t v = q.at(p, r, e);
        if (flag[p] && F[p] < cap) nodes[F[p]] = v;
        if (e == ~0ull) continue;
        const uint64_t base = F[q.Q[r]];
        const uint64_t fd = flag[p] ? p : v < q.n ? t.first(r, v) : ~0ull;
        const uint64_t fs = t.first(r, q.src[e]);
        if (fd >= P || fs >= P) {  // a source that is neither a seed nor an earlier destination
            atomicOr(err, 1u);
            ldst[e] = lsrc[e] = kNoId;
            continue;
        }
        ldst[e] = (uint32_t)(F[fd] - base);
        lsrc[e] = (uint32_t)(F[fs] - base);
    }
}

__global__ void rel_node_off_kernel(const uint64_t* __restrict__ Q, const uint64_t* __restrict__ F, uint32_t R,
                                    uint64_t* __restrict__ node_off) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= R; r += gridDim.x * blockDim.x) node_off[r] = F[Q[r]];
}

// Clears the table, inserts every position, marks the first appearances, numbers them (scan) and
// writes the sampled vertices and the local ids of every edge.
template <class Tab>
hipError_t rel_build(const Tab& t, const RelSeq& q, uint64_t P, uint32_t* d_flag, uint64_t* d_F, uint64_t* d_part,
                     uint64_t* d_total, const uint32_t* d_invalid, uint32_t* d_bad, grass_khop_blocks* out,
                     hipStream_t s) {
    const unsigned gP = (unsigned)std::min<uint64_t>((P + kExBlock - 1) / kExBlock + 1, kExGrid);
    hipLaunchKernelGGL(rel_clear_kernel<Tab>, dim3(kExGrid), dim3(256), 0, s, t);
    if (P) {
        hipLaunchKernelGGL(rel_insert_kernel<Tab>, dim3(gP), dim3(kExBlock), 0, s, q, P, t, d_invalid, d_bad);
        hipLaunchKernelGGL(rel_mark_kernel<Tab>, dim3(gP), dim3(kExBlock), 0, s, q, P, t, d_invalid, d_flag);
    }
    hipError_t e = exclusive_scan(FlagGen{d_flag}, Count{P, nullptr}, P, d_F, d_part, d_total, s);
    if (e != hipSuccess || !P) return e;
    hipLaunchKernelGGL(rel_emit_kernel<Tab>, dim3(gP), dim3(kExBlock), 0, s, q, P, t, (const uint32_t*)d_flag,
                       (const uint64_t*)d_F, out->nodes_dev, out->node_capacity, out->lsrc_dev, out->ldst_dev,
                       d_invalid, d_bad);
    return hipGetLastError();
}

int khop_blocks(grass_graph* g, const grass_khop_request* reqs, uint32_t R, const grass_khop_output* b,
                grass_khop_blocks* out, hipStream_t s) {
    std::lock_guard<std::mutex> lock(g->mu);
    Workspace& ws = g->ws;
    out->nodes = 0;
    const uint32_t H = b->max_hops;
    if (H > GRASS_KHOP_MAX_HOPS || (H && !b->offsets_dev) || b->total > b->capacity || !out->node_off_dev)
        return GRASS_EINVAL;
    if (b->total && (!b->src_dev || !b->dst_dev || !out->lsrc_dev || !out->ldst_dev)) return GRASS_EINVAL;
    if (ws.reserve_stage(R * sizeof(RelReq) + 64) != hipSuccess) return GRASS_ENOMEM;
    RelReq* hreq = (RelReq*)ws.stage();
    uint64_t S = 0;
    for (uint32_t r = 0; r < R; ++r) {
        if (reqs[r].hops > H || (reqs[r].num_seeds && !reqs[r].seeds_dev)) return GRASS_EINVAL;
        hreq[r] = RelReq{reqs[r].seeds_dev, reqs[r].num_seeds, reqs[r].hops};
        S += reqs[r].num_seeds;
    }
    const uint64_t P = S + b->total;
    if (P && (!out->nodes_dev && out->node_capacity)) return GRASS_EINVAL;
    uint64_t slots = 1024;
    while (slots < P + P / 2 + 2) slots <<= 1;  // load <= 2/3: distinct keys <= positions
    const uint64_t nb = scan_blocks(std::max<uint64_t>(std::max<uint64_t>(P, R), 1));
    const size_t need = Carve::need(R * sizeof(RelReq)) + Carve::need(R * 8) + Carve::need((R + 1) * 8) +
                        Carve::need(P * 4 + 4) + Carve::need((P + 1) * 8) + Carve::need((nb + 2) * 8) +
                        Carve::need(64);
    // Packed 8-byte slots when the request, vertex and position bits fit one word; 16-byte slots
    // otherwise, or when GRASS_EXPORT_SLOT=16 asks for them.
    auto bits = [](uint64_t x) {
        uint32_t k = 0;
        while (k < 64 && (x >> k) != 0) ++k;
        return k;
    };
    const uint32_t rb = bits(R - 1), vb = std::max<uint32_t>(bits(g->n ? g->n - 1 : 0), 1u),
                   pb = std::max<uint32_t>(bits(P), 1u);  // positions < 2^pb - 1
    const char* force = getenv("GRASS_EXPORT_SLOT");
    const bool wide = rb + vb + pb > 64 || (force && strcmp(force, "16") == 0);
    if (ws.scratch.reserve(need) != hipSuccess || ws.ht.reserve(slots * (wide ? 16 : 8)) != hipSuccess)
        return GRASS_ENOMEM;
    // this overwrites the k-hop dedup table's memory: the next k-hop batch clears it first
    ws.ht_gen = 0xFFFF;
    Carve cv(ws.scratch.ptr);
    RelReq* d_req = cv.take<RelReq>(R);
    uint64_t* d_len = cv.take<uint64_t>(R);
    uint64_t* d_Q = cv.take<uint64_t>(R + 1);
    uint32_t* d_flag = cv.take<uint32_t>(P + 1);
    uint64_t* d_F = cv.take<uint64_t>(P + 1);
    uint64_t* d_part = cv.take<uint64_t>(nb + 2);
    uint64_t* d_sc = cv.take<uint64_t>(8);  // [0] sequence total, [1] node total, [2] error flags
    const uint32_t lg = (uint32_t)__builtin_ctzll(slots);

    hipError_t e = hipMemcpyAsync(d_req, hreq, R * sizeof(RelReq), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    // the two halves of d_sc[2]: read back together below
    uint32_t* d_bad = (uint32_t*)(d_sc + 2);  // an id that is not a vertex, or an unknown source
    uint32_t* d_invalid = d_bad + 1;          // offsets that do not fit the requests
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, s, d_bad, (uint64_t)2, 0u);
    const unsigned gR = (unsigned)std::min<uint64_t>((R + 255) / 256 + 1, 1024);
    if (H) {
        const uint64_t cells = (uint64_t)H * (R + 1);
        const unsigned gC = (unsigned)std::min<uint64_t>((cells + 255) / 256, 1024);
        hipLaunchKernelGGL(rel_check_kernel, dim3(gC), dim3(256), 0, s, (const uint64_t*)b->offsets_dev,
                           (const RelReq*)d_req, R, H, b->total, d_invalid);
    }
    hipLaunchKernelGGL(rel_len_kernel, dim3(gR), dim3(256), 0, s, (const RelReq*)d_req,
                       (const uint64_t*)b->offsets_dev, R, H, (const uint32_t*)d_invalid, d_len);
    e = exclusive_scan(LenGen{d_len}, Count{R, nullptr}, R, d_Q, d_part, d_sc, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(rel_total_kernel, dim3(1), dim3(64), 0, s, (const uint64_t*)d_Q, R, P, d_invalid);
    RelSeq q{d_req, b->offsets_dev, b->src_dev, b->dst_dev, d_Q, R, H, (uint32_t)g->n};
    if (wide) {
        RelTab16 t{(unsigned long long*)ws.ht.ptr, slots - 1, 64 - lg};
        e = rel_build(t, q, P, d_flag, d_F, d_part, d_sc + 1, d_invalid, d_bad, out, s);
    } else {
        RelTab8 t{(unsigned long long*)ws.ht.ptr, slots - 1, 64 - lg, vb, pb};
        e = rel_build(t, q, P, d_flag, d_F, d_part, d_sc + 1, d_invalid, d_bad, out, s);
    }
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(rel_node_off_kernel, dim3(gR), dim3(256), 0, s, (const uint64_t*)d_Q, (const uint64_t*)d_F, R,
                       out->node_off_dev);
    uint64_t* hres = ws.result_slot();
    hres[2] = 0;
    e = hipMemcpyAsync(hres, d_Q + R, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hres + 1, d_F + P, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hres + 2, d_bad, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    // the offsets are not the batch of these requests (rows out of order or past the total, edges
    // at hops a request does not have, sizes that differ), an id is not a vertex, or an edge's
    // source is not among the request's earlier vertices
    if (hres[0] != P || hres[2] != 0) return GRASS_EINVAL;
    out->nodes = hres[1];
    return out->nodes > out->node_capacity ? GRASS_ECAPACITY : GRASS_OK;
}

// ------------------------------------------------------------------ MDRW local ids

// Rank of x in the ascending a[0, n); kNoId when absent.
__device__ __forceinline__ uint32_t rank_in(const uint32_t* __restrict__ a, uint64_t n, uint32_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == x ? (uint32_t)lo : kNoId;
}

__global__ __launch_bounds__(kExBlock) void mdrw_local_kernel(const uint64_t* __restrict__ eoff,
                                                              const uint32_t* __restrict__ esrc,
                                                              const uint32_t* __restrict__ edst, uint64_t E,
                                                              uint32_t I, const uint64_t* __restrict__ voff,
                                                              const uint32_t* __restrict__ verts,
                                                              uint32_t* __restrict__ lsrc,
                                                              uint32_t* __restrict__ ldst) {
    for (uint64_t e = (uint64_t)blockIdx.x * kExBlock + threadIdx.x; e < E; e += (uint64_t)gridDim.x * kExBlock) {
        const uint32_t i = req_of(eoff, I, e);
        const uint32_t* V = verts + voff[i];
        const uint64_t nv = voff[i + 1] - voff[i];
        lsrc[e] = rank_in(V, nv, esrc[e]);
        ldst[e] = rank_in(V, nv, edst[e]);
    }
}

// ------------------------------------------------------------------ DLPack

struct DlHold {
    grass_dl_managed_tensor mt;
    int64_t shape[1];
    grass_mdrw_result* res;
};

void dl_release(grass_dl_managed_tensor* self) {
    if (!self) return;
    DlHold* h = (DlHold*)self->manager_ctx;
    grass_mdrw_result* res = h->res;
    delete h;
    res->release();
}

}  // namespace
}  // namespace grass

extern "C" {

int grass_khop_blocks_build(grass_graph* g, const grass_khop_request* reqs, uint32_t count,
                            const grass_khop_output* batch, grass_khop_blocks* out, void* stream) {
    if (!g || !batch || !out || (count && !reqs) || count > (1u << 17)) return GRASS_EINVAL;
    if (count == 0) {
        out->nodes = 0;
        return GRASS_OK;
    }
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) return GRASS_EDEVICE;
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    int rc = grass::khop_blocks(g, reqs, count, batch, out, s);
    (void)hipSetDevice(prev);
    return rc;
}

int grass_mdrw_local_edges(grass_graph* g, const grass_mdrw_result* res, int which, uint32_t* lsrc_dev,
                           uint32_t* ldst_dev, void* stream) {
    if (!g || !res || (which != GRASS_MDRW_EDGES_INDUCED && which != GRASS_MDRW_EDGES_STEPS)) return GRASS_EINVAL;
    const grass_mdrw_output& o = res->out;
    const bool ind = which == GRASS_MDRW_EDGES_INDUCED;
    const uint64_t E = ind ? o.induced : o.steps;
    if (E == 0 || o.count == 0) return GRASS_OK;
    if (!lsrc_dev || !ldst_dev) return GRASS_EINVAL;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) return GRASS_EDEVICE;
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    const unsigned grid = (unsigned)std::min<uint64_t>((E + grass::kExBlock - 1) / grass::kExBlock, grass::kExGrid);
    hipLaunchKernelGGL(grass::mdrw_local_kernel, dim3(grid), dim3(grass::kExBlock), 0, s,
                       ind ? o.ind_off_dev : o.step_off_dev, ind ? o.ind_src_dev : o.step_src_dev,
                       ind ? o.ind_dst_dev : o.step_dst_dev, E, o.count, o.vert_off_dev, o.verts_dev, lsrc_dev,
                       ldst_dev);
    const int rc = grass::status_of(hipGetLastError());
    (void)hipSetDevice(prev);
    return rc;
}

int grass_mdrw_to_dlpack(grass_mdrw_result* res, int which, grass_dl_managed_tensor** out) {
    if (!res || !out || which < GRASS_MDRW_STEP_OFF || which > GRASS_MDRW_IND_DST) return GRASS_EINVAL;
    const grass_mdrw_output& o = res->out;
    const uint64_t no = (uint64_t)o.count + 1;
    const void* data[8] = {o.step_off_dev, o.step_src_dev, o.step_dst_dev, o.vert_off_dev,
                           o.verts_dev,    o.ind_off_dev,  o.ind_src_dev,  o.ind_dst_dev};
    const uint64_t len[8] = {no, o.steps, o.steps, no, o.verts, no, o.induced, o.induced};
    const bool offs = which == GRASS_MDRW_STEP_OFF || which == GRASS_MDRW_VERT_OFF || which == GRASS_MDRW_IND_OFF;
    grass::DlHold* h = new (std::nothrow) grass::DlHold();
    if (!h) return GRASS_ENOMEM;
    h->shape[0] = (int64_t)len[which];
    grass_dl_tensor& t = h->mt.dl_tensor;
    t.data = const_cast<void*>(data[which]);
    t.device = grass_dl_device{10, res->device};
    t.ndim = 1;
    t.dtype = grass_dl_dtype{0, (uint8_t)(offs ? 64 : 32), 1};
    t.shape = h->shape;
    t.strides = nullptr;
    t.byte_offset = 0;
    h->mt.manager_ctx = h;
    h->mt.deleter = grass::dl_release;
    res->refs.fetch_add(1, std::memory_order_relaxed);
    h->res = res;
    *out = &h->mt;
    return GRASS_OK;
}

}  // extern "C"
