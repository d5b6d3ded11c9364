// Uniform random walks over the compressed graph.
//
// One lane per walker; every step is one header parse plus one random-access decode of the
// current vertex's record, so no transit grouping, sort or host round trip is needed between
// steps (the reference regroups walkers by transit with a radix sort, run-length encode and
// three D2H copies per step, LE/src/grass.cu:1453-1861).  Step s of walker w draws its
// neighbour index from Philox(seed, {instance, w, s, 0}) (DeepWalkApp::next,
// LE/src/apps/randomwalk/randomWalks.cu:62-79, with cuRAND replaced).
//
// plain:   one launch, one request (grass_random_walk);
// batched: many requests of different walker counts and lengths in one persistent launch;
//          waves pull 64-walker granules of one request from a device queue, longest walks
//          first.  Outputs are addressed by (request, walker), so the schedule never changes
//          a result.
// COUNT variants write nothing and instead sum the algorithmic bytes of the same walks:
//   per walker 4 (start written) [+ 4 start read], per visited vertex 8 (record start) +
//   ceil(header bits / 8), per step ceil(payload bits read / 8) + 4 (id written).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "devutil.h"
#include "philox.h"
#include "runtime.h"

namespace grass {

constexpr int kRwBlock = 256;

__device__ __forceinline__ uint64_t bytes_of_bits(uint32_t bits) { return (bits + 7) >> 3; }

template <int FMT, bool COUNT>
__device__ __forceinline__ uint64_t walk_one(const DevGraph& g, uint32_t v, uint32_t L,
                                             uint64_t seed, uint32_t inst, uint32_t w,
                                             uint32_t* __restrict__ row, uint32_t* __restrict__ err) {
    uint64_t acc = 4;
    if (!COUNT) row[0] = v;
    uint32_t s = 0;
    if (v >= g.n) {
        atomicCAS(err, kErrNone, kErrBadStart);
        if (!COUNT) row[0] = g.n;
    } else {
        for (; s < L; ++s) {
            Hdr h;
            parse_header<FMT>(g.w, g.start[v], h);
            if (COUNT) acc += 8 + bytes_of_bits(h.hbits);
            if (h.d == 0) break;
            U4 r = philox4x32_10(inst, w, s, 0, seed);
            uint32_t k = (uint32_t)uniform_index(r.x, r.y, h.d);
            uint32_t nv, fb = 0;
            if (FMT == kLE) nv = le_get(g.w, h, v, k, COUNT ? &fb : nullptr);
            else nv = gt_get(g.w, h, v, k, COUNT ? &fb : nullptr);
            if (nv >= g.n) {
                atomicCAS(err, kErrNone, kErrBadNeighbor);
                break;
            }
            if (COUNT) acc += bytes_of_bits(fb) + 4;
            v = nv;
            if (!COUNT) row[s + 1] = v;
        }
    }
    if (!COUNT)
        for (; s < L; ++s) row[s + 1] = g.n;
    return acc;
}

template <int FMT, bool COUNT>
__global__ __launch_bounds__(kRwBlock) void rw_kernel(DevGraph g, const uint32_t* __restrict__ starts,
                                                      uint64_t W, uint32_t L, uint64_t seed,
                                                      uint32_t inst, uint32_t* __restrict__ out,
                                                      uint32_t* __restrict__ err,
                                                      unsigned long long* bytes) {
    const uint64_t w = (uint64_t)blockIdx.x * kRwBlock + threadIdx.x;
    uint64_t acc = 0;
    if (w < W) {
        uint32_t v = starts ? starts[w] : (uint32_t)uniform(seed, inst, (uint32_t)w, kTagStart, 0, g.n);
        acc = walk_one<FMT, COUNT>(g, v, L, seed, inst, (uint32_t)w,
                                   COUNT ? nullptr : out + w * (uint64_t)(L + 1), err);
        if (starts) acc += 4;
    }
    if (COUNT) {
        acc = wave_sum(acc);
        if (lane_id() == 0) atomicAdd(bytes, (unsigned long long)acc);
    }
}

struct RwDesc {
    uint64_t walkers;
    uint64_t gran_begin;  // first granule of this request
    const uint32_t* starts;
    uint32_t* out;
    uint32_t length;
    uint32_t instance;
};

template <int FMT, bool COUNT>
__global__ __launch_bounds__(kRwBlock) void rw_batch_kernel(DevGraph g, const RwDesc* __restrict__ reqs,
                                                            uint32_t R, uint64_t ngran,
                                                            unsigned long long* queue, uint64_t seed,
                                                            uint32_t* __restrict__ err,
                                                            unsigned long long* bytes) {
    const int lane = lane_id();
    uint64_t acc = 0;
    for (;;) {
        unsigned long long gi = 0;
        if (lane == 0) gi = atomicAdd(queue, 1ull);
        gi = __shfl(gi, 0, kWave);
        if (gi >= ngran) break;  // every wave exits once the queue is drained
        uint32_t lo = 0, hi = R - 1;  // last request with gran_begin <= gi
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo + 1) / 2;
            if (reqs[mid].gran_begin <= gi) lo = mid;
            else hi = mid - 1;
        }
        const RwDesc d = reqs[lo];
        const uint64_t w = (gi - d.gran_begin) * kWave + lane;
        if (w < d.walkers) {
            uint32_t v = d.starts ? d.starts[w]
                                  : (uint32_t)uniform(seed, d.instance, (uint32_t)w, kTagStart, 0, g.n);
            acc += walk_one<FMT, COUNT>(g, v, d.length, seed, d.instance, (uint32_t)w,
                                        COUNT ? nullptr : d.out + w * (uint64_t)(d.length + 1), err);
            if (d.starts) acc += 4;
        }
    }
    if (COUNT) {
        acc = wave_sum(acc);
        if (lane == 0) atomicAdd(bytes, (unsigned long long)acc);
    }
}

template <bool COUNT>
static void launch_rw(grass_graph* g, dim3 grid, hipStream_t s, const uint32_t* starts, uint64_t W,
                      uint32_t L, uint64_t seed, uint32_t inst, uint32_t* out,
                      unsigned long long* bytes) {
    if (g->fmt == kLE)
        hipLaunchKernelGGL((rw_kernel<kLE, COUNT>), grid, dim3(kRwBlock), 0, s, g->dev(), starts, W,
                           L, seed, inst, out, g->d_err, bytes);
    else
        hipLaunchKernelGGL((rw_kernel<kGT, COUNT>), grid, dim3(kRwBlock), 0, s, g->dev(), starts, W,
                           L, seed, inst, out, g->d_err, bytes);
}

int random_walk(grass_graph* g, const uint32_t* starts, uint64_t W, uint32_t L, uint64_t seed,
                uint32_t inst, uint32_t* out, hipStream_t s, uint64_t* bytes) {
    if (W == 0) {
        if (bytes) *bytes = 0;
        return GRASS_OK;
    }
    if ((!out && !bytes) || W > 0xFFFFFFFFull || L == 0xFFFFFFFFu) return GRASS_EINVAL;
    const dim3 grid((unsigned)((W + kRwBlock - 1) / kRwBlock));
    if (!bytes) {
        launch_rw<false>(g, grid, s, starts, W, L, seed, inst, out, nullptr);
        return status_of(hipGetLastError());
    }
    std::lock_guard<std::mutex> lock(g->mu);
    if (g->scratch.reserve(Carve::need(8)) != hipSuccess) return GRASS_ENOMEM;
    unsigned long long* d_bytes = (unsigned long long*)g->scratch.ptr;
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_bytes, (uint64_t)1, 0ull);
    launch_rw<true>(g, grid, s, starts, W, L, seed, inst, nullptr, d_bytes);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(bytes, d_bytes, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return status_of(e);
}

template <int FMT, bool COUNT>
static void launch_rw_batch(grass_graph* g, hipStream_t s, const RwDesc* d, uint32_t R, uint64_t gran,
                            unsigned long long* queue, uint64_t seed, unsigned long long* bytes) {
    static unsigned grid = 0;
    if (!grid) grid = persistent_grid((const void*)rw_batch_kernel<FMT, COUNT>, kRwBlock);
    hipLaunchKernelGGL((rw_batch_kernel<FMT, COUNT>), dim3(grid), dim3(kRwBlock), 0, s, g->dev(), d,
                       R, gran, queue, seed, g->d_err, bytes);
}

int random_walk_batch(grass_graph* g, const grass_rw_request* reqs, uint32_t count, uint64_t seed,
                      hipStream_t s, uint64_t* bytes) {
    if (bytes) *bytes = 0;
    if (count == 0) return GRASS_OK;
    if (!reqs) return GRASS_EINVAL;
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < count; ++i) {
        if (reqs[i].walkers == 0) continue;
        if ((!reqs[i].out_dev && !bytes) || reqs[i].walkers > 0xFFFFFFFFull ||
            reqs[i].length == 0xFFFFFFFFu)
            return GRASS_EINVAL;
        order.push_back(i);
    }
    if (order.empty()) return GRASS_OK;
    // longest walks first, so the tail of the queue holds the short granules
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return reqs[a].length > reqs[b].length;
    });
    std::vector<RwDesc> desc(order.size());
    uint64_t gran = 0;
    for (size_t j = 0; j < order.size(); ++j) {
        const grass_rw_request& r = reqs[order[j]];
        desc[j] = RwDesc{r.walkers, gran, r.starts_dev, r.out_dev, r.length, r.instance};
        gran += (r.walkers + kWave - 1) / kWave;
    }
    std::lock_guard<std::mutex> lock(g->mu);
    const size_t need = Carve::need(desc.size() * sizeof(RwDesc)) + Carve::need(64);
    hipError_t e = g->scratch.reserve(need);
    if (e != hipSuccess) return GRASS_ENOMEM;
    Carve c(g->scratch.ptr);
    RwDesc* d_desc = c.take<RwDesc>(desc.size());
    unsigned long long* d_ctr = c.take<unsigned long long>(2);  // [0] queue head, [1] bytes
    e = hipMemcpyAsync(d_desc, desc.data(), desc.size() * sizeof(RwDesc), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_ctr, (uint64_t)2, 0ull);
    const uint32_t R = (uint32_t)desc.size();
    if (g->fmt == kLE) {
        if (bytes) launch_rw_batch<kLE, true>(g, s, d_desc, R, gran, d_ctr, seed, d_ctr + 1);
        else launch_rw_batch<kLE, false>(g, s, d_desc, R, gran, d_ctr, seed, nullptr);
    } else {
        if (bytes) launch_rw_batch<kGT, true>(g, s, d_desc, R, gran, d_ctr, seed, d_ctr + 1);
        else launch_rw_batch<kGT, false>(g, s, d_desc, R, gran, d_ctr, seed, nullptr);
    }
    e = hipGetLastError();
    if (e == hipSuccess && bytes) e = hipMemcpyAsync(bytes, d_ctr + 1, 8, hipMemcpyDeviceToHost, s);
    // the descriptors live in shared scratch: finish before another call may reuse it
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return status_of(e);
}

}  // namespace grass
