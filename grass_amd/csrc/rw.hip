// Uniform random walks over the compressed graph.
//
// One lane per walker.  A step reads the current vertex's 16-byte decoded header (built at
// load) and then only the payload fields of the drawn neighbour, so no transit grouping, sort
// or host round trip is needed between steps (the reference regroups walkers by transit with a
// radix sort, run-length encode and three D2H copies per step, LE/src/grass.cu:1453-1861).
// Step s of walker w draws its neighbour index from Philox(seed, {instance, w, s, 0})
// (DeepWalkApp::next, LE/src/apps/randomwalk/randomWalks.cu:62-79, with cuRAND replaced).
//
// The 64 walkers of a wave belong to one request and so have one length: the step loop is
// wave-uniform, which allows staged output: each lane puts its walk values in an LDS tile of
// STAGE columns (8, 16 or 32: a tuning choice) and the wave writes whole row segments of STAGE
// ids (the walk matrix is row-major, so one column written straight from 64 lanes would touch
// 64 lines).  GT accesses go through the block checkpoints built at load, so a step reads fewer
// gaps than the checkpoint interval.
//
// plain:   one launch, one request (grass_random_walk);
// batched: many requests of different walker counts and lengths in one persistent launch;
//          waves pull 64-walker granules of one request from a device queue (`claim` granules
//          per queue access, a tuning choice), longest walks first.  Outputs are addressed by
//          (request, walker), so neither the schedule nor the tuning ever changes a result.
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
constexpr int kRwWaves = kRwBlock / kWave;
// LDS row pitch of a STAGE-column tile (odd: spreads the lanes' rows over the banks)
template <uint32_t STAGE>
constexpr uint32_t pitch() { return STAGE + 1; }

__device__ __forceinline__ uint64_t bytes_of_bits(uint32_t bits) { return (bits + 7) >> 3; }

// Writes staged columns [c0, c0 + ncols) of the wave's rows: STAGE lanes per row, 64 / STAGE rows
// per store.
template <uint32_t STAGE>
__device__ __forceinline__ void flush_tile(const uint32_t* __restrict__ tile, uint32_t* __restrict__ row0,
                                           uint64_t stride, uint32_t nrows, uint32_t c0, uint32_t ncols) {
    wave_sync();
    const int lane = lane_id();
    const uint32_t col = lane % STAGE, rs = lane / STAGE;
    for (uint32_t r = rs; r < (uint32_t)kWave; r += kWave / STAGE)
        if (r < nrows && col < ncols) row0[r * stride + c0 + col] = tile[r * pitch<STAGE>() + col];
    wave_sync();
}

template <int FMT, bool COUNT, uint32_t STAGE>
__device__ __forceinline__ uint64_t walk_wave(const DevGraph& g, bool active, uint32_t v, uint32_t L,
                                              uint64_t seed, uint32_t inst, uint32_t w,
                                              uint32_t* __restrict__ row0, uint32_t nrows,
                                              uint32_t* __restrict__ tile, uint32_t* __restrict__ err) {
    const int lane = lane_id();
    const uint64_t stride = (uint64_t)L + 1;
    uint64_t acc = active ? 4 : 0;
    bool live = active;
    if (active && v >= g.n) {
        atomicCAS(err, kErrNone, kErrBadStart);
        live = false;
    }
    if (!COUNT) {
        tile[lane * pitch<STAGE>()] = live ? v : g.n;
        if (L == 0) flush_tile<STAGE>(tile, row0, stride, nrows, 0, 1);
    }
    for (uint32_t s = 0; s < L; ++s) {
        uint32_t nv = 0;
        if (live) {
            Hdr h;
            load_hdr<FMT>(g, v, h);
            if (COUNT) acc += 8 + bytes_of_bits(header_bits<FMT>(h));
            if (h.d == 0) {
                live = false;
            } else {
                U4 r = philox4x32_10(inst, w, s, 0, seed);
                const uint32_t k = (uint32_t)uniform_index(r.x, r.y, h.d);
                uint32_t fb = 0;
                nv = get_neighbor<FMT>(g, h, v, k, COUNT ? &fb : nullptr);
                if (COUNT) acc += bytes_of_bits(fb) + 4;
            }
        }
        if (live) {
            if (nv >= g.n) {
                atomicCAS(err, kErrNone, kErrBadNeighbor);
                live = false;
            } else {
                v = nv;
            }
        }
        if (!COUNT) {
            const uint32_t col = s + 1, cc = col % STAGE;
            tile[lane * pitch<STAGE>() + cc] = live ? v : g.n;
            if (cc == STAGE - 1 || col == L) flush_tile<STAGE>(tile, row0, stride, nrows, col - cc, cc + 1);
        }
    }
    return acc;
}

template <int FMT, bool COUNT, uint32_t STAGE>
__global__ __launch_bounds__(kRwBlock) void rw_kernel(DevGraph g, const uint32_t* __restrict__ starts,
                                                      uint64_t W, uint32_t L, uint64_t seed,
                                                      uint32_t inst, uint32_t* __restrict__ out,
                                                      uint32_t* __restrict__ err,
                                                      unsigned long long* bytes) {
    __shared__ uint32_t tiles[kRwWaves][kWave * pitch<STAGE>()];
    const uint64_t w = (uint64_t)blockIdx.x * kRwBlock + threadIdx.x;
    const uint64_t w0 = w & ~(uint64_t)(kWave - 1);
    if (w0 >= W) return;  // whole wave past the end
    const bool active = w < W;
    uint32_t v = 0;
    if (active) v = starts ? starts[w] : (uint32_t)uniform(seed, inst, (uint32_t)w, kTagStart, 0, g.n);
    const uint32_t nrows = (uint32_t)std::min<uint64_t>(kWave, W - w0);
    uint64_t acc = walk_wave<FMT, COUNT, STAGE>(g, active, v, L, seed, inst, (uint32_t)w,
                                         COUNT ? nullptr : out + w0 * ((uint64_t)L + 1), nrows,
                                         tiles[threadIdx.x / kWave], err);
    if (COUNT) {
        if (active && starts) acc += 4;
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

template <int FMT, bool COUNT, uint32_t STAGE>
__global__ __launch_bounds__(kRwBlock) void rw_batch_kernel(DevGraph g, const RwDesc* __restrict__ reqs,
                                                            uint32_t R, uint64_t ngran, uint32_t claim,
                                                            unsigned long long* queue, uint64_t seed,
                                                            uint32_t* __restrict__ err,
                                                            unsigned long long* bytes) {
    __shared__ uint32_t tiles[kRwWaves][kWave * pitch<STAGE>()];
    uint32_t* tile = tiles[threadIdx.x / kWave];
    const int lane = lane_id();
    uint64_t acc = 0;
    for (;;) {
        unsigned long long g0 = 0;
        if (lane == 0) g0 = atomicAdd(queue, (unsigned long long)claim);
        g0 = __shfl(g0, 0, kWave);
        if (g0 >= ngran) break;  // every wave exits once the queue is drained
        const uint64_t g1 = g0 + claim < ngran ? g0 + claim : ngran;
        for (uint64_t gi = g0; gi < g1; ++gi) {
            uint32_t lo = 0, hi = R - 1;  // last request with gran_begin <= gi
            while (lo < hi) {
                uint32_t mid = lo + (hi - lo + 1) / 2;
                if (reqs[mid].gran_begin <= gi) lo = mid;
                else hi = mid - 1;
            }
            const RwDesc d = reqs[lo];
            const uint64_t w0 = (gi - d.gran_begin) * kWave;
            const uint64_t w = w0 + lane;
            const bool active = w < d.walkers;
            uint32_t v = 0;
            if (active)
                v = d.starts ? d.starts[w] : (uint32_t)uniform(seed, d.instance, (uint32_t)w, kTagStart, 0, g.n);
            const uint32_t nrows = (uint32_t)std::min<uint64_t>(kWave, d.walkers - w0);
            acc += walk_wave<FMT, COUNT, STAGE>(g, active, v, d.length, seed, d.instance, (uint32_t)w,
                                                COUNT ? nullptr : d.out + w0 * ((uint64_t)d.length + 1), nrows,
                                                tile, err);
            if (COUNT && active && d.starts) acc += 4;
        }
    }
    if (COUNT) {
        acc = wave_sum(acc);
        if (lane == 0) atomicAdd(bytes, (unsigned long long)acc);
    }
}

template <int FMT, bool COUNT>
static void launch_rw_fmt(uint32_t stage, dim3 grid, hipStream_t s, const DevGraph& dg, const uint32_t* starts,
                          uint64_t W, uint32_t L, uint64_t seed, uint32_t inst, uint32_t* out, uint32_t* err,
                          unsigned long long* bytes) {
    if (stage <= 8)
        hipLaunchKernelGGL((rw_kernel<FMT, COUNT, 8>), grid, dim3(kRwBlock), 0, s, dg, starts, W, L, seed, inst,
                           out, err, bytes);
    else if (stage <= 16)
        hipLaunchKernelGGL((rw_kernel<FMT, COUNT, 16>), grid, dim3(kRwBlock), 0, s, dg, starts, W, L, seed, inst,
                           out, err, bytes);
    else
        hipLaunchKernelGGL((rw_kernel<FMT, COUNT, 32>), grid, dim3(kRwBlock), 0, s, dg, starts, W, L, seed, inst,
                           out, err, bytes);
}

template <bool COUNT>
static void launch_rw(grass_graph* g, dim3 grid, hipStream_t s, const uint32_t* starts, uint64_t W,
                      uint32_t L, uint64_t seed, uint32_t inst, uint32_t* out,
                      unsigned long long* bytes) {
    const uint32_t stage = g->tune.rw_stage.load(std::memory_order_relaxed);
    if (g->fmt == kLE)
        launch_rw_fmt<kLE, COUNT>(stage, grid, s, g->dev(), starts, W, L, seed, inst, out, g->d_err, bytes);
    else
        launch_rw_fmt<kGT, COUNT>(stage, grid, s, g->dev(), starts, W, L, seed, inst, out, g->d_err, bytes);
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
    if (g->ws_bytes.reserve(64) != hipSuccess) return GRASS_ENOMEM;
    unsigned long long* d_bytes = (unsigned long long*)g->ws_bytes.ptr;
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_bytes, (uint64_t)1, 0ull);
    launch_rw<true>(g, grid, s, starts, W, L, seed, inst, nullptr, d_bytes);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(bytes, d_bytes, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return status_of(e);
}

template <int FMT, bool COUNT, uint32_t STAGE>
static void launch_rw_batch_stage(grass_graph* g, hipStream_t s, const RwDesc* d, uint32_t R, uint64_t gran,
                                  uint32_t claim, unsigned long long* queue, uint64_t seed, uint32_t* err,
                                  unsigned long long* bytes) {
    static unsigned grid = 0;
    if (!grid) grid = persistent_grid((const void*)rw_batch_kernel<FMT, COUNT, STAGE>, kRwBlock);
    hipLaunchKernelGGL((rw_batch_kernel<FMT, COUNT, STAGE>), dim3(grid), dim3(kRwBlock), 0, s, g->dev(), d, R,
                       gran, claim, queue, seed, err, bytes);
}

template <int FMT, bool COUNT>
static void launch_rw_batch(grass_graph* g, hipStream_t s, const RwDesc* d, uint32_t R, uint64_t gran,
                            unsigned long long* queue, uint64_t seed, uint32_t* err, unsigned long long* bytes) {
    const uint32_t stage = g->tune.rw_stage.load(std::memory_order_relaxed);
    const uint32_t claim = g->tune.rw_claim.load(std::memory_order_relaxed);
    if (stage <= 8) launch_rw_batch_stage<FMT, COUNT, 8>(g, s, d, R, gran, claim, queue, seed, err, bytes);
    else if (stage <= 16) launch_rw_batch_stage<FMT, COUNT, 16>(g, s, d, R, gran, claim, queue, seed, err, bytes);
    else launch_rw_batch_stage<FMT, COUNT, 32>(g, s, d, R, gran, claim, queue, seed, err, bytes);
}

int rw_reserve(Workspace& ws, uint32_t R) {
    if (ws.reserve_stage(R * sizeof(RwDesc)) != hipSuccess ||
        ws.scratch.reserve(Carve::need(R * sizeof(RwDesc)) + Carve::need(64)) != hipSuccess)
        return GRASS_ENOMEM;
    return GRASS_OK;
}

int rw_batch_enqueue(grass_graph* g, const grass_rw_request* reqs, uint32_t count, uint64_t seed,
                     hipStream_t s, Workspace& ws, unsigned long long* bytes_dev) {
    if (count == 0) return GRASS_OK;
    if (!reqs) return GRASS_EINVAL;
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < count; ++i) {
        if (reqs[i].walkers == 0) continue;
        if ((!reqs[i].out_dev && !bytes_dev) || reqs[i].walkers > 0xFFFFFFFFull ||
            reqs[i].length == 0xFFFFFFFFu)
            return GRASS_EINVAL;
        order.push_back(i);
    }
    if (order.empty()) return GRASS_OK;
    // longest walks first, so the tail of the queue holds the short granules
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return reqs[a].length > reqs[b].length;
    });
    const uint32_t R = (uint32_t)order.size();
    // descriptors staged in the workspace's pinned memory (the copy below is truly async)
    if (ws.reserve_stage(R * sizeof(RwDesc)) != hipSuccess) return GRASS_ENOMEM;
    RwDesc* desc = (RwDesc*)ws.stage();
    uint64_t gran = 0;
    for (uint32_t j = 0; j < R; ++j) {
        const grass_rw_request& r = reqs[order[j]];
        desc[j] = RwDesc{r.walkers, gran, r.starts_dev, r.out_dev, r.length, r.instance};
        gran += (r.walkers + kWave - 1) / kWave;
    }
    const size_t need = Carve::need(R * sizeof(RwDesc)) + Carve::need(64);
    hipError_t e = ws.scratch.reserve(need);
    if (e != hipSuccess) return GRASS_ENOMEM;
    Carve c(ws.scratch.ptr);
    RwDesc* d_desc = c.take<RwDesc>(R);
    unsigned long long* d_queue = c.take<unsigned long long>(1);
    e = hipMemcpyAsync(d_desc, desc, R * sizeof(RwDesc), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_queue, (uint64_t)1, 0ull);
    uint32_t* err = err_word(g, ws);
    if (g->fmt == kLE) {
        if (bytes_dev) launch_rw_batch<kLE, true>(g, s, d_desc, R, gran, d_queue, seed, err, bytes_dev);
        else launch_rw_batch<kLE, false>(g, s, d_desc, R, gran, d_queue, seed, err, nullptr);
    } else {
        if (bytes_dev) launch_rw_batch<kGT, true>(g, s, d_desc, R, gran, d_queue, seed, err, bytes_dev);
        else launch_rw_batch<kGT, false>(g, s, d_desc, R, gran, d_queue, seed, err, nullptr);
    }
    return status_of(hipGetLastError());
}

// ------------------------------------------------------------------ batch without allocation
// grass_random_walk_batch_noalloc keeps its working memory in one fixed device array per GPU:
// the work-queue counter, the caller's requests as copied in, and the descriptors built from
// them on the device.  The array's host-side name is not a device address: the host takes the
// address with hipGetSymbolAddress once per device and passes that pointer to the kernels.
constexpr size_t kNaQueueOff = 0;
constexpr size_t kNaReqOff = 256;
constexpr size_t kNaDescOff =
    kNaReqOff + ((GRASS_RW_NOALLOC_MAX_REQUESTS * sizeof(grass_rw_request) + 255) & ~(size_t)255);
constexpr size_t kNaBytes = kNaDescOff + GRASS_RW_NOALLOC_MAX_REQUESTS * sizeof(RwDesc);
static_assert(sizeof(grass_rw_request) % 8 == 0, "request copies are read as an array");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaPrepBlock = 1024;
constexpr uint32_t kNaPerThread = (GRASS_RW_NOALLOC_MAX_REQUESTS + kNaPrepBlock - 1) / kNaPrepBlock;

// Descriptors of the copied requests in their given order: granule ranges by one workgroup's
// exclusive scan of ceil(walkers / 64).  A request without walkers gets an empty range, so the
// batch kernel's search (last request whose range starts at or before the granule) skips it.
// Also zeroes the queue counter (a kernel, not a memset, so the call can be captured).
__global__ __launch_bounds__(kNaPrepBlock) void rw_noalloc_prepare_kernel(
    const grass_rw_request* __restrict__ reqs, uint32_t R, RwDesc* __restrict__ desc,
    unsigned long long* __restrict__ queue) {
    __shared__ uint64_t lds[kNaPrepBlock / kWave];
    const uint32_t t = threadIdx.x, i0 = t * kNaPerThread;
    uint64_t gr[kNaPerThread];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kNaPerThread; ++j) {
        const uint32_t i = i0 + j;
        gr[j] = i < R ? (reqs[i].walkers + kWave - 1) / kWave : 0;
        sum += gr[j];
    }
    uint64_t total = 0;
    uint64_t run = block_exclusive_sum<uint64_t, kNaPrepBlock>(sum, total, lds);
#pragma unroll
    for (uint32_t j = 0; j < kNaPerThread; ++j) {
        const uint32_t i = i0 + j;
        if (i < R) {
            const grass_rw_request r = reqs[i];
            desc[i] = RwDesc{r.walkers, run, r.starts_dev, r.out_dev, r.length, r.instance};
        }
        run += gr[j];
    }
    if (t == 0) *queue = 0;
}

namespace {
std::mutex na_mu;
void* na_ptr[64] = {};  // device address of `scratch` on each GPU (set on the first call there)
}  // namespace

// The current device's copy of `scratch` (nullptr if its address cannot be taken).
static unsigned char* noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        na_ptr[dev] = p;
    }
    return (unsigned char*)na_ptr[dev];
}

int random_walk_batch_noalloc(grass_graph* g, const grass_rw_request* reqs, uint32_t count,
                              uint64_t seed, hipStream_t s) {
    if (count == 0) return GRASS_OK;
    if (!reqs) return GRASS_EINVAL;
    if (count > GRASS_RW_NOALLOC_MAX_REQUESTS) return GRASS_ECAPACITY;
    uint64_t gran = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const grass_rw_request& r = reqs[i];
        if (r.walkers == 0) continue;
        if (!r.out_dev || r.walkers > 0xFFFFFFFFull || r.length == 0xFFFFFFFFu) return GRASS_EINVAL;
        gran += (r.walkers + kWave - 1) / kWave;
    }
    if (gran == 0) return GRASS_OK;
    unsigned char* base = noalloc_scratch();
    if (!base) return GRASS_EDEVICE;
    grass_rw_request* d_reqs = (grass_rw_request*)(base + kNaReqOff);
    RwDesc* d_desc = (RwDesc*)(base + kNaDescOff);
    unsigned long long* d_queue = (unsigned long long*)(base + kNaQueueOff);
    hipError_t e = hipMemcpyAsync(d_reqs, reqs, count * sizeof(grass_rw_request), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(rw_noalloc_prepare_kernel, dim3(1), dim3(kNaPrepBlock), 0, s, d_reqs, count, d_desc,
                       d_queue);
    if (g->fmt == kLE) launch_rw_batch<kLE, false>(g, s, d_desc, count, gran, d_queue, seed, g->d_err, nullptr);
    else launch_rw_batch<kGT, false>(g, s, d_desc, count, gran, d_queue, seed, g->d_err, nullptr);
    return status_of(hipGetLastError());
}

int random_walk_batch(grass_graph* g, const grass_rw_request* reqs, uint32_t count, uint64_t seed,
                      hipStream_t s, uint64_t* bytes) {
    if (bytes) *bytes = 0;
    std::lock_guard<std::mutex> lock(g->mu);
    unsigned long long* d_bytes = nullptr;
    if (bytes) {
        // a counter of its own, ahead of the workspace's scratch that the batch carves
        if (g->ws_bytes.reserve(64) != hipSuccess) return GRASS_ENOMEM;
        d_bytes = (unsigned long long*)g->ws_bytes.ptr;
        hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_bytes, (uint64_t)1, 0ull);
    }
    int rc = rw_batch_enqueue(g, reqs, count, seed, s, g->ws, d_bytes);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (bytes) e = hipMemcpyAsync(g->ws.result_slot(), d_bytes, 8, hipMemcpyDeviceToHost, s);
    // the descriptors live in the graph's workspace: finish before another call may reuse it
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && bytes) *bytes = *g->ws.result_slot();
    return status_of(e);
}

}  // namespace grass
