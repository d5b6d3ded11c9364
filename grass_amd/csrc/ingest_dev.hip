// Device-side ingest: the R-MAT edge generator and graph_sort on the GPU, with the same results
// as grass_rmat_edges and grass_graph_sort on the host (graph_sort: LE/chunk_compress/
// graph_sort.cpp:22-82), so that R-MAT -> graph_sort -> grass_compress_device runs without the
// graph leaving HBM.
//
// graph_sort:
//   1. first[x] = the first stream position (2i for u, 2i + 1 for v of edge i; self loops skipped)
//      where id x appears -- an atomicMin that reads first, since most ids recur.
//   2. The ids that appear, sorted by first position (rocPRIM radix sort of (first, id) pairs),
//      get new ids 0, 1, ... in that order: the reference's numbering by first appearance.
//   3. Every edge that is not a self loop emits the key (a << 32 | b) of its renumbered ends and,
//      for undirected input, (b << 32 | a); a self loop emits ~0 twice (sorted to the end).
//   4. A radix sort of the keys over the bits that can be set, then unique: flags, a scan, and a
//      scatter of the column ids; the row pointers follow from where each source's run starts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/grass_ingest.h"
#include "blob.h"
#include "devutil.h"
#include "philox.h"
#include "runtime.h"
#include "scan.h"

namespace grass {
namespace {

constexpr int kIgBlock = 256;
constexpr unsigned kIgGrid = 8192;

// Edge i: level l picks quadrant (0,0)/(0,1)/(1,0)/(1,1) by word l % 4 of
// Philox(seed, {i_lo, i_hi, l / 4, 0x524D4154}) against the thresholds a, a + b, a + b + c (x 2^32).
__global__ __launch_bounds__(kIgBlock) void rmat_kernel(uint32_t scale, uint64_t m, uint64_t tA, uint64_t tAB,
                                                        uint64_t tABC, uint64_t seed, uint32_t* __restrict__ src,
                                                        uint32_t* __restrict__ dst) {
    for (uint64_t i = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kIgBlock) {
        uint32_t u = 0, v = 0;
        U4 r{0, 0, 0, 0};
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 3) == 0) r = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), l >> 2, 0x524D4154u, seed);
            const uint64_t x = (l & 3) == 0 ? r.x : (l & 3) == 1 ? r.y : (l & 3) == 2 ? r.z : r.w;
            u = (u << 1) | (uint32_t)(x >= tAB);
            v = (v << 1) | (uint32_t)((x >= tA && x < tAB) || x >= tABC);
        }
        src[i] = u;
        dst[i] = v;
    }
}

__global__ __launch_bounds__(kIgBlock) void max_id_kernel(const uint32_t* __restrict__ src,
                                                          const uint32_t* __restrict__ dst, uint64_t m,
                                                          unsigned int* __restrict__ out) {
    __shared__ uint32_t lds[kIgBlock / kWave];
    uint32_t mx = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kIgBlock)
        mx = max(mx, max(src[i], dst[i]));
    mx = block_max<uint32_t, kIgBlock>(mx, lds);
    if (threadIdx.x == 0) atomicMax(out, mx);
}

__device__ __forceinline__ void first_min(unsigned long long* first, uint32_t x, unsigned long long pos) {
    if (__atomic_load_n(&first[x], __ATOMIC_RELAXED) > pos) atomicMin(&first[x], pos);
}

__global__ __launch_bounds__(kIgBlock) void first_pos_kernel(const uint32_t* __restrict__ src,
                                                             const uint32_t* __restrict__ dst, uint64_t m,
                                                             unsigned long long* first) {
    for (uint64_t i = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kIgBlock) {
        const uint32_t u = src[i], v = dst[i];
        if (u == v) continue;
        first_min(first, u, 2 * i);
        first_min(first, v, 2 * i + 1);
    }
}

struct SeenGen {
    const unsigned long long* first;
    __device__ __forceinline__ uint64_t operator()(uint64_t x) const { return first[x] != ~0ull; }
};

// The ids that appear, as (first position, id) pairs at their rank among the ids (any order).
__global__ __launch_bounds__(kIgBlock) void seen_kernel(const unsigned long long* __restrict__ first, uint64_t nid,
                                                        const uint64_t* __restrict__ rank,
                                                        unsigned long long* __restrict__ fkey,
                                                        uint32_t* __restrict__ fid) {
    for (uint64_t x = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; x < nid; x += (uint64_t)gridDim.x * kIgBlock) {
        const unsigned long long f = first[x];
        if (f == ~0ull) continue;
        fkey[rank[x]] = f;
        fid[rank[x]] = (uint32_t)x;
    }
}

__global__ __launch_bounds__(kIgBlock) void newid_kernel(const uint32_t* __restrict__ fid, uint64_t n,
                                                         uint32_t* __restrict__ newid) {
    for (uint64_t k = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kIgBlock)
        newid[fid[k]] = (uint32_t)k;
}

__global__ __launch_bounds__(kIgBlock) void edge_keys_kernel(const uint32_t* __restrict__ src,
                                                             const uint32_t* __restrict__ dst, uint64_t m,
                                                             const uint32_t* __restrict__ newid, int per,
                                                             unsigned long long* __restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kIgBlock) {
        const uint32_t u = src[i], v = dst[i];
        unsigned long long k0 = ~0ull, k1 = ~0ull;
        if (u != v) {
            const unsigned long long a = newid[u], b = newid[v];
            k0 = (a << 32) | b;
            k1 = (b << 32) | a;
        }
        keys[per * i] = k0;
        if (per == 2) keys[2 * i + 1] = k1;
    }
}

struct UniqGen {
    const unsigned long long* keys;
    __device__ __forceinline__ uint64_t operator()(uint64_t j) const {
        const unsigned long long k = keys[j];
        return k != ~0ull && (j == 0 || keys[j - 1] != k);
    }
};

// The distinct keys' column ids at their ranks, and the row pointers: with the source of position
// j taken as n for a ~0 key and for j = P, a position whose source exceeds its predecessor's
// starts rows (predecessor's source, its source] at its rank.  The keys are sorted, so these runs
// cover rows 0..n once.
__global__ __launch_bounds__(kIgBlock) void csr_kernel(const unsigned long long* __restrict__ keys, uint64_t P,
                                                       const uint64_t* __restrict__ rank, uint64_t n,
                                                       uint64_t* __restrict__ rowptr, uint32_t* __restrict__ col) {
    for (uint64_t j = (uint64_t)blockIdx.x * kIgBlock + threadIdx.x; j <= P; j += (uint64_t)gridDim.x * kIgBlock) {
        const unsigned long long k = j < P ? keys[j] : ~0ull;
        const unsigned long long kp = j ? keys[j - 1] : ~0ull;
        if (k != ~0ull && (j == 0 || kp != k)) col[rank[j]] = (uint32_t)k;
        const int64_t a = k != ~0ull ? (int64_t)(k >> 32) : (int64_t)n;
        const int64_t pa = j == 0 ? -1 : (kp != ~0ull ? (int64_t)(kp >> 32) : (int64_t)n);
        for (int64_t x = pa + 1; x <= a; ++x) rowptr[x] = rank[j];
    }
}

int bits_for(uint64_t x) {  // bits needed to hold x
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

int graph_sort_device(const uint32_t* src, const uint32_t* dst, uint64_t m, int directed, hipStream_t s,
                      uint64_t* rowptr, uint64_t rowptr_cap, uint32_t* col, uint64_t col_cap, uint64_t* n_out,
                      uint64_t* m_out) {
    const int per = directed ? 1 : 2;
    const uint64_t P = (uint64_t)per * m;
    DevMem small;
    if (small.alloc(64) != hipSuccess) return GRASS_ENOMEM;
    unsigned int* d_max = small.as<unsigned int>();
    uint64_t* d_tot = (uint64_t*)(small.as<char>() + 8);
    hipLaunchKernelGGL(fill_kernel<unsigned int>, dim3(1), dim3(64), 0, s, d_max, (uint64_t)1, 0u);
    const unsigned gm = (unsigned)std::min<uint64_t>((m + kIgBlock - 1) / kIgBlock + 1, kIgGrid);
    if (m) hipLaunchKernelGGL(max_id_kernel, dim3(gm), dim3(kIgBlock), 0, s, src, dst, m, d_max);
    unsigned int maxid = 0;
    hipError_t e = hipMemcpyAsync(&maxid, d_max, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    if (maxid >= 0x80000000u) return GRASS_EINVAL;
    const uint64_t nid = m ? (uint64_t)maxid + 1 : 0;

    // 1-2: first positions, the ids that appear, their new ids
    DevMem d_first, d_rank, d_fkey, d_fid, d_newid, d_part;
    const uint64_t nbI = scan_blocks(std::max<uint64_t>(nid, 1)), nbP = scan_blocks(std::max<uint64_t>(P, 1));
    if (d_first.alloc(std::max<uint64_t>(nid, 1) * 8) != hipSuccess ||
        d_rank.alloc((nid + 1) * 8) != hipSuccess ||
        d_part.alloc((std::max(nbI, nbP) + 2) * 8) != hipSuccess)
        return GRASS_ENOMEM;
    unsigned long long* first = d_first.as<unsigned long long>();
    uint64_t* rank = d_rank.as<uint64_t>();
    const unsigned gi = (unsigned)std::min<uint64_t>((nid + kIgBlock - 1) / kIgBlock + 1, kIgGrid);
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(gi), dim3(kIgBlock), 0, s, first, nid, ~0ull);
    if (m) hipLaunchKernelGGL(first_pos_kernel, dim3(gm), dim3(kIgBlock), 0, s, src, dst, m, first);
    e = exclusive_scan(SeenGen{first}, Count{nid, nullptr}, nid, rank, d_part.as<uint64_t>(), d_tot, s);
    uint64_t n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, rank + nid, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    *n_out = n;
    if (rowptr_cap < n + 1) return GRASS_ECAPACITY;
    if (d_fkey.alloc(std::max<uint64_t>(n, 1) * 16) != hipSuccess ||
        d_fid.alloc(std::max<uint64_t>(n, 1) * 8) != hipSuccess ||
        d_newid.alloc(std::max<uint64_t>(nid, 1) * 4) != hipSuccess)
        return GRASS_ENOMEM;
    unsigned long long* fk = d_fkey.as<unsigned long long>();
    uint32_t* fid = d_fid.as<uint32_t>();
    hipLaunchKernelGGL(seen_kernel, dim3(gi), dim3(kIgBlock), 0, s, (const unsigned long long*)first, nid,
                       (const uint64_t*)rank, fk, fid);
    {
        rocprim::double_buffer<unsigned long long> kb(fk, fk + std::max<uint64_t>(n, 1));
        rocprim::double_buffer<uint32_t> vb(fid, fid + std::max<uint64_t>(n, 1));
        size_t tb = 0;
        const unsigned eb = (unsigned)std::max(1, bits_for(2 * m));
        e = rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0u, eb, s);
        DevMem tmp;
        if (e == hipSuccess && tmp.alloc(tb) != hipSuccess) return GRASS_ENOMEM;
        if (e == hipSuccess && n) e = rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, (size_t)n, 0u, eb, s);
        if (e != hipSuccess) return status_of(e);
        const unsigned gn = (unsigned)std::min<uint64_t>((n + kIgBlock - 1) / kIgBlock + 1, kIgGrid);
        if (n) hipLaunchKernelGGL(newid_kernel, dim3(gn), dim3(kIgBlock), 0, s, (const uint32_t*)vb.current(), n,
                                  d_newid.as<uint32_t>());
        e = hipStreamSynchronize(s);  // the sort's temporary storage goes with this scope
        if (e != hipSuccess) return status_of(e);
    }
    d_first.reset();
    d_fkey.reset();
    d_fid.reset();
    d_rank.reset();

    // 3-4: keys, sorted over the bits a key of the renumbered ids can set (~0 still sorts last)
    // two buffers of P + 1 keys: the sort's double buffer, and after it the free one holds the
    // ranks of the distinct keys (P + 1 entries)
    DevMem d_keys;
    if (d_keys.alloc((P + 1) * 16) != hipSuccess) return GRASS_ENOMEM;
    unsigned long long* keys = d_keys.as<unsigned long long>();
    if (m) hipLaunchKernelGGL(edge_keys_kernel, dim3(gm), dim3(kIgBlock), 0, s, src, dst, m,
                              (const uint32_t*)d_newid.as<uint32_t>(), per, keys);
    rocprim::double_buffer<unsigned long long> kb(keys, keys + P + 1);
    {
        size_t tb = 0;
        const unsigned eb = 32u + (unsigned)std::max(1, bits_for(n ? n - 1 : 0));
        e = rocprim::radix_sort_keys(nullptr, tb, kb, (size_t)P, 0u, eb, s);
        DevMem tmp;
        if (e == hipSuccess && tmp.alloc(tb) != hipSuccess) return GRASS_ENOMEM;
        if (e == hipSuccess && P) e = rocprim::radix_sort_keys(tmp.p, tb, kb, (size_t)P, 0u, eb, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return status_of(e);
    }
    const unsigned long long* sk = kb.current();
    uint64_t* krank = (uint64_t*)kb.alternate();
    e = exclusive_scan(UniqGen{sk}, Count{P, nullptr}, P, krank, d_part.as<uint64_t>(), d_tot, s);
    uint64_t M = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&M, krank + P, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    *m_out = M;
    if (col_cap < M || (M && !col)) return GRASS_ECAPACITY;
    const unsigned gp = (unsigned)std::min<uint64_t>((P + kIgBlock) / kIgBlock + 1, kIgGrid);
    hipLaunchKernelGGL(csr_kernel, dim3(gp), dim3(kIgBlock), 0, s, sk, P, (const uint64_t*)krank, n, rowptr,
                       col);
    e = hipStreamSynchronize(s);
    return status_of(e);
}

}  // namespace
}  // namespace grass

extern "C" {

int grass_rmat_edges_device(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                            int device, void* stream, uint32_t* src_dev, uint32_t* dst_dev) {
    if (scale == 0 || scale > 31 || edge_factor == 0 || !src_dev || !dst_dev) return GRASS_EINVAL;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GRASS_EINVAL;
    const uint64_t m = (uint64_t)edge_factor << scale;
    const double two32 = 4294967296.0;
    const uint64_t tA = (uint64_t)(a * two32), tAB = (uint64_t)((a + b) * two32),
                   tABC = (uint64_t)((a + b + c) * two32);
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    const unsigned grid = (unsigned)std::min<uint64_t>((m + grass::kIgBlock - 1) / grass::kIgBlock, grass::kIgGrid);
    hipLaunchKernelGGL(grass::rmat_kernel, dim3(grid), dim3(grass::kIgBlock), 0, (hipStream_t)stream, scale, m, tA,
                       tAB, tABC, seed, src_dev, dst_dev);
    const int rc = grass::status_of(hipGetLastError());
    (void)hipSetDevice(prev);
    return rc;
}

int grass_graph_sort_device(const uint32_t* src_dev, const uint32_t* dst_dev, uint64_t m, int directed, int device,
                            void* stream, uint64_t* rowptr_dev, uint64_t rowptr_capacity, uint32_t* col_dev,
                            uint64_t col_capacity, uint64_t* n, uint64_t* m_out) {
    if (!n || !m_out || !rowptr_dev || (m && (!src_dev || !dst_dev)) || m >= (1ull << 34)) return GRASS_EINVAL;
    *n = *m_out = 0;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    const int rc = grass::graph_sort_device(src_dev, dst_dev, m, directed, (hipStream_t)stream, rowptr_dev,
                                            rowptr_capacity, col_dev, col_capacity, n, m_out);
    (void)hipSetDevice(prev);
    return rc;
}

int grass_compress_on_device(const grass_csr* csr, int format, int device, grass_blob** graph, grass_blob** offset) {
    if (!csr || !graph || !offset || csr->n == 0) return GRASS_EINVAL;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    int rc = GRASS_OK;
    {
        grass::DevMem rp, cl;
        if (rp.alloc((csr->n + 1) * 8) != hipSuccess || cl.alloc(std::max<uint64_t>(csr->m, 1) * 4) != hipSuccess)
            rc = GRASS_ENOMEM;
        hipError_t e = hipSuccess;
        if (!rc) e = hipMemcpy(rp.p, csr->rowptr.data(), (csr->n + 1) * 8, hipMemcpyHostToDevice);
        if (!rc && e == hipSuccess && csr->m) e = hipMemcpy(cl.p, csr->col.data(), csr->m * 4, hipMemcpyHostToDevice);
        if (!rc && e != hipSuccess) rc = grass::status_of(e);
        if (!rc)
            rc = grass_compress_device(rp.as<uint64_t>(), cl.as<uint32_t>(), csr->n, format, device, nullptr, graph,
                                       offset);
    }
    (void)hipSetDevice(prev);
    return rc;
}

int grass_graph_sort_on_device(const uint32_t* src, const uint32_t* dst, uint64_t m, int directed, int device,
                               grass_csr** out) {
    if (!out || (m && (!src || !dst)) || m >= (1ull << 34)) return GRASS_EINVAL;
    *out = nullptr;
    uint32_t maxid = 0;
    for (uint64_t i = 0; i < m; ++i) maxid = std::max(maxid, std::max(src[i], dst[i]));
    if (maxid >= 0x80000000u) return GRASS_EINVAL;
    const uint64_t rcap = (uint64_t)maxid + 2, ccap = std::max<uint64_t>(directed ? m : 2 * m, 1);
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    int rc = GRASS_OK;
    grass_csr* csr = new (std::nothrow) grass_csr();
    if (!csr) rc = GRASS_ENOMEM;
    {
        grass::DevMem ds, dd, rp, cl;
        if (!rc && (ds.alloc(std::max<uint64_t>(m, 1) * 4) != hipSuccess ||
                    dd.alloc(std::max<uint64_t>(m, 1) * 4) != hipSuccess || rp.alloc(rcap * 8) != hipSuccess ||
                    cl.alloc(ccap * 4) != hipSuccess))
            rc = GRASS_ENOMEM;
        hipError_t e = hipSuccess;
        if (!rc && m) e = hipMemcpy(ds.p, src, m * 4, hipMemcpyHostToDevice);
        if (!rc && m && e == hipSuccess) e = hipMemcpy(dd.p, dst, m * 4, hipMemcpyHostToDevice);
        if (!rc && e != hipSuccess) rc = grass::status_of(e);
        uint64_t n = 0, mo = 0;
        if (!rc)
            rc = grass_graph_sort_device(ds.as<uint32_t>(), dd.as<uint32_t>(), m, directed, device, nullptr,
                                         rp.as<uint64_t>(), rcap, cl.as<uint32_t>(), ccap, &n, &mo);
        if (!rc) {
            try {
                csr->n = n;
                csr->m = mo;
                csr->rowptr.resize(n + 1);
                csr->col.resize(mo);
            } catch (const std::bad_alloc&) {
                rc = GRASS_ENOMEM;
            }
        }
        if (!rc) e = hipMemcpy(csr->rowptr.data(), rp.p, (n + 1) * 8, hipMemcpyDeviceToHost);
        if (!rc && e == hipSuccess && mo) e = hipMemcpy(csr->col.data(), cl.p, mo * 4, hipMemcpyDeviceToHost);
        if (!rc && e != hipSuccess) rc = grass::status_of(e);
    }
    (void)hipSetDevice(prev);
    if (rc) {
        delete csr;
        return rc;
    }
    *out = csr;
    return GRASS_OK;
}

}  // extern "C"
