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
    __device__ __forceinline__ uint32_t at(const uint32_t* __restrict__ w, uint32_t k, int& head) const {
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
        const uint32_t raw = read_bits(w, bo + off, hd ? f1 : abs ? 32u : f2);
        head = hd || abs;
        return hd ? (blk >= fp ? v + raw + 1 : v - raw - 1) : abs ? raw : raw + 1;
    }
};

// Segmented inclusive sum over runs of at most `span` lanes (a lane with head != 0 starts a
// run): log2(span) shuffle steps instead of six.
__device__ __forceinline__ uint32_t short_segmented_sum(uint32_t x, int head, uint32_t span) {
    const int lane = lane_id();
    if (span == 4) {  // the default checkpoint interval: two fixed steps
#pragma unroll
        for (int off = 1; off < 4; off <<= 1) {
            const uint32_t px = __shfl_up(x, off, kWave);
            const int ph = __shfl_up(head, off, kWave);
            if (lane >= off && !head) {
                x += px;
                head = ph;
            }
        }
        return x;
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

template <int FMT>
__global__ __launch_bounds__(kDecBlock) void decode_csr_kernel(DevGraph g,
                                                               const uint64_t* __restrict__ rowptr,
                                                               const uint32_t* __restrict__ wstart,
                                                               uint32_t* __restrict__ col,
                                                               uint64_t m, uint32_t per_wave) {
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
        Hdr h;
        load_hdr<kGT>(g, vlo, h);
        uint32_t blk, i;
        gt_locate(h, (uint32_t)(e0 - rowptr[vlo]), blk, i);
        eb = e0 - (i & (span - 1));  // back to the block head or checkpoint
    }
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
            if (FMT == kGT) x = gl.at(g.w, (uint32_t)(ee - rc), head);
            else x = element_uniform<FMT>(g, uh, hv, (uint32_t)(ee - rc), d1, d0, head);
        } else {
            // Owner of slot ee: one coalesced load of the 64 row ends after `cur`, then a 6-step
            // search over them in registers; slots more than 64 vertices past `cur` (only
            // possible across vertices without edges) fall back to a binary search of rowptr.
            // (row ends as 32-bit offsets from the start of `cur`'s list, which is at or before
            // every slot of the iteration; one list is < 2^31 entries, larger offsets saturate)
            const uint64_t bj = (uint64_t)cur + 1 + lane <= g.n ? rowptr[cur + 1 + lane] : ~0ull;
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
            uint32_t v;
            uint64_t vstart;
            if (pos < (uint32_t)kWave) {
                v = cur + pos;
                vstart = rc + (pos ? prev : 0u);
            } else {
                v = owner_of(rowptr, cur + kWave, vhi, ee);
                vstart = rowptr[v];
            }
            const uint32_t nc = __builtin_amdgcn_readfirstlane(__shfl(v, kWave - 1, kWave));
            if (nc != cur) {
                cur = nc;
                rc = rowptr[cur];
                rn = rowptr[cur + 1];
            }
            const uint32_t k = (uint32_t)(ee - vstart);
            Hdr h;
            load_hdr<FMT>(g, v, h);
            if (FMT == kLE) {
                x = le_get(g.w, h, v, k);
            } else {
                uint32_t blk, i;
                gt_locate(h, k, blk, i);
                x = gt_element(g, h, v, blk, i, head);
            }
        }
        if (FMT == kGT) {
            if (lane == 0 && !head) {
                x += carry;
                head = 1;
            }
            x = short_segmented_sum(x, head, span);
            carry = __shfl(x, kWave - 1, kWave);
        }
        if (act && e >= e0) col[e] = x;
    }
}

int decode_csr(grass_graph* g, uint64_t* rowptr, uint32_t* col, hipStream_t s) {
    const uint64_t n = g->n, m = g->m;
    const uint64_t nb = scan_blocks(n);
    {
        std::lock_guard<std::mutex> lock(g->mu);
        hipError_t e = g->ws.scratch.reserve((nb + 4) * 8);
        if (e != hipSuccess) return GRASS_ENOMEM;
        uint64_t* part = (uint64_t*)g->ws.scratch.ptr;
        e = exclusive_scan(DegGen{g->d_deg}, Count{n, nullptr}, n, rowptr, part, part + nb, s);
        if (e != hipSuccess) return status_of(e);
        // the scratch is reused by the next call on any stream: finish the scan first
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return status_of(e);
    }
    if (m == 0 || !col) return GRASS_OK;
    const uint32_t per_wave = g->tune.decode_edges_per_wave.load(std::memory_order_relaxed);
    const uint64_t waves = (m + per_wave - 1) / per_wave;
    const uint64_t blocks = (waves + kDecBlock / kWave - 1) / (kDecBlock / kWave);
    if (blocks > 0x7fffffffull) return GRASS_EINVAL;
    // The waves' first vertices depend only on the graph and per_wave: built once per per_wave
    // value from these row pointers, complete before it is published, never changed after.
    uint32_t slot = 0;
    while ((512u << slot) < per_wave && slot + 1 < kWaveStartSlots) ++slot;
    if ((512u << slot) != per_wave) return GRASS_EINVAL;
    const uint32_t* wstart;
    {
        std::lock_guard<std::mutex> lock(g->mu);
        if (!g->d_wstart[slot]) {
            uint32_t* t = nullptr;
            if (hipMalloc(&t, (waves + 1) * 4) != hipSuccess) return GRASS_ENOMEM;
            const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 16384);
            hipLaunchKernelGGL(wave_start_kernel, dim3(grid), dim3(256), 0, s, (const uint64_t*)rowptr,
                               (uint32_t)n, per_wave, waves, t);
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
    if (g->fmt == kLE)
        hipLaunchKernelGGL(decode_csr_kernel<kLE>, dim3((unsigned)blocks), dim3(kDecBlock), 0, s,
                           g->dev(), (const uint64_t*)rowptr, wstart, col, m, per_wave);
    else
        hipLaunchKernelGGL(decode_csr_kernel<kGT>, dim3((unsigned)blocks), dim3(kDecBlock), 0, s,
                           g->dev(), (const uint64_t*)rowptr, wstart, col, m, per_wave);
    return status_of(hipGetLastError());
}

}  // namespace grass
