// Compressed-graph loader: files -> pinned host staging -> HBM, then on the device:
//   1. byte swap of the .graph and .offset images into big-endian word views,
//   2. record-length extraction from .offset fused into a reduce-then-scan (bit starts),
//   3. a validator that parses every record header, checks header + payload == record
//      length, and writes the degree array and the decoded-header cache,
//   4. GT: the checkpointed payload stream the samplers decode (gt_expand_kernel).
// Replaces loadGraph (LE/src/grass.cu:1015-1095), copyPartitionToGPU (LE/src/utils.hpp:328-353)
// with its decode_offset kernel (LE/src/bcg.cuh:284-329) and CUB scan.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "devutil.h"
#include "runtime.h"
#include "scan.h"

namespace grass {

__global__ void bswap_words_kernel(uint32_t* w, uint64_t nwords) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) w[i] = __builtin_bswap32(w[i]);
}

// Length field u of .offset: bits [8 + u*ubl, 8 + (u+1)*ubl) of the file, ubl in [1, 64]
// (LE/src/bcg.cuh:293-318 reads the same field across up to three words).
struct OffsetFieldGen {
    const uint32_t* ow;  // big-endian word view of the whole .offset file, padded
    uint32_t ubl;
    __device__ __forceinline__ uint64_t operator()(uint64_t u) const {
        uint64_t b = 8 + u * ubl;
        const uint32_t* p = ow + (b >> 5);
        uint32_t s = (uint32_t)(b & 31);
        uint64_t hi = ((uint64_t)p[0] << 32) | p[1];
        uint64_t win = s ? (hi << s) | (p[2] >> (32 - s)) : hi;
        return win >> (64 - ubl);
    }
};

constexpr int kValBlock = 256;

// stats: [0] sum of degrees, [1] max degree, [2] malformed records, [3] first malformed vertex
template <int FMT>
__global__ __launch_bounds__(kValBlock) void validate_kernel(const uint32_t* __restrict__ w,
                                                             const uint64_t* __restrict__ start,
                                                             uint32_t n, uint64_t gbits,
                                                             uint32_t* __restrict__ deg,
                                                             uint4* __restrict__ hdr,
                                                             unsigned long long* stats) {
    __shared__ unsigned long long lds[kValBlock / kWave];
    unsigned long long sum = 0, mx = 0, bad = 0;
    for (uint32_t v = blockIdx.x * kValBlock + threadIdx.x; v < n; v += gridDim.x * kValBlock) {
        const uint64_t b0 = start[v], end = start[v + 1];
        uint64_t b = b0;
        // Every read below stays within [b0, end + 95] bits, inside the padded word view,
        // whatever the .offset held (a corrupt one may even wrap the 64-bit scan).
        bool ok = end > b0 && end <= gbits;
        Hdr h;
        h.d = h.f1 = h.f2 = h.fp = h.C = h.q = h.R = 0;
        ok = ok && gamma_checked(w, b, h.d) && b <= end;
        if (FMT == kLE) {
            if (ok && h.d >= 1) ok = gamma_checked(w, b, h.f1) && b <= end;
            if (ok && h.d == 1) ok = h.f1 < n;
            if (ok && h.d >= 2) {
                ok = gamma_checked(w, b, h.f2) && b <= end && gamma_checked(w, b, h.fp) && b <= end;
                if (ok) {
                    le_geometry(h.d, h.C, h.q, h.R);
                    ok = h.f1 >= 1 && h.f1 <= 31 && h.f2 <= 31 && h.fp <= h.C + 1;
                }
            }
        } else {
            ok = ok && gamma_checked(w, b, h.f1) && b <= end && gamma_checked(w, b, h.f2) &&
                 b <= end && gamma_checked(w, b, h.fp) && b <= end;
            if (ok && h.d) {
                gt_geometry(h.d, h.C, h.q, h.R);
                ok = h.f1 >= 1 && h.f1 <= 31 && h.f2 <= 31 && h.fp <= h.C &&
                     (h.f2 >= 1 || h.d == h.C);
            }
        }
        if (ok) {
            h.pay = b;
            ok = (b + payload_bits(FMT, h) == end) && h.d < 0x80000000u;
        }
        if (!ok) {
            ++bad;
            atomicMin(&stats[3], (unsigned long long)v);
            h.d = h.f1 = h.f2 = h.fp = h.C = 0;
            h.pay = b0;
        }
        deg[v] = h.d;
        hdr[v] = pack_hdr(h, FMT);
        sum += h.d;
        mx = h.d > mx ? h.d : mx;
    }
    sum = block_sum<unsigned long long, kValBlock>(sum, lds);
    mx = block_max<unsigned long long, kValBlock>(mx, lds);
    bad = block_sum<unsigned long long, kValBlock>(bad, lds);
    if (threadIdx.x == 0) {
        atomicAdd(&stats[0], sum);
        atomicMax(&stats[1], mx);
        if (bad) atomicAdd(&stats[2], bad);
    }
}

// ------------------------------------------------------------------ GT payload stream

// Words of vertex v's payload in the device GT stream (P = 1 << shift).
struct GtxWordsGen {
    const uint4* hdr;
    uint32_t shift;
    __device__ __forceinline__ uint64_t operator()(uint64_t v) const {
        Hdr h;
        unpack_hdr<kGT>(hdr[v], h);
        return (gtx_payload_bits(h, shift) + 31) >> 5;
    }
};

constexpr int kExpBlock = 256;
constexpr uint32_t kExpWords = kWave + 2;  // words that one pass of 64 fields (<= 32 bits) can touch

// One wave per GT vertex: decodes the .graph record 64 positions at a time (heads absolute,
// bodies as gaps, a segmented prefix sum carrying the running id from pass to pass) and writes
// the vertex's payload in the device stream at word woff[v]: heads and bodies copied, the 32-bit
// id at every checkpoint position.  A pass assembles its fields in LDS and stores its complete
// words; a last word that the pass leaves partial is carried into the next pass.  A vertex's
// stream words are its own (its payload starts on a word), so every word is stored once and
// nothing is ORed into memory.  Finally the header's payload bit is moved to the stream.
__global__ __launch_bounds__(kExpBlock) void gt_expand_kernel(DevGraph g, const uint64_t* __restrict__ woff,
                                                              uint4* __restrict__ hdr, uint32_t* __restrict__ out) {
    __shared__ uint32_t buf[kExpBlock / kWave][kExpWords];
    const int lane = lane_id();
    uint32_t* B = buf[threadIdx.x / kWave];
    const uint32_t sh = g.cp_shift, P = 1u << sh;
    const uint64_t nw = (uint64_t)gridDim.x * kExpBlock / kWave;
    for (uint64_t v = ((uint64_t)blockIdx.x * kExpBlock + threadIdx.x) / kWave; v < g.n; v += nw) {
        uint4 e = hdr[v];
        Hdr h;
        unpack_hdr<kGT>(e, h);  // payload bit in .graph
        Hdr hx = h;
        hx.pay = woff[v] << 5;  // payload bit in the stream
        uint32_t carry = 0;
        uint32_t pend = 0;        // the partial last word of the previous pass
        uint64_t pend_w = ~0ull;  // its index, ~0 when that pass ended on a word boundary
        for (uint32_t base = 0; base < h.d; base += kWave) {
            const uint32_t k = base + lane;
            const bool act = k < h.d;
            uint32_t blk, i;
            gt_locate(h, act ? k : h.d - 1, blk, i);
            const uint64_t bo = gt_file_block_bit(h, blk);
            uint32_t raw, x, wdt;
            int head;
            if (i == 0) {
                raw = read_bits(g.w, bo, h.f1);
                x = blk >= h.fp ? (uint32_t)v + raw + 1 : (uint32_t)v - raw - 1;
                wdt = h.f1;
                head = 1;
            } else {
                raw = read_bits0(g.w, bo + h.f1 + (uint64_t)(i - 1) * h.f2, h.f2);
                x = raw + 1;
                wdt = h.f2;
                head = 0;
            }
            if (lane == 0 && !head) {
                x += carry;
                head = 1;
            }
            x = wave_segmented_sum(x, head);
            carry = __shfl(x, kWave - 1, kWave);
            if (i && (i & (P - 1)) == 0) {
                raw = x;
                wdt = 32;
            }
            const uint64_t at = gtx_block_bit(hx, blk, sh) + gtx_field(hx, i, sh);
            // words [w0, w0 + nwords) hold this pass's fields; the last active lane ends them
            const uint32_t la = min((uint32_t)kWave - 1, h.d - 1 - base);
            const uint64_t w0 = __shfl(at, 0, kWave) >> 5;
            const uint64_t end = __shfl(at + wdt, (int)la, kWave);
            const uint32_t nwords = (uint32_t)(((end - 1) >> 5) - w0 + 1);
            for (uint32_t j = lane; j < kExpWords; j += kWave) B[j] = (j == 0 && pend_w == w0) ? pend : 0u;
            wave_sync();
            if (act) put_field(B, w0, at, wdt, raw);
            wave_sync();
            const bool open = (end & 31) != 0;
            const uint32_t full = open ? nwords - 1 : nwords;
            for (uint32_t j = lane; j < full; j += kWave) out[w0 + j] = B[j];
            pend = B[nwords - 1];
            pend_w = open ? w0 + nwords - 1 : ~0ull;
            wave_sync();
        }
        if (pend_w != ~0ull && lane == 0) out[pend_w] = pend;
        if (lane == 0) {
            e.x = (uint32_t)hx.pay;
            e.y = (e.y & ~63u) | (uint32_t)(hx.pay >> 32);
            hdr[v] = e;
        }
    }
}

// Host threads that copy a chunk into pinned memory: GRASS_LOAD_THREADS in [1, 64], else up to 8.
unsigned load_threads() {
    if (const char* e = getenv("GRASS_LOAD_THREADS")) {
        const unsigned long t = strtoul(e, nullptr, 10);
        if (t >= 1 && t <= 64) return (unsigned)t;
    }
    return std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
}

// memcpy split over up to `threads` threads (slices of at least 4 MB); a thread that cannot be
// started leaves its slice to the calling thread.
void parallel_copy(void* dst, const uint8_t* src, uint64_t len, unsigned threads) {
    const uint64_t t = std::min<uint64_t>(threads, std::max<uint64_t>(1, len >> 22));
    if (t <= 1) {
        memcpy(dst, src, len);
        return;
    }
    const uint64_t step = ((len + t - 1) / t + 4095) & ~4095ull;
    std::vector<std::thread> pool;
    std::vector<uint64_t> mine = {0};
    for (uint64_t lo = step; lo < len; lo += step) {
        const uint64_t n = std::min(step, len - lo);
        try {
            pool.emplace_back([=] { memcpy((uint8_t*)dst + lo, src + lo, n); });
        } catch (...) {
            mine.push_back(lo);
        }
    }
    for (uint64_t lo : mine) memcpy((uint8_t*)dst + lo, src + lo, std::min(step, len - lo));
    for (std::thread& th : pool) th.join();
}

namespace {

constexpr uint64_t kChunk = 64ull << 20;  // pinned staging chunk

// Streams `src` into device memory through two pinned chunks so that the copy of chunk i+1
// overlaps the DMA of chunk i.  A source held in host memory is copied with several threads.
int stage_to_device(ByteSource& src, char* dst, hipStream_t stream, uint64_t* tail,
                    unsigned tail_len, uint8_t* tail_out) {
    const uint64_t total = src.size();
    if (total == 0) return GRASS_OK;
    if (const uint8_t* dev = src.device_data()) {  // already in device memory: one copy, no staging
        const uint64_t tl = std::min<uint64_t>(tail_len, total);
        hipError_t e = hipMemcpyAsync(dst, dev, total, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess && tail_out && tl)
            e = hipMemcpyAsync(tail_out, dev + total - tl, tl, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (tail) *tail = tl;
        return e == hipSuccess ? GRASS_OK : GRASS_EDEVICE;
    }
    const uint64_t chunk = std::min<uint64_t>(kChunk, total);
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = GRASS_OK;
    for (int i = 0; i < 2 && rc == GRASS_OK; ++i) {
        if (hipHostMalloc(&pin[i], chunk, hipHostMallocDefault) != hipSuccess) rc = GRASS_ENOMEM;
        else if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
            rc = GRASS_EDEVICE;
    }
    uint64_t off = 0;
    int slot = 0;
    bool used[2] = {false, false};
    const uint8_t* mem = src.data();
    const unsigned threads = load_threads();
    while (rc == GRASS_OK && off < total) {
        uint64_t len = std::min(chunk, total - off);
        if (used[slot] && hipEventSynchronize(ev[slot]) != hipSuccess) { rc = GRASS_EDEVICE; break; }
        if (mem) parallel_copy(pin[slot], mem + off, len, threads);
        else if (!src.read(pin[slot], len)) { rc = GRASS_EIO; break; }
        // keep the last tail_len bytes of the stream on the host
        if (tail_out && off + len > total - std::min<uint64_t>(tail_len, total)) {
            uint64_t t0 = total - std::min<uint64_t>(tail_len, total);
            for (uint64_t i = std::max(off, t0); i < off + len; ++i)
                tail_out[i - t0] = ((uint8_t*)pin[slot])[i - off];
        }
        if (hipMemcpyAsync(dst + off, pin[slot], len, hipMemcpyHostToDevice, stream) != hipSuccess ||
            hipEventRecord(ev[slot], stream) != hipSuccess) {
            rc = GRASS_EDEVICE;
            break;
        }
        used[slot] = true;
        off += len;
        slot ^= 1;
    }
    if (tail) *tail = std::min<uint64_t>(tail_len, total);
    if (hipStreamSynchronize(stream) != hipSuccess && rc == GRASS_OK) rc = GRASS_EDEVICE;
    for (int i = 0; i < 2; ++i) {
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        if (pin[i]) (void)hipHostFree(pin[i]);
    }
    return rc;
}

// Field u of .offset read on the host from the file's last `tlen` bytes.
uint64_t host_field(const uint8_t* tailb, uint64_t tlen, uint64_t file_bytes, uint64_t u,
                    uint32_t ubl) {
    uint64_t b = 8 + u * ubl;  // bit position in the file
    uint64_t first = file_bytes - tlen;
    uint64_t v = 0;
    for (uint32_t t = 0; t < ubl; ++t, ++b) {
        uint64_t byte = b >> 3;
        uint32_t bit = byte >= first && byte < file_bytes ? (tailb[byte - first] >> (7 - (b & 7))) & 1 : 0;
        v = (v << 1) | bit;
    }
    return v;
}

}  // namespace

int build_gt_stream(grass_graph* g, uint64_t* words);

int load_graph(grass_graph* g, ByteSource& gsrc, ByteSource& osrc) {
    const uint64_t gbytes = gsrc.size(), obytes = osrc.size();
    if (obytes < 2 || gbytes < 1) return GRASS_EFORMAT;
    if (8 * gbytes >= kMaxPayloadBits) return GRASS_EFORMAT;  // payload bits must fit the header cache
    hipStream_t s = g->stream;
    const uint64_t gwords = (gbytes + 3) / 4 + kPadWords;
    const uint64_t owords = (obytes + 3) / 4 + kPadWords;
    uint32_t* d_off = nullptr;
    if (hipMalloc(&g->d_words, gwords * 4) != hipSuccess) return GRASS_ENOMEM;
    if (hipMalloc(&d_off, owords * 4) != hipSuccess) return GRASS_ENOMEM;
    int rc = GRASS_OK;
    auto fail = [&](int code) {
        (void)hipFree(d_off);
        return code;
    };
    if (hipMemsetAsync(g->d_words, 0, gwords * 4, s) != hipSuccess ||
        hipMemsetAsync(d_off, 0, owords * 4, s) != hipSuccess)
        return fail(GRASS_EDEVICE);
    rc = stage_to_device(gsrc, (char*)g->d_words, s, nullptr, 0, nullptr);
    if (rc) return fail(rc);
    uint8_t tailb[64];
    uint64_t tlen = 0;
    rc = stage_to_device(osrc, (char*)d_off, s, &tlen, sizeof(tailb), tailb);
    if (rc) return fail(rc);

    // ubl = byte 0 of .offset; n = floor(8 (bytes - 1) / ubl) (LE/src/grass.cu:1071,1088),
    // minus trailing zero-length fields, which can only be padding (a record is >= 1 bit).
    uint8_t ubl = 0;
    if (hipMemcpy(&ubl, d_off, 1, hipMemcpyDeviceToHost) != hipSuccess) return fail(GRASS_EDEVICE);
    if (ubl == 0 || ubl > 64) return fail(GRASS_EFORMAT);
    uint64_t n = ((obytes - 1) * 8) / ubl;
    uint64_t tail_fields_max = std::min<uint64_t>(n, 7);
    for (uint64_t t = 0; t < tail_fields_max && n > 0; ++t) {
        if (8 + (n - 1) * (uint64_t)ubl < 8 * (obytes - tlen)) break;  // field not in the tail
        if (host_field(tailb, tlen, obytes, n - 1, ubl) != 0) break;
        --n;
    }
    if (n == 0 || n >= 0x80000000ull) return fail(GRASS_EFORMAT);
    g->n = n;
    g->ubl = ubl;
    g->graph_bytes = gbytes;
    g->offset_bytes = obytes;

    const unsigned bsg = 2048;
    hipLaunchKernelGGL(bswap_words_kernel, dim3(bsg), dim3(256), 0, s, g->d_words, gwords);
    hipLaunchKernelGGL(bswap_words_kernel, dim3(bsg), dim3(256), 0, s, d_off, owords);

    if (hipMalloc(&g->d_start, (n + 1) * 8) != hipSuccess) return fail(GRASS_ENOMEM);
    if (hipMalloc(&g->d_deg, n * 4) != hipSuccess) return fail(GRASS_ENOMEM);
    const uint64_t hdr_bytes = n * 16;
    if (hipMalloc(&g->d_hdr, hdr_bytes) != hipSuccess) return fail(GRASS_ENOMEM);
    const uint64_t nb = scan_blocks(n);
    uint64_t* d_tmp = nullptr;  // partials[nb], total, stats[4]
    if (hipMalloc(&d_tmp, (nb + 8) * 8) != hipSuccess) return fail(GRASS_ENOMEM);
    uint64_t* d_total = d_tmp + nb;
    unsigned long long* d_stats = (unsigned long long*)(d_tmp + nb + 2);
    unsigned long long init[4] = {0, 0, 0, ~0ull};
    hipError_t e = hipMemcpyAsync(d_stats, init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = exclusive_scan(OffsetFieldGen{d_off, ubl}, Count{n, nullptr}, n, g->d_start, d_tmp,
                           d_total, s);
    if (e == hipSuccess) {
        unsigned grid = (unsigned)std::min<uint64_t>((n + kValBlock - 1) / kValBlock, 16384);
        if (g->fmt == kLE)
            hipLaunchKernelGGL(validate_kernel<kLE>, dim3(grid), dim3(kValBlock), 0, s, g->d_words,
                               g->d_start, (uint32_t)n, 8 * gbytes, g->d_deg, g->d_hdr, d_stats);
        else
            hipLaunchKernelGGL(validate_kernel<kGT>, dim3(grid), dim3(kValBlock), 0, s, g->d_words,
                               g->d_start, (uint32_t)n, 8 * gbytes, g->d_deg, g->d_hdr, d_stats);
        e = hipGetLastError();
    }
    uint64_t total = 0;
    unsigned long long stats[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(stats, d_stats, sizeof(stats), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_tmp);
    (void)hipFree(d_off);
    if (e != hipSuccess) return status_of(e);
    // records must cover the .graph bits up to < 8 bits of final padding
    if (total > 8 * gbytes || 8 * gbytes - total >= 8) return GRASS_EFORMAT;
    if (stats[2] != 0) return GRASS_EFORMAT;
    g->total_bits = total;
    g->m = stats[0];
    g->max_degree = stats[1];
    uint64_t xwords = 0;
    if (g->fmt == kGT) {
        int rc2 = build_gt_stream(g, &xwords);
        if (rc2) return rc2;
    }
    // What stays resident is what the kernels read: the payload words (LE: the .graph image; GT:
    // the payload stream, which replaces the image), the header cache and the degrees.  The record
    // starts are folded into the header cache.
    (void)hipFree(g->d_start);
    g->d_start = nullptr;
    if (g->fmt == kGT) {
        (void)hipFree(g->d_words);
        g->d_words = nullptr;
    }
    g->device_bytes = (g->fmt == kGT ? xwords : gwords) * 4 + n * 4 + hdr_bytes;
    return GRASS_OK;
}

// Builds the device GT payload stream (g->d_gtx) from .graph and points the headers at it.
int build_gt_stream(grass_graph* g, uint64_t* words) {
    hipStream_t s = g->stream;
    const uint64_t n = g->n, nb = scan_blocks(n);
    uint64_t* d_woff = nullptr;  // n + 1 word offsets, partials, total
    if (hipMalloc(&d_woff, (n + 1 + nb + 2) * 8) != hipSuccess) return GRASS_ENOMEM;
    hipError_t e = exclusive_scan(GtxWordsGen{g->d_hdr, g->cp_shift}, Count{n, nullptr}, n, d_woff,
                                  d_woff + n + 1, d_woff + n + 1 + nb, s);
    uint64_t total = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&total, d_woff + n, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    uint32_t* out = nullptr;
    int rc = status_of(e);
    // payload bits must fit the header cache
    if (!rc && 32 * total >= kMaxPayloadBits) rc = GRASS_EFORMAT;
    if (!rc && hipMalloc(&out, (total + kPadWords) * 4) != hipSuccess) rc = GRASS_ENOMEM;
    if (!rc) {
        e = hipMemsetAsync(out, 0, (total + kPadWords) * 4, s);
        if (e == hipSuccess) {
            // g->dev() still decodes .graph here: d_gtx is set once the stream is complete
            hipLaunchKernelGGL(gt_expand_kernel, dim3(4096), dim3(kExpBlock), 0, s, g->dev(),
                               (const uint64_t*)d_woff, g->d_hdr, out);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        rc = status_of(e);
    }
    (void)hipFree(d_woff);
    if (rc) {
        if (out) (void)hipFree(out);
        return rc;
    }
    g->d_gtx = out;
    *words = total + kPadWords;
    return GRASS_OK;
}

}  // namespace grass
