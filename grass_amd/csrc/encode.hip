// Device-side LE/GT encoder: the .graph and .offset images that grass_compress writes on the host
// (the reference's "normal" compressors, LE/chunk_compress/compressor.hpp:49-71,288-403 and
// GT/chunk_compress/compressor.hpp:49-62,245-349), computed from a CSR resident in device memory.
//
// 1. enc_plan_kernel, one wave per vertex: the record's field widths (LE rock_bit, sea_bit,
//    first_pos; GT head_bit, body_bit, first_pos) as maxima over the list, and its length.
// 2. An exclusive scan of the lengths: the bit where each record starts.
// 3. enc_write_kernel, one wave per vertex: the fields sit at offsets that follow from the widths
//    (LE: rocks, then signed seas; GT: per block the head, then the bodies), so the header's gamma
//    codes and 64 fields at a time are assembled in LDS and stored; only the words a record shares
//    with its neighbours are ORed into memory.
// 4. enc_offset_kernel: the .offset image, byte 0 = ubl then n length fields of ubl bits.
// The stream is MSB-first in big-endian-bit words; .graph words are stored byte-swapped, so that
// memory holds the file's bytes, and the .offset words are swapped after assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "../../include/grass_ingest.h"
#include "blob.h"
#include "devutil.h"
#include "runtime.h"
#include "scan.h"

namespace grass {
namespace {

constexpr int kEncBlock = 256;
constexpr int kEncWaves = kEncBlock / kWave;
constexpr uint32_t kEncWords = kWave + 2;  // words one pass of 64 fields of <= 32 bits can touch

// get_bit_num (compressor.hpp:16-22): bit length of |x|, 1 for 0.
__device__ __forceinline__ uint32_t bit_num(int64_t x) {
    if (x < 0) x = -x;
    return x == 0 ? 1u : 64u - (uint32_t)__clzll((unsigned long long)x);
}

// Length of the Elias-gamma code of x + 1 (compressor.hpp:411-417).
__device__ __forceinline__ uint32_t gamma_bits(uint64_t x) {
    return 2u * (63u - (uint32_t)__clzll((unsigned long long)(x + 1))) + 1u;
}

// GT head field as the reference writes it: |N - v| - 1, a self loop (-1) coming out as 1
// (GT/chunk_compress/compressor.hpp:407-419 keeps the low bit of -1).
__device__ __forceinline__ uint32_t head_field(int64_t head) { return head >= 0 ? (uint32_t)head : 1u; }

// Position k < d of an LE list -> chunk c, index i in the chunk (i = 0: the chunk's rock); the
// last position is rock C.  Chunks hold q + 1 positions (the first R) or q.
__device__ __forceinline__ void le_pos(uint32_t k, uint32_t q, uint32_t R, uint32_t& c, uint32_t& i) {
    const uint32_t big = R * (q + 1);
    if (k < big) {
        c = k / (q + 1);
        i = k - c * (q + 1);
    } else {
        const uint32_t kk = k - big;
        c = R + kk / q;
        i = kk - (c - R) * q;
    }
}

// Position k < d of a GT list -> block b, index i in the block.
__device__ __forceinline__ void gt_pos(uint32_t k, uint32_t q, uint32_t R, uint32_t& b, uint32_t& i) {
    le_pos(k, q, R, b, i);  // the same split: R blocks of q + 1 positions, then blocks of q
}

// Residual of LE sea (c, i >= 1) at list position k: its distance from the chunk's linear
// interpolation between rocks c and c + 1 (compressor.hpp:329-390).
__device__ __forceinline__ int64_t le_residual(const uint32_t* N, uint32_t k, uint32_t i, uint32_t L) {
    const uint32_t s0 = k - i;
    const int64_t rl = N[s0], rr = N[s0 + L];
    const int64_t dis = (rr - rl) / (int64_t)L;
    return (int64_t)N[k] - (rl + (int64_t)i * dis);
}

struct LenGen {
    const uint64_t* len;
    __device__ __forceinline__ uint64_t operator()(uint64_t v) const { return len[v]; }
};

template <int FMT>
__global__ __launch_bounds__(kEncBlock) void enc_plan_kernel(const uint64_t* __restrict__ rowptr,
                                                             const uint32_t* __restrict__ col, uint64_t n,
                                                             uint64_t* __restrict__ len, uint2* __restrict__ widths,
                                                             unsigned long long* maxlen) {
    const int lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * kEncWaves;
    unsigned long long most = 0;
    for (uint64_t v = ((uint64_t)blockIdx.x * kEncBlock + threadIdx.x) / kWave; v < n; v += nw) {
        const uint64_t base = rowptr[v];
        const uint32_t d = (uint32_t)(rowptr[v + 1] - base);
        const uint32_t* N = col + base;
        const int64_t vv = (int64_t)v;
        uint32_t f1 = 0, f2 = 0, fp = 0;
        uint64_t bits;
        if (FMT == kLE) {
            if (d == 0) {
                bits = gamma_bits(0);
            } else if (d == 1) {
                bits = gamma_bits(1) + gamma_bits(N[0]);
            } else {
                uint32_t C, q, R;
                le_geometry(d, C, q, R);
                int64_t a = (int64_t)N[0] - vv, z = (int64_t)N[d - 1] - vv;
                a = a < 0 ? -a : a;
                z = z < 0 ? -z : z;
                f1 = bit_num(a > z ? a : z);
                uint32_t sb = 0, fpl = 0;
                for (uint32_t k = lane; k < d; k += kWave) {
                    uint32_t c, i;
                    le_pos(k, q, R, c, i);
                    if (i == 0) {
                        if ((int64_t)N[k] < vv) fpl = max(fpl, c + 1);
                    } else {
                        sb = max(sb, bit_num(le_residual(N, k, i, q + (c < R))));
                    }
                }
                f2 = wave_max(sb);
                fp = wave_max(fpl);
                bits = (uint64_t)gamma_bits(d) + gamma_bits(f1) + gamma_bits(f2) + gamma_bits(fp) +
                       (uint64_t)(C + 1) * f1 + (uint64_t)(d - 1 - C) * (f2 + 1);
            }
        } else {
            uint32_t C = 0, q = 0, R = 0;
            if (d) gt_geometry(d, C, q, R);
            uint32_t hb = 0, bb = 0, fpl = 0;
            for (uint32_t k = lane; k < d; k += kWave) {
                uint32_t b, i;
                gt_pos(k, q, R, b, i);
                if (i == 0) {
                    const int64_t x = (int64_t)N[k] - vv;
                    if (x < 0) fpl = max(fpl, b + 1);
                    hb = max(hb, bit_num((x < 0 ? -x : x) - 1));
                } else {
                    bb = max(bb, bit_num((int64_t)N[k] - (int64_t)N[k - 1] - 1));
                }
            }
            f1 = wave_max(hb);
            f2 = wave_max(bb);
            fp = wave_max(fpl);
            bits = (uint64_t)gamma_bits(d) + gamma_bits(f1) + gamma_bits(f2) + gamma_bits(fp) + (uint64_t)C * f1 +
                   (uint64_t)(d - C) * f2;
        }
        if (lane == 0) {
            len[v] = bits;
            widths[v] = make_uint2(f1 | (f2 << 8), fp);
        }
        most = max(most, (unsigned long long)bits);
    }
    if (lane == 0 && most) atomicMax(maxlen, most);
}

// A record is written pass by pass.  Each pass assembles its words in LDS (B[0] = the word holding
// the pass's first bit) and stores them; only the record's first word and its last, partial word
// are shared with the neighbouring records and ORed into memory (`out` starts zeroed).  A partial
// last word of a pass that does not end the record is carried into the next pass instead of
// being written twice.
constexpr uint32_t kRecWords = kEncWords + 6;  // the first pass also holds the header (<= 96 bits)

struct RecState {   // wave-uniform
    uint32_t pend;  // the carried partial word
    uint64_t pend_w;  // its index, ~0 when nothing is carried
    bool pend_first;  // it is the record's first word
};

// Elias-gamma code of x + 1 (x + 1 < 2^32) at stream bit `at` of the pass buffer: L zero bits,
// then x + 1 in L + 1 bits.
__device__ __forceinline__ void put_gamma(uint32_t* B, uint64_t w0, uint64_t& at, uint32_t x) {
    const uint32_t y = x + 1;
    const uint32_t L = 31u - (uint32_t)__clz(y);
    put_field(B, w0, at + L, L + 1, y);
    at += 2 * L + 1;
}

// One pass over stream bits [from, ...): the header codes hv[0..nh) at `from` (first pass only,
// written by lane 0), then up to 64 fields, lane l holding field l at `at` (act: the lane has one,
// `last` the highest lane that has one; nf = 0: no fields).
__device__ __forceinline__ void rec_pass(uint32_t* out, uint32_t* B, int lane, uint64_t from, bool first, bool final,
                                         const uint32_t* hv, int nh, bool nf, bool act, uint64_t at, uint32_t wdt,
                                         uint32_t val, uint32_t last, RecState& st) {
    const uint64_t w0 = from >> 5;
    const bool carried = st.pend_w == w0;
    for (uint32_t j = lane; j < kRecWords; j += kWave) B[j] = (j == 0 && carried) ? st.pend : 0u;
    wave_sync();
    uint64_t to = from;
    if (first) {
        for (int i = 0; i < nh; ++i) {
            if (lane == 0) {
                uint64_t h = to;
                put_gamma(B, w0, h, hv[i]);
            }
            to += gamma_bits(hv[i]);
        }
    }
    if (nf) {
        if (act) put_field(B, w0, at, wdt, wdt < 32 ? val & ((1u << wdt) - 1) : val);
        to = __shfl(at + wdt, (int)last, kWave);
    }
    wave_sync();
    const uint32_t nwords = to > from ? (uint32_t)(((to - 1) >> 5) - w0 + 1) : 0u;
    const bool open = (to & 31) != 0;
    const bool shared0 = first || (carried && st.pend_first);
    for (uint32_t j = lane; j < nwords; j += kWave) {
        const bool tail = j == nwords - 1 && open;
        if (tail && !final) continue;  // carried into the next pass
        // stored byte-swapped: the words become the file's bytes (a swap commutes with the OR)
        const uint32_t x = __builtin_bswap32(B[j]);
        if ((j == 0 && shared0) || tail) atomicOr(&out[w0 + j], x);
        else out[w0 + j] = x;
    }
    if (nwords) {
        st.pend = B[nwords - 1];
        st.pend_w = open && !final ? w0 + nwords - 1 : ~0ull;
        st.pend_first = nwords == 1 && shared0;
    }
    wave_sync();
}

template <int FMT>
__global__ __launch_bounds__(kEncBlock) void enc_write_kernel(const uint64_t* __restrict__ rowptr,
                                                              const uint32_t* __restrict__ col, uint64_t n,
                                                              const uint64_t* __restrict__ start,
                                                              const uint2* __restrict__ widths,
                                                              uint32_t* __restrict__ out) {
    __shared__ uint32_t buf[kEncWaves][kRecWords];
    const int lane = lane_id();
    uint32_t* B = buf[threadIdx.x / kWave];
    const uint64_t nw = (uint64_t)gridDim.x * kEncWaves;
    for (uint64_t v = ((uint64_t)blockIdx.x * kEncBlock + threadIdx.x) / kWave; v < n; v += nw) {
        const uint64_t base = rowptr[v];
        const uint32_t d = (uint32_t)(rowptr[v + 1] - base);
        const uint32_t* N = col + base;
        const int64_t vv = (int64_t)v;
        const uint2 wd = widths[v];
        const uint32_t f1 = wd.x & 0xFFu, f2 = wd.x >> 8, fp = wd.y;
        const uint64_t at0 = start[v];
        RecState st{0u, ~0ull, false};
        // header: gamma codes (LE: d [rock_bit | the neighbour] [sea_bit first_pos]; GT: d
        // head_bit body_bit first_pos)
        uint32_t hv[4] = {d, f1, f2, fp};
        int nh = 4;
        if (FMT == kLE && d <= 1) {
            hv[1] = d ? N[0] : 0u;
            nh = d ? 2 : 1;
        }
        uint64_t pay = at0;
        for (int i = 0; i < nh; ++i) pay += gamma_bits(hv[i]);
        if (FMT == kLE) {
            if (d <= 1) {
                rec_pass(out, B, lane, at0, true, true, hv, nh, false, false, 0, 0, 0, 0, st);
                continue;
            }
            uint32_t C, q, R;
            le_geometry(d, C, q, R);
            const uint32_t S = d - 1 - C, bigS = R * q;
            // rocks 0..C: |N - v| in rock_bit bits
            for (uint32_t c0 = 0; c0 <= C; c0 += kWave) {
                const uint32_t c = c0 + lane;
                const bool act = c <= C;
                const uint32_t cc = act ? c : C;
                const uint32_t k = cc * q + min(cc, R);
                const int64_t x = (int64_t)N[k] - vv;
                const uint64_t at = pay + (uint64_t)cc * f1;
                rec_pass(out, B, lane, c0 ? pay + (uint64_t)c0 * f1 : at0, c0 == 0, S == 0 && c0 + kWave > C, hv, nh,
                         true, act, at, f1, (uint32_t)(x < 0 ? -x : x), min((uint32_t)kWave - 1, C - c0), st);
            }
            // seas, chunk by chunk: |residual| in sea_bit bits, then its sign
            const uint64_t seas = pay + (uint64_t)(C + 1) * f1;
            for (uint32_t j0 = 0; j0 < S; j0 += kWave) {
                const uint32_t j = j0 + lane;
                const bool act = j < S;
                const uint32_t jj = act ? j : S - 1;
                uint32_t c, i;
                if (jj < bigS) {
                    c = jj / q;
                    i = jj - c * q + 1;
                } else {  // q >= 2 here: the chunks past R hold q - 1 seas
                    c = R + (jj - bigS) / (q - 1);
                    i = (jj - bigS) - (c - R) * (q - 1) + 1;
                }
                const uint32_t k = c * q + min(c, R) + i;
                const int64_t res = le_residual(N, k, i, q + (c < R));
                const uint32_t mag = (uint32_t)(res < 0 ? -res : res);
                rec_pass(out, B, lane, seas + (uint64_t)j0 * (f2 + 1), false, j0 + kWave >= S, hv, nh, true, act,
                         seas + (uint64_t)jj * (f2 + 1), f2 + 1, (mag << 1) | (res < 0 ? 1u : 0u),
                         min((uint32_t)kWave - 1, S - 1 - j0), st);
            }
        } else {
            if (d == 0) {
                rec_pass(out, B, lane, at0, true, true, hv, nh, false, false, 0, 0, 0, 0, st);
                continue;
            }
            uint32_t C, q, R;
            gt_geometry(d, C, q, R);
            for (uint32_t k0 = 0; k0 < d; k0 += kWave) {
                const uint32_t k = k0 + lane;
                const bool act = k < d;
                const uint32_t kk = act ? k : d - 1;
                uint32_t b, i;
                gt_pos(kk, q, R, b, i);
                const uint64_t bo = pay + (uint64_t)b * (f1 + (uint64_t)(q - 1) * f2) + (uint64_t)min(b, R) * f2;
                uint32_t val;
                if (i == 0) {
                    const int64_t x = (int64_t)N[kk] - vv;
                    val = head_field((x < 0 ? -x : x) - 1);
                } else {
                    val = (uint32_t)((int64_t)N[kk] - (int64_t)N[kk - 1] - 1);
                }
                const uint64_t at = i ? bo + f1 + (uint64_t)(i - 1) * f2 : bo;
                const uint64_t from = k0 ? __shfl(at, 0, kWave) : at0;
                rec_pass(out, B, lane, from, k0 == 0, k0 + kWave >= d, hv, nh, true, act, at, i ? f2 : f1, val,
                         min((uint32_t)kWave - 1, d - 1 - k0), st);
            }
        }
    }
}

// .offset: field v = the length of record v in ubl bits at bit 8 + v ubl, in 64-bit words
// (compressor.hpp:217-285, 392-403).
__global__ __launch_bounds__(256) void enc_offset_kernel(const uint64_t* __restrict__ len, uint64_t n, uint32_t ubl,
                                                         unsigned long long* __restrict__ ow) {
    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (uint64_t)gridDim.x * 256) {
        const uint64_t x = len[v], at = 8 + v * ubl;
        const uint64_t w = at >> 6;
        const uint32_t s = (uint32_t)(at & 63);
        if (s + ubl <= 64) {
            atomicOr(&ow[w], (unsigned long long)(x << (64 - s - ubl)));
        } else {
            atomicOr(&ow[w], (unsigned long long)(x >> (s + ubl - 64)));
            atomicOr(&ow[w + 1], (unsigned long long)(x << (128 - s - ubl)));
        }
    }
}

__global__ void bswap64_kernel(unsigned long long* w, uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = __builtin_bswap64(w[i]);
}

// Device -> host through two pinned chunks: the threaded copy out of one chunk overlaps the DMA
// into the other.
int fetch_to_host(uint8_t* dst, const void* src, uint64_t len, hipStream_t s) {
    if (len == 0) return GRASS_OK;
    const uint64_t chunk = std::min<uint64_t>(64ull << 20, len);
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = GRASS_OK;
    for (int i = 0; i < 2 && !rc; ++i) {
        if (hipHostMalloc(&pin[i], chunk, hipHostMallocDefault) != hipSuccess) rc = GRASS_ENOMEM;
        else if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) rc = GRASS_EDEVICE;
    }
    const unsigned threads = load_threads();
    const uint64_t nch = (len + chunk - 1) / chunk;
    auto issue = [&](uint64_t c) {
        const uint64_t off = c * chunk, n = std::min(chunk, len - off);
        return hipMemcpyAsync(pin[c & 1], (const char*)src + off, n, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipEventRecord(ev[c & 1], s) == hipSuccess;
    };
    if (!rc && !issue(0)) rc = GRASS_EDEVICE;
    for (uint64_t c = 0; c < nch && !rc; ++c) {
        if (c + 1 < nch && !issue(c + 1)) {
            rc = GRASS_EDEVICE;
            break;
        }
        if (hipEventSynchronize(ev[c & 1]) != hipSuccess) {
            rc = GRASS_EDEVICE;
            break;
        }
        const uint64_t off = c * chunk;
        parallel_copy(dst + off, (const uint8_t*)pin[c & 1], std::min(chunk, len - off), threads);
    }
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = GRASS_EDEVICE;
    for (int i = 0; i < 2; ++i) {
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        if (pin[i]) (void)hipHostFree(pin[i]);
    }
    return rc;
}

// The .graph and .offset images in device memory (the files' bytes), complete when this returns.
struct Images {
    DevMem g, o;
    uint64_t gbytes = 0, obytes = 0;
};

int encode_images(const uint64_t* rowptr, const uint32_t* col, uint64_t n, int format, hipStream_t s,
                  Images& im) {
    const bool le = format == GRASS_FMT_LE;
    const uint64_t nb = scan_blocks(n);
    DevMem tmp;  // len[n], start[n + 1], widths[n], partials, total, maxlen
    Carve need(nullptr);
    auto plan = [&](Carve& c, uint64_t** len, uint64_t** start, uint2** widths, uint64_t** part,
                    unsigned long long** most) {
        *len = c.take<uint64_t>(n);
        *start = c.take<uint64_t>(n + 1);
        *widths = c.take<uint2>(n);
        *part = c.take<uint64_t>(nb + 2);
        *most = c.take<unsigned long long>(1);
    };
    uint64_t *d_len, *d_start, *d_part;
    uint2* d_w;
    unsigned long long* d_max;
    plan(need, &d_len, &d_start, &d_w, &d_part, &d_max);
    if (tmp.alloc(need.off + 256) != hipSuccess) return GRASS_ENOMEM;
    Carve c(tmp.p);
    plan(c, &d_len, &d_start, &d_w, &d_part, &d_max);
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const unsigned grid = (unsigned)std::max(cus, 1) * 16;
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_max, (uint64_t)1, 0ull);
    if (le)
        hipLaunchKernelGGL(enc_plan_kernel<kLE>, dim3(grid), dim3(kEncBlock), 0, s, rowptr, col, n, d_len, d_w, d_max);
    else
        hipLaunchKernelGGL(enc_plan_kernel<kGT>, dim3(grid), dim3(kEncBlock), 0, s, rowptr, col, n, d_len, d_w, d_max);
    hipError_t e = exclusive_scan(LenGen{d_len}, Count{n, nullptr}, n, d_start, d_part, d_part + nb + 1, s);
    uint64_t hv[2] = {0, 0};  // total bits, longest record
    if (e == hipSuccess) e = hipMemcpyAsync(&hv[0], d_start + n, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&hv[1], d_max, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    const uint64_t total = hv[0], maxlen = hv[1];
    if (total >= (1ull << 40)) return GRASS_EINVAL;  // past what a reader's 40-bit offsets address
    // .graph
    const uint64_t gwords = (total + 31) / 32 + 2;
    DevMem& gbuf = im.g;
    if (gbuf.alloc(gwords * 4) != hipSuccess) return GRASS_ENOMEM;
    uint32_t* d_g = (uint32_t*)gbuf.p;
    e = hipMemsetAsync(d_g, 0, gwords * 4, s);
    if (e != hipSuccess) return status_of(e);
    if (le)
        hipLaunchKernelGGL(enc_write_kernel<kLE>, dim3(grid), dim3(kEncBlock), 0, s, rowptr, col, n,
                           (const uint64_t*)d_start, (const uint2*)d_w, d_g);
    else
        hipLaunchKernelGGL(enc_write_kernel<kGT>, dim3(grid), dim3(kEncBlock), 0, s, rowptr, col, n,
                           (const uint64_t*)d_start, (const uint2*)d_w, d_g);
    // .offset: ubl = get_bit_num(longest record), then the n lengths
    uint32_t ubl = 1;
    while (ubl < 64 && (maxlen >> ubl)) ++ubl;
    const uint64_t obits = 8 + n * (uint64_t)ubl;
    const uint64_t owords = (obits + 63) / 64 + 1;
    DevMem& obuf = im.o;
    if (obuf.alloc(owords * 8) != hipSuccess) return GRASS_ENOMEM;
    unsigned long long* d_o = (unsigned long long*)obuf.p;
    e = hipMemsetAsync(d_o, 0, owords * 8, s);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, s, d_o, (uint64_t)1,
                       (unsigned long long)ubl << 56);
    hipLaunchKernelGGL(enc_offset_kernel, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 16384)), dim3(256), 0, s,
                       (const uint64_t*)d_len, n, ubl, d_o);
    hipLaunchKernelGGL(bswap64_kernel, dim3(256), dim3(256), 0, s, d_o, owords);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    im.gbytes = (total + 7) / 8;
    im.obytes = (obits + 7) / 8;
    return GRASS_OK;
}

int compress_device(const uint64_t* rowptr, const uint32_t* col, uint64_t n, int format, hipStream_t s,
                    grass_blob* gb, grass_blob* ob) {
    Images im;
    int rc = encode_images(rowptr, col, n, format, s, im);
    if (rc) return rc;
    if (!gb->resize(im.gbytes) || !ob->resize(im.obytes)) return GRASS_ENOMEM;
    rc = fetch_to_host(gb->data(), im.g.p, gb->size(), s);
    return rc ? rc : fetch_to_host(ob->data(), im.o.p, ob->size(), s);
}

}  // namespace
}  // namespace grass

extern "C" {

int grass_compress_device(const uint64_t* rowptr_dev, const uint32_t* col_dev, uint64_t n, int format, int device,
                          void* stream, grass_blob** graph, grass_blob** offset) {
    if (!rowptr_dev || !graph || !offset || n == 0 || n >= 0x80000000ull ||
        (format != GRASS_FMT_LE && format != GRASS_FMT_GT))
        return GRASS_EINVAL;
    *graph = *offset = nullptr;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    grass_blob* gb = new (std::nothrow) grass_blob();
    grass_blob* ob = new (std::nothrow) grass_blob();
    int rc = gb && ob ? grass::compress_device(rowptr_dev, col_dev, n, format, (hipStream_t)stream, gb, ob)
                      : GRASS_ENOMEM;
    (void)hipSetDevice(prev);
    if (rc) {
        delete gb;
        delete ob;
        return rc;
    }
    *graph = gb;
    *offset = ob;
    return GRASS_OK;
}

int grass_compress_device_open(const uint64_t* rowptr_dev, const uint32_t* col_dev, uint64_t n, int format,
                               int device, void* stream, grass_graph** out) {
    if (!rowptr_dev || !out || n == 0 || n >= 0x80000000ull || (format != GRASS_FMT_LE && format != GRASS_FMT_GT))
        return GRASS_EINVAL;
    *out = nullptr;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(device)) return GRASS_EDEVICE;
    int rc;
    {
        grass::Images im;
        rc = grass::encode_images(rowptr_dev, col_dev, n, format, (hipStream_t)stream, im);
        if (!rc) {
            grass::DeviceSource gs(im.g.p, im.gbytes), os(im.o.p, im.obytes);
            rc = grass::open_sources(gs, os, format, device, out);
        }
    }
    (void)hipSetDevice(prev);
    return rc;
}

}  // extern "C"
