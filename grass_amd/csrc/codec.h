// Device codec for the GraSS compressed graph (LE and GT variants).
//
// The .graph file is one MSB-first bit stream: the records of vertices 0..n-1 back to back.
// In HBM it is held as big-endian 32-bit words (word k = bytes 4k..4k+3, first byte in the
// top bits), the view the reference builds on the host (LE/src/grass.cu:1036-1050), padded
// with kPadWords zero words so that every window read below stays inside the allocation.
//
// Record layout (LE/chunk_compress/compressor.hpp:288-390, GT/chunk_compress/compressor.hpp:
// 245-349), header fields are Elias-gamma codes of value+1:
//   LE: g(d) [g(rock_bit) [g(sea_bit) g(first_pos)]] rocks... seas...
//       d = 1 stores the neighbour itself in the rock_bit slot.
//   GT: g(d) g(head_bit) g(body_bit) g(first_pos) then per block: head, bodies (interleaved).
// The header parse follows LE/src/bcg.cuh:131-152 and GT/src/bcg.cuh:110-136; random access
// follows LE/src/bcg.cuh:196-261 and GT/src/bcg.cuh:139-168.  Chunk geometry uses an exact
// integer square root (the reference decoder's sqrt((float)x) disagrees with its encoder's
// sqrt(double) for degrees above ~8.4M).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grass {

enum Format : int { kLE = 0, kGT = 1 };

constexpr int kPadWords = 16;  // zero words after the last .graph word (a header parse may run
                                // <= 252 bits past a corrupt record start before the validator stops)

struct DevGraph {
    // Payload words the samplers decode (big-endian word view, padded): the .graph image itself
    // for LE; for GT the checkpointed payload stream built at load (see gtx_block_bit).
    const uint32_t* __restrict__ w;
    const uint32_t* __restrict__ deg;    // degree of v, decoded and validated at load
    const uint4* __restrict__ hdr;       // decoded header of v (see pack_hdr), built at load
    uint32_t n;
    int fmt;
    uint32_t cp_shift;  // GT checkpoint interval = 1 << cp_shift in-block positions
};

// Decoded record header.  f1/f2 = rock_bit/sea_bit (LE) or head_bit/body_bit (GT).
struct Hdr {
    uint64_t pay;  // bit where the payload starts
    uint32_t d;
    uint32_t f1, f2, fp;
    uint32_t C, q, R;  // LE: chunks, q = floor((d-1)/C); GT: blocks, q = floor(d/C); R remainder
};

// GT checkpoints.  The GT payload stream on the device keeps the .graph blocks and their order
// but stores, at every in-block position that is a nonzero multiple of P = 1 << cp_shift, the
// absolute id (32 bits) in place of the gap.  A GT access at in-block position i then reads one
// contiguous span, the absolute value at position i & ~(P - 1) (a checkpoint, or the block
// head) and at most P - 1 gaps after it, instead of the head and i gaps (the reference's
// O(sqrt d) walk, GT/src/bcg.cuh:154-168): one memory line where a separate checkpoint array
// needed two.  Every vertex's payload starts on a 32-bit word.
constexpr uint32_t kGtCheckShiftDefault = 2;

// Decoded-header cache: one 16-byte entry per vertex, so that a sampling step reads one aligned
// dwordx4 instead of walking four dependent gamma codes through the record, and needs no
// division to recover the chunk geometry.
//   x = payload bit [31:0]          (LE d == 1: the neighbour itself)
//   y = payload bit [37:32] | f1 << 6 | f2 << 11 | q << 16
//   z = d
//   w = C | first_pos << 16          (C, q, first_pos <= sqrt(2^31) + 2 < 2^16)
// Payload bits must be < 2^38 (a payload below 32 GiB); f1, f2 <= 31.
constexpr uint64_t kMaxPayloadBits = 1ull << 38;

__host__ __device__ __forceinline__ uint4 pack_hdr(const Hdr& h, int fmt) {
    uint4 e;
    e.z = h.d;
    if (fmt == kLE && h.d == 1) {
        e.x = h.f1;
        e.y = 0;
        e.w = 0;
        return e;
    }
    e.x = (uint32_t)h.pay;
    e.y = ((uint32_t)(h.pay >> 32) & 63u) | (h.f1 << 6) | (h.f2 << 11) | (h.q << 16);
    e.w = h.C | (h.fp << 16);
    return e;
}

template <int FMT>
__device__ __forceinline__ void unpack_hdr(uint4 e, Hdr& h) {
    h.d = e.z;
    if (FMT == kLE && h.d == 1) {
        h.f1 = e.x;
        h.f2 = h.fp = h.C = h.q = h.R = 0;
        h.pay = 0;
        return;
    }
    h.pay = ((uint64_t)(e.y & 63u) << 32) | e.x;
    h.f1 = (e.y >> 6) & 31u;
    h.f2 = (e.y >> 11) & 31u;
    h.q = e.y >> 16;
    h.C = e.w & 0xFFFFu;
    h.fp = e.w >> 16;
    const uint32_t D = FMT == kLE ? h.d - 1 : h.d;
    h.R = D - h.C * h.q;
}

// Header bits of a well-formed record (a field holding x is the Elias-gamma code of x + 1,
// 2 floor(log2(x + 1)) + 1 bits): LE g(d) [g(rock_bit) [g(sea_bit) g(first_pos)]] (d = 1
// stores the neighbour in the rock_bit field), GT all four fields.
template <int FMT>
__device__ __forceinline__ uint32_t header_bits(const Hdr& h) {
    const uint32_t b = 32 - __clz(h.d + 1);  // bits of d + 1
    uint32_t t = 2 * b - 1;
    if (FMT == kGT || h.d >= 1) t += 2 * (32 - __clz(h.f1 + 1)) - 1;
    if (FMT == kGT || h.d >= 2) t += 2 * (32 - __clz(h.f2 + 1)) - 1 + 2 * (32 - __clz(h.fp + 1)) - 1;
    return t;
}

// The payload bit of a GT entry points into the device's GT payload stream once the loader
// has built it (into .graph before that).
template <int FMT>
__device__ __forceinline__ void load_hdr(const DevGraph& g, uint32_t v, Hdr& h) {
    unpack_hdr<FMT>(g.hdr[v], h);
}

__host__ __device__ __forceinline__ uint32_t isqrt_u32(uint32_t x) {
    uint32_t s = (uint32_t)sqrtf((float)x);
    while ((uint64_t)s * s > x) --s;
    while ((uint64_t)(s + 1) * (s + 1) <= x) ++s;
    return s;
}

// Chunk geometry of an LE record with d >= 2 (LE/chunk_compress/compressor.hpp:61-67).
__host__ __device__ __forceinline__ void le_geometry(uint32_t d, uint32_t& C, uint32_t& q,
                                                     uint32_t& R) {
    uint32_t D = d - 1, s = isqrt_u32(D), r = D - s * s;
    C = s + (r > s);
    q = D / C;
    R = D - C * q;
}

// Block geometry of a GT record (GT/chunk_compress/compressor.hpp:49-62).
__host__ __device__ __forceinline__ void gt_geometry(uint32_t d, uint32_t& C, uint32_t& q,
                                                     uint32_t& R) {
    uint32_t s = isqrt_u32(d);
    if (s * s == d) {
        C = q = s;
        R = 0;
    } else {
        C = s + 1;
        q = d / C;
        R = d - C * q;
    }
}

// 32-bit window at bit b (the funnel shift of LE/src/bcg.cuh:66-72).
__device__ __forceinline__ uint32_t peek32(const uint32_t* __restrict__ w, uint64_t b) {
    const uint32_t* p = w + (b >> 5);
    uint64_t x = ((uint64_t)p[0] << 32) | p[1];
    return (uint32_t)(x >> (32 - (uint32_t)(b & 31)));
}

// 64-bit window at bit b (three words).
__device__ __forceinline__ uint64_t peek64(const uint32_t* __restrict__ w, uint64_t b) {
    const uint32_t* p = w + (b >> 5);
    const uint32_t s = (uint32_t)(b & 31);
    uint64_t x = ((uint64_t)p[0] << 32) | p[1];
    return s ? (x << s) | (p[2] >> (32 - s)) : x;
}

// nbits in [1, 32].
__device__ __forceinline__ uint32_t read_bits(const uint32_t* __restrict__ w, uint64_t b,
                                              uint32_t nbits) {
    return peek32(w, b) >> (32 - nbits);
}

// nbits in [0, 32].
__device__ __forceinline__ uint32_t read_bits0(const uint32_t* __restrict__ w, uint64_t b,
                                               uint32_t nbits) {
    return nbits ? read_bits(w, b, nbits) : 0u;
}

// Elias-gamma value at b, b advanced past it (LE/src/bcg.cuh:74-98), for the load-time
// validator: false on an all-zero 32-bit window (a code the reference's 32-bit clz cannot
// represent).  Samplers never parse headers: they read the decoded-header cache.
__device__ __forceinline__ bool gamma_checked(const uint32_t* __restrict__ w, uint64_t& b,
                                              uint32_t& val) {
    uint32_t x = peek32(w, b);
    if (x == 0) return false;
    uint32_t L = __clz(x);
    uint32_t y = peek32(w, b + L);
    b += 2 * L + 1;
    val = (y >> (31 - L)) - 1u;
    return true;
}

// Payload length implied by a header; the validator requires header + payload == record.
__host__ __device__ __forceinline__ uint64_t payload_bits(int fmt, const Hdr& h) {
    if (fmt == kLE) {
        if (h.d <= 1) return 0;
        return (uint64_t)(h.C + 1) * h.f1 + (uint64_t)(h.d - 1 - h.C) * (h.f2 + 1);
    }
    if (h.d == 0) return 0;
    return (uint64_t)h.C * h.f1 + (uint64_t)(h.d - h.C) * h.f2;
}

// floor(n / d) for 1 <= d < 2^31 (a chunk or block length, at most a degree): a float reciprocal
// estimate corrected by one step.  While n < 2^20 the estimate's relative error (rounding of d,
// the reciprocal, the product: under 2^-22) is under 0.25 in absolute terms, so the truncated
// estimate is the quotient or one off it and one signed-remainder step fixes it; larger numerators
// take the integer division.  A per-lane division whose divisor differs from lane to lane, at a
// fraction of the integer sequence's cost.
__device__ __forceinline__ uint32_t udiv_small(uint32_t n, uint32_t d) {
    if (n >= (1u << 20)) return n / d;
    uint32_t q = (uint32_t)((float)n * __builtin_amdgcn_rcpf((float)d));
    const int32_t r = (int32_t)(n - q * d);
    if (r < 0) --q;
    else if (r >= (int32_t)d) ++q;
    return q;
}

// Position k of an LE record -> (chunk c, index i in chunk, chunk length len).
__device__ __forceinline__ void le_locate(const Hdr& h, uint32_t k, uint32_t& c, uint32_t& i,
                                          uint32_t& len) {
    const uint32_t big = h.R * (h.q + 1);
    const bool lo = k < big;
    len = lo ? h.q + 1 : h.q;
    const uint32_t kk = lo ? k : k - big;
    c = udiv_small(kk, len);
    i = kk - c * len;
    if (!lo) c += h.R;
}

// Neighbour k of vertex v, LE (LE/src/bcg.cuh:196-261); 0 <= k < d.
__device__ __forceinline__ uint32_t le_get(const uint32_t* __restrict__ w, const Hdr& h,
                                           uint32_t v, uint32_t k, uint32_t* fbits = nullptr) {
    if (h.d == 1) {
        if (fbits) *fbits = 0;
        return h.f1;
    }
    uint32_t c, i, len;
    le_locate(h, k, c, i, len);
    const uint32_t rb = h.f1;
    // rocks c and c+1 are adjacent fields of rb <= 31 bits: one 64-bit window holds both
    const uint64_t win = peek64(w, h.pay + (uint64_t)c * rb);
    uint32_t r1 = (uint32_t)(win >> (64 - rb));
    uint32_t a = c < h.fp ? v - r1 : v + r1;
    if (i == 0) {
        if (fbits) *fbits = rb;
        return a;
    }
    uint32_t r2 = (uint32_t)((win << rb) >> (64 - rb));
    uint32_t z = c + 1 < h.fp ? v - r2 : v + r2;
    uint32_t j = (h.q - 1) * c + (c < h.R ? c : h.R) + i - 1;
    uint64_t so = h.pay + (uint64_t)(h.C + 1) * rb + (uint64_t)j * (h.f2 + 1);
    uint32_t x = read_bits(w, so, h.f2 + 1);
    uint32_t mag = x >> 1;
    if (fbits) *fbits = 2 * rb + h.f2 + 1;
    return udiv_small(z - a, len) * i + a + ((x & 1) ? 0u - mag : mag);
}

// Position k of a GT record -> (block, index i in block).
__device__ __forceinline__ void gt_locate(const Hdr& h, uint32_t k, uint32_t& blk, uint32_t& i) {
    const uint32_t big = (h.q + 1) * h.R;
    const bool lo = k < big;
    const uint32_t len = lo ? h.q + 1 : h.q;
    const uint32_t kk = lo ? k : k - big;
    const uint32_t b = udiv_small(kk, len);
    i = kk - b * len;
    blk = lo ? b : b + h.R;
}

// Bit where block blk of a GT record starts in .graph (GT/src/bcg.cuh:161).
__device__ __forceinline__ uint64_t gt_file_block_bit(const Hdr& h, uint32_t blk) {
    return h.pay + (uint64_t)blk * (h.f2 * (h.q - 1) + h.f1) +
           (uint64_t)(blk < h.R ? blk : h.R) * h.f2;
}

// Device GT payload stream geometry (P = 1 << sh): bit of in-block position i from the block
// start, i.e. the head, i - 1 fields and one 32-bit checkpoint per P of them.
__host__ __device__ __forceinline__ uint64_t gtx_field(const Hdr& h, uint32_t i, uint32_t sh) {
    return i ? h.f1 + (uint64_t)(i - 1) * h.f2 + (uint64_t)((i - 1) >> sh) * (32 - h.f2) : 0;
}

// Bit where block blk starts: R blocks of q + 1 positions, then blocks of q.
__host__ __device__ __forceinline__ uint64_t gtx_block_bit(const Hdr& h, uint32_t blk, uint32_t sh) {
    const uint32_t big = blk < h.R ? blk : h.R;
    return h.pay + (uint64_t)big * gtx_field(h, h.q + 1, sh) + (uint64_t)(blk - big) * gtx_field(h, h.q, sh);
}

// Payload bits of a GT vertex in the device stream.
__host__ __device__ __forceinline__ uint64_t gtx_payload_bits(const Hdr& h, uint32_t sh) {
    return h.d ? (uint64_t)h.R * gtx_field(h, h.q + 1, sh) + (uint64_t)(h.C - h.R) * gtx_field(h, h.q, sh) : 0;
}

// Absolute value of the head of block blk.
__device__ __forceinline__ uint32_t gt_head(const uint32_t* __restrict__ w, const Hdr& h,
                                            uint32_t v, uint32_t blk, uint64_t bo) {
    uint32_t head = read_bits(w, bo, h.f1);
    return blk >= h.fp ? v + head + 1 : v - head - 1;
}

// Sum of `count` consecutive width-bit fields starting at bit b (width in [1, 32]); one load
// per 32 bits consumed.
__device__ __forceinline__ uint32_t sum_fields(const uint32_t* __restrict__ w, uint64_t b,
                                               uint32_t count, uint32_t width) {
    const uint32_t* p = w + (b >> 5);
    uint32_t s = (uint32_t)(b & 31);
    uint64_t buf = (((uint64_t)p[0] << 32) | p[1]) << s;
    int avail = 64 - (int)s;
    p += 2;
    uint32_t sum = 0;
    for (uint32_t t = 0; t < count; ++t) {
        if (avail < (int)width) {
            buf |= (uint64_t)(*p++) << (32 - avail);
            avail += 32;
        }
        sum += (uint32_t)(buf >> (64 - width));
        buf <<= width;
        avail -= (int)width;
    }
    return sum;
}

// GT random access in the device stream (GT/src/bcg.cuh:154-168 reads the head and i gaps):
// neighbour k of v is the absolute value at in-block position i & ~(P - 1) (the block head, or
// a checkpoint) plus the at most P - 1 gaps (body + 1) stored right after it.  fbits receives
// the bits the same access costs in the .graph layout with a separate checkpoint table: 32 for
// a checkpoint (head_bit for a head) plus the gap fields.
// The absolute field and the gaps after it are read with one 16-byte load (four words from the
// field's first word; the stream is padded, so the words past its end exist): 128 bits hold the
// field's offset in its word (< 32), the field and `rem` gaps when they add up to 128 or less --
// a checkpoint and three gaps of up to 21 bits.  Longer spans read the gaps word by word.
struct __attribute__((packed, aligned(4))) Words4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ uint32_t gt_get_cp(const DevGraph& g, const Hdr& h, uint32_t v, uint32_t k,
                                              uint32_t* fbits = nullptr) {
    uint32_t blk, i;
    gt_locate(h, k, blk, i);
    const uint32_t sh = g.cp_shift;
    const uint64_t bo = gtx_block_bit(h, blk, sh);
    const uint32_t b = (i >> sh) << sh;  // 0 when the vertex has no checkpoints (q < P)
    const uint32_t rem = i - b;
    const uint64_t at = b ? bo + gtx_field(h, b, sh) : bo;  // the checkpoint, or the block head
    const uint32_t aw = b ? 32u : h.f1;
    const Words4 q = *(const Words4*)(g.w + (at >> 5));
    uint64_t hi = ((uint64_t)q.x << 32) | q.y, lo = ((uint64_t)q.z << 32) | q.w;
    const uint32_t s = (uint32_t)(at & 31);
    const uint32_t a = (uint32_t)((hi << s) >> (64 - aw));
    uint32_t val = b ? a : (blk >= h.fp ? v + a + 1 : v - a - 1);
    if (rem) {
        const uint32_t f2 = h.f2, sft = s + aw;  // 1 <= sft <= 63
        if (sft + rem * f2 <= 128) {
            hi = (hi << sft) | (lo >> (64 - sft));
            lo <<= sft;
            uint32_t sum = rem;
            if (f2)  // (a well-formed block with gaps has f2 >= 1; 0-bit fields read as 0)
                for (uint32_t t = 0; t < rem; ++t) {
                    sum += (uint32_t)(hi >> (64 - f2));
                    hi = (hi << f2) | (lo >> (64 - f2));
                    lo <<= f2;
                }
            val += sum;
        } else {
            val += sum_fields(g.w, at + aw, rem, f2) + rem;
        }
    }
    if (fbits) *fbits = aw + rem * h.f2;
    return val;
}

// Position i of block blk of a GT vertex in the device stream: the absolute id (head = 1) at
// the block head and at checkpoints, else the gap + 1 (head = 0); a segmented prefix sum over
// consecutive positions turns them into ids.
__device__ __forceinline__ uint32_t gt_element(const DevGraph& g, const Hdr& h, uint32_t v, uint32_t blk,
                                               uint32_t i, int& head) {
    const uint32_t sh = g.cp_shift;
    const uint64_t bo = gtx_block_bit(h, blk, sh);
    // one load whatever the kind of field, so that lanes of different kinds do not diverge
    const bool hd = i == 0, abs = !hd && (i & ((1u << sh) - 1)) == 0;
    const uint32_t raw = read_bits(g.w, bo + gtx_field(h, i, sh), hd ? h.f1 : abs ? 32u : h.f2);
    head = hd || abs;
    return hd ? (blk >= h.fp ? v + raw + 1 : v - raw - 1) : abs ? raw : raw + 1;
}

// Division by a divisor every lane shares, as a multiply-high and one correction: with
// m = floor(2^32 / d) the estimate floor(n * m / 2^32) is floor(n / d) or one less, n < 2^32.
struct UDiv {
    uint32_t d = 1, m = 0;
    __device__ __forceinline__ void init(uint32_t dd) {
        d = dd;
        // floor(2^32 / d) from a 32-bit division: floor((2^32 - 1) / d), plus one when d divides
        // 2^32 (a power of two), so no 64-bit division is emulated
        m = 0;
        if (dd > 1) {
            m = 0xFFFFFFFFu / dd;
            if ((uint64_t)(m + 1) * dd == 0x100000000ull) ++m;
        }
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        if (d == 1) return n;
        uint32_t q = __umulhi(n, m);
        if (n - q * d >= d) ++q;
        return q;
    }
};

// Element k of vertex v when the whole wave is inside v's list (h uniform): the record
// geometry's divisions go through the shared UDivs d1 = q + 1 and d0 = q.  LE returns the
// neighbour; GT returns an absolute id (block head or checkpoint, head = 1) or the gap + 1
// (head = 0).
template <int FMT>
__device__ __forceinline__ uint32_t element_uniform(const DevGraph& g, const Hdr& h, uint32_t v, uint32_t k,
                                                    const UDiv& d1, const UDiv& d0, int& head) {
    head = 1;
    const uint32_t big = (h.q + 1) * h.R;
    uint32_t c, i, len;
    if (k < big) {
        c = d1.div(k);
        i = k - c * (h.q + 1);
        len = h.q + 1;
    } else {
        const uint32_t kk = k - big;
        const uint32_t c2 = d0.div(kk);
        i = kk - c2 * h.q;
        c = c2 + h.R;
        len = h.q;
    }
    if (FMT == kLE) {
        if (h.d == 1) return h.f1;
        const uint32_t rb = h.f1;
        const uint64_t win = peek64(g.w, h.pay + (uint64_t)c * rb);
        const uint32_t r1 = (uint32_t)(win >> (64 - rb));
        const uint32_t a = c < h.fp ? v - r1 : v + r1;
        if (i == 0) return a;
        const uint32_t r2 = (uint32_t)((win << rb) >> (64 - rb));
        const uint32_t z = c + 1 < h.fp ? v - r2 : v + r2;
        const uint32_t j = (h.q - 1) * c + (c < h.R ? c : h.R) + i - 1;
        const uint64_t so = h.pay + (uint64_t)(h.C + 1) * rb + (uint64_t)j * (h.f2 + 1);
        const uint32_t x = read_bits(g.w, so, h.f2 + 1);
        const uint32_t mag = x >> 1;
        const uint32_t step = (len == h.q + 1 ? d1 : d0).div(z - a);  // (z - a) / len
        return step * i + a + ((x & 1) ? 0u - mag : mag);
    } else {
        return gt_element(g, h, v, c, i, head);
    }
}

// Neighbour k (< d) of vertex v with header h.
template <int FMT>
__device__ __forceinline__ uint32_t get_neighbor(const DevGraph& g, const Hdr& h, uint32_t v, uint32_t k,
                                                 uint32_t* fbits = nullptr) {
    return FMT == kLE ? le_get(g.w, h, v, k, fbits) : gt_get_cp(g, h, v, k, fbits);
}

}  // namespace grass
