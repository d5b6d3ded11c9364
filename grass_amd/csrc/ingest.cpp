// Host ingest: R-MAT generator, graph_sort preprocessing and the LE / GT compressors.
//
// The reference tools (LE/chunk_compress/{graph_sort.cpp,compressor.hpp}, GT/chunk_compress/
// compressor.hpp) are single-threaded text pipelines holding a vector<bool> per vertex and two
// 16M-entry gamma tables (~55 GB and hours at R-MAT scale 26).  This keeps their output bytes
// and their numbering exactly, but works on flat arrays with std::thread:
//   graph_sort: first-appearance numbering, then a bucketed counting sort by source with a
//               per-bucket sort + unique (no global sort of all pairs);
//   compress:   pass 1 computes every record's bit length, an exclusive scan places the
//               records, pass 2 writes them into a shared word buffer (atomic OR only on the
//               two words a thread shares with its neighbours).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <ctype.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/grass.h"
#include "../../include/grass_ingest.h"
#include "blob.h"
#include "codec.h"
#include "philox.h"

struct grass_edges {
    std::vector<uint32_t> src, dst;
};

namespace grass {
namespace {

int threads_of(int nthreads) {
    if (nthreads > 0) return nthreads;
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0) hc = 1;
    return (int)std::min(hc, 16u);  // one GPU's share of a host
}

// f(t, lo, hi) over nt static chunks of [0, n).
template <class F>
void parallel_chunks(int nt, uint64_t n, F f) {
    if (nt <= 1 || n < 4096) {
        f(0, 0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        uint64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=, &f] { f(t, lo, hi); });
    }
    for (auto& x : th) x.join();
}

// f(i) for i in [0, n), dynamically scheduled in grains.
template <class F>
void parallel_dynamic(int nt, uint64_t n, uint64_t grain, F f) {
    std::atomic<uint64_t> next{0};
    auto body = [&] {
        for (;;) {
            uint64_t lo = next.fetch_add(grain);
            if (lo >= n) break;
            uint64_t hi = std::min(n, lo + grain);
            for (uint64_t i = lo; i < hi; ++i) f(i);
        }
    };
    if (nt <= 1) {
        body();
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(body);
    for (auto& x : th) x.join();
}

// get_bit_num (LE/chunk_compress/compressor.hpp:16-22): bit length of |x|, 1 for 0.
inline int bit_num(int64_t x) {
    if (x < 0) x = -x;
    if (x == 0) return 1;
    return 64 - __builtin_clzll((uint64_t)x);
}

// Length of the Elias-gamma code of x + 1 (compressor.hpp:411-417).
inline uint64_t gamma_len(uint64_t x) { return 2 * (uint64_t)(63 - __builtin_clzll(x + 1)) + 1; }

// MSB-first bit writer into big-endian-bit 64-bit words.  Words [lo_w, hi_w] may be shared with
// the neighbouring threads' ranges and are updated atomically.
struct BitWriter {
    uint64_t* w;
    uint64_t pos;
    uint64_t lo_w, hi_w;
    inline void orw(uint64_t idx, uint64_t x) {
        if (idx == lo_w || idx == hi_w) __atomic_fetch_or(&w[idx], x, __ATOMIC_RELAXED);
        else w[idx] |= x;
    }
    // x < 2^nb, nb in [0, 64]
    inline void put(uint64_t x, int nb) {
        if (nb == 0) return;
        uint64_t idx = pos >> 6;
        int off = (int)(pos & 63);
        if (off + nb <= 64) {
            orw(idx, x << (64 - off - nb));
        } else {
            int spill = off + nb - 64;
            orw(idx, x >> spill);
            orw(idx + 1, x << (64 - spill));
        }
        pos += nb;
    }
    inline void gamma(uint64_t x) {
        uint64_t y = x + 1;
        int L = 63 - __builtin_clzll(y);
        put(0, L);
        put(y, L + 1);
    }
};

struct LeRec {
    uint32_t C = 0, q = 0, R = 0, rb = 0, sb = 0, fp = 0;
    uint64_t bits = 0;
};

// LE record geometry and field widths (compressor.hpp:49-71, 288-327).
LeRec le_plan(uint32_t v, const uint32_t* N, uint32_t d) {
    LeRec r;
    if (d == 0) {
        r.bits = gamma_len(0);
        return r;
    }
    if (d == 1) {
        r.bits = gamma_len(1) + gamma_len(N[0]);
        return r;
    }
    le_geometry(d, r.C, r.q, r.R);
    const int64_t vv = v;
    r.rb = (uint32_t)bit_num(std::max(std::llabs((int64_t)N[0] - vv), std::llabs((int64_t)N[d - 1] - vv)));
    uint64_t cur = 0;
    int sb = 0;
    uint32_t fp = 0;
    for (uint32_t c = 0; c < r.C; ++c) {
        uint32_t inner = (r.q - 1) + (c < r.R);
        int64_t rl = N[cur], rr = N[cur + inner + 1];
        int64_t dis = (rr - rl) / (int64_t)(inner + 1);
        if (rl < vv) fp = c + 1;
        for (uint32_t i = 1; i <= inner; ++i)
            sb = std::max(sb, bit_num((int64_t)N[cur + i] - (rl + (int64_t)i * dis)));
        cur += inner + 1;
    }
    if ((int64_t)N[cur] < vv) fp = r.C + 1;
    r.sb = (uint32_t)sb;
    r.fp = fp;
    r.bits = gamma_len(d) + gamma_len(r.rb) + gamma_len(r.sb) + gamma_len(r.fp) +
             (uint64_t)(r.C + 1) * r.rb + (uint64_t)(d - 1 - r.C) * (r.sb + 1);
    return r;
}

// LE record bits (compressor.hpp:329-390 with GATHER_HEAD: rocks, then signed seas).
void le_write(BitWriter& bw, uint32_t v, const uint32_t* N, uint32_t d, const LeRec& r) {
    bw.gamma(d);
    if (d == 0) return;
    if (d == 1) {
        bw.gamma(N[0]);
        return;
    }
    bw.gamma(r.rb);
    bw.gamma(r.sb);
    bw.gamma(r.fp);
    const int64_t vv = v;
    uint64_t cur = 0;
    for (uint32_t c = 0; c <= r.C; ++c) {
        bw.put((uint64_t)std::llabs((int64_t)N[cur] - vv), (int)r.rb);
        if (c < r.C) cur += (r.q - 1) + (c < r.R) + 1;
    }
    cur = 0;
    for (uint32_t c = 0; c < r.C; ++c) {
        uint32_t inner = (r.q - 1) + (c < r.R);
        int64_t rl = N[cur], rr = N[cur + inner + 1];
        int64_t dis = (rr - rl) / (int64_t)(inner + 1);
        for (uint32_t i = 1; i <= inner; ++i) {
            int64_t res = (int64_t)N[cur + i] - (rl + (int64_t)i * dis);
            bw.put((uint64_t)std::llabs(res), (int)r.sb);
            bw.put(res < 0 ? 1 : 0, 1);
        }
        cur += inner + 1;
    }
}

struct GtRec {
    uint32_t C = 0, q = 0, R = 0, hb = 0, bb = 0, fp = 0;
    uint64_t bits = 0;
};

// GT head value as the reference writes it: |N - v| - 1, where a self loop (-1) comes out
// as 1 (append_bit of -1 keeps its lowest bit, GT/chunk_compress/compressor.hpp:407-419).
inline uint64_t gt_head_field(int64_t head) { return head >= 0 ? (uint64_t)head : 1; }

// GT record geometry and widths (GT/chunk_compress/compressor.hpp:49-62, 245-282).
GtRec gt_plan(uint32_t v, const uint32_t* N, uint32_t d) {
    GtRec r;
    if (d) gt_geometry(d, r.C, r.q, r.R);
    const int64_t vv = v;
    int hb = 0, bb = 0;
    uint64_t cur = 0;
    for (uint32_t b = 0; b < r.C; ++b) {
        uint32_t len = r.q + (b < r.R);
        int64_t x = (int64_t)N[cur] - vv;
        if (x < 0) r.fp = b + 1;
        hb = std::max(hb, bit_num(std::llabs(x) - 1));
        for (uint32_t i = 1; i < len; ++i)
            bb = std::max(bb, bit_num((int64_t)N[cur + i] - (int64_t)N[cur + i - 1] - 1));
        cur += len;
    }
    r.hb = (uint32_t)hb;
    r.bb = (uint32_t)bb;
    r.bits = gamma_len(d) + gamma_len(r.hb) + gamma_len(r.bb) + gamma_len(r.fp) +
             (uint64_t)r.C * r.hb + (uint64_t)(d - r.C) * r.bb;
    return r;
}

// GT record bits (GT/chunk_compress/compressor.hpp:297-349, heads and bodies interleaved).
void gt_write(BitWriter& bw, uint32_t v, const uint32_t* N, uint32_t d, const GtRec& r) {
    bw.gamma(d);
    bw.gamma(r.hb);
    bw.gamma(r.bb);
    bw.gamma(r.fp);
    const int64_t vv = v;
    uint64_t cur = 0;
    for (uint32_t b = 0; b < r.C; ++b) {
        uint32_t len = r.q + (b < r.R);
        bw.put(gt_head_field(std::llabs((int64_t)N[cur] - vv) - 1), (int)r.hb);
        for (uint32_t i = 1; i < len; ++i)
            bw.put((uint64_t)((int64_t)N[cur + i] - (int64_t)N[cur + i - 1] - 1), (int)r.bb);
        cur += len;
    }
}

// Big-endian-bit words -> bytes.
void words_to_bytes(const std::vector<uint64_t>& w, uint64_t nbytes, grass_blob& out,
                    int nt) {
    if (!out.resize(nbytes)) throw std::bad_alloc();
    uint8_t* o = out.data();
    parallel_chunks(nt, nbytes, [&](int, uint64_t lo, uint64_t hi) {
        for (uint64_t j = lo; j < hi; ++j) o[j] = (uint8_t)(w[j >> 3] >> (56 - 8 * (j & 7)));
    });
}

// Whitespace-separated unsigned integers of buf[lo, hi) (lo and hi on token boundaries).
// Parsing stops at the first token that is not a number below 2^32 - 1; *stop then holds its
// position.  Tokens are what the reference's `fin >> u >> v` (graph_sort.cpp:45) and
// fscanf("%ld %ld") (compressor.hpp:101) read.
void parse_tokens(const char* buf, uint64_t lo, uint64_t hi, std::vector<uint32_t>& out, uint64_t* stop) {
    *stop = UINT64_MAX;
    uint64_t i = lo;
    for (;;) {
        while (i < hi && isspace((unsigned char)buf[i])) ++i;
        if (i >= hi) return;
        const uint64_t t0 = i;
        uint64_t v = 0;
        bool ok = buf[i] >= '0' && buf[i] <= '9';
        while (i < hi && !isspace((unsigned char)buf[i])) {
            const char ch = buf[i];
            if (ch < '0' || ch > '9') ok = false;
            else if (ok) v = v * 10 + (uint64_t)(ch - '0');
            if (v >= 0xFFFFFFFFull) ok = false;
            ++i;
        }
        if (!ok) {
            *stop = t0;
            return;
        }
        out.push_back((uint32_t)v);
    }
}

}  // namespace
}  // namespace grass

using namespace grass;

extern "C" {

int grass_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                     uint64_t seed, uint32_t* src, uint32_t* dst, int nthreads) {
    if (scale == 0 || scale > 31 || edge_factor == 0 || !src || !dst) return GRASS_EINVAL;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GRASS_EINVAL;
    const uint64_t m = (uint64_t)edge_factor << scale;
    const double two32 = 4294967296.0;
    const uint64_t tA = (uint64_t)(a * two32), tAB = (uint64_t)((a + b) * two32),
                   tABC = (uint64_t)((a + b + c) * two32);
    const int nt = threads_of(nthreads);
    parallel_chunks(nt, m, [&](int, uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t u = 0, v = 0;
            U4 r{0, 0, 0, 0};
            for (uint32_t l = 0; l < scale; ++l) {
                if ((l & 3) == 0) r = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), l >> 2, 0x524D4154u, seed);
                uint64_t x = (l & 3) == 0 ? r.x : (l & 3) == 1 ? r.y : (l & 3) == 2 ? r.z : r.w;
                uint32_t bu = x >= tAB;
                uint32_t bv = (x >= tA && x < tAB) || x >= tABC;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            src[i] = u;
            dst[i] = v;
        }
    });
    return GRASS_OK;
}

int grass_graph_sort(const uint32_t* src, const uint32_t* dst, uint64_t m, int directed,
                     int nthreads, grass_csr** out) {
    if (!out || (m && (!src || !dst))) return GRASS_EINVAL;
    *out = nullptr;
    const int nt = threads_of(nthreads);
    std::vector<uint32_t> tmax(nt, 0);
    parallel_chunks(nt, m, [&](int t, uint64_t lo, uint64_t hi) {
        uint32_t mx = 0;
        for (uint64_t i = lo; i < hi; ++i) mx = std::max(mx, std::max(src[i], dst[i]));
        tmax[t] = mx;
    });
    const uint64_t maxid = m ? *std::max_element(tmax.begin(), tmax.end()) : 0;
    if (maxid >= 0x80000000ull) return GRASS_EINVAL;
    grass_csr* csr = new (std::nothrow) grass_csr();
    if (!csr) return GRASS_ENOMEM;
    try {
        // ids in order of first appearance, u before v, self loops skipped (graph_sort.cpp:47-52)
        std::vector<uint32_t> newid(maxid + 1, UINT32_MAX);
        uint32_t tot = 0;
        for (uint64_t i = 0; i < m; ++i) {
            uint32_t u = src[i], v = dst[i];
            if (u == v) continue;
            if (newid[u] == UINT32_MAX) newid[u] = tot++;
            if (newid[v] == UINT32_MAX) newid[v] = tot++;
        }
        const uint64_t n = tot;
        // bucket by source: B buckets of 2^shift vertices
        int shift = 0;
        while ((n >> shift) > 65535) ++shift;
        const uint64_t B = (n >> shift) + 1;
        const int per = directed ? 1 : 2;
        std::vector<uint64_t> hist((uint64_t)nt * B, 0);
        parallel_chunks(nt, m, [&](int t, uint64_t lo, uint64_t hi) {
            uint64_t* h = &hist[(uint64_t)t * B];
            for (uint64_t i = lo; i < hi; ++i) {
                uint32_t u = src[i], v = dst[i];
                if (u == v) continue;
                uint32_t a = newid[u], b = newid[v];
                h[a >> shift]++;
                if (per == 2) h[b >> shift]++;
            }
        });
        // offsets: bucket-major, thread-minor (keeps each thread's writes in stream order)
        std::vector<uint64_t> boff(B + 1, 0);
        {
            uint64_t acc = 0;
            for (uint64_t bk = 0; bk < B; ++bk) {
                boff[bk] = acc;
                for (int t = 0; t < nt; ++t) {
                    uint64_t x = hist[(uint64_t)t * B + bk];
                    hist[(uint64_t)t * B + bk] = acc;
                    acc += x;
                }
            }
            boff[B] = acc;
        }
        std::vector<uint64_t> keys(boff[B]);
        parallel_chunks(nt, m, [&](int t, uint64_t lo, uint64_t hi) {
            uint64_t* h = &hist[(uint64_t)t * B];
            for (uint64_t i = lo; i < hi; ++i) {
                uint32_t u = src[i], v = dst[i];
                if (u == v) continue;
                uint64_t a = newid[u], b = newid[v];
                keys[h[a >> shift]++] = (a << 32) | b;
                if (per == 2) keys[h[b >> shift]++] = (b << 32) | a;
            }
        });
        std::vector<uint32_t>().swap(newid);
        std::vector<uint64_t>().swap(hist);
        // per bucket: sort, unique, degrees
        std::vector<uint64_t> deg(n + 1, 0);
        std::vector<uint64_t> bend(B, 0);  // end of the unique prefix of each bucket
        parallel_dynamic(nt, B, 4, [&](uint64_t bk) {
            uint64_t* p = keys.data() + boff[bk];
            uint64_t len = boff[bk + 1] - boff[bk];
            std::sort(p, p + len);
            uint64_t u = len ? (uint64_t)(std::unique(p, p + len) - p) : 0;
            bend[bk] = boff[bk] + u;
            for (uint64_t j = 0; j < u; ++j) deg[p[j] >> 32]++;  // vertices of one bucket only
        });
        csr->n = n;
        csr->rowptr.assign(n + 1, 0);
        for (uint64_t v = 0; v < n; ++v) csr->rowptr[v + 1] = csr->rowptr[v] + deg[v];
        csr->m = csr->rowptr[n];
        csr->col.resize(csr->m);
        parallel_dynamic(nt, B, 4, [&](uint64_t bk) {
            const uint64_t* p = keys.data() + boff[bk];
            uint64_t u = bend[bk] - boff[bk];
            if (!u) return;
            uint64_t pos = csr->rowptr[p[0] >> 32];
            for (uint64_t j = 0; j < u; ++j) csr->col[pos + j] = (uint32_t)p[j];
        });
    } catch (const std::bad_alloc&) {
        delete csr;
        return GRASS_ENOMEM;
    }
    *out = csr;
    return GRASS_OK;
}

int grass_csr_from_sorted_edges(const uint32_t* src, const uint32_t* dst, uint64_t m,
                                grass_csr** out) {
    if (!out || (m && (!src || !dst))) return GRASS_EINVAL;
    *out = nullptr;
    grass_csr* csr = new (std::nothrow) grass_csr();
    if (!csr) return GRASS_ENOMEM;
    try {
        uint64_t n = 0;
        for (uint64_t i = 0; i < m; ++i) {
            n = std::max<uint64_t>(n, (uint64_t)std::max(src[i], dst[i]) + 1);
            if (i && (src[i] < src[i - 1] || (src[i] == src[i - 1] && dst[i] < dst[i - 1]))) {
                delete csr;
                return GRASS_EINVAL;  // the compressor asserts sorted input (compressor.hpp:109-113)
            }
        }
        if (n >= 0x80000000ull) {
            delete csr;
            return GRASS_EINVAL;
        }
        csr->n = n;
        csr->rowptr.assign(n + 1, 0);
        csr->col.reserve(m);
        for (uint64_t i = 0; i < m; ++i) {
            if (i && src[i] == src[i - 1] && dst[i] == dst[i - 1]) continue;  // adjacent duplicate
            csr->rowptr[src[i] + 1]++;
            csr->col.push_back(dst[i]);
        }
        for (uint64_t v = 0; v < n; ++v) csr->rowptr[v + 1] += csr->rowptr[v];
        csr->m = csr->col.size();
    } catch (const std::bad_alloc&) {
        delete csr;
        return GRASS_ENOMEM;
    }
    *out = csr;
    return GRASS_OK;
}

int grass_csr_info(const grass_csr* csr, uint64_t* n, uint64_t* m) {
    if (!csr) return GRASS_EINVAL;
    if (n) *n = csr->n;
    if (m) *m = csr->m;
    return GRASS_OK;
}

int grass_csr_arrays(const grass_csr* csr, const uint64_t** rowptr, const uint32_t** col) {
    if (!csr) return GRASS_EINVAL;
    if (rowptr) *rowptr = csr->rowptr.data();
    if (col) *col = csr->col.data();
    return GRASS_OK;
}

void grass_csr_free(grass_csr* csr) { delete csr; }

int grass_compress(const grass_csr* csr, int format, int nthreads, grass_blob** graph,
                   grass_blob** offset) {
    if (!csr || !graph || !offset || (format != GRASS_FMT_LE && format != GRASS_FMT_GT))
        return GRASS_EINVAL;
    *graph = *offset = nullptr;
    const uint64_t n = csr->n;
    if (n == 0) return GRASS_EINVAL;
    for (uint64_t v = 0; v < n; ++v)
        if (csr->rowptr[v + 1] - csr->rowptr[v] >= 0x80000000ull) return GRASS_EINVAL;
    const int nt = threads_of(nthreads);
    grass_blob* gb = new (std::nothrow) grass_blob();
    grass_blob* ob = new (std::nothrow) grass_blob();
    if (!gb || !ob) {
        delete gb;
        delete ob;
        return GRASS_ENOMEM;
    }
    try {
        const uint64_t* rp = csr->rowptr.data();
        const uint32_t* col = csr->col.data();
        // pass 1: record lengths
        std::vector<uint64_t> start(n + 1, 0);
        parallel_dynamic(nt, n, 4096, [&](uint64_t v) {
            const uint32_t* N = col + rp[v];
            uint32_t d = (uint32_t)(rp[v + 1] - rp[v]);
            start[v + 1] = format == GRASS_FMT_LE ? le_plan((uint32_t)v, N, d).bits
                                                  : gt_plan((uint32_t)v, N, d).bits;
        });
        uint64_t maxlen = 0;
        for (uint64_t v = 0; v < n; ++v) maxlen = std::max(maxlen, start[v + 1]);
        for (uint64_t v = 0; v < n; ++v) start[v + 1] += start[v];
        const uint64_t total = start[n];
        // pass 2: records into a shared word buffer, one contiguous vertex range per thread
        std::vector<uint64_t> w(total / 64 + 2, 0);
        parallel_chunks(nt, n, [&](int, uint64_t lo, uint64_t hi) {
            if (lo >= hi) return;
            BitWriter bw{w.data(), start[lo], start[lo] >> 6, start[hi] ? (start[hi] - 1) >> 6 : 0};
            for (uint64_t v = lo; v < hi; ++v) {
                const uint32_t* N = col + rp[v];
                uint32_t d = (uint32_t)(rp[v + 1] - rp[v]);
                if (format == GRASS_FMT_LE) le_write(bw, (uint32_t)v, N, d, le_plan((uint32_t)v, N, d));
                else gt_write(bw, (uint32_t)v, N, d, gt_plan((uint32_t)v, N, d));
            }
        });
        words_to_bytes(w, (total + 7) / 8, *gb, nt);
        // .offset: byte 0 = ubl, then n fields of ubl bits holding the record lengths
        // (LE/chunk_compress/compressor.hpp:217-285, 392-403)
        const int ubl = bit_num((int64_t)maxlen);
        const uint64_t obits = 8 + n * (uint64_t)ubl;
        std::vector<uint64_t> ow(obits / 64 + 2, 0);
        parallel_chunks(nt, n, [&](int, uint64_t lo, uint64_t hi) {
            if (lo >= hi) return;
            uint64_t b0 = 8 + lo * ubl, b1 = 8 + hi * ubl;
            BitWriter bw{ow.data(), b0, b0 >> 6, (b1 - 1) >> 6};
            for (uint64_t v = lo; v < hi; ++v) bw.put(start[v + 1] - start[v], ubl);
        });
        ow[0] |= (uint64_t)ubl << 56;
        words_to_bytes(ow, (obits + 7) / 8, *ob, nt);
    } catch (const std::bad_alloc&) {
        delete gb;
        delete ob;
        return GRASS_ENOMEM;
    }
    *graph = gb;
    *offset = ob;
    return GRASS_OK;
}

int grass_blob_data(const grass_blob* blob, const uint8_t** data, uint64_t* size) {
    if (!blob) return GRASS_EINVAL;
    if (data) *data = blob->data();
    if (size) *size = blob->size();
    return GRASS_OK;
}

void grass_blob_free(grass_blob* blob) { delete blob; }

int grass_write_files(const char* prefix, const grass_blob* graph, const grass_blob* offset) {
    if (!prefix || !graph || !offset) return GRASS_EINVAL;
    std::string p(prefix);
    const grass_blob* blobs[2] = {graph, offset};
    const char* ext[2] = {".graph", ".offset"};
    for (int i = 0; i < 2; ++i) {
        FILE* f = fopen((p + ext[i]).c_str(), "wb");
        if (!f) return GRASS_EIO;
        size_t sz = blobs[i]->size();
        bool ok = fwrite(blobs[i]->data(), 1, sz, f) == sz;
        ok = (fclose(f) == 0) && ok;
        if (!ok) return GRASS_EIO;
    }
    return GRASS_OK;
}

int grass_read_edge_list(const char* path, uint64_t skip_lines, int nthreads, grass_edges** out) {
    if (!path || !out) return GRASS_EINVAL;
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return GRASS_EIO;
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    const bool err = ferror(f) != 0;
    fclose(f);
    if (err) return GRASS_EIO;
    uint64_t pos = 0;
    const uint64_t n = buf.size();
    for (uint64_t l = 0; l < skip_lines && pos < n; ++l) {
        while (pos < n && buf[pos] != '\n') ++pos;
        if (pos < n) ++pos;
    }
    // chunks cut at whitespace so that no token straddles two chunks
    const int nt = threads_of(nthreads);
    std::vector<uint64_t> cut(nt + 1, n);
    cut[0] = pos;
    for (int t = 1; t < nt; ++t) {
        uint64_t c = std::max(cut[t - 1], pos + (n - pos) * t / nt);
        while (c < n && !isspace((unsigned char)buf[c])) ++c;
        cut[t] = c;
    }
    std::vector<std::vector<uint32_t>> tok(nt);
    std::vector<uint64_t> stop(nt, UINT64_MAX);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] { parse_tokens(buf.data(), cut[t], cut[t + 1], tok[t], &stop[t]); });
        for (auto& x : th) x.join();
    }
    std::unique_ptr<grass_edges> e(new (std::nothrow) grass_edges());
    if (!e) return GRASS_ENOMEM;
    uint64_t total = 0;
    int last = nt - 1;
    for (int t = 0; t < nt; ++t) {
        total += tok[t].size();
        if (stop[t] != UINT64_MAX) {
            last = t;
            break;
        }
    }
    const uint64_t m = total / 2;  // an unpaired last token is ignored, as the reference does
    e->src.resize(m);
    e->dst.resize(m);
    uint64_t k = 0;
    for (int t = 0; t <= last; ++t)
        for (uint32_t v : tok[t]) {
            if (k >= 2 * m) break;
            (k & 1 ? e->dst : e->src)[k >> 1] = v;
            ++k;
        }
    *out = e.release();
    return GRASS_OK;
}

int grass_edges_data(const grass_edges* e, const uint32_t** src, const uint32_t** dst, uint64_t* m) {
    if (!e) return GRASS_EINVAL;
    if (src) *src = e->src.data();
    if (dst) *dst = e->dst.data();
    if (m) *m = e->src.size();
    return GRASS_OK;
}

void grass_edges_free(grass_edges* e) { delete e; }

int grass_write_edge_list(const char* path, const grass_csr* csr, int nthreads) {
    if (!path || !csr) return GRASS_EINVAL;
    const int nt = threads_of(nthreads);
    std::vector<std::string> part(nt);
    parallel_chunks(nt, csr->n, [&](int t, uint64_t lo, uint64_t hi) {
        std::string& o = part[t];
        char line[32];
        for (uint64_t u = lo; u < hi; ++u)
            for (uint64_t k = csr->rowptr[u]; k < csr->rowptr[u + 1]; ++k) {
                int len = snprintf(line, sizeof(line), "%llu %u\n", (unsigned long long)u, csr->col[k]);
                o.append(line, (size_t)len);
            }
    });
    FILE* f = fopen(path, "wb");
    if (!f) return GRASS_EIO;
    bool ok = true;
    for (const std::string& p : part) ok = ok && fwrite(p.data(), 1, p.size(), f) == p.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? GRASS_OK : GRASS_EIO;
}

}  // extern "C"
