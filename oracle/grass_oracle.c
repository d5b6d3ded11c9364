/*
 * grass_oracle.c -- CPU oracle (test infrastructure only; see grass_oracle.h).
 *
 * Every decoding function restates the reference decoder; the file:line it follows is cited
 * at each function.  Deliberate differences from the reference, all of which the reference's
 * own encoder makes unobservable on valid files:
 *   - chunk/block geometry uses an exact integer square root where the reference decoder
 *     uses sqrt((float)x) (LE/src/bcg.cuh:142, GT/src/bcg.cuh:120); the encoder uses
 *     sqrt(double) (LE/chunk_compress/compressor.hpp:62, GT/...:50), which is exact below
 *     2^31, so the integer root is what the files were written with;
 *   - reads past the last word see zero padding (the reference reads one word past the end,
 *     LE/src/bcg.cuh:66-72 with no padding at LE/src/grass.cu:1046-1050);
 *   - arithmetic is 64-bit where the reference uses int32 (identical for ids < 2^31).
 */
#define _GNU_SOURCE
#include "grass_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct or_graph {
    uint32_t *w;        /* big-endian word view of .graph, zero padded */
    uint64_t nwords;    /* words holding file bytes (rounded up) */
    uint64_t nbits;     /* 8 * .graph bytes */
    uint64_t *start;    /* n + 1 record bit starts */
    uint64_t n;
    int fmt;
    uint32_t cp_interval; /* GT byte model: the device's checkpoint interval (0 = none) */
};

/* ---------------------------------------------------------------- bit stream primitives */

/* LE/src/grass.cu:1036-1050: bytes packed big-endian into 32-bit words, last word zero-filled. */
void or_pack_words(const uint8_t *bytes, uint64_t nbytes, uint32_t *words) {
    uint64_t nw = (nbytes + 3) / 4;
    for (uint64_t k = 0; k < nw + 4; ++k) words[k] = 0;
    for (uint64_t i = 0; i < nbytes; ++i)
        words[i >> 2] |= (uint32_t)bytes[i] << (24 - 8 * (i & 3));
}

/* 32-bit window starting at bit b (LE/src/bcg.cuh:66-72, __funnelshift_l of two words). */
static inline uint32_t peek32(const uint32_t *w, uint64_t b) {
    uint64_t k = b >> 5;
    unsigned s = (unsigned)(b & 31);
    uint64_t x = ((uint64_t)w[k] << 32) | w[k + 1];
    return (uint32_t)(x >> (32 - s));
}

/* nbits in [0, 32] MSB-first (LE/src/bcg.cuh:82-86 "decode"). */
static inline uint32_t readb(const uint32_t *w, uint64_t b, unsigned nbits) {
    if (nbits == 0) return 0;
    return peek32(w, b) >> (32 - nbits);
}

/* Elias-gamma of x + 1 (LE/src/bcg.cuh:74-98; encoder LE/chunk_compress/compressor.hpp:411-417).
 * Returns 0 on an all-zero window, which the reference's 32-bit clz cannot decode. */
static inline int gamma_read(const uint32_t *w, uint64_t *b, uint64_t *val) {
    uint32_t x = peek32(w, *b);
    if (x == 0) return 0;
    unsigned L = (unsigned)__builtin_clz(x);
    *b += L;
    *val = (uint64_t)readb(w, *b, L + 1) - 1;
    *b += L + 1;
    return 1;
}

static inline uint64_t isqrt64(uint64_t x) {
    uint64_t r = 0;
    for (int bit = 31; bit >= 0; --bit) {
        uint64_t t = r | (1ull << bit);
        if (t * t <= x) r = t;
    }
    return r;
}

/* ------------------------------------------------------------------------- offsets */

/* LE/src/grass.cu:1063-1088 (ubl, n) and LE/src/bcg.cuh:284-329 (one field per vertex). */
int or_parse_offsets(const uint8_t *off, uint64_t nbytes, uint32_t *ubl_out, uint64_t *n_formula,
                     uint64_t *lens) {
    if (!off || nbytes < 1) return OR_EFORMAT;
    uint32_t ubl = off[0];
    if (ubl == 0 || ubl > 64) return OR_EFORMAT;
    uint64_t n = ((nbytes - 1) * 8) / ubl;
    if (ubl_out) *ubl_out = ubl;
    if (n_formula) *n_formula = n;
    if (!lens) return OR_OK;
    uint64_t fb = nbytes - 1;
    uint32_t *w = (uint32_t *)malloc(((fb + 3) / 4 + 4) * sizeof(uint32_t));
    if (!w) return OR_ENOMEM;
    or_pack_words(off + 1, fb, w);
    for (uint64_t u = 0; u < n; ++u) {
        uint64_t b = u * ubl;
        uint64_t v = 0;
        unsigned left = ubl;
        while (left) {
            unsigned take = left > 32 ? 32 : left;
            v = (v << take) | readb(w, b, take);
            b += take;
            left -= take;
        }
        lens[u] = v;
    }
    free(w);
    return OR_OK;
}

int or_open(const uint8_t *graph, uint64_t graph_bytes, const uint8_t *offset,
            uint64_t offset_bytes, int fmt, or_graph **out) {
    if (!out || (fmt != OR_LE && fmt != OR_GT)) return OR_EINVAL;
    *out = NULL;
    uint32_t ubl;
    uint64_t nf;
    int rc = or_parse_offsets(offset, offset_bytes, &ubl, &nf, NULL);
    if (rc) return rc;
    uint64_t *lens = (uint64_t *)malloc((nf + 1) * sizeof(uint64_t));
    if (!lens) return OR_ENOMEM;
    rc = or_parse_offsets(offset, offset_bytes, &ubl, &nf, lens);
    if (rc) { free(lens); return rc; }
    uint64_t n = nf;
    while (n > 0 && lens[n - 1] == 0) --n;            /* padding fields (zero length) */
    if (n == 0 || n >= (1ull << 31)) { free(lens); return OR_EFORMAT; }
    or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
    if (!g) { free(lens); return OR_ENOMEM; }
    g->n = n;
    g->fmt = fmt;
    g->cp_interval = 4;
    g->nbits = graph_bytes * 8;
    g->nwords = (graph_bytes + 3) / 4;
    g->w = (uint32_t *)malloc((g->nwords + 4) * sizeof(uint32_t));
    g->start = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    if (!g->w || !g->start) { free(lens); or_close(g); return OR_ENOMEM; }
    or_pack_words(graph, graph_bytes, g->w);
    /* exclusive scan (LE/src/utils.hpp:337-342); a zero-length record is malformed */
    uint64_t acc = 0;
    for (uint64_t v = 0; v < n; ++v) {
        if (lens[v] == 0) { free(lens); or_close(g); return OR_EFORMAT; }
        g->start[v] = acc;
        acc += lens[v];
    }
    g->start[n] = acc;
    free(lens);
    if (acc > g->nbits || g->nbits - acc >= 8) { or_close(g); return OR_EFORMAT; }
    *out = g;
    return OR_OK;
}

void or_set_gt_checkpoint(or_graph *g, uint32_t interval) {
    if (g) g->cp_interval = interval;
}

void or_close(or_graph *g) {
    if (!g) return;
    free(g->w);
    free(g->start);
    free(g);
}

uint64_t or_num_vertices(const or_graph *g) { return g->n; }
uint64_t or_total_bits(const or_graph *g) { return g->start[g->n]; }
uint64_t or_record_start(const or_graph *g, uint64_t v) { return g->start[v]; }

/* ------------------------------------------------------------------------- headers */

typedef struct {
    uint64_t v, d, f1, f2, fp, C, q, R, payload, hbits;
} hdr_t;

/* LE header: LE/src/bcg.cuh:131-152.  GT header: GT/src/bcg.cuh:110-136. */
static int parse_header(const or_graph *g, uint64_t v, hdr_t *h) {
    uint64_t b = g->start[v], b0 = b;
    uint64_t end = g->start[v + 1];
    memset(h, 0, sizeof(*h));
    h->v = v;
    if (!gamma_read(g->w, &b, &h->d)) return OR_EFORMAT;
    if (g->fmt == OR_LE) {
        if (h->d >= 1 && !gamma_read(g->w, &b, &h->f1)) return OR_EFORMAT;
        if (h->d >= 2) {
            if (!gamma_read(g->w, &b, &h->f2)) return OR_EFORMAT;
            if (!gamma_read(g->w, &b, &h->fp)) return OR_EFORMAT;
            uint64_t D = h->d - 1, s = isqrt64(D), r = D - s * s;
            h->C = s + (r > s);
            h->q = D / h->C;
            h->R = D - h->C * h->q;
            if (h->f1 == 0 || h->f1 > 32 || h->f2 > 32 || h->fp > h->C + 1) return OR_EFORMAT;
        }
        if (h->d == 1 && h->f1 >= g->n) return OR_EFORMAT;
    } else {
        if (!gamma_read(g->w, &b, &h->f1)) return OR_EFORMAT;
        if (!gamma_read(g->w, &b, &h->f2)) return OR_EFORMAT;
        if (!gamma_read(g->w, &b, &h->fp)) return OR_EFORMAT;
        uint64_t s = isqrt64(h->d);
        if (s * s == h->d) {
            h->C = h->q = s;
            h->R = 0;
        } else {
            h->C = s + 1;
            h->q = h->d / h->C;
            h->R = h->d - h->C * h->q;
        }
        if (h->f1 > 32 || h->f2 > 32 || h->fp > h->C) return OR_EFORMAT;
        if (h->d > 0 && h->f1 == 0) return OR_EFORMAT;
    }
    h->payload = b;
    h->hbits = b - b0;
    if (b > end) return OR_EFORMAT;
    return OR_OK;
}

int or_header(const or_graph *g, uint64_t v, int64_t out[9]) {
    if (!g || v >= g->n) return OR_EINVAL;
    hdr_t h;
    int rc = parse_header(g, v, &h);
    if (rc) return rc;
    out[0] = (int64_t)h.d; out[1] = (int64_t)h.f1; out[2] = (int64_t)h.f2;
    out[3] = (int64_t)h.fp; out[4] = (int64_t)h.C; out[5] = (int64_t)h.q;
    out[6] = (int64_t)h.R; out[7] = (int64_t)h.payload; out[8] = (int64_t)h.hbits;
    return OR_OK;
}

int64_t or_degree(const or_graph *g, uint64_t v) {
    if (!g || v >= g->n) return -1;
    hdr_t h;
    if (parse_header(g, v, &h)) return -1;
    return (int64_t)h.d;
}

/* LE/src/bcg.cuh:196-261 (get_vertex). */
static int64_t le_get(const or_graph *g, const hdr_t *h, uint64_t k, uint64_t *fbits) {
    if (h->d == 0 || k >= h->d) return -1;
    if (h->d == 1) { if (fbits) *fbits = 0; return (int64_t)h->f1; }
    const uint64_t rb = h->f1, sb = h->f2, q = h->q, R = h->R;
    uint64_t len = q + 1, c, i;
    if (k >= R * (q + 1)) {
        k -= R * (q + 1);
        len = q;
        c = k / len + R;
        i = k % len;
    } else {
        c = k / len;
        i = k % len;
    }
    int64_t v = (int64_t)h->v;
    int64_t r1 = readb(g->w, h->payload + c * rb, (unsigned)rb);
    r1 = c < h->fp ? v - r1 : v + r1;
    if (i == 0) { if (fbits) *fbits = rb; return r1; }
    int64_t r2 = readb(g->w, h->payload + (c + 1) * rb, (unsigned)rb);
    r2 = c + 1 < h->fp ? v - r2 : v + r2;
    uint64_t so = h->payload + (h->C + 1) * rb + ((q - 1) * c + (c < R ? c : R) + i - 1) * (sb + 1);
    int64_t sea = readb(g->w, so, (unsigned)sb);
    int neg = (int)readb(g->w, so + sb, 1);
    if (fbits) *fbits = 2 * rb + sb + 1;
    return ((r2 - r1) / (int64_t)len) * (int64_t)i + r1 + (neg ? -sea : sea);
}

/* GT/src/bcg.cuh:139-168 (decode_head, decode_body, get_vertex). */
static int64_t gt_get(const or_graph *g, const hdr_t *h, uint64_t k, uint64_t *fbits) {
    if (h->d == 0 || k >= h->d) return -1;
    const uint64_t hb = h->f1, bb = h->f2, q = h->q, R = h->R;
    uint64_t blk, i;
    if (k < (q + 1) * R) {
        blk = k / (q + 1);
        i = k % (q + 1);
    } else {
        k -= (q + 1) * R;
        blk = k / q + R;
        i = k % q;
    }
    uint64_t b = h->payload + blk * (bb * q - bb + hb) + (blk < R ? blk : R) * bb;
    int64_t head = readb(g->w, b, (unsigned)hb);
    b += hb;
    int64_t val = blk >= h->fp ? (int64_t)h->v + head + 1 : (int64_t)h->v - head - 1;
    for (uint64_t t = 0; t < i; ++t, b += bb) val += (int64_t)readb(g->w, b, (unsigned)bb) + 1;
    /* Bits the device reads: it keeps the absolute id at every P-th in-block position (a 32-bit
     * checkpoint built at load, P = cp_interval), so element i costs that checkpoint plus i mod P
     * gaps once i >= P, and the head plus i gaps before. */
    const uint64_t P = g->cp_interval;
    if (fbits) *fbits = (P && i >= P) ? 32 + (i % P) * bb : hb + i * bb;
    return val;
}

static inline int64_t get_h(const or_graph *g, const hdr_t *h, uint64_t k, uint64_t *fbits) {
    return g->fmt == OR_LE ? le_get(g, h, k, fbits) : gt_get(g, h, k, fbits);
}

int64_t or_get(const or_graph *g, uint64_t v, uint64_t k) {
    if (!g || v >= g->n) return -1;
    hdr_t h;
    if (parse_header(g, v, &h)) return -1;
    return get_h(g, &h, k, NULL);
}

int64_t or_payload_bits(const or_graph *g, uint64_t v, uint64_t k) {
    if (!g || v >= g->n) return -1;
    hdr_t h;
    if (parse_header(g, v, &h)) return -1;
    uint64_t fb = 0;
    if (get_h(g, &h, k, &fb) < 0) return -1;
    return (int64_t)fb;
}

/* Whole-list decode.  LE uses random access per element; GT walks blocks sequentially
 * (head then bodies, GT/src/bcg.cuh:139-151), which visits each field once. */
static void decode_list(const or_graph *g, const hdr_t *h, uint32_t *out) {
    if (g->fmt == OR_LE) {
        for (uint64_t k = 0; k < h->d; ++k) out[k] = (uint32_t)le_get(g, h, k, NULL);
        return;
    }
    const uint64_t hb = h->f1, bb = h->f2;
    uint64_t b = h->payload, k = 0;
    for (uint64_t blk = 0; blk < h->C; ++blk) {
        uint64_t len = h->q + (blk < h->R);
        int64_t head = readb(g->w, b, (unsigned)hb);
        b += hb;
        int64_t val = blk >= h->fp ? (int64_t)h->v + head + 1 : (int64_t)h->v - head - 1;
        out[k++] = (uint32_t)val;
        for (uint64_t t = 1; t < len; ++t, b += bb) {
            val += (int64_t)readb(g->w, b, (unsigned)bb) + 1;
            out[k++] = (uint32_t)val;
        }
    }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int or_decode_csr(const or_graph *g, uint64_t *rowptr, uint32_t *col, int nthreads) {
    if (!g || !rowptr) return OR_EINVAL;
    set_threads(nthreads);
    const int64_t n = (int64_t)g->n;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t v = 0; v < n; ++v) {
        hdr_t h;
        if (parse_header(g, (uint64_t)v, &h)) { bad = 1; rowptr[v + 1] = 0; continue; }
        rowptr[v + 1] = h.d;
    }
    if (bad) return OR_EFORMAT;
    rowptr[0] = 0;
    for (int64_t v = 0; v < n; ++v) rowptr[v + 1] += rowptr[v];
    if (!col) return OR_OK;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; ++v) {
        hdr_t h;
        parse_header(g, (uint64_t)v, &h);
        decode_list(g, &h, col + rowptr[v]);
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------- Philox */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* Philox4x32-10 as published by Salmon, Moraes, Dror and Shaw (SC'11): ten rounds of
 * two 32x32->64 multiplies, Weyl key schedule between rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

static inline void philox_seed(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                               uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

uint64_t or_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                    uint64_t bound) {
    uint32_t r[4];
    philox_seed(seed, c0, c1, c2, c3, r);
    return mulhi64(((uint64_t)r[1] << 32) | r[0], bound);
}

#define TAG_START 0xFFFFFFFFu
#define TAG_MDRW_STEP 0xFFFFFFFEu

/* ------------------------------------------------------------------------- random walk */

static inline uint64_t ceil8(uint64_t bits) { return (bits + 7) >> 3; }

/* The reference's DeepWalk app: one uniform neighbour per step
 * (LE/src/apps/randomwalk/randomWalks.cu:62-79), a dead end ends the walk
 * (randomWalks.cu:66-68, padded with the INVALID_VERTEX = n sentinel of LE/src/grass.cu:1182). */
int or_rw(const or_graph *g, uint64_t seed, uint32_t instance, const uint32_t *starts,
          uint64_t W, uint32_t L, uint32_t *out, int nthreads, uint64_t *bytes) {
    if (!g || !out) return OR_EINVAL;
    const uint64_t n = g->n;
    if (starts)
        for (uint64_t w = 0; w < W; ++w)
            if (starts[w] >= n) return OR_EINVAL;
    set_threads(nthreads);
    uint64_t total = 0;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : total) reduction(| : bad)
    for (int64_t wi = 0; wi < (int64_t)W; ++wi) {
        uint64_t w = (uint64_t)wi;
        uint32_t *row = out + w * (uint64_t)(L + 1);
        uint64_t v = starts ? starts[w] : or_uniform(seed, instance, (uint32_t)w, TAG_START, 0, n);
        row[0] = (uint32_t)v;
        uint64_t acc = 4 + (starts ? 4 : 0);
        uint32_t s = 0;
        for (; s < L; ++s) {
            hdr_t h;
            if (parse_header(g, v, &h)) { bad = 1; break; }
            acc += 8 + ceil8(h.hbits);
            if (h.d == 0) break;
            uint64_t k = or_uniform(seed, instance, (uint32_t)w, s, 0, h.d);
            uint64_t fb = 0;
            int64_t nv = get_h(g, &h, k, &fb);
            if (nv < 0 || (uint64_t)nv >= n) { bad = 1; break; }
            acc += ceil8(fb) + 4;
            v = (uint64_t)nv;
            row[s + 1] = (uint32_t)v;
        }
        for (; s < L; ++s) row[s + 1] = (uint32_t)n;
        total += acc;
    }
    if (bytes) *bytes = total;
    return bad ? OR_EFORMAT : OR_OK;
}

/* ------------------------------------------------------------------------- k-hop */

typedef struct {
    uint32_t *keys;
    uint64_t cap;     /* power of two */
    uint64_t count;
} hset_t;

static int hset_init(hset_t *s, uint64_t expect) {
    uint64_t cap = 16;
    while (cap < 2 * expect + 2) cap <<= 1;
    s->keys = (uint32_t *)malloc(cap * sizeof(uint32_t));
    if (!s->keys) return 0;
    memset(s->keys, 0xFF, cap * sizeof(uint32_t));
    s->cap = cap;
    s->count = 0;
    return 1;
}

/* Inserts x (x != 0xFFFFFFFF); returns 1 when x was not yet present. */
static int hset_insert(hset_t *s, uint32_t x) {
    uint64_t m = s->cap - 1, i = ((uint64_t)x * 0x9E3779B97F4A7C15ull) >> 40 & m;
    for (;;) {
        if (s->keys[i] == x) return 0;
        if (s->keys[i] == 0xFFFFFFFFu) { s->keys[i] = x; s->count++; return 1; }
        i = (i + 1) & m;
    }
}

struct or_khop_result {
    uint32_t R;
    uint64_t *edge_off;   /* R + 1 */
    uint64_t *hop_count;  /* R * OR_MAX_HOPS */
    uint32_t *src, *dst;
    uint64_t bytes;
};

typedef struct {
    uint32_t *src, *dst;
    uint64_t len, cap;
    uint64_t hop_count[OR_MAX_HOPS];
    uint64_t bytes;
    int err;
} khop_buf;

static int kb_push(khop_buf *b, uint32_t s, uint32_t d) {
    if (b->len == b->cap) {
        uint64_t nc = b->cap ? 2 * b->cap : 1024;
        uint32_t *ns = (uint32_t *)realloc(b->src, nc * sizeof(uint32_t));
        if (!ns) return 0;
        b->src = ns;
        uint32_t *nd = (uint32_t *)realloc(b->dst, nc * sizeof(uint32_t));
        if (!nd) return 0;
        b->dst = nd;
        b->cap = nc;
    }
    b->src[b->len] = s;
    b->dst[b->len] = d;
    b->len++;
    return 1;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* Floyd's algorithm (Bentley & Floyd, CACM 1987) for a uniform c-subset of [0, d):
 * for j = d-c .. d-1 draw t in [0, j]; take t unless already taken, else take j.
 * Draw number jj = j - (d - c) uses philox(seed, {inst, vertex, hop, jj}).  The subset is
 * returned sorted.  c == d takes every position without drawing. */
static void floyd_subset(uint64_t seed, uint32_t inst, uint32_t vtx, uint32_t hop, uint64_t d,
                         uint64_t c, uint64_t *sel) {
    if (c == d) {
        for (uint64_t j = 0; j < d; ++j) sel[j] = j;
        return;
    }
    uint64_t cnt = 0;
    for (uint64_t j = d - c; j < d; ++j) {
        uint64_t t = or_uniform(seed, inst, vtx, hop, (uint32_t)(j - (d - c)), j + 1);
        int dup = 0;
        for (uint64_t a = 0; a < cnt; ++a)
            if (sel[a] == t) { dup = 1; break; }
        sel[cnt++] = dup ? j : t;
    }
    qsort(sel, cnt, sizeof(uint64_t), cmp_u64);
}

static void khop_one(const or_graph *g, uint64_t seed, const uint32_t *seeds, uint64_t ns,
                     uint32_t inst, uint32_t hops, const uint32_t *fan, khop_buf *out) {
    /* frontier 0 = seeds, first appearance order */
    uint32_t *front = (uint32_t *)malloc((ns + 1) * sizeof(uint32_t));
    uint64_t nf = 0;
    hset_t hs;
    if (!front || !hset_init(&hs, ns)) { out->err = OR_ENOMEM; free(front); return; }
    for (uint64_t i = 0; i < ns; ++i) {
        if (seeds[i] >= g->n) { out->err = OR_EINVAL; break; }
        if (hset_insert(&hs, seeds[i])) front[nf++] = seeds[i];
    }
    free(hs.keys);
    uint64_t sel[64];
    for (uint32_t h = 0; h < hops && !out->err; ++h) {
        uint64_t e0 = out->len;
        uint64_t c_max = fan[h];
        if (c_max > 64) c_max = 64;
        for (uint64_t i = 0; i < nf; ++i) {
            hdr_t hd;
            uint32_t f = front[i];
            if (parse_header(g, f, &hd)) { out->err = OR_EFORMAT; break; }
            out->bytes += 8 + ceil8(hd.hbits) + 4;
            uint64_t c = hd.d < c_max ? hd.d : c_max;
            if (c == 0) continue;
            floyd_subset(seed, inst, f, h, hd.d, c, sel);
            for (uint64_t a = 0; a < c; ++a) {
                uint64_t fb = 0;
                int64_t x = get_h(g, &hd, sel[a], &fb);
                if (x < 0 || (uint64_t)x >= g->n) { out->err = OR_EFORMAT; break; }
                out->bytes += ceil8(fb) + 8;
                if (!kb_push(out, f, (uint32_t)x)) { out->err = OR_ENOMEM; break; }
            }
        }
        out->hop_count[h] = out->len - e0;
        if (h + 1 == hops || out->err) break;
        /* next frontier: destinations of this hop, deduplicated, first appearance order */
        uint64_t ne = out->len - e0;
        uint32_t *nfront = (uint32_t *)malloc((ne + 1) * sizeof(uint32_t));
        if (!nfront || !hset_init(&hs, ne)) { out->err = OR_ENOMEM; free(nfront); break; }
        uint64_t nn = 0;
        for (uint64_t e = e0; e < out->len; ++e)
            if (hset_insert(&hs, out->dst[e])) nfront[nn++] = out->dst[e];
        free(hs.keys);
        free(front);
        front = nfront;
        nf = nn;
    }
    free(front);
}

int or_khop(const or_graph *g, uint64_t seed, uint32_t R, const uint64_t *seed_off,
            const uint32_t *seeds, const uint32_t *inst, const uint32_t *hops,
            const uint32_t *fanout, int nthreads, or_khop_result **out) {
    if (!g || !out || !seed_off || !inst || !hops || !fanout) return OR_EINVAL;
    for (uint32_t r = 0; r < R; ++r) {
        if (hops[r] > OR_MAX_HOPS || seed_off[r + 1] < seed_off[r]) return OR_EINVAL;
        for (uint32_t h = 0; h < hops[r]; ++h)
            if (fanout[(uint64_t)r * OR_MAX_HOPS + h] > 64) return OR_EINVAL;
    }
    set_threads(nthreads);
    khop_buf *bufs = (khop_buf *)calloc(R ? R : 1, sizeof(khop_buf));
    if (!bufs) return OR_ENOMEM;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r = 0; r < (int64_t)R; ++r)
        khop_one(g, seed, seeds + seed_off[r], seed_off[r + 1] - seed_off[r], inst[r], hops[r],
                 fanout + (uint64_t)r * OR_MAX_HOPS, &bufs[r]);
    int err = 0;
    or_khop_result *res = (or_khop_result *)calloc(1, sizeof(or_khop_result));
    if (!res) err = OR_ENOMEM;
    uint64_t total = 0;
    for (uint32_t r = 0; r < R; ++r) {
        if (bufs[r].err && !err) err = bufs[r].err;
        total += bufs[r].len;
    }
    if (!err) {
        res->R = R;
        res->edge_off = (uint64_t *)calloc(R + 1, sizeof(uint64_t));
        res->hop_count = (uint64_t *)calloc((uint64_t)R * OR_MAX_HOPS + 1, sizeof(uint64_t));
        res->src = (uint32_t *)malloc((total + 1) * sizeof(uint32_t));
        res->dst = (uint32_t *)malloc((total + 1) * sizeof(uint32_t));
        if (!res->edge_off || !res->hop_count || !res->src || !res->dst) err = OR_ENOMEM;
    }
    if (!err) {
        uint64_t off = 0;
        for (uint32_t r = 0; r < R; ++r) {
            res->edge_off[r] = off;
            if (bufs[r].len) {
                memcpy(res->src + off, bufs[r].src, bufs[r].len * sizeof(uint32_t));
                memcpy(res->dst + off, bufs[r].dst, bufs[r].len * sizeof(uint32_t));
            }
            memcpy(res->hop_count + (uint64_t)r * OR_MAX_HOPS, bufs[r].hop_count,
                   sizeof(bufs[r].hop_count));
            off += bufs[r].len;
            res->bytes += bufs[r].bytes;
        }
        res->edge_off[R] = off;
    }
    for (uint32_t r = 0; r < R; ++r) { free(bufs[r].src); free(bufs[r].dst); }
    free(bufs);
    if (err) { or_khop_result_free(res); return err; }
    *out = res;
    return OR_OK;
}

void or_khop_result_sizes(const or_khop_result *res, uint64_t *edge_off, uint64_t *hop_count,
                          uint64_t *total, uint64_t *bytes) {
    if (edge_off) memcpy(edge_off, res->edge_off, (res->R + 1) * sizeof(uint64_t));
    if (hop_count) memcpy(hop_count, res->hop_count, (uint64_t)res->R * OR_MAX_HOPS * sizeof(uint64_t));
    if (total) *total = res->edge_off[res->R];
    if (bytes) *bytes = res->bytes;
}

void or_khop_result_edges(const or_khop_result *res, uint32_t *src, uint32_t *dst) {
    uint64_t t = res->edge_off[res->R];
    if (src) memcpy(src, res->src, t * sizeof(uint32_t));
    if (dst) memcpy(dst, res->dst, t * sizeof(uint32_t));
}

void or_khop_result_free(or_khop_result *res) {
    if (!res) return;
    free(res->edge_off);
    free(res->hop_count);
    free(res->src);
    free(res->dst);
    free(res);
}

/* ------------------------------------------------------------------------- MDRW */

struct or_mdrw_result {
    uint32_t I;
    uint64_t *step_off, *vert_off, *ind_off;
    uint32_t *step_src, *step_dst, *verts, *ind_src, *ind_dst;
    uint64_t bytes;
};

typedef struct {
    uint32_t *ssrc, *sdst;
    uint64_t nsteps;
    uint32_t *verts;
    uint64_t nverts;
    khop_buf ind;
    uint64_t bytes;
    int err;
} mdrw_buf;

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* Frontier sampling (Ribeiro & Towsley, IMC 2010): the reference's multi-RW app keeps a set
 * of roots and moves one per step (LE/src/apps/multiRW/multiRW.cu:22-45,74-97); here the
 * moving root is drawn proportionally to degree, as its comment at multiRW.cu:77 intends. */
static void mdrw_one(const or_graph *g, uint64_t seed, uint32_t inst, uint32_t m, uint32_t target,
                     mdrw_buf *out) {
    const uint64_t n = g->n;
    uint64_t steps = target > m ? target - m : 0;
    uint32_t *F = (uint32_t *)malloc((m + 1) * sizeof(uint32_t));
    uint64_t *wt = (uint64_t *)malloc((m + 1) * sizeof(uint64_t));
    hdr_t *H = (hdr_t *)malloc((m + 1) * sizeof(hdr_t));
    out->ssrc = (uint32_t *)malloc((steps + 1) * sizeof(uint32_t));
    out->sdst = (uint32_t *)malloc((steps + 1) * sizeof(uint32_t));
    uint32_t *all = (uint32_t *)malloc((m + steps + 1) * sizeof(uint32_t));
    if (!F || !wt || !H || !out->ssrc || !out->sdst || !all) { out->err = OR_ENOMEM; goto done; }
    uint64_t W = 0, na = 0;
    for (uint32_t j = 0; j < m; ++j) {
        F[j] = (uint32_t)or_uniform(seed, inst, j, TAG_START, 0, n);
        if (parse_header(g, F[j], &H[j])) { out->err = OR_EFORMAT; goto done; }
        wt[j] = H[j].d;
        W += wt[j];
        all[na++] = F[j];
        out->bytes += 8 + ceil8(H[j].hbits) + 4;
    }
    uint64_t t = 0;
    for (; t < steps && W > 0; ++t) {
        uint32_t r[4];
        philox_seed(seed, inst, (uint32_t)t, TAG_MDRW_STEP, 0, r);
        uint64_t x = mulhi64(((uint64_t)r[1] << 32) | r[0], W);
        uint32_t j = 0;
        uint64_t acc = 0;
        for (; j < m; ++j) {
            acc += wt[j];
            if (acc > x) break;
        }
        uint64_t k = mulhi64(((uint64_t)r[3] << 32) | r[2], H[j].d);
        uint64_t fb = 0;
        int64_t v = get_h(g, &H[j], k, &fb);
        if (v < 0 || (uint64_t)v >= n) { out->err = OR_EFORMAT; goto done; }
        out->ssrc[t] = F[j];
        out->sdst[t] = (uint32_t)v;
        out->bytes += ceil8(fb) + 8;
        W -= wt[j];
        F[j] = (uint32_t)v;
        if (parse_header(g, (uint64_t)v, &H[j])) { out->err = OR_EFORMAT; goto done; }
        out->bytes += 8 + ceil8(H[j].hbits);
        wt[j] = H[j].d;
        W += wt[j];
        all[na++] = (uint32_t)v;
    }
    out->nsteps = t;
    qsort(all, na, sizeof(uint32_t), cmp_u32);
    uint64_t nu = 0;
    for (uint64_t i = 0; i < na; ++i)
        if (nu == 0 || all[nu - 1] != all[i]) all[nu++] = all[i];
    out->verts = all;
    out->nverts = nu;
    all = NULL;
    /* induced edges: every adjacency entry (x, y) with x and y in the sample */
    for (uint64_t i = 0; i < nu && !out->err; ++i) {
        hdr_t hd;
        uint32_t x = out->verts[i];
        if (parse_header(g, x, &hd)) { out->err = OR_EFORMAT; break; }
        uint32_t *lst = (uint32_t *)malloc((hd.d + 1) * sizeof(uint32_t));
        if (!lst) { out->err = OR_ENOMEM; break; }
        decode_list(g, &hd, lst);
        for (uint64_t a = 0; a < hd.d; ++a) {
            if (bsearch(&lst[a], out->verts, nu, sizeof(uint32_t), cmp_u32))
                if (!kb_push(&out->ind, x, lst[a])) { out->err = OR_ENOMEM; break; }
        }
        out->bytes += 8 + ceil8(g->start[x + 1] - g->start[x]);
        free(lst);
    }
    out->bytes += out->ind.len * 8 + nu * 4;
done:
    free(F);
    free(wt);
    free(H);
    free(all);
}

int or_mdrw(const or_graph *g, uint64_t seed, uint32_t I, const uint32_t *inst,
            const uint32_t *fsize, uint32_t target, int nthreads, or_mdrw_result **out) {
    if (!g || !out || !inst || !fsize) return OR_EINVAL;
    for (uint32_t i = 0; i < I; ++i)
        if (fsize[i] == 0) return OR_EINVAL;
    set_threads(nthreads);
    mdrw_buf *bufs = (mdrw_buf *)calloc(I ? I : 1, sizeof(mdrw_buf));
    if (!bufs) return OR_ENOMEM;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)I; ++i) mdrw_one(g, seed, inst[i], fsize[i], target, &bufs[i]);
    int err = 0;
    uint64_t ts = 0, tv = 0, ti = 0;
    for (uint32_t i = 0; i < I; ++i) {
        if (bufs[i].err && !err) err = bufs[i].err;
        ts += bufs[i].nsteps;
        tv += bufs[i].nverts;
        ti += bufs[i].ind.len;
    }
    or_mdrw_result *res = NULL;
    if (!err) {
        res = (or_mdrw_result *)calloc(1, sizeof(or_mdrw_result));
        if (!res) err = OR_ENOMEM;
    }
    if (!err) {
        res->I = I;
        res->step_off = (uint64_t *)calloc(I + 1, sizeof(uint64_t));
        res->vert_off = (uint64_t *)calloc(I + 1, sizeof(uint64_t));
        res->ind_off = (uint64_t *)calloc(I + 1, sizeof(uint64_t));
        res->step_src = (uint32_t *)malloc((ts + 1) * sizeof(uint32_t));
        res->step_dst = (uint32_t *)malloc((ts + 1) * sizeof(uint32_t));
        res->verts = (uint32_t *)malloc((tv + 1) * sizeof(uint32_t));
        res->ind_src = (uint32_t *)malloc((ti + 1) * sizeof(uint32_t));
        res->ind_dst = (uint32_t *)malloc((ti + 1) * sizeof(uint32_t));
        if (!res->step_off || !res->vert_off || !res->ind_off || !res->step_src ||
            !res->step_dst || !res->verts || !res->ind_src || !res->ind_dst)
            err = OR_ENOMEM;
    }
    if (!err) {
        uint64_t so = 0, vo = 0, io = 0;
        for (uint32_t i = 0; i < I; ++i) {
            mdrw_buf *b = &bufs[i];
            res->step_off[i] = so;
            res->vert_off[i] = vo;
            res->ind_off[i] = io;
            if (b->nsteps) {
                memcpy(res->step_src + so, b->ssrc, b->nsteps * sizeof(uint32_t));
                memcpy(res->step_dst + so, b->sdst, b->nsteps * sizeof(uint32_t));
            }
            if (b->nverts) memcpy(res->verts + vo, b->verts, b->nverts * sizeof(uint32_t));
            if (b->ind.len) {
                memcpy(res->ind_src + io, b->ind.src, b->ind.len * sizeof(uint32_t));
                memcpy(res->ind_dst + io, b->ind.dst, b->ind.len * sizeof(uint32_t));
            }
            so += b->nsteps;
            vo += b->nverts;
            io += b->ind.len;
            res->bytes += b->bytes;
        }
        res->step_off[I] = so;
        res->vert_off[I] = vo;
        res->ind_off[I] = io;
    }
    for (uint32_t i = 0; i < I; ++i) {
        free(bufs[i].ssrc);
        free(bufs[i].sdst);
        free(bufs[i].verts);
        free(bufs[i].ind.src);
        free(bufs[i].ind.dst);
    }
    free(bufs);
    if (err) { or_mdrw_result_free(res); return err; }
    *out = res;
    return OR_OK;
}

void or_mdrw_result_sizes(const or_mdrw_result *res, uint64_t *step_off, uint64_t *vert_off,
                          uint64_t *ind_off, uint64_t *bytes) {
    if (step_off) memcpy(step_off, res->step_off, (res->I + 1) * sizeof(uint64_t));
    if (vert_off) memcpy(vert_off, res->vert_off, (res->I + 1) * sizeof(uint64_t));
    if (ind_off) memcpy(ind_off, res->ind_off, (res->I + 1) * sizeof(uint64_t));
    if (bytes) *bytes = res->bytes;
}

void or_mdrw_result_data(const or_mdrw_result *res, uint32_t *step_src, uint32_t *step_dst,
                         uint32_t *verts, uint32_t *ind_src, uint32_t *ind_dst) {
    uint32_t I = res->I;
    if (step_src) memcpy(step_src, res->step_src, res->step_off[I] * sizeof(uint32_t));
    if (step_dst) memcpy(step_dst, res->step_dst, res->step_off[I] * sizeof(uint32_t));
    if (verts) memcpy(verts, res->verts, res->vert_off[I] * sizeof(uint32_t));
    if (ind_src) memcpy(ind_src, res->ind_src, res->ind_off[I] * sizeof(uint32_t));
    if (ind_dst) memcpy(ind_dst, res->ind_dst, res->ind_off[I] * sizeof(uint32_t));
}

void or_mdrw_result_free(or_mdrw_result *res) {
    if (!res) return;
    free(res->step_off);
    free(res->vert_off);
    free(res->ind_off);
    free(res->step_src);
    free(res->step_dst);
    free(res->verts);
    free(res->ind_src);
    free(res->ind_dst);
    free(res);
}
