/*
 * selftest.c -- TEST INFRASTRUCTURE ONLY.  Drives every function of the oracle once over a
 * compressed graph given as files, for a build under AddressSanitizer + UndefinedBehavior-
 * Sanitizer (make -C oracle sanitize): decode to CSR, random access of every (v, k), walks,
 * k-hop and MDRW batches, and a set of corrupted inputs that must be rejected without any
 * out-of-bounds access.  Exit status 0 when every check passes.
 *
 *   selftest <prefix> <le|gt>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grass_oracle.h"

static uint8_t *slurp(const char *path, uint64_t *n) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    uint8_t *b = (uint8_t *)malloc(sz > 0 ? (size_t)sz : 1);
    if (b && sz > 0 && fread(b, 1, (size_t)sz, f) != (size_t)sz) {
        free(b);
        b = NULL;
    }
    fclose(f);
    *n = (uint64_t)(sz > 0 ? sz : 0);
    return b;
}

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "selftest: check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: selftest <prefix> <le|gt>\n");
        return 2;
    }
    const int fmt = strcmp(argv[2], "gt") == 0 ? OR_GT : OR_LE;
    char path[4096];
    uint64_t gn = 0, on = 0;
    snprintf(path, sizeof(path), "%s.graph", argv[1]);
    uint8_t *gb = slurp(path, &gn);
    snprintf(path, sizeof(path), "%s.offset", argv[1]);
    uint8_t *ob = slurp(path, &on);
    CHECK(gb && ob);

    or_graph *g = NULL;
    CHECK(or_open(gb, gn, ob, on, fmt, &g) == OR_OK);
    const uint64_t n = or_num_vertices(g);
    uint64_t *rowptr = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    CHECK(or_decode_csr(g, rowptr, NULL, 2) == OR_OK);
    const uint64_t m = rowptr[n];
    uint32_t *col = (uint32_t *)malloc((m + 1) * sizeof(uint32_t));
    CHECK(or_decode_csr(g, rowptr, col, 2) == OR_OK);
    for (uint64_t v = 0; v < n; ++v)
        for (uint64_t k = rowptr[v]; k < rowptr[v + 1]; ++k)
            CHECK(or_get(g, v, k - rowptr[v]) == (int64_t)col[k]);

    const uint64_t W = 500;
    const uint32_t L = 20;
    uint32_t *walks = (uint32_t *)malloc(W * (L + 1) * sizeof(uint32_t));
    uint64_t bytes = 0;
    CHECK(or_rw(g, 7, 1, NULL, W, L, walks, 2, &bytes) == OR_OK);

    uint32_t seeds[64], inst[3] = {1, 2, 3}, hops[3] = {2, 3, 1};
    uint64_t soff[4] = {0, 10, 40, 64};
    uint32_t fan[3 * OR_MAX_HOPS] = {0};
    fan[0] = 25, fan[1] = 10;
    fan[OR_MAX_HOPS] = 10, fan[OR_MAX_HOPS + 1] = 10, fan[OR_MAX_HOPS + 2] = 5;
    fan[2 * OR_MAX_HOPS] = 64;
    for (int i = 0; i < 64; ++i) seeds[i] = (uint32_t)((i * 2654435761u) % (n ? n : 1));
    or_khop_result *kr = NULL;
    CHECK(or_khop(g, 5, 3, soff, seeds, inst, hops, fan, 2, &kr) == OR_OK);
    uint64_t eoff[4], hc[3 * OR_MAX_HOPS], tot = 0, kb = 0;
    or_khop_result_sizes(kr, eoff, hc, &tot, &kb);
    uint32_t *ks = (uint32_t *)malloc((tot + 1) * 4), *kd = (uint32_t *)malloc((tot + 1) * 4);
    or_khop_result_edges(kr, ks, kd);
    or_khop_result_free(kr);

    uint32_t fs[3] = {1, 40, 300}, mi[3] = {4, 5, 6};
    or_mdrw_result *mr = NULL;
    CHECK(or_mdrw(g, 9, 3, mi, fs, 400, 2, &mr) == OR_OK);
    uint64_t so[4], vo[4], io[4], mb = 0;
    or_mdrw_result_sizes(mr, so, vo, io, &mb);
    uint32_t *a = (uint32_t *)malloc((so[3] + 1) * 4), *b = (uint32_t *)malloc((so[3] + 1) * 4);
    uint32_t *vv = (uint32_t *)malloc((vo[3] + 1) * 4);
    uint32_t *is = (uint32_t *)malloc((io[3] + 1) * 4), *id = (uint32_t *)malloc((io[3] + 1) * 4);
    or_mdrw_result_data(mr, a, b, vv, is, id);
    or_mdrw_result_free(mr);
    or_close(g);

    /* corrupted inputs: every one is rejected (or decodes) without touching memory it must not */
    const uint64_t cuts[4] = {gn / 2, gn > 1 ? gn - 1 : 0, 1, 0};
    for (int i = 0; i < 4; ++i) {
        or_graph *h = NULL;
        if (or_open(gb, cuts[i], ob, on, fmt, &h) == OR_OK) or_close(h);
    }
    uint8_t *ob2 = (uint8_t *)malloc(on ? on : 1);
    memcpy(ob2, ob, on);
    if (on) ob2[0] = 0; /* ubl = 0 */
    or_graph *h = NULL;
    CHECK(or_open(gb, gn, ob2, on, fmt, &h) != OR_OK);
    if (on) ob2[0] = 70; /* ubl > 64 */
    CHECK(or_open(gb, gn, ob2, on, fmt, &h) != OR_OK);
    uint8_t *gb2 = (uint8_t *)malloc(gn ? gn : 1);
    memset(gb2, 0, gn);
    if (or_open(gb2, gn, ob, on, fmt, &h) == OR_OK) or_close(h);

    free(gb2), free(ob2), free(a), free(b), free(vv), free(is), free(id), free(ks), free(kd);
    free(walks), free(col), free(rowptr), free(gb), free(ob);
    printf("selftest ok: n=%llu m=%llu walks=%llu khop=%llu mdrw steps=%llu\n", (unsigned long long)n,
           (unsigned long long)m, (unsigned long long)W, (unsigned long long)tot, (unsigned long long)so[3]);
    return 0;
}
