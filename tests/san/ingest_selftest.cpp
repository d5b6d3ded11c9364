// Drives the threaded host ingest (grass_amd/csrc/ingest.cpp) for the sanitizer builds of
// `make -C grass_amd sanitize`: R-MAT edges, graph_sort (directed and undirected), both
// compressors, and the edge-list reader and writer, each run with several threads and checked
// against the same call on one thread (the outputs do not depend on the thread count).
//   ingest_selftest <scale> <tmpdir>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/grass.h"
#include "../../include/grass_ingest.h"

static int fail(const char* what) {
    fprintf(stderr, "ingest selftest: %s\n", what);
    return 1;
}

static bool same_csr(const grass_csr* a, const grass_csr* b) {
    uint64_t na, ma, nb, mb;
    const uint64_t *ra, *rb;
    const uint32_t *ca, *cb;
    grass_csr_info(a, &na, &ma);
    grass_csr_info(b, &nb, &mb);
    if (na != nb || ma != mb) return false;
    grass_csr_arrays(a, &ra, &ca);
    grass_csr_arrays(b, &rb, &cb);
    return memcmp(ra, rb, (na + 1) * 8) == 0 && memcmp(ca, cb, ma * 4) == 0;
}

static bool same_blob(const grass_blob* a, const grass_blob* b) {
    const uint8_t *pa, *pb;
    uint64_t la, lb;
    grass_blob_data(a, &pa, &la);
    grass_blob_data(b, &pb, &lb);
    return la == lb && memcmp(pa, pb, la) == 0;
}

int main(int argc, char** argv) {
    if (argc < 3) return fail("usage: ingest_selftest <scale> <tmpdir>");
    const uint32_t scale = (uint32_t)atoi(argv[1]);
    const std::string dir = argv[2];
    const uint64_t m = 16ull << scale;
    std::vector<uint32_t> src(m), dst(m), src1(m), dst1(m);
    if (grass_rmat_edges(scale, 16, 0.57, 0.19, 0.19, 42, src.data(), dst.data(), 4)) return fail("rmat");
    if (grass_rmat_edges(scale, 16, 0.57, 0.19, 0.19, 42, src1.data(), dst1.data(), 1)) return fail("rmat 1");
    if (src != src1 || dst != dst1) return fail("rmat differs with the thread count");
    for (int directed = 0; directed < 2; ++directed) {
        grass_csr *c4 = nullptr, *c1 = nullptr;
        if (grass_graph_sort(src.data(), dst.data(), m, directed, 4, &c4)) return fail("graph_sort");
        if (grass_graph_sort(src.data(), dst.data(), m, directed, 1, &c1)) return fail("graph_sort 1");
        if (!same_csr(c4, c1)) return fail("graph_sort differs with the thread count");
        for (int fmt = 0; fmt < 2; ++fmt) {
            grass_blob *g4, *o4, *g1, *o1;
            if (grass_compress(c4, fmt, 4, &g4, &o4)) return fail("compress");
            if (grass_compress(c4, fmt, 1, &g1, &o1)) return fail("compress 1");
            if (!same_blob(g4, g1) || !same_blob(o4, o1)) return fail("compress differs with the thread count");
            const std::string prefix = dir + "/san_" + std::to_string(directed) + "_" + std::to_string(fmt);
            if (grass_write_files(prefix.c_str(), g4, o4)) return fail("write_files");
            grass_blob_free(g4);
            grass_blob_free(o4);
            grass_blob_free(g1);
            grass_blob_free(o1);
        }
        // the sorted edge list written with threads and read back with threads
        const std::string txt = dir + "/san_edges_" + std::to_string(directed) + ".txt";
        if (grass_write_edge_list(txt.c_str(), c4, 4)) return fail("write_edge_list");
        grass_edges* e = nullptr;
        if (grass_read_edge_list(txt.c_str(), 0, 4, &e)) return fail("read_edge_list");
        const uint32_t *es, *ed;
        uint64_t em;
        grass_edges_data(e, &es, &ed, &em);
        grass_csr* back = nullptr;
        if (grass_csr_from_sorted_edges(es, ed, em, &back)) return fail("csr_from_sorted_edges");
        uint64_t n4, m4, nb, mb;
        grass_csr_info(c4, &n4, &m4);
        grass_csr_info(back, &nb, &mb);
        if (mb != m4) return fail("edge list read back differs");
        const uint64_t *r4, *rb;
        const uint32_t *k4, *kb;
        grass_csr_arrays(c4, &r4, &k4);
        grass_csr_arrays(back, &rb, &kb);
        if (memcmp(k4, kb, m4 * 4) != 0) return fail("edge list columns differ");
        grass_edges_free(e);
        grass_csr_free(back);
        grass_csr_free(c4);
        grass_csr_free(c1);
    }
    printf("ingest selftest ok\n");
    return 0;
}
