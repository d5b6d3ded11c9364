// graph_sort: the reference's preprocessing tool (LE/chunk_compress/graph_sort.cpp:22-82) with the
// same command line and output: drops self loops, numbers vertices by first appearance, for `u`
// (undirected) makes every pair (min, max) and adds its mirror, sorts, removes duplicates, and
// writes "u v" lines.  Runs on the host's threads (grass_graph_sort), or with `--device N` on GPU N
// (grass_graph_sort_on_device; the same output).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/grass.h"
#include "../../include/grass_ingest.h"

int main(int argc, char* argv[]) {
    int device = -1;
    for (int i = 1; i + 1 < argc; ++i)  // `--device N` anywhere; the rest is the reference's command line
        if (!strcmp(argv[i], "--device")) {
            device = atoi(argv[i + 1]);
            for (int j = i; j + 2 < argc; ++j) argv[j] = argv[j + 2];
            argc -= 2;
            break;
        }
    if (argc < 4) {
        printf("incorrect arguments.\n");
        printf("<input_path> <output_path> <u/d> [<skip_line>]\n");
        return 1;
    }
    const int directed = argv[3][0] == 'd';
    const int skip = argc >= 5 ? atoi(argv[4]) : 0;
    printf("Skip %d Lines\n", skip);
    grass_edges* e = nullptr;
    int rc = grass_read_edge_list(argv[1], skip > 0 ? (uint64_t)skip : 0, 0, &e);
    if (rc) {
        fprintf(stderr, "%s: %s\n", argv[1], grass_strerror(rc));
        return 1;
    }
    const uint32_t *src, *dst;
    uint64_t m;
    grass_edges_data(e, &src, &dst, &m);
    grass_csr* csr = nullptr;
    rc = device >= 0 ? grass_graph_sort_on_device(src, dst, m, directed, device, &csr)
                     : grass_graph_sort(src, dst, m, directed, 0, &csr);
    grass_edges_free(e);
    if (!rc) rc = grass_write_edge_list(argv[2], csr, 0);
    if (rc) {
        fprintf(stderr, "graph_sort: %s\n", grass_strerror(rc));
        grass_csr_free(csr);
        return 1;
    }
    uint64_t n, mm;
    grass_csr_info(csr, &n, &mm);
    printf("Graph has %llu vertices, %llu edges.\n", (unsigned long long)n, (unsigned long long)mm);
    grass_csr_free(csr);
    return 0;
}
