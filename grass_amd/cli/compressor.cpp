// compressor: the reference's chunk_compress tool (LE/chunk_compress/compress.cpp,
// GT/chunk_compress/compress.cpp) with the same command line, `<input_path> <output_path>
// [<encode_type>]`, writing <output_path>.graph and <output_path>.offset byte-identical to it.
// The reference builds one binary per codec; here `--format le|gt` (default le) selects it, and
// `--device N` encodes on GPU N (the same bytes).
// Only the "normal" encoding exists (the GT "head_bit" mode writes files its own decoder cannot
// read).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/grass.h"
#include "../../include/grass_ingest.h"

int main(int argc, char* argv[]) {
    int fmt = GRASS_FMT_LE;
    int device = -1;
    const char* pos[3] = {nullptr, nullptr, nullptr};
    int np = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--format") && i + 1 < argc) {
            const char* f = argv[++i];
            if (!strcmp(f, "le")) fmt = GRASS_FMT_LE;
            else if (!strcmp(f, "gt")) fmt = GRASS_FMT_GT;
            else np = 99;
        } else if (!strcmp(argv[i], "--device") && i + 1 < argc) {
            device = atoi(argv[++i]);
        } else if (np < 3) {
            pos[np++] = argv[i];
        } else {
            np = 99;
        }
    }
    if (np != 2 && np != 3) {
        printf("incorrect arguments.\n");
        printf("<input_path> <output_path> [<encode_type>]   (--format le|gt, --device N)\n");
        return 1;
    }
    if (np == 3 && strcmp(pos[2], "normal") != 0) {
        fprintf(stderr, "encode type %s is not supported (only \"normal\")\n", pos[2]);
        return 1;
    }
    grass_edges* e = nullptr;
    int rc = grass_read_edge_list(pos[0], 0, 0, &e);
    if (rc) {
        fprintf(stderr, "%s: %s\n", pos[0], grass_strerror(rc));
        return 1;
    }
    const uint32_t *src, *dst;
    uint64_t m;
    grass_edges_data(e, &src, &dst, &m);
    grass_csr* csr = nullptr;
    rc = grass_csr_from_sorted_edges(src, dst, m, &csr);
    grass_edges_free(e);
    if (rc) {
        fprintf(stderr, "%s: %s (the edge list must be sorted by (u, v))\n", pos[0], grass_strerror(rc));
        return 1;
    }
    printf("%s graph loaded.\n", pos[0]);
    grass_blob *gb = nullptr, *ob = nullptr;
    rc = device >= 0 ? grass_compress_on_device(csr, fmt, device, &gb, &ob)
                     : grass_compress(csr, fmt, 0, &gb, &ob);
    if (!rc) rc = grass_write_files(pos[1], gb, ob);
    grass_blob_free(gb);
    grass_blob_free(ob);
    grass_csr_free(csr);
    if (rc) {
        fprintf(stderr, "compressor: %s\n", grass_strerror(rc));
        return 1;
    }
    printf("CGR generation completed.\n");
    return 0;
}
