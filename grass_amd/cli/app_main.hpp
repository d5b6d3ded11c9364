// Command-line front end shared by the rw / ks / mrw tools.
//
// Same options and behaviour as the reference's appMain (LE/src/main.cu:5-40): `-g <cgc_path>`
// names the <cgc_path>.graph / <cgc_path>.offset pair, `-n <runs>` repeats the sampling (default
// 1); with no option at all, or without -g, the usage text is printed and the exit status is 0.
// Added options: --format le|gt (the reference picks the codec by which binary is built),
// --seed, --samples (alias --walkers; default: App::numSamples(n)), --batch (samples per library
// call), --device, --json (one JSON line instead of the timing lines), --quiet, --out FILE (the
// samples, in FileSink's format).
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/grass_app.hpp"

namespace grass_app {

inline void print_usage(const char* prog) {
    printf("usage: \n\n");
    printf("-h --help              Prints this help\n");
    printf("-g --graph-file        File containing graph\n");
    printf("-n --nruns             Number of runs\n");
    printf("   --format le|gt      Codec of the compressed graph (default le)\n");
    printf("   --seed S            Philox seed (default 42)\n");
    printf("   --samples N         Samples per run (alias --walkers; default one per vertex)\n");
    printf("   --batch B           Samples per library call (default 65536)\n");
    printf("   --device D          HIP device ordinal (default 0)\n");
    printf("   --json              Print one JSON line\n");
    printf("   --quiet             Omit the reference's timing lines\n");
    printf("   --out FILE          Write the samples to FILE (binary, one record per sample)\n");
    (void)prog;
}

template <class App>
int appMain(int argc, char* argv[]) {
    RunOptions opt;
    const char* graph = nullptr;
    const char* out_path = nullptr;
    int nruns = 1;
    if (argc < 2) {
        print_usage(argv[0]);
        return 0;
    }
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto value = [&](const char* name) -> const char* {
            if (i + 1 >= argc) {
                fprintf(stderr, "%s needs a value\n", name);
                exit(2);
            }
            return argv[++i];
        };
        if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
            print_usage(argv[0]);
            return 0;
        } else if (!strcmp(a, "-g") || !strcmp(a, "--graph-file")) {
            graph = value(a);
        } else if (!strcmp(a, "-n") || !strcmp(a, "--nruns")) {
            nruns = atoi(value(a));
        } else if (!strcmp(a, "--format")) {
            const char* f = value(a);
            if (!strcmp(f, "le")) opt.format = GRASS_FMT_LE;
            else if (!strcmp(f, "gt")) opt.format = GRASS_FMT_GT;
            else {
                fprintf(stderr, "--format must be le or gt\n");
                return 2;
            }
        } else if (!strcmp(a, "--seed")) {
            opt.seed = strtoull(value(a), nullptr, 0);
        } else if (!strcmp(a, "--samples") || !strcmp(a, "--walkers")) {
            opt.samples = strtoull(value(a), nullptr, 0);
        } else if (!strcmp(a, "--batch")) {
            opt.batch = strtoull(value(a), nullptr, 0);
        } else if (!strcmp(a, "--device")) {
            opt.device = atoi(value(a));
        } else if (!strcmp(a, "--json")) {
            opt.json = true;
        } else if (!strcmp(a, "--quiet")) {
            opt.quiet = true;
        } else if (!strcmp(a, "--out")) {
            out_path = value(a);
        } else {
            fprintf(stderr, "unknown option %s\n", a);
            print_usage(argv[0]);
            return 2;
        }
    }
    if (!graph) {
        print_usage(argv[0]);
        return 0;
    }
    if (nruns < 1) nruns = 1;
    // like the reference, a graph that cannot be loaded is reported but not an exit failure
    // (LE/src/main.cu:38-40 ignores nextdoor's result); a sampling error is
    std::unique_ptr<FileSink> sink;
    if (out_path) {
        sink.reset(new FileSink(out_path));
        if (!sink->ok()) {
            fprintf(stderr, "cannot write %s\n", out_path);
            return 1;
        }
        opt.sink = sink.get();
    }
    RunReport rep;
    nextdoor<App>(graph, nruns, opt, &rep);
    return (rep.status == GRASS_OK || rep.status == GRASS_EIO) ? 0 : 1;
}

}  // namespace grass_app
