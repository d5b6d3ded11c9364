// grass_app.hpp -- the App plugin interface and its driver, over the C ABI of grass.h.
//
// The reference's samplers are compile-time Apps (RandomWalkApp / DeepWalkApp,
// LE/src/apps/randomwalk/randomWalks.cu:3-79; KHopApp, LE/src/apps/khop/khop.cu:1-69;
// MultiRWApp, LE/src/apps/multiRW/multiRW.cu:13-97) run by nextdoor<SampleType, App>(graph_file,
// nruns) (LE/src/grass.cu:1924-1992) from appMain (LE/src/main.cu:5-40).  This header keeps the
// same names and members with the same meaning -- steps(), stepSize(k), numSamples(n),
// initialSample(i, n), initialSampleSize(), hasExplicitTransits() -- so an App written for the
// reference ports by copying those members.  What differs:
//   * The per-step `next` / `stepTransits` device functions are the library's own kernels, picked
//     by the App's `kind` (uniform walk step; k-hop draw without replacement; degree-proportional
//     frontier move); every draw comes from Philox keyed by (seed, instance, walker, step).
//   * Samples are returned (the reference discards them unless built with GET_SAMPLE).
//   * nextdoor() prints the reference's timing lines (LE/src/grass.cu:1865,1984-1987) and then
//     its own throughput line.
// Host-only C++17; link with libgrass_amd.so.
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <string>
#include <vector>

#include "grass.h"

namespace grass_app {

enum class Kind { Walk, KHop, MultiRW };

// ------------------------------------------------------------------ the reference's Apps

// LE/src/apps/randomwalk/randomWalks.cu:3-61: 80 steps of one neighbour, one sample per vertex
// starting at i % n.
struct RandomWalkApp {
    static constexpr Kind kind = Kind::Walk;
    int steps() const { return 80; }
    int stepSize(int) const { return 1; }
    uint64_t numSamples(uint64_t n_nodes) const { return n_nodes; }
    std::vector<uint32_t> initialSample(uint64_t sampleIdx, uint64_t n_nodes) const {
        return {(uint32_t)(sampleIdx % n_nodes)};
    }
    int initialSampleSize() const { return 1; }
    bool hasExplicitTransits() const { return false; }
};
using DeepWalkApp = RandomWalkApp;  // LE/src/apps/randomwalk/main.cu:5

// LE/src/apps/khop/khop.cu:1-69: two hops of 10 then 25 neighbours, one sample per vertex.  Here
// the neighbours of one vertex are distinct (drawn without replacement) and each hop's frontier
// is deduplicated.
struct KHopApp {
    static constexpr Kind kind = Kind::KHop;
    int steps() const { return 2; }
    int stepSize(int k) const { return k == 0 ? 10 : 25; }
    uint64_t numSamples(uint64_t n_nodes) const { return n_nodes; }
    std::vector<uint32_t> initialSample(uint64_t sampleIdx, uint64_t n_nodes) const {
        return {(uint32_t)(sampleIdx % n_nodes)};
    }
    int initialSampleSize() const { return 1; }
    bool hasExplicitTransits() const { return false; }
};

// LE/src/apps/multiRW/multiRW.cu:1-97: 80 roots per sample, 100 steps, one sample per vertex.
// Here the roots are Philox-drawn, the moving root is chosen proportionally to degree, and the
// sample is returned with its induced edges.
struct MultiRWApp {
    static constexpr Kind kind = Kind::MultiRW;
    static constexpr uint32_t kNumRootVertices = 80;  // NUM_ROOT_VERTICES, multiRW.cu:3
    int steps() const { return 100; }
    int stepSize(int) const { return 1; }
    uint64_t numSamples(uint64_t n_nodes) const { return n_nodes; }
    int initialSampleSize() const { return 1; }
    bool hasExplicitTransits() const { return true; }
};

// ------------------------------------------------------------------ driver

struct RunOptions {
    int format = GRASS_FMT_LE;
    int device = 0;
    uint64_t seed = 42;
    uint64_t samples = 0;   // 0: App::numSamples(n)
    uint64_t batch = 1u << 16;  // samples per library call
    bool json = false;
    bool quiet = false;     // no reference timing lines
};

struct RunReport {
    int status = GRASS_OK;
    uint64_t n = 0, samples = 0, sampled_edges = 0;
    double load_s = 0, run_s = 0;  // run_s: all nruns
    int nruns = 0;
};

namespace detail {

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DevBuf {
    grass_graph* g = nullptr;
    void* p = nullptr;
    ~DevBuf() {
        if (g && p) grass_device_free(g, p);
    }
    int alloc(grass_graph* graph, uint64_t bytes) {
        g = graph;
        return grass_device_alloc(graph, bytes, &p);
    }
};

// One pass of the App over samples [s0, s1); adds the sampled edges.
template <class App>
int run_batch(grass_graph* g, const App& app, uint64_t n, uint64_t s0, uint64_t s1, uint64_t seed,
              uint64_t* edges);

template <>
inline int run_batch<RandomWalkApp>(grass_graph* g, const RandomWalkApp& app, uint64_t n, uint64_t s0,
                                    uint64_t s1, uint64_t seed, uint64_t* edges) {
    const uint64_t W = s1 - s0;
    const uint32_t L = (uint32_t)app.steps();
    std::vector<uint32_t> starts(W);
    for (uint64_t i = 0; i < W; ++i) starts[i] = app.initialSample(s0 + i, n)[0];
    DevBuf st, out;
    int rc = st.alloc(g, W * 4);
    if (!rc) rc = out.alloc(g, W * (L + 1ull) * 4);
    if (!rc) rc = grass_copy_to_device(g, st.p, starts.data(), W * 4);
    if (!rc) rc = grass_random_walk(g, (const uint32_t*)st.p, W, L, seed, (uint32_t)s0, (uint32_t*)out.p, nullptr);
    if (!rc) rc = grass_sync(g, nullptr);
    if (!rc) *edges += W * L;
    return rc;
}

template <>
inline int run_batch<KHopApp>(grass_graph* g, const KHopApp& app, uint64_t n, uint64_t s0, uint64_t s1,
                              uint64_t seed, uint64_t* edges) {
    const uint64_t R = s1 - s0;
    std::vector<uint32_t> seeds(R);
    for (uint64_t i = 0; i < R; ++i) seeds[i] = app.initialSample(s0 + i, n)[0];
    DevBuf sd;
    int rc = sd.alloc(g, R * 4);
    if (!rc) rc = grass_copy_to_device(g, sd.p, seeds.data(), R * 4);
    if (rc) return rc;
    std::vector<grass_khop_request> reqs(R);
    for (uint64_t i = 0; i < R; ++i) {
        grass_khop_request& q = reqs[i];
        q = grass_khop_request{};
        q.seeds_dev = (const uint32_t*)sd.p + i;
        q.num_seeds = 1;
        q.instance = (uint32_t)(s0 + i);
        q.hops = (uint32_t)app.steps();
        for (int h = 0; h < app.steps(); ++h) q.fanout[h] = (uint32_t)app.stepSize(h);
    }
    uint64_t cap = 0;
    rc = grass_khop_bound(g, reqs.data(), (uint32_t)R, &cap);
    DevBuf src, dst, off;
    if (!rc) rc = src.alloc(g, cap * 4);
    if (!rc) rc = dst.alloc(g, cap * 4);
    if (!rc) rc = off.alloc(g, (R + 1) * GRASS_KHOP_MAX_HOPS * 8);
    if (rc) return rc;
    grass_khop_output o{};
    o.src_dev = (uint32_t*)src.p;
    o.dst_dev = (uint32_t*)dst.p;
    o.capacity = cap;
    o.offsets_dev = (uint64_t*)off.p;
    rc = grass_khop_batch(g, reqs.data(), (uint32_t)R, seed, &o, nullptr);
    if (!rc) rc = grass_sync(g, nullptr);
    if (!rc) *edges += o.total;
    return rc;
}

template <>
inline int run_batch<MultiRWApp>(grass_graph* g, const MultiRWApp& app, uint64_t, uint64_t s0, uint64_t s1,
                                 uint64_t seed, uint64_t* edges) {
    const uint64_t I = s1 - s0;
    std::vector<grass_mdrw_request> reqs(I);
    for (uint64_t i = 0; i < I; ++i) reqs[i] = grass_mdrw_request{(uint32_t)(s0 + i), MultiRWApp::kNumRootVertices};
    grass_mdrw_result* res = nullptr;
    int rc = grass_mdrw(g, reqs.data(), (uint32_t)I, MultiRWApp::kNumRootVertices + (uint32_t)app.steps(), seed,
                        nullptr, &res);
    if (!rc) {
        grass_mdrw_output o;
        rc = grass_mdrw_output_get(res, &o);
        if (!rc) *edges += o.steps;
    }
    grass_mdrw_free(res);
    return rc;
}

}  // namespace detail

// Runs App over numSamples(n) samples, nruns times, in batches of opt.batch samples.  Returns
// false when the graph cannot be loaded (the reference's nextdoor returns false the same way,
// LE/src/grass.cu:1928-1930).
template <class App>
bool nextdoor(const char* graph_file, int nruns, const RunOptions& opt = RunOptions(), RunReport* report = nullptr) {
    RunReport rep;
    rep.nruns = nruns;
    const double t_load0 = detail::now_s();
    grass_graph* g = nullptr;
    rep.status = grass_open(graph_file, opt.format, opt.device, &g);
    if (rep.status != GRASS_OK) {
        fprintf(stderr, "Open graph file failed! (%s)\n", grass_strerror(rep.status));
        if (report) *report = rep;
        return false;
    }
    rep.load_s = detail::now_s() - t_load0;
    grass_graph_info info;
    grass_info(g, &info);
    rep.n = info.n;
    App app;
    rep.samples = opt.samples ? opt.samples : app.numSamples(info.n);
    const uint64_t batch = opt.batch ? opt.batch : rep.samples;
    const double t0 = detail::now_s();
    for (int r = 0; r < nruns && rep.status == GRASS_OK; ++r) {
        const double tr = detail::now_s();
        for (uint64_t s0 = 0; s0 < rep.samples && rep.status == GRASS_OK; s0 += batch) {
            const uint64_t s1 = s0 + batch < rep.samples ? s0 + batch : rep.samples;
            rep.status = detail::run_batch<App>(g, app, info.n, s0, s1, opt.seed + (uint64_t)r, &rep.sampled_edges);
        }
        if (!opt.quiet && !opt.json)
            printf("Transit Parallel: End to end time %g secs\n", detail::now_s() - tr);
    }
    rep.run_s = detail::now_s() - t0;
    grass_close(g);
    if (rep.status != GRASS_OK) {
        fprintf(stderr, "sampling failed: %s\n", grass_strerror(rep.status));
    } else if (opt.json) {
        printf("{\"samples\": %llu, \"runs\": %d, \"sampled_edges\": %llu, \"load_s\": %.6f, \"run_s\": %.6f, "
               "\"edges_per_s\": %.1f}\n",
               (unsigned long long)rep.samples, nruns, (unsigned long long)rep.sampled_edges, rep.load_s, rep.run_s,
               rep.run_s > 0 ? rep.sampled_edges / rep.run_s : 0.0);
    } else {
        if (!opt.quiet) {
            printf("EndToEndTime: %g\n", rep.load_s + rep.run_s);
            printf("transferEndToEndTime: %g\n", rep.load_s + rep.run_s);
            printf("transferDataTime: %g\n", rep.load_s);
            printf("runSamplingTime: %g\n", rep.run_s);
        }
        printf("samples %llu  sampled edges %llu  %.3f G edges/s\n", (unsigned long long)rep.samples,
               (unsigned long long)rep.sampled_edges, rep.run_s > 0 ? rep.sampled_edges / rep.run_s / 1e9 : 0.0);
    }
    if (report) *report = rep;
    return rep.status == GRASS_OK;
}

}  // namespace grass_app
