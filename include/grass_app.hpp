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
//   * Samples can be returned: RunOptions::sink receives every batch's samples in host memory
//     (FileSink writes them to a file, the tools' --out).  The reference sizes a per-sample buffer
//     (getFinalSampleSize, LE/src/grass.cu:1115-1126) but never fills it (:219-227, :1228-1229).
//   * nextdoor() prints the reference's timing lines (LE/src/grass.cu:1865,1984-1987) and then
//     its own throughput line.
// Host-only C++17; link with libgrass_amd.so.
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <memory>
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

// ------------------------------------------------------------------ samples out

// An MDRW batch's result in host memory (the arrays of grass_mdrw_fetch; offsets hold count + 1
// entries, indexed by the sample's position in the batch).
struct MdrwHost {
    uint32_t count = 0;
    std::vector<uint64_t> step_off, vert_off, ind_off;
    std::vector<uint32_t> step_src, step_dst, verts, ind_src, ind_dst;
};

// Receives each batch's samples in host memory, in sample order (RunOptions::sink).  A callback
// that returns a status other than GRASS_OK stops the run with it.  Sample i of run r was drawn
// with Philox seed `RunOptions::seed + r`; its keys are those of the library call that ran its
// batch [s0, s1):
//   walks   grass_random_walk over starts initialSample(i) = i % n, walker i - s0 of instance s0;
//   k-hop   one grass_khop_request per sample, seed i % n, instance i, fanouts stepSize(0..);
//   MultiRW one grass_mdrw_request per sample, instance i, 80 roots, target 80 + steps().
struct SampleSink {
    virtual ~SampleSink() {}
    // rows[(i - s0) * (length + 1) + j] = vertex j of sample i's walk (column 0 its start), padded
    // with n after a vertex without neighbours.
    virtual int walks(int run, uint64_t s0, uint64_t s1, uint32_t length, const uint32_t* rows) {
        (void)run, (void)s0, (void)s1, (void)length, (void)rows;
        return GRASS_OK;
    }
    // The edges of sample i at hop h are [off[h * (s1 - s0 + 1) + i - s0], off[... + 1]) of src / dst.
    virtual int khop(int run, uint64_t s0, uint64_t s1, uint32_t hops, const uint64_t* off, const uint32_t* src,
                     const uint32_t* dst) {
        (void)run, (void)s0, (void)s1, (void)hops, (void)off, (void)src, (void)dst;
        return GRASS_OK;
    }
    virtual int mdrw(int run, uint64_t s0, uint64_t s1, const MdrwHost& res) {
        (void)run, (void)s0, (void)s1, (void)res;
        return GRASS_OK;
    }
};

// Writes the samples to a file, one record per sample in sample order, runs one after the other;
// little-endian, no header:
//   walks:   uint32 row[length + 1];
//   k-hop:   per hop: uint64 count, uint32 src[count], uint32 dst[count];
//   MultiRW: uint64 steps, uint32 step_src[steps], uint32 step_dst[steps],
//            uint64 verts, uint32 verts[verts] (ascending),
//            uint64 induced, uint32 ind_src[induced], uint32 ind_dst[induced].
class FileSink : public SampleSink {
  public:
    explicit FileSink(const char* path) : f_(fopen(path, "wb")) {}
    ~FileSink() override {
        if (f_) fclose(f_);
    }
    bool ok() const { return f_ != nullptr; }
    int walks(int, uint64_t s0, uint64_t s1, uint32_t length, const uint32_t* rows) override {
        return put(rows, 4, (s1 - s0) * ((uint64_t)length + 1));
    }
    int khop(int, uint64_t s0, uint64_t s1, uint32_t hops, const uint64_t* off, const uint32_t* src,
             const uint32_t* dst) override {
        const uint64_t R = s1 - s0;
        for (uint64_t i = 0; i < R; ++i)
            for (uint32_t h = 0; h < hops; ++h) {
                const uint64_t a = off[h * (R + 1) + i], b = off[h * (R + 1) + i + 1], c = b - a;
                int rc = put(&c, 8, 1);
                if (!rc) rc = put(src + a, 4, c);
                if (!rc) rc = put(dst + a, 4, c);
                if (rc) return rc;
            }
        return GRASS_OK;
    }
    int mdrw(int, uint64_t, uint64_t, const MdrwHost& r) override {
        for (uint32_t i = 0; i < r.count; ++i) {
            const uint64_t s = r.step_off[i], sc = r.step_off[i + 1] - s;
            const uint64_t v = r.vert_off[i], vc = r.vert_off[i + 1] - v;
            const uint64_t e = r.ind_off[i], ec = r.ind_off[i + 1] - e;
            int rc = put(&sc, 8, 1);
            if (!rc) rc = put(r.step_src.data() + s, 4, sc);
            if (!rc) rc = put(r.step_dst.data() + s, 4, sc);
            if (!rc) rc = put(&vc, 8, 1);
            if (!rc) rc = put(r.verts.data() + v, 4, vc);
            if (!rc) rc = put(&ec, 8, 1);
            if (!rc) rc = put(r.ind_src.data() + e, 4, ec);
            if (!rc) rc = put(r.ind_dst.data() + e, 4, ec);
            if (rc) return rc;
        }
        return GRASS_OK;
    }

  private:
    int put(const void* p, size_t size, uint64_t count) {
        if (!count) return GRASS_OK;
        return f_ && fwrite(p, size, count, f_) == count ? GRASS_OK : GRASS_EIO;
    }
    FILE* f_;
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
    SampleSink* sink = nullptr;  // receives the samples (nullptr: they stay on the device and go)
};

struct RunReport {
    int status = GRASS_OK;
    uint64_t n = 0, samples = 0, sampled_edges = 0;
    double load_s = 0, run_s = 0;  // run_s: all nruns, sampling only
    double out_s = 0;              // copies to host memory and the sink's callbacks
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

// Copies the batch's samples to host memory and hands them to the sink (timed as out_s).
struct Out {
    SampleSink* sink;
    int run;
    double* out_s;
};

// One pass of the App over samples [s0, s1); adds the sampled edges.
template <class App>
int run_batch(grass_graph* g, const App& app, uint64_t n, uint64_t s0, uint64_t s1, uint64_t seed,
              uint64_t* edges, const Out& out);

template <>
inline int run_batch<RandomWalkApp>(grass_graph* g, const RandomWalkApp& app, uint64_t n, uint64_t s0,
                                    uint64_t s1, uint64_t seed, uint64_t* edges, const Out& o) {
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
    if (!rc && o.sink) {
        const double t = now_s();
        std::vector<uint32_t> rows(W * (L + 1ull));
        rc = grass_copy_to_host(g, rows.data(), out.p, rows.size() * 4);
        if (!rc) rc = o.sink->walks(o.run, s0, s1, L, rows.data());
        *o.out_s += now_s() - t;
    }
    return rc;
}

template <>
inline int run_batch<KHopApp>(grass_graph* g, const KHopApp& app, uint64_t n, uint64_t s0, uint64_t s1,
                              uint64_t seed, uint64_t* edges, const Out& ob) {
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
    if (!rc && ob.sink) {
        const double t = now_s();
        std::vector<uint64_t> hoff((uint64_t)o.max_hops * (R + 1));
        std::vector<uint32_t> hs(o.total), hd(o.total);
        rc = grass_copy_to_host(g, hoff.data(), off.p, hoff.size() * 8);
        if (!rc) rc = grass_copy_to_host(g, hs.data(), src.p, o.total * 4);
        if (!rc) rc = grass_copy_to_host(g, hd.data(), dst.p, o.total * 4);
        if (!rc) rc = ob.sink->khop(ob.run, s0, s1, o.max_hops, hoff.data(), hs.data(), hd.data());
        *ob.out_s += now_s() - t;
    }
    return rc;
}

template <>
inline int run_batch<MultiRWApp>(grass_graph* g, const MultiRWApp& app, uint64_t, uint64_t s0, uint64_t s1,
                                 uint64_t seed, uint64_t* edges, const Out& ob) {
    const uint64_t I = s1 - s0;
    std::vector<grass_mdrw_request> reqs(I);
    for (uint64_t i = 0; i < I; ++i) reqs[i] = grass_mdrw_request{(uint32_t)(s0 + i), MultiRWApp::kNumRootVertices};
    grass_mdrw_result* res = nullptr;
    int rc = grass_mdrw(g, reqs.data(), (uint32_t)I, MultiRWApp::kNumRootVertices + (uint32_t)app.steps(), seed,
                        nullptr, &res);
    grass_mdrw_output o{};
    if (!rc) rc = grass_mdrw_output_get(res, &o);
    if (!rc) *edges += o.steps;
    if (!rc && ob.sink) {
        const double t = now_s();
        MdrwHost h;
        h.count = o.count;
        h.step_off.resize(o.count + 1ull);
        h.vert_off.resize(o.count + 1ull);
        h.ind_off.resize(o.count + 1ull);
        h.step_src.resize(o.steps);
        h.step_dst.resize(o.steps);
        h.verts.resize(o.verts);
        h.ind_src.resize(o.induced);
        h.ind_dst.resize(o.induced);
        rc = grass_mdrw_fetch(res, h.step_off.data(), h.step_src.data(), h.step_dst.data(), h.vert_off.data(),
                              h.verts.data(), h.ind_off.data(), h.ind_src.data(), h.ind_dst.data());
        if (!rc) rc = ob.sink->mdrw(ob.run, s0, s1, h);
        *ob.out_s += now_s() - t;
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
        const double tr = detail::now_s(), out0 = rep.out_s;
        for (uint64_t s0 = 0; s0 < rep.samples && rep.status == GRASS_OK; s0 += batch) {
            const uint64_t s1 = s0 + batch < rep.samples ? s0 + batch : rep.samples;
            rep.status = detail::run_batch<App>(g, app, info.n, s0, s1, opt.seed + (uint64_t)r, &rep.sampled_edges,
                                                detail::Out{opt.sink, r, &rep.out_s});
        }
        if (!opt.quiet && !opt.json)
            printf("Transit Parallel: End to end time %g secs\n", detail::now_s() - tr - (rep.out_s - out0));
    }
    rep.run_s = detail::now_s() - t0 - rep.out_s;
    grass_close(g);
    if (rep.status != GRASS_OK) {
        fprintf(stderr, "sampling failed: %s\n", grass_strerror(rep.status));
    } else if (opt.json) {
        printf("{\"samples\": %llu, \"runs\": %d, \"sampled_edges\": %llu, \"load_s\": %.6f, \"run_s\": %.6f, "
               "\"out_s\": %.6f, \"edges_per_s\": %.1f}\n",
               (unsigned long long)rep.samples, nruns, (unsigned long long)rep.sampled_edges, rep.load_s, rep.run_s,
               rep.out_s, rep.run_s > 0 ? rep.sampled_edges / rep.run_s : 0.0);
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
