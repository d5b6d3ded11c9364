// Captured batches: a fixed set of random-walk and k-hop requests recorded once into a HIP graph
// and replayed as one submission.
//
// A steady-state sampling loop (the same request shapes every step, new seed vertices written
// into the same device buffers) pays the ~30 kernel launches of a 2-hop pipeline plus the walk
// launch every step.  A plan runs the batch once eagerly (which sizes every workspace, so that
// nothing allocates while capturing), then captures it with the walks and the k-hop pipeline on
// two branches forked from the launch stream and joined back into it, and instantiates the graph.  Each replay starts with a clear of the k-hop
// dedup table (the captured pipeline uses the table's first generations), so replays are
// independent of each other and of any eager call.
#include <hip/hip_runtime.h>
#include <string.h>

#include <new>
#include <vector>

#include "runtime.h"

using namespace grass;

struct grass_plan {
    grass_graph* g = nullptr;
    std::vector<grass_rw_request> rw;
    std::vector<grass_khop_request> kh;
    grass_khop_output kout{};
    uint64_t seed = 0;
    Workspace ws_rw, ws_kh;
    hipStream_t s_main = nullptr, s_side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    MdrwCapture* md = nullptr;  // an MDRW batch (grass_plan_create_mdrw)

    ~grass_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (s_main) (void)hipStreamSynchronize(s_main);
        if (s_side) (void)hipStreamSynchronize(s_side);
        for (hipEvent_t e : {fork, join})
            if (e) (void)hipEventDestroy(e);
        if (s_main) (void)hipStreamDestroy(s_main);
        if (s_side) (void)hipStreamDestroy(s_side);
        ws_rw.release();
        ws_kh.release();
        if (md) mdrw_capture_free(md);
    }

    // The batch: walks on s_main, the k-hop pipeline on s_side forked from and joined back into
    // s_main, so that in the graph the two run side by side.
    int enqueue() {
        if (md) return mdrw_capture_enqueue(g, md, s_main);
        if (!rw.empty()) {
            int rc = rw_batch_enqueue(g, rw.data(), (uint32_t)rw.size(), seed, s_main, ws_rw, nullptr);
            if (rc) return rc;
        }
        if (!kh.empty()) {
            if (hipEventRecord(fork, s_main) != hipSuccess || hipStreamWaitEvent(s_side, fork, 0) != hipSuccess)
                return GRASS_EDEVICE;
            ws_kh.ht_gen = 0xFFFF;  // start every run with a cleared table
            grass_khop_output o = kout;
            int rc = khop_enqueue(g, kh.data(), (uint32_t)kh.size(), seed, &o, s_side, ws_kh);
            if (rc) return rc;
            if (hipEventRecord(join, s_side) != hipSuccess || hipStreamWaitEvent(s_main, join, 0) != hipSuccess)
                return GRASS_EDEVICE;
        }
        return GRASS_OK;
    }
};

namespace {

// Captures p->enqueue() on s_main into p->graph and instantiates it.
int capture(grass_plan* p) {
    if (hipStreamBeginCapture(p->s_main, hipStreamCaptureModeRelaxed) != hipSuccess) return GRASS_EDEVICE;
    int erc = p->enqueue();
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(p->s_main, &gr);
    p->graph = gr;
    if (erc) return erc;
    if (e != hipSuccess || !gr) return GRASS_EDEVICE;
    if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) return GRASS_EDEVICE;
    return GRASS_OK;
}

int make_streams(grass_plan* p) {
    if (hipStreamCreateWithFlags(&p->s_main, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&p->s_side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->join, hipEventDisableTiming) != hipSuccess)
        return GRASS_EDEVICE;
    return GRASS_OK;
}

}  // namespace

extern "C" {

int grass_plan_create_mdrw(grass_graph* g, const grass_mdrw_request* reqs, uint32_t count, uint32_t target,
                           uint64_t seed, grass_plan** out) {
    if (!g || !out || (count && !reqs)) return GRASS_EINVAL;
    *out = nullptr;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) return GRASS_EDEVICE;
    grass_plan* p = new (std::nothrow) grass_plan();
    int rc = p ? GRASS_OK : GRASS_ENOMEM;
    if (!rc) {
        p->g = g;
        p->seed = seed;
        rc = make_streams(p);
    }
    // 1. one eager run (synchronous): validates the requests, sizes the outputs and working memory
    if (!rc) rc = mdrw_capture_prepare(g, reqs, count, target, seed, p->s_main, &p->md);
    // 2. capture the same batch into those buffers
    if (!rc) rc = capture(p);
    (void)hipSetDevice(prev);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return GRASS_OK;
}

int grass_plan_mdrw_output(const grass_plan* p, grass_mdrw_output* out) {
    if (!p || !out || !p->md) return GRASS_EINVAL;
    *out = mdrw_capture_result(p->md)->out;
    return GRASS_OK;
}

int grass_plan_create(grass_graph* g, const grass_rw_request* rw, uint32_t nrw, const grass_khop_request* kh,
                      uint32_t nkh, const grass_khop_output* kout, uint64_t seed, grass_plan** out) {
    if (!g || !out || (nrw && !rw) || (nkh && (!kh || !kout))) return GRASS_EINVAL;
    *out = nullptr;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(g->device)) return GRASS_EDEVICE;
    grass_plan* p = new (std::nothrow) grass_plan();
    int rc = p ? GRASS_OK : GRASS_ENOMEM;
    if (!rc) {
        p->g = g;
        p->rw.assign(rw, rw + nrw);
        p->kh.assign(kh, kh + nkh);
        if (kout) p->kout = *kout;
        p->seed = seed;
        rc = make_streams(p);
    }
    // 1. one eager run: validates the requests and sizes every workspace
    if (!rc) rc = p->enqueue();
    if (!rc && hipStreamSynchronize(p->s_main) != hipSuccess) rc = GRASS_EDEVICE;
    if (!rc && !p->kh.empty() && *p->ws_kh.result_slot() > p->kout.capacity) rc = GRASS_ECAPACITY;
    // 2. capture (nothing allocates now: the workspaces already fit)
    if (!rc) rc = capture(p);
    (void)hipSetDevice(prev);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return GRASS_OK;
}

int grass_plan_launch(grass_plan* p, void* stream) {
    if (!p) return GRASS_EINVAL;
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (!grass::enter_device(p->g->device)) return GRASS_EDEVICE;
    hipStream_t s = stream ? (hipStream_t)stream : p->g->stream;
    hipError_t e = hipGraphLaunch(p->exec, s);
    (void)hipSetDevice(prev);
    return status_of(e);
}

int grass_plan_khop_total(const grass_plan* p, uint64_t* total) {
    if (!p || !total) return GRASS_EINVAL;
    *total = p->kh.empty() ? 0 : *const_cast<grass_plan*>(p)->ws_kh.result_slot();
    return GRASS_OK;
}

int grass_plan_destroy(grass_plan* p) {
    if (!p) return GRASS_OK;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->g->device);
    delete p;
    (void)hipSetDevice(prev);
    return GRASS_OK;
}

}  // extern "C"
