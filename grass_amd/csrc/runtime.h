// Internal host-side types shared by the HIP translation units of libgrass_amd.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>

#include "../../include/grass.h"
#include "codec.h"

namespace grass {

// Device scratch that grows on demand and is reused across calls (no per-call hipMalloc in
// steady state).  Guarded by the owning graph's mutex.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            bytes = 0;
        }
        size_t cap = need + need / 4 + (1u << 20);
        hipError_t e = hipMalloc(&ptr, cap);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Pinned host memory that grows on demand (descriptor staging, results read back).
struct PinnedBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (ptr) {
            (void)hipHostFree(ptr);
            ptr = nullptr;
            bytes = 0;
        }
        size_t cap = need + need / 2 + 4096;
        hipError_t e = hipHostMalloc(&ptr, cap, hipHostMallocDefault);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// An owned device allocation, freed with its owner.
struct DevMem {
    void* p = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes > 256 ? bytes : 256); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    template <typename T>
    T* as() const { return (T*)p; }
};

// Working memory of one in-flight batch: device scratch for descriptors, frontiers and scans,
// a private k-hop dedup table (generation-tagged, so it is reused across batches without
// clearing) and pinned host memory whose first kResultBytes receive the batch's results (the
// k-hop edge total) and whose remainder stages descriptors.  Batches that may be in flight
// together use different workspaces; a workspace is reused only after its batch completed.
struct Workspace {
    static constexpr size_t kResultBytes = 64;
    Scratch scratch;
    Scratch ht;
    uint64_t ht_slots = 0;
    uint32_t ht_gen = 0xFFFF;
    PinnedBuf host;
    // Error word of the kernels this workspace's batches launch (first problem met, kErr*);
    // nullptr: they record into the graph's shared word (the synchronous entry points).
    uint32_t* err = nullptr;
    uint64_t* result_slot() { return (uint64_t*)host.ptr; }
    void* stage() { return (char*)host.ptr + kResultBytes; }
    hipError_t reserve_stage(size_t bytes) { return host.reserve(kResultBytes + bytes); }
    void release() {
        scratch.release();
        ht.release();
        host.release();
        if (err) (void)hipFree(err);
        err = nullptr;
        ht_slots = 0;
        ht_gen = 0xFFFF;
    }
};

// Launch-shape choices that never change a result (grass_tuning; read at every launch).
struct TuneState {
    std::atomic<uint32_t> rw_stage{32};              // walk columns staged per LDS flush: 8, 16, 32
    std::atomic<uint32_t> rw_claim{1};               // 64-walker granules per queue access: 1, 2, 4
    std::atomic<uint32_t> khop_lane_scale{1};        // k-hop lane layout: 0 packed, 1 auto, 2 wide groups
    std::atomic<uint32_t> decode_edges_per_wave{1024};  // full-decode slots per wave
};

// Bump allocator over a Scratch region (256-byte aligned slices).
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* p) : base((char*)p) {}
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base + off);
        off += count * sizeof(T);
        return p;
    }
    static size_t need(size_t bytes) { return ((bytes + 255) & ~(size_t)255) + 256; }
};

// Error word layout in device memory (kernels record the first problem they meet).
enum : uint32_t {
    kErrNone = 0,
    kErrBadNeighbor = 1,  // a decoded id >= n: the graph decodes to garbage
    kErrBadStart = 2,     // a caller-supplied vertex id >= n
};

constexpr uint32_t kWaveStartSlots = 5;  // decode slots per wave: 512, 1024, ..., 8192

}  // namespace grass

struct grass_graph {
    int device = 0;
    int fmt = 0;
    uint64_t n = 0, m = 0, max_degree = 0;
    uint64_t graph_bytes = 0, offset_bytes = 0, total_bits = 0;
    uint32_t ubl = 0;
    uint64_t device_bytes = 0;
    uint32_t* d_words = nullptr;  // .graph words: LE payload; GT: used at load only
    uint64_t* d_start = nullptr;  // record bit starts: used at load only
    uint32_t* d_deg = nullptr;
    uint4* d_hdr = nullptr;  // decoded-header cache, 16 B per vertex
    uint32_t* d_gtx = nullptr;  // GT: checkpointed payload stream the samplers decode
    uint32_t cp_shift = grass::kGtCheckShiftDefault;  // GT checkpoint interval 1 << cp_shift
    uint32_t* d_err = nullptr;  // [0] first error kind, [1] detail
    // full decode: first vertex of every wave, one table per slots-per-wave value 512 << i (GT
    // with P <= 4: of every wave's first run, the waves taking (512 << i) / P runs each)
    uint32_t* d_wstart[grass::kWaveStartSlots] = {};
    uint64_t* d_roff = nullptr;  // full decode, GT with P <= 4: runs before each vertex (n + 1)
    uint64_t dec_runs = 0;       // and their total
    grass::Scratch dec_part;     // full decode: the row-pointer scan's partials (under mu)
    hipEvent_t dec_ev = nullptr;  // recorded after each full decode's scan
    hipStream_t stream = nullptr;
    std::mutex mu;
    grass::Workspace ws;     // working memory of the synchronous entry points (under mu)
    grass::Scratch ws_bytes; // device counter of the byte-counting entry points (under mu)
    grass::TuneState tune;

    grass::DevGraph dev() const {
        grass::DevGraph d;
        d.w = d_gtx ? d_gtx : d_words;
        d.deg = d_deg;
        d.hdr = d_hdr;
        d.n = (uint32_t)n;
        d.fmt = fmt;
        d.cp_shift = cp_shift;
        return d;
    }
};

// An MDRW batch's output.  Its owner and every exported DLPack tensor of it hold a reference;
// the device memory goes with the last one.
struct grass_mdrw_result {
    grass_mdrw_output out{};
    grass::DevMem steps, verts, ind, offs;
    int device = 0;
    std::atomic<int> refs{1};
    void release() {
        if (refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete this;
    }
};

namespace grass {

int status_of(hipError_t e);

// The error word the kernels of a batch on `ws` record into.
inline uint32_t* err_word(grass_graph* g, const Workspace& ws) { return ws.err ? ws.err : g->d_err; }

// Status of a recorded error kind (kErrNone -> GRASS_OK).
inline int status_of_err(uint32_t kind) {
    return kind == kErrNone ? GRASS_OK : kind == kErrBadStart ? GRASS_EINVAL : GRASS_EFORMAT;
}

// Makes `dev` current for an entry point.  HIP keeps the thread's last failed runtime call as its
// "last error", and the launch checks here read that state after each launch; a failure the caller
// made and ignored before entering the library must not be reported as one of this call's
// launches, so the state is cleared on entry.
inline bool enter_device(int dev) {
    if (hipSetDevice(dev) != hipSuccess) return false;
    (void)hipGetLastError();
    return true;
}

// Workgroups that fill the current device once at the kernel's occupancy (persistent grids).
inline unsigned persistent_grid(const void* kernel, int block) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0);
    if (cus <= 0) cus = 1;
    if (per <= 0) per = 1;
    return (unsigned)(cus * per);
}

// loader.hip
struct ByteSource {
    virtual ~ByteSource() {}
    virtual uint64_t size() const = 0;
    // The whole stream in host memory (a mapping or a caller's buffer), or nullptr when it can
    // only be read in order.
    virtual const uint8_t* data() const { return nullptr; }
    // The whole stream in device memory of the graph's device (then data() and read() are unused).
    virtual const uint8_t* device_data() const { return nullptr; }
    // Reads len bytes at the current position; false on a short read.
    virtual bool read(void* dst, uint64_t len) = 0;
};
int load_graph(grass_graph* g, ByteSource& graph_src, ByteSource& offset_src);
// capi.cpp: a graph on `device` from two sources (grass_open and its variants)
int open_sources(ByteSource& graph_src, ByteSource& offset_src, int format, int device, grass_graph** out);
// Images already in device memory.
struct DeviceSource : ByteSource {
    const uint8_t* p;
    uint64_t sz;
    DeviceSource(const void* ptr, uint64_t n) : p((const uint8_t*)ptr), sz(n) {}
    uint64_t size() const override { return sz; }
    const uint8_t* device_data() const override { return p; }
    bool read(void*, uint64_t) override { return false; }
};
// Host threads for copies between pageable and pinned memory (GRASS_LOAD_THREADS, default <= 8).
unsigned load_threads();
void parallel_copy(void* dst, const uint8_t* src, uint64_t len, unsigned threads);

// decode.hip
int decode_csr(grass_graph* g, uint64_t* rowptr, uint32_t* col, hipStream_t stream);

// rw.hip
// bytes != nullptr: write nothing, return the algorithmic byte count of the same walks
int random_walk(grass_graph* g, const uint32_t* starts, uint64_t walkers, uint32_t length,
                uint64_t seed, uint32_t instance, uint32_t* out, hipStream_t stream,
                uint64_t* bytes);
int random_walk_batch(grass_graph* g, const grass_rw_request* reqs, uint32_t count, uint64_t seed,
                      hipStream_t stream, uint64_t* bytes);
// Enqueues a batch on `stream` with `ws` and returns without waiting; `ws` must not be reused
// before the stream has passed the batch.
int rw_batch_enqueue(grass_graph* g, const grass_rw_request* reqs, uint32_t count, uint64_t seed,
                     hipStream_t stream, Workspace& ws, unsigned long long* bytes_dev);
// Sizes a workspace for any walk batch of at most R requests (as khop_reserve).
int rw_reserve(Workspace& ws, uint32_t R);
// grass_random_walk_batch_noalloc: enqueues only, with the library's fixed device array
int random_walk_batch_noalloc(grass_graph* g, const grass_rw_request* reqs, uint32_t count,
                              uint64_t seed, hipStream_t stream);

// khop.hip
int khop_bound(const grass_graph* g, const grass_khop_request* reqs, uint32_t count,
               uint64_t* edges, uint64_t* items, uint64_t* dedup, uint32_t* max_hops);
int khop_batch(grass_graph* g, const grass_khop_request* reqs, uint32_t count, uint64_t seed,
               grass_khop_output* out, hipStream_t stream);
// As rw_batch_enqueue; the batch's edge total reaches *ws.result_slot() (pinned host memory)
// when the stream passes the batch.

int khop_enqueue(grass_graph* g, const grass_khop_request* reqs, uint32_t count, uint64_t seed,
                 grass_khop_output* out, hipStream_t stream, Workspace& ws);
// Sizes a workspace ahead of time so that no later batch of at most R requests whose seed total,
// frontier bound and dedup bound (khop_bound's items / dedup) are at most S, I and D grows it: a
// growth frees and reallocates memory, which waits for the whole device.
int khop_reserve(const grass_graph* g, Workspace& ws, uint32_t R, uint64_t S, uint64_t I, uint64_t D);

// mdrw.hip: an MDRW batch kept for capture.  _prepare runs it once eagerly on `stream` (which
// sizes its outputs and working memory); _enqueue enqueues the same batch again into those
// buffers without allocating or reading anything back (so it can be captured).
struct MdrwCapture;
int mdrw_capture_prepare(grass_graph* g, const grass_mdrw_request* reqs, uint32_t count, uint32_t target,
                         uint64_t seed, hipStream_t stream, MdrwCapture** out);
int mdrw_capture_enqueue(grass_graph* g, MdrwCapture* c, hipStream_t stream);
const grass_mdrw_result* mdrw_capture_result(const MdrwCapture* c);
void mdrw_capture_free(MdrwCapture* c);

}  // namespace grass
