// C ABI entry points of libgrass_amd.so (declared in include/grass.h).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <new>
#include <string>

#include "runtime.h"

namespace grass {

int status_of(hipError_t e) {
    switch (e) {
        case hipSuccess: return GRASS_OK;
        case hipErrorOutOfMemory: return GRASS_ENOMEM;
        case hipErrorInvalidValue: return GRASS_EINVAL;
        default: return GRASS_EDEVICE;
    }
}

namespace {

// A file mapped read-only, so that the loader copies from the page cache with several threads;
// read() (pread) serves files that cannot be mapped.
struct FileSource : ByteSource {
    int fd = -1;
    uint64_t sz = 0, pos = 0;
    const uint8_t* map = nullptr;
    explicit FileSource(const std::string& path) {
        fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
        struct stat st;
        if (fd < 0 || fstat(fd, &st) != 0) return;
        sz = (uint64_t)st.st_size;
        if (sz == 0) return;
        void* p = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p != MAP_FAILED) {
            map = (const uint8_t*)p;
            (void)madvise(p, sz, MADV_SEQUENTIAL);  // advice values are not flags: one call each
            (void)madvise(p, sz, MADV_WILLNEED);
        }
    }
    ~FileSource() override {
        if (map) munmap((void*)map, sz);
        if (fd >= 0) ::close(fd);
    }
    bool ok() const { return fd >= 0; }
    uint64_t size() const override { return sz; }
    const uint8_t* data() const override { return map; }
    bool read(void* dst, uint64_t len) override {
        uint8_t* d = (uint8_t*)dst;
        while (len) {
            const ssize_t got = pread(fd, d, len, (off_t)pos);
            if (got <= 0) return false;
            d += got;
            pos += (uint64_t)got;
            len -= (uint64_t)got;
        }
        return true;
    }
};

struct MemSource : ByteSource {
    const uint8_t* p;
    uint64_t sz, pos = 0;
    MemSource(const void* ptr, uint64_t n) : p((const uint8_t*)ptr), sz(n) {}
    uint64_t size() const override { return sz; }
    const uint8_t* data() const override { return p; }
    bool read(void* dst, uint64_t len) override {
        if (pos + len > sz) return false;
        memcpy(dst, p + pos, len);
        pos += len;
        return true;
    }
};

// Makes `dev` current for the scope of an API call and restores the caller's device.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = grass::enter_device(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void free_graph(grass_graph* g) {
    if (!g) return;
    if (g->d_words) (void)hipFree(g->d_words);
    if (g->d_start) (void)hipFree(g->d_start);
    if (g->d_deg) (void)hipFree(g->d_deg);
    if (g->d_hdr) (void)hipFree(g->d_hdr);
    if (g->d_gtx) (void)hipFree(g->d_gtx);
    for (uint32_t* t : g->d_wstart)
        if (t) (void)hipFree(t);
    if (g->d_roff) (void)hipFree(g->d_roff);
    if (g->d_err) (void)hipFree(g->d_err);
    g->ws.release();
    g->ws_bytes.release();
    g->dec_part.release();
    if (g->dec_ev) (void)hipEventDestroy(g->dec_ev);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}


hipStream_t pick(grass_graph* g, void* stream) { return stream ? (hipStream_t)stream : g->stream; }

}  // namespace

int open_sources(ByteSource& gs, ByteSource& os, int format, int device, grass_graph** out) {
    if (!out) return GRASS_EINVAL;
    *out = nullptr;
    if (format != GRASS_FMT_LE && format != GRASS_FMT_GT) return GRASS_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GRASS_EDEVICE;
    if (device < 0 || device >= ndev) return GRASS_EINVAL;
    DeviceGuard guard(device);
    if (!guard.ok) return GRASS_EDEVICE;
    grass_graph* g = new (std::nothrow) grass_graph();
    if (!g) return GRASS_ENOMEM;
    g->device = device;
    g->fmt = format;
    // GT checkpoint interval: GRASS_GT_CHECKPOINT = a power of two in [2, 1024] overrides the default
    if (const char* e = getenv("GRASS_GT_CHECKPOINT")) {
        const unsigned long p = strtoul(e, nullptr, 10);
        for (uint32_t sh = 1; sh <= 10; ++sh)
            if (p == (1ul << sh)) g->cp_shift = sh;
    }
    int rc = GRASS_OK;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) rc = GRASS_EDEVICE;
    if (!rc && hipMalloc(&g->d_err, 64) != hipSuccess) rc = GRASS_ENOMEM;
    if (!rc && hipMemset(g->d_err, 0, 64) != hipSuccess) rc = GRASS_EDEVICE;
    if (!rc) rc = load_graph(g, gs, os);
    if (rc) {
        free_graph(g);
        return rc;
    }
    *out = g;
    return GRASS_OK;
}

}  // namespace grass

using namespace grass;

extern "C" {

const char* grass_strerror(int status) {
    switch (status) {
        case GRASS_OK: return "success";
        case GRASS_EINVAL: return "invalid argument";
        case GRASS_EIO: return "cannot read the .graph/.offset files";
        case GRASS_EFORMAT: return "malformed compressed graph";
        case GRASS_ENOMEM: return "out of memory";
        case GRASS_EDEVICE: return "HIP runtime error";
        case GRASS_ECAPACITY: return "output or admission capacity exceeded";
        case GRASS_ECANCELLED: return "request cancelled";
        case GRASS_ETIMEOUT: return "timed out";
        case GRASS_ENOTREADY: return "not ready";
        default: return "unknown status";
    }
}

int grass_open(const char* prefix, int format, int device, grass_graph** out) {
    if (!prefix || !out) return GRASS_EINVAL;
    std::string p(prefix);
    FileSource gs(p + ".graph"), os(p + ".offset");
    if (!gs.ok() || !os.ok()) {
        *out = nullptr;
        return GRASS_EIO;
    }
    return open_sources(gs, os, format, device, out);
}

int grass_open_device_images(const void* graph_dev, uint64_t graph_bytes, const void* offset_dev,
                             uint64_t offset_bytes, int format, int device, grass_graph** out) {
    if (!graph_dev || !offset_dev || !out) return GRASS_EINVAL;
    DeviceSource gs(graph_dev, graph_bytes), os(offset_dev, offset_bytes);
    return open_sources(gs, os, format, device, out);
}

int grass_open_memory(const void* graph, uint64_t graph_bytes, const void* offset,
                      uint64_t offset_bytes, int format, int device, grass_graph** out) {
    if (!graph || !offset || !out) return GRASS_EINVAL;
    MemSource gs(graph, graph_bytes), os(offset, offset_bytes);
    return open_sources(gs, os, format, device, out);
}

int grass_close(grass_graph* g) {
    if (!g) return GRASS_OK;
    DeviceGuard guard(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    free_graph(g);
    return GRASS_OK;
}

int grass_info(const grass_graph* g, grass_graph_info* info) {
    if (!g || !info) return GRASS_EINVAL;
    info->n = g->n;
    info->m = g->m;
    info->max_degree = g->max_degree;
    info->graph_bytes = g->graph_bytes;
    info->offset_bytes = g->offset_bytes;
    info->total_bits = g->total_bits;
    info->ubl = g->ubl;
    info->format = g->fmt;
    info->device = g->device;
    info->device_bytes = g->device_bytes;
    info->gt_checkpoint = g->fmt == GRASS_FMT_GT ? (1u << g->cp_shift) : 0;
    return GRASS_OK;
}

int grass_device_arrays(const grass_graph* g, const uint32_t** words, const uint64_t** starts,
                        const uint32_t** degree) {
    if (!g) return GRASS_EINVAL;
    if (words) *words = g->dev().w;
    if (starts) *starts = g->d_start;
    if (degree) *degree = g->d_deg;
    return GRASS_OK;
}

int grass_device_alloc(grass_graph* g, uint64_t bytes, void** ptr_dev) {
    if (!g || !ptr_dev) return GRASS_EINVAL;
    *ptr_dev = nullptr;
    DeviceGuard guard(g->device);
    if (!guard.ok) return GRASS_EDEVICE;
    return hipMalloc(ptr_dev, bytes ? bytes : 1) == hipSuccess ? GRASS_OK : GRASS_ENOMEM;
}

int grass_device_free(grass_graph* g, void* ptr_dev) {
    if (!g) return GRASS_EINVAL;
    if (!ptr_dev) return GRASS_OK;
    DeviceGuard guard(g->device);
    return hipFree(ptr_dev) == hipSuccess ? GRASS_OK : GRASS_EDEVICE;
}

int grass_copy_to_host(grass_graph* g, void* host, const void* src_dev, uint64_t bytes) {
    if (!g || (bytes && (!host || !src_dev))) return GRASS_EINVAL;
    if (!bytes) return GRASS_OK;
    DeviceGuard guard(g->device);
    return hipMemcpy(host, src_dev, bytes, hipMemcpyDeviceToHost) == hipSuccess ? GRASS_OK : GRASS_EDEVICE;
}

int grass_copy_to_device(grass_graph* g, void* dst_dev, const void* host, uint64_t bytes) {
    if (!g || (bytes && (!host || !dst_dev))) return GRASS_EINVAL;
    if (!bytes) return GRASS_OK;
    DeviceGuard guard(g->device);
    return hipMemcpy(dst_dev, host, bytes, hipMemcpyHostToDevice) == hipSuccess ? GRASS_OK : GRASS_EDEVICE;
}

int grass_sync(grass_graph* g, void* stream) {
    if (!g) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    hipStream_t s = pick(g, stream);
    if (hipStreamSynchronize(s) != hipSuccess) return GRASS_EDEVICE;
    uint32_t err[2] = {0, 0};
    if (hipMemcpy(err, g->d_err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess) return GRASS_EDEVICE;
    if (err[0] == kErrNone) return GRASS_OK;
    if (hipMemset(g->d_err, 0, sizeof(err)) != hipSuccess) return GRASS_EDEVICE;
    return err[0] == kErrBadStart ? GRASS_EINVAL : GRASS_EFORMAT;
}

int grass_decode_csr(grass_graph* g, uint64_t* rowptr_dev, uint32_t* col_dev, void* stream) {
    if (!g || !rowptr_dev) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    return decode_csr(g, rowptr_dev, col_dev, pick(g, stream));
}

int grass_random_walk(grass_graph* g, const uint32_t* starts_dev, uint64_t walkers, uint32_t length,
                      uint64_t seed, uint32_t instance, uint32_t* out_dev, void* stream) {
    if (!g) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    if (!out_dev) return GRASS_EINVAL;
    return random_walk(g, starts_dev, walkers, length, seed, instance, out_dev, pick(g, stream),
                       nullptr);
}

int grass_random_walk_batch(grass_graph* g, const grass_rw_request* reqs, uint32_t count,
                            uint64_t seed, void* stream) {
    if (!g) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    for (uint32_t i = 0; reqs && i < count; ++i)
        if (reqs[i].walkers && !reqs[i].out_dev) return GRASS_EINVAL;
    return random_walk_batch(g, reqs, count, seed, pick(g, stream), nullptr);
}

int grass_random_walk_batch_noalloc(grass_graph* g, const grass_rw_request* reqs, uint32_t count,
                                    uint64_t seed, void* stream) {
    if (!g) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    if (!guard.ok) return GRASS_EDEVICE;
    return random_walk_batch_noalloc(g, reqs, count, seed, pick(g, stream));
}

int grass_khop_bound(const grass_graph* g, const grass_khop_request* reqs, uint32_t count,
                     uint64_t* max_edges) {
    if (!g || (count && !reqs) || !max_edges) return GRASS_EINVAL;
    return khop_bound(g, reqs, count, max_edges, nullptr, nullptr, nullptr);
}

int grass_khop_batch(grass_graph* g, const grass_khop_request* reqs, uint32_t count, uint64_t seed,
                     grass_khop_output* out, void* stream) {
    if (!g || (count && !reqs) || !out) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    return khop_batch(g, reqs, count, seed, out, pick(g, stream));
}

int grass_random_walk_bytes(grass_graph* g, const uint32_t* starts_dev, uint64_t walkers,
                            uint32_t length, uint64_t seed, uint32_t instance, void* stream,
                            uint64_t* bytes) {
    if (!g || !bytes) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    return random_walk(g, starts_dev, walkers, length, seed, instance, nullptr, pick(g, stream),
                       bytes);
}

int grass_random_walk_batch_bytes(grass_graph* g, const grass_rw_request* reqs, uint32_t count,
                                  uint64_t seed, void* stream, uint64_t* bytes) {
    if (!g || !bytes) return GRASS_EINVAL;
    DeviceGuard guard(g->device);
    return random_walk_batch(g, reqs, count, seed, pick(g, stream), bytes);
}

}  // extern "C"
