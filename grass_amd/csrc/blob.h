// Host objects handed out through the ingest ABI (include/grass_ingest.h): byte images
// (grass_blob_data) and CSRs (grass_csr_arrays).
#pragma once
#include <stdint.h>

#include <memory>
#include <new>
#include <vector>

struct grass_csr {
    uint64_t n = 0, m = 0;
    std::vector<uint64_t> rowptr;
    std::vector<uint32_t> col;
};

struct grass_blob {
    std::unique_ptr<uint8_t[]> buf;
    uint64_t len = 0;
    // Storage for n bytes, left uninitialised (every writer fills all of it); false when it
    // cannot be allocated.
    bool resize(uint64_t n) {
        buf.reset(new (std::nothrow) uint8_t[n ? n : 1]);
        len = buf ? n : 0;
        return buf != nullptr;
    }
    uint8_t* data() const { return buf.get(); }
    uint64_t size() const { return len; }
};
