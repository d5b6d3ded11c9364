// Byte image handed out through the ingest ABI (include/grass_ingest.h: grass_blob_data).
#pragma once
#include <stdint.h>

#include <vector>

struct grass_blob {
    std::vector<uint8_t> bytes;
};
