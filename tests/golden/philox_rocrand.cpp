// Writes Philox4x32-10 vectors from rocRAND's open-source host/device implementation
// (rocrand_philox4x32_10.h, ten_rounds) so the oracle's Philox is checked against an
// independent implementation.  Build and run (host code only, no GPU needed):
//   hipcc -O1 -o /tmp/philox_rocrand tests/golden/philox_rocrand.cpp && /tmp/philox_rocrand > tests/golden/philox_rocrand.txt
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    Probe p;
    uint64_t s = 0x0123456789abcdefull;
    auto next = [&]() { s = s * 6364136223846793005ull + 1442695040888963407ull; return (uint32_t)(s >> 32); };
    printf("# ctr0 ctr1 ctr2 ctr3 key0 key1 -> out0 out1 out2 out3 (hex)\n");
    for (int i = 0; i < 64; ++i) {
        uint4 c = {next(), next(), next(), next()};
        uint2 k = {next(), next()};
        if (i == 0) { c = {0, 0, 0, 0}; k = {0, 0}; }
        if (i == 1) { c = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}; k = {0xffffffffu, 0xffffffffu}; }
        if (i == 2) { c = {7, 3, 0xFFFFFFFFu, 0}; k = {42, 0}; }
        uint4 r = p.run(c, k);
        printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c.x, c.y, c.z, c.w, k.x, k.y, r.x, r.y, r.z, r.w);
    }
    return 0;
}
