// Random-access ceilings of the memory system (libgrass_calib.so, measurement only).
//
// The sampling kernels do not stream: a random-walk step reads one random 16-byte header and one
// random span of the payload stream, so their HBM traffic is a count of random 64-byte lines,
// not a stream of bytes.  These kernels measure what random lines the GPU can deliver, so that a
// sampling kernel's measured traffic can be read against a random-access ceiling as well as
// against the 8 TB/s streaming peak.  Every lane reads 4 bytes from each line it touches (the
// memory system moves the whole 64-byte line):
//   mode 0  independent lines: `steps` loads per lane whose addresses do not depend on loaded data;
//   mode 1  dependent chain: each load's address depends on the previous load's value;
//   mode 2  walk-shaped chain: per step two dependent loads, the second one's address depending on
//           the first one's value and the next step's on both (a walk's header, then its payload).
// Addresses are a multiply-xorshift hash of (lane, step, loaded values) over `lines` lines.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t mix(uint64_t x) {
    x ^= x >> 31;
    x *= 0x9E3779B97F4A7C15ull;
    x ^= x >> 29;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 32;
    return x;
}

__device__ __forceinline__ uint64_t line_of(uint64_t h, uint64_t lines) {
    return (uint64_t)(((unsigned __int128)h * lines) >> 64);
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void gather_kernel(const uint32_t* __restrict__ buf, uint64_t lines,
                                                        uint32_t steps, uint64_t lanes, uint64_t seed,
                                                        unsigned long long* __restrict__ sink) {
    uint32_t acc = 0;
    for (uint64_t lane = (uint64_t)blockIdx.x * kBlock + threadIdx.x; lane < lanes;
         lane += (uint64_t)gridDim.x * kBlock) {
        uint64_t x = mix(lane * 0x100000001B3ull + seed);
        if (MODE == 0) {
            for (uint32_t s = 0; s < steps; s += 8) {
                uint32_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint64_t h = mix(x + (uint64_t)(s + j) * 0xD6E8FEB86659FD93ull);
                    v[j] = (s + j < steps) ? buf[line_of(h, lines) * 16] : 0u;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) acc += v[j];
            }
        } else {
            for (uint32_t s = 0; s < steps; ++s) {
                const uint32_t a = buf[line_of(x, lines) * 16];
                x = mix(x ^ a);
                if (MODE == 2) {
                    const uint32_t b = buf[line_of(x, lines) * 16];
                    x = mix(x ^ b);
                }
            }
            acc += (uint32_t)x;
        }
    }
    if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);  // keeps the loads; practically never taken
}

}  // namespace

extern "C" {

// Runs one launch of `mode` over a device buffer of lines x 64 bytes and returns its duration
// (HIP events on `stream`) in *ms and the lines it read in *lines_read.  The grid is the
// persistent shape the sampling kernels use: every CU filled once at the kernel's occupancy.
int grass_calib_gather(const void* buf_dev, uint64_t lines, uint32_t mode, uint32_t steps, uint64_t lanes,
                       uint64_t seed, void* stream, void* sink_dev, double* ms, uint64_t* lines_read) {
    if (!buf_dev || !lines || mode > 2 || !sink_dev || !ms || !lines_read) return 1;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const void* k = mode == 0 ? (const void*)gather_kernel<0> : mode == 1 ? (const void*)gather_kernel<1>
                                                                          : (const void*)gather_kernel<2>;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, kBlock, 0);
    uint64_t grid = (uint64_t)(cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    const uint64_t need = (lanes + kBlock - 1) / kBlock;
    if (grid > need) grid = need ? need : 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 2;
    (void)hipEventRecord(e0, s);
    const uint32_t* b = (const uint32_t*)buf_dev;
    unsigned long long* sk = (unsigned long long*)sink_dev;
    if (mode == 0) hipLaunchKernelGGL(gather_kernel<0>, dim3((unsigned)grid), dim3(kBlock), 0, s, b, lines, steps, lanes, seed, sk);
    else if (mode == 1) hipLaunchKernelGGL(gather_kernel<1>, dim3((unsigned)grid), dim3(kBlock), 0, s, b, lines, steps, lanes, seed, sk);
    else hipLaunchKernelGGL(gather_kernel<2>, dim3((unsigned)grid), dim3(kBlock), 0, s, b, lines, steps, lanes, seed, sk);
    (void)hipEventRecord(e1, s);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return 3;
    *ms = t;
    *lines_read = lanes * (uint64_t)steps * (mode == 2 ? 2 : 1);
    return 0;
}

}  // extern "C"
