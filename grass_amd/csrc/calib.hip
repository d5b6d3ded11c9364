// Random-access ceilings and synchronisation costs of the GPU (libgrass_calib.so, measurement only).
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

// One grid-wide barrier: every workgroup arrives on a monotonic counter (barrier k is complete when
// it reaches nblocks * (k + 1)); the last one to arrive publishes generation k + 1, the others
// wait for it.  The wait is bounded (about a second): a workgroup that gives up reports it and
// returns, so the kernel always ends.
__device__ __forceinline__ bool grid_barrier(unsigned* cnt, unsigned* gen, unsigned nblocks, unsigned& my) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        const unsigned g = my;
        __threadfence();
        const unsigned arrived = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks * (g + 1) - 1) {
            __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <= g) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1u << 24)) {
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
        my = g + 1;
    }
    __syncthreads();
    return ok;
}

// `iters` rounds of (one global store per thread, one grid barrier); words: state[0] counter,
// state[1] generation, state[2] workgroups that gave up waiting.
__global__ __launch_bounds__(kBlock) void barrier_kernel(unsigned* state, unsigned nblocks, unsigned iters,
                                                         unsigned* __restrict__ data, uint32_t words) {
    unsigned my = 0;
    for (unsigned i = 0; i < iters; ++i) {
        data[(blockIdx.x * kBlock + threadIdx.x + i * 97u) % words] = i;
        if (!grid_barrier(state, state + 1, nblocks, my)) {
            if (threadIdx.x == 0) atomicAdd(state + 2, 1u);
            return;
        }
    }
}

// One global store per thread: a kernel whose cost is its launch.
__global__ __launch_bounds__(kBlock) void store_kernel(unsigned* __restrict__ data, uint32_t words, unsigned i) {
    data[(blockIdx.x * kBlock + threadIdx.x + i * 97u) % words] = i;
}

__global__ void zero_kernel(unsigned* p, unsigned n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0;
}

}  // namespace

extern "C" {

// What it costs to make every workgroup of a launch wait for every other one: `iters` grid barriers
// inside one launch of `blocks` workgroups (at most what the device holds at once) against `iters`
// dependent launches of the same workgroups on one stream.  scratch_dev: 16 bytes + data_words x 4
// bytes of device memory.  Returns 0, or 4 if a barrier wait gave up (*barrier_us is then invalid).
int grass_calib_sync_cost(void* scratch_dev, uint32_t data_words, uint32_t blocks, uint32_t iters, void* stream,
                          double* barrier_us, double* boundary_us) {
    if (!scratch_dev || !data_words || !blocks || !iters || !barrier_us || !boundary_us) return 1;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)barrier_kernel, kBlock, 0);
    if ((uint64_t)blocks > (uint64_t)(cus > 0 ? cus : 1) * (per > 0 ? per : 1)) return 1;  // all must be resident
    unsigned* state = (unsigned*)scratch_dev;
    unsigned* data = state + 4;
    hipEvent_t e0, e1, e2;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess)
        return 2;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, state, 4u);
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(barrier_kernel, dim3(blocks), dim3(kBlock), 0, s, state, blocks, iters, data, data_words);
    (void)hipEventRecord(e1, s);
    for (uint32_t i = 0; i < iters; ++i)
        hipLaunchKernelGGL(store_kernel, dim3(blocks), dim3(kBlock), 0, s, data, data_words, i);
    (void)hipEventRecord(e2, s);
    hipError_t e = hipEventSynchronize(e2);
    float t1 = 0, t2 = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t1, e0, e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t2, e1, e2);
    unsigned gave_up = 0;
    if (e == hipSuccess) e = hipMemcpy(&gave_up, state + 2, 4, hipMemcpyDeviceToHost);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (e != hipSuccess) return 3;
    *barrier_us = t1 * 1e3 / iters;
    *boundary_us = t2 * 1e3 / iters;
    return gave_up ? 4 : 0;
}


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
