// vx_microbench.hip -- ceilings for the measurement report: random-address atomic
// throughput in HBM, the operations the interner's probe, commit and GC passes issue.
//
// kind 0: 64-bit atomicCAS(expected 0) on uniformly random slots of a table
// kind 1: 32-bit atomicAdd on uniformly random counters
// kind 2: 32-bit atomicAdd on 64 hot counters (the contended case)
// kind 3: one 32-bit atomicAdd per wavefront on a single counter (a serial point such
//         as an allocation cursor that every block of a launch updates)
// kind 4: one 32-bit atomicAdd per wavefront on 64 counters 128 bytes apart (striped)
// Every lane issues `per_lane` operations; addresses come from a splitmix64 stream
// of the global lane index, so runs are reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vx_device.h"

namespace vx {

__global__ void k_atomic_bench(void* buf, uint64_t slots, uint32_t per_lane, int kind, uint64_t* sink) {
    uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x = mix64(lane * 0x9E3779B97F4A7C15ull + 1);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < per_lane; i++) {
        x = mix64(x + 0x9E3779B97F4A7C15ull);
        if (kind == 0) {
            uint64_t* t = (uint64_t*)buf;
            acc += atomicCAS((unsigned long long*)&t[x & (slots - 1)], 0ull, (unsigned long long)(x | 1));
        } else if (kind == 1) {
            uint32_t* c = (uint32_t*)buf;
            acc += atomicAdd(&c[x & (slots - 1)], 1u);
        } else if (kind == 2) {
            uint32_t* c = (uint32_t*)buf;
            acc += atomicAdd(&c[x & 63u], 1u);
        } else {
            uint32_t* c = (uint32_t*)buf;
            uint32_t r = 0;
            if ((threadIdx.x & 63u) == 0) r = atomicAdd(&c[kind == 3 ? 0u : 32u * (blockIdx.x & 63u)], 1u);
            acc += __shfl(r, 0);
        }
    }
    if (acc == 0x5A5A5A5A5A5A5A5Aull) sink[0] = acc;  // keeps the returns live
}

// A plain streaming copy (16 B per lane per step): the reference point that scales
// the memory counters (bytes read and written are exactly `bytes` each).
__global__ void k_copy_bench(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace vx
