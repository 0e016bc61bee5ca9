// bench_memsys.hip -- device memory-rate probe of bench.py --workload memsys (bench_memsys.so),
// the ceilings the hash kernels are judged against: at the size of a table, how many random
// 128-byte bucket reads, random 64-bit CAS and random 8-byte stores per second the device
// sustains. A measurement tool for the benchmark, not part of the table library.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ inline uint32_t mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// one lane reads one whole 128-byte bucket (8 x 16 B) per op, as the insert's single-lane view
__global__ void read_bucket_kernel(const uint4 *__restrict__ buf, uint64_t nb, uint64_t ops, uint64_t *sink)
{
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ops; i += stride) {
        const uint4 *p = buf + (uint64_t)(mix((uint32_t)i * 2654435761u + 17u) % nb) * 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint4 v = p[j];
            acc += v.x ^ v.w;
        }
    }
    if (acc == 0x9e3779b97f4a7c15ull) sink[0] = acc;  // keeps the loads
}

// one 64-bit CAS on a random slot per op
__global__ void cas_kernel(unsigned long long *__restrict__ buf, uint64_t nslots, uint64_t ops, uint64_t *sink)
{
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ops; i += stride) {
        unsigned long long exp = 0;
        acc += __hip_atomic_compare_exchange_strong(buf + mix((uint32_t)i * 2246822519u + 5u) % nslots, &exp,
                                                    (unsigned long long)i + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    }
    if (acc == 0x9e3779b97f4a7c15ull) sink[0] = acc;
}

// one lane reads the first 64 bytes of a random bucket (4 x 16 B) per op
__global__ void read_half_kernel(const uint4 *__restrict__ buf, uint64_t nb, uint64_t ops, uint64_t *sink)
{
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ops; i += stride) {
        const uint4 *p = buf + (uint64_t)(mix((uint32_t)i * 2654435761u + 17u) % nb) * 8;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint4 v = p[j];
            acc += v.x ^ v.w;
        }
    }
    if (acc == 0x9e3779b97f4a7c15ull) sink[0] = acc;
}

// read a random bucket, then CAS one of its slots from the value read (an insert exchange)
__global__ void read_cas_kernel(unsigned long long *__restrict__ buf, uint64_t nb, uint64_t ops, uint64_t *sink)
{
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ops; i += stride) {
        const uint64_t b = mix((uint32_t)i * 2654435761u + 17u) % nb;
        const uint4 *p = reinterpret_cast<const uint4 *>(buf + b * 16);
        uint32_t h = 0;
        unsigned long long first = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint4 v = p[j];
            h += v.x ^ v.z;
            if (j == 0) first = ((unsigned long long)v.y << 32) | v.x;
        }
        unsigned long long exp = first;
        acc += __hip_atomic_compare_exchange_strong(buf + b * 16 + (h & 0u), &exp, exp + 1, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (acc == 0x9e3779b97f4a7c15ull) sink[0] = acc;
}

// one plain 8-byte store to a random slot per op
__global__ void store_kernel(unsigned long long *__restrict__ buf, uint64_t nslots, uint64_t ops)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ops; i += stride)
        buf[mix((uint32_t)i * 2246822519u + 5u) % nslots] = i;
}

// one 64-bit atomic add per wave onto one of `spread` counters 256 B apart (spread 1: a single
// address), the pattern of per-wave tallies and append indexes
__global__ void wave_atomic_kernel(unsigned long long *__restrict__ ctr, uint32_t spread)
{
    const uint32_t w = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if ((threadIdx.x & 63u) == 0) atomicAdd(ctr + (uint64_t)(w % spread) * 32, 1ull);
}

}  // namespace

extern "C" {

// Rates over a zeroed device buffer of `bytes` (rounded down to whole 128-byte buckets), `ops`
// operations per pattern, best of `reps`: out[0] bucket reads/s, out[1] CAS/s, out[2] stores/s,
// out[3] half-bucket (64 B) reads/s, out[4] read-a-bucket-then-CAS-a-slot/s, out[5] per-wave
// atomic adds/s onto a single address, out[6] the same spread over 64 addresses 256 B apart
// (65,536 waves, one add each, for these two).
// Returns 0 or a hipError_t code.
int dyc_memsys_probe(uint64_t bytes, uint64_t ops, int reps, double *out)
{
    const uint64_t nb = bytes / 128, nslots = nb * 16;
    if (nb < 128 || ops == 0 || !out) return (int)hipErrorInvalidValue;  // the counters need 16 KB
    unsigned long long *buf = nullptr;
    uint64_t *sink = nullptr;
    hipError_t e = hipMalloc(&buf, nb * 128);
    if (e != hipSuccess) return (int)e;
    if ((e = hipMalloc(&sink, 64)) != hipSuccess || (e = hipMemset(buf, 0, nb * 128)) != hipSuccess) {
        (void)hipFree(buf);
        return (int)e;
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    const int grid = 16384;  // 65,536 waves
    for (int k = 0; k < 7; ++k) {
        float best = 1e30f;
        for (int r = 0; r < reps + 1; ++r) {  // the first run warms up
            (void)hipEventRecord(a, 0);
            if (k == 0) read_bucket_kernel<<<grid, 256>>>(reinterpret_cast<const uint4 *>(buf), nb, ops, sink);
            else if (k == 1) cas_kernel<<<grid, 256>>>(buf, nslots, ops, sink);
            else if (k == 2) store_kernel<<<grid, 256>>>(buf, nslots, ops);
            else if (k == 3) read_half_kernel<<<grid, 256>>>(reinterpret_cast<const uint4 *>(buf), nb, ops, sink);
            else if (k == 4) read_cas_kernel<<<grid, 256>>>(buf, nb, ops, sink);
            else wave_atomic_kernel<<<grid, 256>>>(buf, k == 5 ? 1u : 64u);
            (void)hipEventRecord(b, 0);
            if ((e = hipEventSynchronize(b)) != hipSuccess) break;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            if (r > 0 && ms < best) best = ms;
        }
        out[k] = (k < 5 ? (double)ops : grid * 4.0) / (best * 1e-3);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(sink);
    (void)hipFree(buf);
    return (int)e;
}

}  // extern "C"
