// dyc_wave.hpp -- wave64 helpers shared by the kernel files of libdyc (internal).
#pragma once

#include "dyc_kernels.hpp"

namespace dyc {

constexpr int kWave = 64;

__device__ inline uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

__device__ inline uint64_t shfl64(uint64_t v, int src)
{
    uint32_t lo = __shfl((uint32_t)v, src);
    uint32_t hi = __shfl((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// 64-bit compare-and-swap at agent scope; on failure `expected` receives the current value.
__device__ inline bool cas_slot(uint64_t *p, uint64_t &expected, uint64_t desired)
{
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// 64-bit compare-and-swap at agent scope that returns the value it found (== expected on
// success). Nothing waits for the result until the caller compares it, so several can be in flight.
__device__ inline uint64_t cas_old(uint64_t *p, uint64_t expected, uint64_t desired)
{
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (unsigned long long)shfl64(v, __lane_id() ^ off);
    return v;
}

// This wave's index in the grid, as a wave-uniform (scalar) value.
__device__ inline uint32_t wave_slot()
{
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
}

// Add v (from one lane) to the tally `field` (a member of *C) in the stripe of wave slot `slot`.
__device__ inline void tally_add(Counters *C, unsigned long long *field, unsigned long long v, uint32_t slot)
{
    unsigned long long *base = reinterpret_cast<unsigned long long *>(C);
    atomicAdd(base + kTallyStrideWords * (1u + slot % kTallyStripes) + (field - base), v);
}

__device__ inline void wave_tally(Counters *C, unsigned long long *field, unsigned long long v, uint32_t slot)
{
    v = wave_sum(v);
    if (__lane_id() == 0 && v) tally_add(C, field, v, slot);
}

__device__ inline void wave_tally(Counters *C, unsigned long long *field, unsigned long long v)
{
    wave_tally(C, field, v, wave_slot());
}

// The wave's largest v into the tally `field` of the wave's stripe (a field the fold takes the
// maximum of, not the sum).
__device__ inline void wave_tally_max(Counters *C, unsigned long long *field, unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long x = shfl64(v, __lane_id() ^ off);
        v = v > x ? v : x;
    }
    if (__lane_id() == 0 && v) {
        unsigned long long *base = reinterpret_cast<unsigned long long *>(C);
        atomicMax(base + kTallyStrideWords * (1u + wave_slot() % kTallyStripes) + (field - base), v);
    }
}


// Blocks for work_waves waves of work under a launch shape (capped; 0 max_blocks = default_cap).
inline int grid_for(uint64_t work_waves, LaunchShape sh, uint64_t default_cap = 1ull << 20)
{
    const uint64_t wpb = (uint64_t)sh.block_threads / kWave;
    uint64_t blocks = (work_waves + wpb - 1) / wpb;
    if (blocks < 1) blocks = 1;
    const uint64_t cap = sh.max_blocks > 0 ? (uint64_t)sh.max_blocks : default_cap;
    if (blocks > cap) blocks = cap;
    const uint64_t hw_cap = ((1ull << 32) - 1) / (uint64_t)sh.block_threads;
    if (blocks > hw_cap) blocks = hw_cap;
    return (int)blocks;
}

}  // namespace dyc
