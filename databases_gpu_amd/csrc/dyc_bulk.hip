// dyc_bulk.hip -- bulk insert by bucket fill (insert_order 3).
//
// The eviction insert spends one random 64-bit CAS per exchange, and the device's CAS rate is
// the insert's ceiling however local the accesses are. A large batch can place most of its keys
// without any global atomic instead:
//   1. pack: each key's target bucket (the start rule's table, cuckoo.cu:336-341) names a
//      chunk of kChunkBuckets consecutive buckets of one subtable; write (chunk, entry) pairs;
//   2. sort the pairs by chunk (rocPRIM radix sort over the chunk id's bits);
//   3. bounds: the first sorted index of every chunk;
//   4. fill: one workgroup per chunk owns its buckets for the launch. It takes the buckets'
//      free-slot masks into LDS, counts the entries per bucket, reserves one range of the
//      overflow list for the chunk's excess, then gives each entry the next free slot of its
//      bucket (an LDS counter) or appends it to the overflow list. A dense round fills a copy
//      of the chunk in LDS and writes it back whole; a sparse one stores each entry directly;
//   5. the overflow is packed again for the key's other table (round 2) and filled the same way;
//   6. what still overflows goes through the eviction insert (launch_insert_entries).
// The placement rules that define a key's buckets are unchanged; only the order in which free
// slots are taken differs, and the slot layout after a parallel insert is not deterministic
// anyway. A placement by fill counts as one exchange in the kick tally.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "dyc_kernels.hpp"
#include "dyc_wave.hpp"

namespace dyc {

namespace {

__device__ inline uint32_t start_table(const Desc &D, uint32_t k)
{
    const uint32_t p = pair_of(D, k);
    return (k & 1u) ? pos2(D, p) : pos1(D, p);
}

// Round 1 (IN_ENTRIES false): keys/vals, target = the start table; reserved keys go to chunk 0
// and are rejected by its fill. Round 2 (IN_ENTRIES true): overflow entries, target = the other
// table of the key's pair.
template <bool IN_ENTRIES>
__global__ __launch_bounds__(256) void bulk_pack_kernel(Desc D, Chunks Ch, const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ vals,
                                                        const uint64_t *__restrict__ in, uint64_t n,
                                                        uint32_t *__restrict__ cid, uint64_t *__restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t e;
        uint32_t k, c = 0;
        if constexpr (IN_ENTRIES) {
            e = in[i];
            k = key_of(e);
            const uint32_t p = pair_of(D, k);
            const uint32_t t = other_pos(D, p, (k & 1u) ? pos2(D, p) : pos1(D, p));
            c = chunk_base(Ch, t) + (bucket_of(D, k, t) >> kChunkShift);
        } else {
            k = keys[i];
            e = pack(k, vals[i]);
            if (!reserved_key(k)) {
                const uint32_t t = start_table(D, k);
                c = chunk_base(Ch, t) + (bucket_of(D, k, t) >> kChunkShift);
            }
        }
        cid[i] = c;
        out[i] = e;
    }
}

// start[c] = the first index of chunk c in the sorted chunk ids, for c in [0, total]; index i
// writes the starts of the chunks in (cid[i - 1], cid[i]] (with cid[-1] = -1, cid[n] = total).
// A thread takes four consecutive indexes with one 16-byte load (cid is 16-byte aligned).
__global__ __launch_bounds__(256) void bulk_bounds_kernel(const uint32_t *__restrict__ cid, uint64_t n,
                                                          uint32_t total, uint32_t *__restrict__ start)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t quads = n / 4 + 1;  // the last quad holds index n (and the tail)
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < quads; g += stride) {
        const uint64_t i0 = g * 4;
        uint32_t v[4];
        if (i0 + 4 <= n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(cid + i0);
            v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = i0 + u < n ? cid[i0 + u] : total;
        }
        int64_t prev = i0 == 0 ? -1 : (int64_t)min(cid[i0 - 1], total);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i0 + u > n) break;
            const int64_t cur = i0 + u == n ? (int64_t)total : (int64_t)min(v[u], total);
            for (int64_t c = prev + 1; c <= cur; ++c) start[c] = (uint32_t)(i0 + u);
            prev = cur;
        }
    }
}

__device__ inline uint32_t nth_set_bit(uint32_t m, uint32_t r)
{
    for (uint32_t q = 0; q < r; ++q) m &= m - 1u;
    return __builtin_ctz(m);
}

// One workgroup per chunk. Buckets of the chunk are written only by this workgroup during the
// launch (every entry of the chunk targets one of them), so slots are taken with LDS counters.
// DENSE (a round with at least two entries per bucket): the chunk's 512 buckets are copied into
// LDS, filled there and written back whole, so every global access is a coalesced 16-byte
// load or store. Otherwise the entries are counted first, the free-slot masks of the buckets
// that received any are read coalesced (8 lanes per bucket, folded by ballot), and each placed
// entry is one store into its slot.
template <bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void bulk_fill_kernel(Desc D, Chunks Ch, const uint32_t *__restrict__ start,
                                                        const uint64_t *__restrict__ ent, uint64_t n,
                                                        uint64_t *__restrict__ ovf, uint64_t ovf_cap,
                                                        unsigned long long *__restrict__ ovf_n,
                                                        Counters *__restrict__ C)
{
    constexpr uint32_t kImg = DENSE ? kChunkBuckets * kSlots : 2;
    __shared__ uint64_t img[kImg];           // DENSE: the chunk's buckets
    __shared__ uint32_t fm[kChunkBuckets];   // free-slot mask of each bucket
    __shared__ uint32_t cnt[kChunkBuckets];  // entries per bucket, then slots taken
    __shared__ uint32_t wred[THREADS / kWave];
    __shared__ unsigned long long obase;
    __shared__ uint32_t onext;
    const uint32_t c = blockIdx.x;
    // clamped to the entries, so a range is never read past them whatever the bounds hold
    const uint32_t hi = (uint32_t)min((uint64_t)start[c + 1], n);
    const uint32_t lo = min(start[c], hi);
    unsigned long long placed = 0, rejected = 0;
    if (lo < hi) {  // block-uniform
        uint32_t t = 0;
        for (uint32_t q = 1; q < D.d; ++q)
            if (Ch.base[q] <= c) t = q;
        const uint32_t nb = selt(t, D.nb);
        const uint32_t b0 = (c - chunk_base(Ch, t)) << kChunkShift;
        const uint32_t len = min(kChunkBuckets, nb - b0);
        uint64_t *const sub = selt(t, D.sub);
        const uint4 *const src = reinterpret_cast<const uint4 *>(sub + (uint64_t)b0 * kSlots);
        constexpr uint32_t kParts = kChunkBuckets * kSlots / 2;  // 16-byte parts of the chunk
        if constexpr (DENSE) {
            uint4 *const im = reinterpret_cast<uint4 *>(img);
            for (uint32_t q = threadIdx.x; q < kParts; q += blockDim.x)
                im[q] = q < len * (kSlots / 2) ? src[q] : make_uint4(~0u, ~0u, ~0u, ~0u);  // past the end: full
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) {
                uint32_t m = 0;
#pragma unroll
                for (int q = 0; q < kSlots; ++q) m |= (uint32_t)(img[j * kSlots + q] == 0ull) << q;
                fm[j] = m;
                cnt[j] = 0;
            }
        } else {
            for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) cnt[j] = 0;
        }
        if (threadIdx.x == 0) onext = 0;
        __syncthreads();
        // count: entries per bucket (an entry whose bucket falls outside the chunk, which the
        // pack rule excludes, would go to the overflow list). Entries are loaded kBatch at a
        // time before any is used, so a thread waits for memory once per batch.
        constexpr uint32_t kBatch = 8;
        uint32_t stray = 0;
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBatch * blockDim.x) {
            uint64_t eb[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                eb[u] = i < hi ? ent[i] : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t k = key_of(eb[u]);
                if (i0 + u * blockDim.x >= hi || reserved_key(k)) continue;
                const uint32_t b = bucket_of(D, k, t);
                if (b >= b0 && b - b0 < len) atomicAdd(&cnt[b - b0], 1u);
                else ++stray;
            }
        }
        __syncthreads();
        if constexpr (!DENSE) {
            // free masks of the buckets that received entries only (a bucket's 8 parts are read
            // by 8 consecutive lanes and folded by ballot; the others read as full)
            // lane l of a wave reads part l % 8 of bucket (q / 8); a bucket's 8 parts sit in 8
            // consecutive lanes, whose two ballot bits per part make its 16-bit mask
            constexpr uint32_t kPer = kParts / THREADS;  // parts per thread, all loaded before use
            uint4 v[kPer];
#pragma unroll
            for (uint32_t u = 0; u < kPer; ++u) {
                const uint32_t q = threadIdx.x + u * THREADS;
                v[u] = q < len * (kSlots / 2) && cnt[q / 8] ? src[q] : make_uint4(1u, 1u, 1u, 1u);
            }
#pragma unroll
            for (uint32_t u = 0; u < kPer; ++u) {
                const uint32_t q = threadIdx.x + u * THREADS;
                const bool e0 = (v[u].x | v[u].y) == 0u, e1 = (v[u].z | v[u].w) == 0u;
                const uint64_t m0 = __ballot(e0), m1 = __ballot(e1);
                const uint32_t lane = __lane_id();
                if ((lane & 7u) == 0) {
                    const uint32_t a = (uint32_t)(m0 >> lane) & 0xffu, b = (uint32_t)(m1 >> lane) & 0xffu;
                    uint32_t m = 0;
#pragma unroll
                    for (int p = 0; p < 8; ++p) m |= (((a >> p) & 1u) << (2 * p)) | (((b >> p) & 1u) << (2 * p + 1));
                    fm[q / 8] = m;
                }
            }
            __syncthreads();
        }
        uint32_t excess = stray;
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) {
            const uint32_t f = (uint32_t)__builtin_popcount(fm[j]);
            excess += cnt[j] > f ? cnt[j] - f : 0u;
        }
        excess = (uint32_t)wave_sum((unsigned long long)excess);
        if (__lane_id() == 0) wred[threadIdx.x / kWave] = excess;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < blockDim.x / kWave; ++w) tot += wred[w];
            obase = tot ? atomicAdd(ovf_n, (unsigned long long)tot) : 0ull;
        }
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) cnt[j] = 0;
        __syncthreads();
        // place
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBatch * blockDim.x) {
            uint64_t eb[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                eb[u] = i < hi ? ent[i] : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                if (i0 + u * blockDim.x >= hi) continue;
                const uint64_t e = eb[u];
                const uint32_t k = key_of(e);
                if (reserved_key(k)) {
                    ++rejected;
                    continue;
                }
                const uint32_t b = bucket_of(D, k, t);
                bool fits = false;
                if (b >= b0 && b - b0 < len) {
                    const uint32_t j = b - b0;
                    const uint32_t r = atomicAdd(&cnt[j], 1u);
                    const uint32_t m = fm[j];
                    if (r < (uint32_t)__builtin_popcount(m)) {
                        const uint32_t slot = nth_set_bit(m, r);
                        if constexpr (DENSE) img[j * kSlots + slot] = e;
                        else sub[(uint64_t)b * kSlots + slot] = e;
                        fits = true;
                        ++placed;
                    }
                }
                if (!fits) {
                    // the counted excess reserved exactly these places; the list holds every entry
                    // of the batch, so the check fails only if the chunk ranges overlapped
                    const unsigned long long at = obase + atomicAdd(&onext, 1u);
                    if (at < ovf_cap) ovf[at] = e;
                }
            }
        }
        if constexpr (DENSE) {
            __syncthreads();
            uint4 *const dst = reinterpret_cast<uint4 *>(sub + (uint64_t)b0 * kSlots);
            const uint4 *const im = reinterpret_cast<const uint4 *>(img);
            for (uint32_t q = threadIdx.x; q < len * (kSlots / 2); q += blockDim.x) dst[q] = im[q];
        }
    }
    wave_tally(C, &C->kicks, placed);
    wave_tally(C, &C->rejected, rejected);
}

}  // namespace

Chunks make_chunks(const Desc &D)
{
    Chunks Ch{};
    uint32_t acc = 0;
    for (uint32_t t = 0; t < kMaxSubtables; ++t) {
        Ch.base[t] = acc;
        if (t < D.d) acc += (D.nb[t] + kChunkBuckets - 1) / kChunkBuckets;
    }
    Ch.base[kMaxSubtables] = acc;
    Ch.total = acc;
    Ch.bits = 1;
    while (Ch.bits < 32 && (1ull << Ch.bits) < acc) ++Ch.bits;
    return Ch;
}

hipError_t launch_bulk_pack(const Desc &D, const Chunks &Ch, const uint32_t *keys, const uint32_t *vals,
                            const uint64_t *entries, uint64_t n, uint32_t *cid, uint64_t *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, LaunchShape{256, 0}, 16384);
    if (entries)
        bulk_pack_kernel<true><<<grid, 256, 0, st>>>(D, Ch, nullptr, nullptr, entries, n, cid, out);
    else
        bulk_pack_kernel<false><<<grid, 256, 0, st>>>(D, Ch, keys, vals, nullptr, n, cid, out);
    return hipGetLastError();
}

hipError_t bulk_sort_bytes(uint64_t n, uint32_t bits, size_t *bytes)
{
    rocprim::double_buffer<uint32_t> kb(nullptr, nullptr);
    rocprim::double_buffer<uint64_t> vb(nullptr, nullptr);
    *bytes = 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, kb, vb, (unsigned int)n, 0u, bits);
}

hipError_t bulk_sort(void *tmp, size_t tmp_bytes, uint32_t *cid0, uint32_t *cid1, uint64_t *ent0, uint64_t *ent1,
                     uint64_t n, uint32_t bits, hipStream_t st, int *sorted_in)
{
    *sorted_in = 0;
    if (n == 0) return hipSuccess;
    rocprim::double_buffer<uint32_t> kb(cid0, cid1);
    rocprim::double_buffer<uint64_t> vb(ent0, ent1);
    const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (unsigned int)n, 0u, bits, st);
    if (e != hipSuccess) return e;
    *sorted_in = kb.current() == cid0 ? 0 : 1;
    if ((vb.current() == ent0 ? 0 : 1) != *sorted_in) return hipErrorUnknown;
    return hipSuccess;
}

hipError_t bulk_sort_wide_bytes(uint64_t n, uint32_t bits, size_t *bytes)
{
    rocprim::double_buffer<uint32_t> kb(nullptr, nullptr);
    rocprim::double_buffer<WideEntry> vb(nullptr, nullptr);
    *bytes = 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, kb, vb, (unsigned int)n, 0u, bits);
}

hipError_t bulk_sort_wide(void *tmp, size_t tmp_bytes, uint32_t *cid0, uint32_t *cid1, WideEntry *ent0,
                          WideEntry *ent1, uint64_t n, uint32_t bits, hipStream_t st, int *sorted_in)
{
    *sorted_in = 0;
    if (n == 0) return hipSuccess;
    rocprim::double_buffer<uint32_t> kb(cid0, cid1);
    rocprim::double_buffer<WideEntry> vb(ent0, ent1);
    const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, kb, vb, (unsigned int)n, 0u, bits, st);
    if (e != hipSuccess) return e;
    *sorted_in = kb.current() == cid0 ? 0 : 1;
    if ((vb.current() == ent0 ? 0 : 1) != *sorted_in) return hipErrorUnknown;
    return hipSuccess;
}

hipError_t launch_bulk_bounds(const uint32_t *cid, uint64_t n, uint32_t total, uint32_t *start, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n / 4 + 1 + kWave - 1) / kWave, LaunchShape{256, 0}, 16384);
    bulk_bounds_kernel<<<grid, 256, 0, st>>>(cid, n, total, start);
    return hipGetLastError();
}

hipError_t launch_bulk_fill(const Desc &D, const Chunks &Ch, const uint32_t *cid, const uint64_t *ent, uint64_t n,
                            uint32_t *start, uint64_t *ovf, uint64_t ovf_cap, unsigned long long *ovf_n, Counters *C,
                            hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipError_t e = launch_bulk_bounds(cid, n, Ch.total, start, st);
    if (e != hipSuccess) return e;
    if (n >= (1ull << 32) || Ch.total == 0) return hipErrorInvalidValue;
    uint64_t buckets = 0;
    for (uint32_t t = 0; t < D.d; ++t) buckets += D.nb[t];
    // dense: copying every chunk through LDS costs less than scattered stores; its 68 KB of LDS
    // allow two workgroups per CU, so they are 512 threads wide to keep more loads in flight
    if (n >= 2 * buckets)
        bulk_fill_kernel<true, 512><<<Ch.total, 512, 0, st>>>(D, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C);
    else
        bulk_fill_kernel<false, 256><<<Ch.total, 256, 0, st>>>(D, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C);
    return hipGetLastError();
}

}  // namespace dyc
