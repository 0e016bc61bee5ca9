// dyc_bulk.hip -- bulk insert by bucket fill (insert_order 3).
//
// The eviction insert spends one random 64-bit CAS per exchange, and the device's CAS rate is
// the insert's ceiling however local the accesses are. A large batch can place most of its keys
// without any global atomic instead:
//   1. pack: each key's target bucket (the start rule's table, cuckoo.cu:336-341) names a
//      chunk of kChunkBuckets consecutive buckets of one subtable; write the chunk ids and count
//      their lowest digit per block;
//   2. partition the entries by chunk (dyc_part.hpp: a least-significant-digit radix partition,
//      8 bits of chunk id per pass, with no wait between blocks), which also gives the first
//      index of every chunk;
//   3. (the first index of every chunk comes with the partition's last pass);
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

#include "dyc_kernels.hpp"
#include "dyc_part.hpp"
#include "dyc_wave.hpp"

namespace dyc {

namespace {

// Chunk ids of the batch in input order, and each block's counts of digit P (the partition's
// first pass). Round 1 (IN_ENTRIES false): keys, target = fill_target (pos1, the bucket a find
// reads first, for as many keys as equal subtable shares allow); reserved keys go to chunk 0 and
// are rejected by its fill. Round 2 (IN_ENTRIES true): overflow entries, target = the other
// bucket of the pair.
template <bool IN_ENTRIES, typename CidT>
__global__ __launch_bounds__(kCountThreads) void bulk_pack_kernel(Desc D, Chunks Ch, const uint32_t *__restrict__ keys,
                                                                  const uint64_t *__restrict__ in, PartGrid G,
                                                                  RadixPass P, CidT *__restrict__ cid,
                                                                  uint32_t *__restrict__ partial)
{
    __shared__ uint32_t h[kMaxRadix];
    // per subtable: buckets, FastDiv multiplier and shifts, first chunk -- one LDS read per key
    // in place of a selection chain over the descriptor's arrays for each of them
    __shared__ uint4 sub_tab[kMaxSubtables];
    __shared__ uint32_t tsel[32];  // the round's subtable by key parity and pair (fill_round_sub)
    const uint32_t R = 1u << P.bits;
    for (uint32_t d = threadIdx.x; d < R; d += kCountThreads) h[d] = 0;
    if (threadIdx.x < 32) tsel[threadIdx.x] = fill_round_sub(D, threadIdx.x, IN_ENTRIES);
    if (threadIdx.x < kMaxSubtables) {
        const uint32_t t = threadIdx.x;
        sub_tab[t] = make_uint4(selt(t, D.nb), selt(t, D.m), selt(t, D.sh), chunk_base(Ch, t));
    }
    __syncthreads();
    const uint64_t lo = count_begin(G, blockIdx.x);
    const uint64_t hi = count_end(G, lo);
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)kCountThreads * kCountItems) {
        uint32_t k[kCountItems];
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i < hi) {
                if constexpr (IN_ENTRIES) k[u] = key_of(in[i]);
                else k[u] = keys[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i >= hi) continue;
            uint32_t c = 0;
            if (IN_ENTRIES || !reserved_key(k[u])) {
                const uint32_t p = pair_of(D, k[u]);
                const uint32_t t = tsel[((k[u] & 1u) << 4) | p];
                const uint4 q = sub_tab[t];
                c = q.w + (fast_mod_sh(hash_t(k[u], t), q.x, q.y, q.z) >> kChunkShift);
            }
            cid[i] = (CidT)c;
            atomicAdd(&h[digit_of(c, P)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kCountThreads) partial[(uint64_t)blockIdx.x * R + d] = h[d];
}

// One workgroup per chunk. Buckets of the chunk are written only by this workgroup during the
// launch (every entry of the chunk targets one of them), so slots are taken with LDS counters.
// DENSE (a round with at least two entries per bucket): the chunk's 512 buckets are copied into
// LDS, filled there and written back whole, so every global access is a coalesced 16-byte
// load or store. Otherwise the entries are counted first, the free-slot masks of the buckets
// that received any are read coalesced (8 lanes per bucket, folded by ballot), and each placed
// entry is one store into its slot.
// The fill of one chunk by the whole workgroup: buckets [b0, b0 + len) of subtable S,
// entries ent[lo, hi) (every one targets S). Entries that do not fit are appended to
// ovf[0, ovf_cap) at the index *ovf_n hands out (one atomic for the chunk's whole excess).
// Placements and rejected reserved keys are added to this lane's `placed` and `rejected`.
// Every thread of the workgroup calls it with the same arguments.
template <bool DENSE, int THREADS>
__device__ inline void fill_chunk(const SubRef &S, uint32_t b0, uint32_t len,
                                  const uint64_t *__restrict__ ent, uint32_t lo, uint32_t hi,
                                  uint64_t *__restrict__ ovf, uint64_t ovf_cap, unsigned long long *__restrict__ ovf_n,
                                  bool zeroed, unsigned long long &placed, unsigned long long &rejected)
{
    constexpr uint32_t kImg = DENSE ? kChunkBuckets * kSlots : 2;
    __shared__ uint64_t img[kImg];           // DENSE: the chunk's buckets
    __shared__ uint32_t fm[kChunkBuckets];   // free-slot mask of each bucket
    __shared__ uint32_t cnt[kChunkBuckets];  // entries per bucket, then slots taken
    __shared__ uint32_t wred[THREADS / kWave];
    __shared__ unsigned long long obase;
    __shared__ uint32_t onext;
    uint64_t *const sub = S.keys;
    const uint4 *const src = reinterpret_cast<const uint4 *>(sub + (uint64_t)b0 * kSlots);
    constexpr uint32_t kParts = kChunkBuckets * kSlots / 2;  // 16-byte parts of the chunk
    if constexpr (DENSE) {
        uint4 *const im = reinterpret_cast<uint4 *>(img);
        if (zeroed) {
            // every slot of the table is known to be zero (the caller's host state says so):
            // the chunk's image starts empty in LDS without reading it
            for (uint32_t q = threadIdx.x; q < kParts; q += THREADS) im[q] = make_uint4(0u, 0u, 0u, 0u);
            for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += THREADS) {
                fm[j] = j < len ? 0xffffu : 0u;  // past the end: full
                cnt[j] = 0;
            }
        } else {
            // the chunk into LDS; a bucket's 8 parts sit in 8 consecutive lanes, whose two ballot
            // bits per part make its free-slot mask (no strided LDS reads)
            constexpr uint32_t kPer = kParts / THREADS;
            uint4 v[kPer];
#pragma unroll
            for (uint32_t u = 0; u < kPer; ++u) {
                const uint32_t q = threadIdx.x + u * THREADS;
                v[u] = q < len * (kSlots / 2) ? src[q] : make_uint4(~0u, ~0u, ~0u, ~0u);  // past the end: full
            }
            const uint32_t lane = __lane_id();
#pragma unroll
            for (uint32_t u = 0; u < kPer; ++u) {
                const uint32_t q = threadIdx.x + u * THREADS;
                im[q] = v[u];
                const uint64_t m0 = __ballot((v[u].x | v[u].y) == 0u), m1 = __ballot((v[u].z | v[u].w) == 0u);
                if ((lane & 7u) == 0) {
                    const uint32_t a = (uint32_t)(m0 >> lane) & 0xffu, b = (uint32_t)(m1 >> lane) & 0xffu;
                    uint32_t m = 0;
#pragma unroll
                    for (int p = 0; p < 8; ++p) m |= (((a >> p) & 1u) << (2 * p)) | (((b >> p) & 1u) << (2 * p + 1));
                    fm[q / 8] = m;
                    cnt[q / 8] = 0;
                }
            }
        }
    } else {
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) cnt[j] = 0;
    }
    if (threadIdx.x == 0) onext = 0;
    __syncthreads();
    // count: entries per bucket (an entry whose bucket falls outside the chunk, which the
    // pack rule excludes, would go to the overflow list). Entries are loaded kBatch at a
    // time before any is used, so a thread waits for memory once per batch. A chunk with at
    // most kHold batches per thread keeps them in registers for the placement, so each entry
    // is read from memory once.
    constexpr uint32_t kBatch = 8;
    constexpr uint32_t kHold = DENSE && THREADS < 1024 ? 2 : 1;  // about 8K entries of a dense chunk
    const bool held = hi - lo <= kHold * kBatch * THREADS;  // block-uniform
    uint64_t hb[kHold][kBatch];
    // held entries also keep their place from the count: the bucket (9 bits) and the rank
    // among the chunk's entries of that bucket (the count's atomic returns it), or kReject /
    // kStray. A bucket's first popcount(free mask) ranks take its free slots.
    constexpr uint32_t kReject = ~0u, kStray = ~0u - 1u;
    static_assert(kChunkBuckets <= 512 && kHold * kBatch * THREADS < (1u << 22), "rank and bucket fit 32 bits");
    uint32_t hp[kHold][kBatch];
    uint32_t stray = 0;
    auto load_batch = [&](uint32_t i0, uint64_t (&eb)[kBatch]) {
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t i = i0 + u * THREADS;
            eb[u] = i < hi ? ent[i] : 0ull;
        }
    };
    auto count_batch = [&](uint32_t i0, const uint64_t (&eb)[kBatch]) {
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t k = key_of(eb[u]);
            if (i0 + u * THREADS >= hi || reserved_key(k)) continue;
            const uint32_t b = bucket_of_uniform(S, k);
            if (b >= b0 && b - b0 < len) atomicAdd(&cnt[b - b0], 1u);
            else ++stray;
        }
    };
    if (held) {
#pragma unroll
        for (uint32_t h = 0; h < kHold; ++h) load_batch(lo + threadIdx.x + h * kBatch * THREADS, hb[h]);
#pragma unroll
        for (uint32_t h = 0; h < kHold; ++h) {
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t k = key_of(hb[h][u]);
                uint32_t at = kReject;  // also past the range (never used)
                if (lo + threadIdx.x + (h * kBatch + u) * THREADS < hi && !reserved_key(k)) {
                    const uint32_t b = bucket_of_uniform(S, k);
                    if (b >= b0 && b - b0 < len) {
                        at = (atomicAdd(&cnt[b - b0], 1u) << 9) | (b - b0);
                    } else {
                        at = kStray;
                        ++stray;
                    }
                }
                hp[h][u] = at;
            }
        }
    } else {
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBatch * THREADS) {
            uint64_t eb[kBatch];
            load_batch(i0, eb);
            count_batch(i0, eb);
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
    if (!held)  // the placement counts again
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) cnt[j] = 0;
    __syncthreads();
    auto overflow = [&](uint64_t e) {
        // the counted excess reserved exactly these places; the list holds every entry of the
        // batch, so the check fails only if the chunk ranges overlapped
        const unsigned long long at = obase + atomicAdd(&onext, 1u);
        if (at < ovf_cap) ovf[at] = e;
    };
    // place
    auto place_batch = [&](uint32_t i0, const uint64_t (&eb)[kBatch]) {
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (i0 + u * THREADS >= hi) continue;
            const uint64_t e = eb[u];
            const uint32_t k = key_of(e);
            if (reserved_key(k)) {
                ++rejected;
                continue;
            }
            const uint32_t b = bucket_of_uniform(S, k);
            bool fits = false;
            if (b >= b0 && b - b0 < len) {
                const uint32_t j = b - b0;
                const uint32_t r = atomicAdd(&cnt[j], 1u);
                const uint32_t m = fm[j];
                if (r < (uint32_t)__builtin_popcount(m)) {
                    const uint32_t slot = nth_set_bit16(m, r);
                    if constexpr (DENSE) img[j * kSlots + slot] = e;
                    else sub[(uint64_t)b * kSlots + slot] = e;
                    fits = true;
                    ++placed;
                }
            }
            if (!fits) overflow(e);
        }
    };
    if (held) {
#pragma unroll
        for (uint32_t h = 0; h < kHold; ++h) {
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                if (lo + threadIdx.x + (h * kBatch + u) * THREADS >= hi) continue;
                const uint32_t at = hp[h][u];
                if (at == kReject) {
                    ++rejected;
                    continue;
                }
                if (at != kStray) {
                    const uint32_t j = at & 511u, r = at >> 9;
                    const uint32_t m = fm[j];
                    if (r < (uint32_t)__builtin_popcount(m)) {
                        const uint32_t slot = nth_set_bit16(m, r);
                        if constexpr (DENSE) img[j * kSlots + slot] = hb[h][u];
                        else sub[(uint64_t)(b0 + j) * kSlots + slot] = hb[h][u];
                        ++placed;
                        continue;
                    }
                }
                overflow(hb[h][u]);
            }
        }
    } else {
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBatch * THREADS) {
            uint64_t eb[kBatch];
            load_batch(i0, eb);
            place_batch(i0, eb);
        }
    }
    if constexpr (DENSE) {
        __syncthreads();
        uint4 *const dst = reinterpret_cast<uint4 *>(sub + (uint64_t)b0 * kSlots);
        const uint4 *const im = reinterpret_cast<const uint4 *>(img);
        for (uint32_t q = threadIdx.x; q < len * (kSlots / 2); q += blockDim.x) dst[q] = im[q];
    }
}

// One workgroup per chunk of one table (chunks numbered subtable by subtable, Ch); the entries
// of chunk c are ent[start[c], start[c + 1]).
template <bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void bulk_fill_kernel(Desc D, Chunks Ch, const uint32_t *__restrict__ start,
                                                        const uint64_t *__restrict__ ent, uint64_t n,
                                                        uint64_t *__restrict__ ovf, uint64_t ovf_cap,
                                                        unsigned long long *__restrict__ ovf_n,
                                                        Counters *__restrict__ C, bool zeroed)
{
    const uint32_t c = blockIdx.x;
    // clamped to the entries, so a range is never read past them whatever the bounds hold
    const uint32_t hi = (uint32_t)min((uint64_t)start[c + 1], n);
    const uint32_t lo = min(start[c], hi);
    unsigned long long placed = 0, rejected = 0;
    if (lo < hi) {  // block-uniform
        uint32_t t = 0;
        for (uint32_t q = 1; q < D.d; ++q)
            if (Ch.base[q] <= c) t = q;
        const uint32_t b0 = (c - chunk_base(Ch, t)) << kChunkShift;
        const uint32_t len = min(kChunkBuckets, selt(t, D.nb) - b0);
        fill_chunk<DENSE, THREADS>(sub_ref(D, t), b0, len, ent, lo, hi, ovf, ovf_cap, ovf_n, zeroed, placed,
                                   rejected);
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

namespace {

// The 8-byte entries' partition: 4096-entry tiles (256 threads), two blocks per CU.
constexpr int kBulkSortThreads = 256;
constexpr uint32_t kBulkSortTile = kBulkSortThreads * kSortItems;

template <typename CidT>
hipError_t bulk_partition(const Desc &D, const Chunks &Ch, const uint32_t *keys, const uint32_t *vals, int in,
                          uint64_t n, const BulkSpace &ws, uint32_t max_blocks, hipStream_t st, int *out)
{
    RadixPass P[kMaxRadixPasses];
    const uint32_t np = radix_passes(Ch.bits, P);
    const PartGrid G = part_grid(n, max_blocks, kBulkSortTile);
    CidT *const cid[2] = {static_cast<CidT *>(ws.cid[0]), static_cast<CidT *>(ws.cid[1])};
    uint64_t *const ent[2] = {static_cast<uint64_t *>(ws.ent[0]), static_cast<uint64_t *>(ws.ent[1])};
    if (keys) {
        bulk_pack_kernel<false, CidT><<<G.blocks * G.sub, kCountThreads, 0, st>>>(D, Ch, keys, nullptr, G, P[0], cid[0],
                                                                          ws.partial);
        return partition_passes<uint64_t, KeyValSource, CidT, kBulkSortThreads>(
            KeyValSource{keys, vals}, 0, ent, cid, G, P, np, ws.partial, ws.total, ws.start, Ch.total, st, out);
    }
    bulk_pack_kernel<true, CidT><<<G.blocks * G.sub, kCountThreads, 0, st>>>(D, Ch, nullptr, ent[in], G, P[0], cid[0],
                                                                     ws.partial);
    return partition_passes<uint64_t, EntrySource<uint64_t>, CidT, kBulkSortThreads>(
        EntrySource<uint64_t>{ent[in]}, 1 - in, ent, cid, G, P, np, ws.partial, ws.total, ws.start, Ch.total, st, out);
}

}  // namespace

size_t bulk_cid_bytes(const Chunks &Ch) { return Ch.total <= 65536u ? 2 : 4; }

hipError_t launch_bulk_partition(const Desc &D, const Chunks &Ch, const uint32_t *keys, const uint32_t *vals, int in,
                                 uint64_t n, const BulkSpace &ws, uint32_t max_blocks, hipStream_t st, int *out)
{
    *out = keys ? 0 : 1 - in;
    if (n == 0) return hipSuccess;
    if (n >= (1ull << 32) || Ch.total == 0 || Ch.bits > kMaxRadixPasses * kMaxRadixBits) return hipErrorInvalidValue;
    if (bulk_cid_bytes(Ch) == 2)
        return bulk_partition<uint16_t>(D, Ch, keys, vals, in, n, ws, max_blocks, st, out);
    return bulk_partition<uint32_t>(D, Ch, keys, vals, in, n, ws, max_blocks, st, out);
}

// ------------------------------------------------------------------------------ many tables
// The group's batch is grouped by table already, so one pass partitions each table's entries by
// its own chunks (an 8-bit local chunk id), and the fill, the overflow regions and the per-table
// counts follow the table of each chunk.

namespace {

__device__ inline uint32_t sub_chunks(uint32_t nb) { return (nb + kChunkBuckets - 1) >> kChunkShift; }

template <bool IN_ENTRIES>
__global__ __launch_bounds__(kCountThreads) void group_bulk_pack_kernel(const Desc *__restrict__ dir,
                                                                        const GroupItem *__restrict__ blocks,
                                                                        const uint32_t *__restrict__ keys,
                                                                        const uint64_t *__restrict__ in,
                                                                        uint8_t *__restrict__ lcid,
                                                                        uint32_t *__restrict__ partial)
{
    __shared__ uint32_t h[kGroupBulkChunks];
    // per subtable: buckets, FastDiv multiplier and shifts, the table's first local chunk of it
    __shared__ uint4 sub_tab[kMaxSubtables];
    __shared__ uint32_t tsel[32];  // the round's subtable by key parity and pair (fill_round_sub)
    const GroupItem item = blocks[blockIdx.x];
    if (item.len == 0) {  // block-uniform: a clipped block of round 2 counts nothing
        for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads)
            partial[(uint64_t)blockIdx.x * kGroupBulkChunks + d] = 0;
        return;
    }
    const Desc D = dir[item.table];  // by value: loaded once, not per use through memory
    for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads) h[d] = 0;
    if (threadIdx.x < 32) tsel[threadIdx.x] = fill_round_sub(D, threadIdx.x, IN_ENTRIES);
    if (threadIdx.x < kMaxSubtables) {
        const uint32_t t = threadIdx.x;
        uint32_t first = 0;
        for (uint32_t q = 0; q < t && q < D.d; ++q) first += sub_chunks(selt(q, D.nb));
        sub_tab[t] = make_uint4(selt(t, D.nb), selt(t, D.m), selt(t, D.sh), first);
    }
    __syncthreads();
    const uint64_t lo = item.begin, hi = item.begin + item.len;
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)kCountThreads * kCountItems) {
        uint32_t k[kCountItems];
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if constexpr (IN_ENTRIES) k[u] = i < hi ? key_of(in[i]) : 0u;
            else k[u] = i < hi ? keys[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i >= hi) continue;
            uint32_t lc = 0;
            if (IN_ENTRIES || !reserved_key(k[u])) {
                const uint32_t p = pair_of(D, k[u]);
                const uint32_t t = tsel[((k[u] & 1u) << 4) | p];
                const uint4 q = sub_tab[t];
                lc = q.w + (fast_mod_sh(hash_t(k[u], t), q.x, q.y, q.z) >> kChunkShift);
            }
            lcid[i] = (uint8_t)lc;  // < kGroupBulkChunks: the host admits tables of at most that many
            atomicAdd(&h[lc & (kGroupBulkChunks - 1)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads)
        partial[(uint64_t)blockIdx.x * kGroupBulkChunks + d] = h[d];
}

__device__ inline const Desc &desc_of(const Desc &D) { return D; }
__device__ inline const Desc &desc_of(const WideDesc &W) { return W.D; }
__device__ inline SubRef sub_of(const Desc &D, uint32_t t) { return sub_ref(D, t); }
__device__ inline SubRef sub_of(const WideDesc &W, uint32_t t) { return sub_ref(W.D, t, selt(t, W.val)); }

// One workgroup per table, one thread per local chunk.
template <typename Dir>
__global__ __launch_bounds__(kGroupBulkChunks) void group_bulk_scan_kernel(const Dir *__restrict__ dir,
                                                                           const BulkTable *__restrict__ tabs,
                                                                           uint32_t *__restrict__ partial,
                                                                           GroupChunk *__restrict__ chunks)
{
    __shared__ uint32_t tmp[kScanWaves];
    const BulkTable T = tabs[blockIdx.x];
    const uint32_t d = threadIdx.x;
    uint32_t run = 0;  // this chunk's entries in the table's blocks before block b
    for (uint32_t b = T.first_block; b < T.first_block + T.blocks; ++b) {
        uint32_t *const p = partial + (uint64_t)b * kGroupBulkChunks + d;
        const uint32_t x = *p;
        *p = run;
        run += x;
    }
    uint32_t all;
    const uint32_t base = (uint32_t)T.off + block_exclusive_scan(run, tmp, &all);  // the batch has < 2^32 entries
    for (uint32_t b = T.first_block; b < T.first_block + T.blocks; ++b) partial[(uint64_t)b * kGroupBulkChunks + d] += base;
    if (d < T.chunks) {
        // local chunk d: subtable t, chunk lc of it (T.chunks sums the subtables' chunks)
        const Dir E = dir[T.table];
        const Desc &D = desc_of(E);
        uint32_t lc = d, t = 0;
        for (; t + 1 < D.d; ++t) {
            const uint32_t k = sub_chunks(selt(t, D.nb));
            if (lc < k) break;
            lc -= k;
        }
        GroupChunk R;
        R.S = sub_of(E, t);
        R.ovf_off = T.off;
        R.ovf_cap = T.n;
        R.b0 = lc << kChunkShift;
        R.len = min(kChunkBuckets, R.S.nb - R.b0);
        R.lo = base;
        R.cnt = run;
        R.j = blockIdx.x;
        R.table = T.table;
        chunks[T.chunk0 + d] = R;
    }
}

template <bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void group_bulk_fill_kernel(const GroupChunk *__restrict__ chunks,
                                                                  const uint64_t *__restrict__ ent,
                                                                  uint64_t *__restrict__ ovf,
                                                                  unsigned long long *__restrict__ ovf_n,
                                                                  unsigned long long *__restrict__ table_counts,
                                                                  Counters *__restrict__ C, bool zeroed)
{
    const GroupChunk R = chunks[blockIdx.x];
    unsigned long long placed = 0, rejected = 0;
    if (R.cnt)  // block-uniform
        fill_chunk<DENSE, THREADS>(R.S, R.b0, R.len, ent, R.lo, R.lo + R.cnt, ovf + R.ovf_off, R.ovf_cap, ovf_n + R.j,
                                   zeroed, placed, rejected);
    const unsigned long long wp = wave_sum(placed);
    if (__lane_id() == 0 && wp) atomicAdd(&table_counts[R.table], wp);
    wave_tally(C, &C->kicks, placed);
    wave_tally(C, &C->rejected, rejected);
}

__global__ __launch_bounds__(256) void group_bulk_clip_kernel(const GroupItem *__restrict__ blocks,
                                                              const uint32_t *__restrict__ btab, uint32_t n_blocks,
                                                              const BulkTable *__restrict__ tabs, uint32_t n_tabs,
                                                              const unsigned long long *__restrict__ ovf_n,
                                                              GroupItem *__restrict__ out, BulkTable *__restrict__ tabs_out)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n_blocks) {
        const GroupItem b = blocks[x];
        const uint32_t j = btab[x];
        const uint64_t lim = tabs[j].off + min(ovf_n[j], (unsigned long long)tabs[j].n);
        const uint64_t len = b.begin < lim ? min((uint64_t)b.len, lim - b.begin) : 0ull;
        out[x] = GroupItem{b.begin, (uint32_t)len, b.table};
    }
    if (tabs_out && x < n_tabs) {
        BulkTable T = tabs[x];
        T.n = (uint32_t)min(ovf_n[x], (unsigned long long)T.n);
        tabs_out[x] = T;
    }
}

}  // namespace

hipError_t launch_group_bulk_clip(const GroupItem *blocks, const uint32_t *btab, uint32_t n_blocks,
                                  const BulkTable *tabs, uint32_t n_tabs, const unsigned long long *ovf_n,
                                  GroupItem *out, BulkTable *tabs_out, hipStream_t st)
{
    const uint32_t n = std::max(n_blocks, tabs_out ? n_tabs : 0u);
    if (n == 0) return hipSuccess;
    group_bulk_clip_kernel<<<(n + 255) / 256, 256, 0, st>>>(blocks, btab, n_blocks, tabs, n_tabs, ovf_n, out, tabs_out);
    return hipGetLastError();
}

hipError_t launch_group_bulk_pack(const Desc *dir, const GroupItem *blocks, uint32_t n_blocks, const uint32_t *keys,
                                  const uint64_t *entries, uint8_t *lcid, uint32_t *partial, hipStream_t st)
{
    if (n_blocks == 0) return hipSuccess;
    if (entries)
        group_bulk_pack_kernel<true><<<n_blocks, kCountThreads, 0, st>>>(dir, blocks, nullptr, entries, lcid, partial);
    else
        group_bulk_pack_kernel<false><<<n_blocks, kCountThreads, 0, st>>>(dir, blocks, keys, nullptr, lcid, partial);
    return hipGetLastError();
}

template <typename Dir>
static hipError_t group_bulk_scan(const Dir *dir, const BulkTable *tabs, uint32_t n_tabs, uint32_t *partial,
                                  GroupChunk *chunks, hipStream_t st)
{
    if (n_tabs == 0) return hipSuccess;
    group_bulk_scan_kernel<Dir><<<n_tabs, kGroupBulkChunks, 0, st>>>(dir, tabs, partial, chunks);
    return hipGetLastError();
}

hipError_t launch_group_bulk_scan(const Desc *dir, const BulkTable *tabs, uint32_t n_tabs, uint32_t *partial,
                                  GroupChunk *chunks, hipStream_t st)
{
    return group_bulk_scan(dir, tabs, n_tabs, partial, chunks, st);
}

hipError_t launch_group_bulk_scan(const WideDesc *dir, const BulkTable *tabs, uint32_t n_tabs, uint32_t *partial,
                                  GroupChunk *chunks, hipStream_t st)
{
    return group_bulk_scan(dir, tabs, n_tabs, partial, chunks, st);
}

hipError_t launch_group_bulk_scatter(const GroupItem *blocks, uint32_t n_blocks, const uint32_t *keys,
                                     const uint32_t *vals, const uint64_t *entries, const uint8_t *lcid,
                                     const uint32_t *partial, uint64_t n, uint64_t *ent, hipStream_t st)
{
    if (n_blocks == 0) return hipSuccess;
    if (n >= (1ull << 32)) return hipErrorInvalidValue;
    const SegLayout lay{blocks, partial, n};
    uint8_t *const no_cid = nullptr;
    if (entries)
        part_scatter_kernel<uint64_t, EntrySource<uint64_t>, uint8_t, kBulkSortThreads, -1, SegLayout>
            <<<n_blocks, kBulkSortThreads, 0, st>>>(EntrySource<uint64_t>{entries}, lcid, lay, RadixPass{0, 8}, ent,
                                                    no_cid);
    else
        part_scatter_kernel<uint64_t, KeyValSource, uint8_t, kBulkSortThreads, -1, SegLayout>
            <<<n_blocks, kBulkSortThreads, 0, st>>>(KeyValSource{keys, vals}, lcid, lay, RadixPass{0, 8}, ent, no_cid);
    return hipGetLastError();
}

hipError_t launch_group_bulk_fill(const GroupChunk *chunks, uint32_t n_chunks, const uint64_t *ent, uint64_t *ovf,
                                  unsigned long long *ovf_n, unsigned long long *table_counts, Counters *C, bool dense,
                                  bool zeroed, hipStream_t st)
{
    if (n_chunks == 0) return hipSuccess;
    if (dense)
        group_bulk_fill_kernel<true, 1024>
            <<<n_chunks, 1024, 0, st>>>(chunks, ent, ovf, ovf_n, table_counts, C, zeroed);
    else
        group_bulk_fill_kernel<false, 256><<<n_chunks, 256, 0, st>>>(chunks, ent, ovf, ovf_n, table_counts, C, false);
    return hipGetLastError();
}

hipError_t launch_bulk_fill(const Desc &D, const Chunks &Ch, const uint32_t *start, const uint64_t *ent, uint64_t n,
                            uint64_t *ovf, uint64_t ovf_cap, unsigned long long *ovf_n, Counters *C, bool zeroed,
                            hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (n >= (1ull << 32) || Ch.total == 0) return hipErrorInvalidValue;
    uint64_t buckets = 0;
    for (uint32_t t = 0; t < D.d; ++t) buckets += D.nb[t];
    // dense: copying every chunk through LDS costs less than scattered stores; its 68 KB of LDS
    // allow two workgroups per CU, so they are 1024 threads wide (48 VGPRs: 8 waves per SIMD) to
    // keep more loads in flight (7.50 -> 7.40 ms for the 2^28-key insert against 512 threads)
    if (n >= 2 * buckets)
        bulk_fill_kernel<true, 1024><<<Ch.total, 1024, 0, st>>>(D, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C, zeroed);
    else
        bulk_fill_kernel<false, 256><<<Ch.total, 256, 0, st>>>(D, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C, false);
    return hipGetLastError();
}

}  // namespace dyc
