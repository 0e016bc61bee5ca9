// dyc_part.hpp -- partition of a bulk-fill batch by table chunk (internal to libdyc).
//
// The bulk fill (dyc_bulk.hip, dyc_wide.hip) needs its batch grouped by chunk: every entry of
// chunk c contiguous, and the first index of every chunk. This is a least-significant-digit
// radix partition of the entries by chunk id, at most 8 bits of the chunk id per pass:
//   pack   (in the table's source file): one pass over the batch computes every entry's chunk
//          id -- the hash is the expensive part, so it is done once -- writes the ids in input
//          order and counts the first digit per block;
//   colscan: per digit, the exclusive prefix of the per-block counts over the blocks, and the
//          digit's total;
//   scatter: each block ranks its range tile by tile (stably: by digit, then input index) and
//          writes each tile's entries to their global places, digit run by digit run, so that
//          consecutive lanes store consecutive words. The first pass reads the caller's batch
//          itself; the last writes no chunk ids;
//   count:  the next digit per block, from the chunk ids alone. Before the last pass it also
//          counts every chunk: its input is then sorted by the bits below the last digit, so a
//          block's range holds few of their values and the counts gather in LDS;
//   segment scan: the first index of every chunk from those counts.
// Blocks own fixed contiguous ranges of whole tiles, the same in every launch of a pass, so the
// prefix of block b is final before its scatter starts: no launch waits on another block.
// Chunk ids are carried as 16-bit words when the table has at most 2^16 chunks.
#pragma once

#include <algorithm>

#include "dyc_kernels.hpp"
#include "dyc_wave.hpp"

namespace dyc {

constexpr int kSortItems = 16;  // entries per thread per scatter tile
constexpr uint32_t kMaxRadixBits = 8;
constexpr uint32_t kMaxRadix = 1u << kMaxRadixBits;
constexpr uint32_t kMaxRadixPasses = 4;
constexpr uint32_t kMaxSortBlocks = 2048;  // colscan gives each of its 256 threads up to 8 blocks
static_assert(kPartialWords == (uint64_t)kMaxSortBlocks * kMaxRadix, "BulkSpace::partial holds every block's counts");
static_assert(kMaxRadixWords == kMaxRadix, "BulkSpace::total holds every digit's total");
// The counting launches (pack, count) stream their block's range with more waves per block
// than the scatter, whose LDS limits it to one or two blocks per CU.
constexpr int kCountThreads = 512;
constexpr int kCountItems = 8;
// Values of the bits below the last digit that one block of the last count bins in LDS; a
// block's entries beyond them (small batches) count with global atomics.
constexpr uint32_t kLowSpan = 8;

struct RadixPass {
    uint32_t shift, bits;  // digit = (cid >> shift) & ((1 << bits) - 1)
};

// Scatter block b of a pass owns tiles [b * per, (b + 1) * per) of the batch, `tile` entries each
// (the last ones may be partial or empty). The counting launches split each block's range into
// `sub` equal parts, one counting block each (their count rows b * sub + s follow the batch's
// order, so the exclusive prefix of row b * sub is scatter block b's): they need more waves than
// the scatter's LDS leaves room for.
struct PartGrid {
    uint64_t n;
    uint32_t blocks, per, tile, sub;
};
constexpr uint32_t kCountBlocks = 1024;  // counting blocks aimed for

inline PartGrid part_grid(uint64_t n, uint32_t max_blocks, uint32_t tile)
{
    const uint64_t tiles = (n + tile - 1) / tile;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(tiles, 1), std::min(max_blocks, kMaxSortBlocks));
    const uint32_t per = (uint32_t)std::max<uint64_t>((tiles + blocks - 1) / blocks, 1);
    blocks = (uint32_t)std::max<uint64_t>((tiles + per - 1) / per, 1);  // no block without a tile
    uint32_t sub = std::max<uint32_t>(1, std::min(kCountBlocks, kMaxSortBlocks) / blocks);
    while (sub > 1 && (tile * per) % sub) --sub;  // equal parts
    return PartGrid{n, blocks, per, tile, sub};
}

// The digit passes of a partition over `bits` bits of chunk id: as few passes as 8-bit digits
// allow, with the bits spread evenly over them (lowest digit first). Returns the pass count.
inline uint32_t radix_passes(uint32_t bits, RadixPass *out)
{
    const uint32_t n = std::max<uint32_t>(1, (bits + kMaxRadixBits - 1) / kMaxRadixBits);
    uint32_t shift = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t b = (bits - shift + (n - p) - 1) / (n - p);
        out[p] = RadixPass{shift, std::max<uint32_t>(b, 1)};
        shift += b;
    }
    return n;
}

__device__ inline uint32_t digit_of(uint32_t cid, RadixPass P) { return (cid >> P.shift) & ((1u << P.bits) - 1u); }

// First and one past the last index of scatter block b's range.
__device__ inline uint64_t part_begin(const PartGrid &G, uint32_t b) { return (uint64_t)b * G.per * G.tile; }
__device__ inline uint64_t part_end(const PartGrid &G, uint64_t lo)
{
    const uint64_t hi = lo + (uint64_t)G.per * G.tile;
    return hi < G.n ? hi : G.n;
}
// The same for counting block q (part q % sub of scatter block q / sub).
__device__ inline uint64_t count_begin(const PartGrid &G, uint32_t q) { return (uint64_t)q * (G.per * G.tile / G.sub); }
__device__ inline uint64_t count_end(const PartGrid &G, uint64_t lo)
{
    const uint64_t hi = lo + G.per * G.tile / G.sub;
    return hi < G.n ? hi : G.n;
}

// Where a scatter pass reads its entries: an entry array, or the caller's key and value arrays
// (the first pass of an insert batch).
template <typename Entry>
struct EntrySource {
    const Entry *e;
    __device__ Entry load(uint64_t i) const { return e[i]; }
};
struct KeyValSource {
    const uint32_t *k, *v;
    __device__ uint64_t load(uint64_t i) const { return pack(k[i], v[i]); }
};
struct WideKeyValSource {
    const uint64_t *k, *v;
    __device__ WideEntry load(uint64_t i) const { return WideEntry{k[i], v[i]}; }
};

namespace {

// Exclusive prefix sum over the block's threads (one value each, at most kScanWaves waves);
// *total receives the sum. `tmp` holds kScanWaves words of LDS. Every thread of the block must
// call it.
constexpr int kScanWaves = 16;
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *tmp, uint32_t *total)
{
    const uint32_t lane = __lane_id(), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == kWave - 1) tmp[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t q = 0; q < nw; ++q) {
        const uint32_t s = tmp[q];
        before += q < w ? s : 0u;
        all += s;
    }
    __syncthreads();  // tmp may be reused by the caller
    *total = all;
    return before + x - v;
}

// Per digit d (one block each): partial[b * R + d] := sum of partial[b' * R + d] over b' < b, and
// total[d] := the sum over all blocks.
__global__ __launch_bounds__(256) void part_colscan_kernel(uint32_t *__restrict__ partial, uint32_t blocks, uint32_t R,
                                                           uint32_t *__restrict__ total)
{
    __shared__ uint32_t tmp[kScanWaves];
    const uint32_t d = blockIdx.x;
    constexpr uint32_t kPer = kMaxSortBlocks / 256;
    const uint32_t b0 = threadIdx.x * kPer;
    uint32_t v[kPer], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        v[j] = b0 + j < blocks ? partial[(uint64_t)(b0 + j) * R + d] : 0u;
        sum += v[j];
    }
    uint32_t all;
    uint32_t run = block_exclusive_scan(sum, tmp, &all);
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        if (b0 + j < blocks) partial[(uint64_t)(b0 + j) * R + d] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) total[d] = all;
}

// Per-block counts of digit P of cid[i] over each block's range: partial[b * R + d]. With
// chunks != nullptr (the last pass, whose input is sorted by the bits of cid below P.shift) it
// also adds every chunk's count to chunks[cid]: a block bins the first kLowSpan values of those
// low bits that its range holds, by (low value, digit), in LDS.
template <typename CidT>
__global__ __launch_bounds__(kCountThreads) void part_count_kernel(const CidT *__restrict__ cid, PartGrid G, RadixPass P,
                                                                   uint32_t *__restrict__ partial,
                                                                   uint32_t *__restrict__ chunks)
{
    __shared__ uint32_t h[kMaxRadix];
    __shared__ uint32_t hc[kLowSpan * kMaxRadix];
    const uint32_t R = 1u << P.bits;
    const uint32_t low_mask = (1u << P.shift) - 1u;
    for (uint32_t d = threadIdx.x; d < R; d += kCountThreads) h[d] = 0;
    if (chunks)
        for (uint32_t x = threadIdx.x; x < kLowSpan * kMaxRadix; x += kCountThreads) hc[x] = 0;
    __syncthreads();
    const uint64_t lo = count_begin(G, blockIdx.x);
    const uint64_t hi = count_end(G, lo);
    const uint32_t l0 = lo < hi ? ((uint32_t)cid[lo] & low_mask) : 0u;
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)kCountThreads * kCountItems) {
        uint32_t c[kCountItems];
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            c[u] = i < hi ? (uint32_t)cid[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            if (i0 + (uint64_t)u * kCountThreads >= hi) continue;
            const uint32_t d = digit_of(c[u], P);
            atomicAdd(&h[d], 1u);
            if (chunks) {
                const uint32_t j = (c[u] & low_mask) - l0;  // the input is sorted by these bits
                if (j < kLowSpan) atomicAdd(&hc[j * kMaxRadix + d], 1u);
                else atomicAdd(&chunks[c[u]], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kCountThreads) partial[(uint64_t)blockIdx.x * R + d] = h[d];
    if (chunks)
        for (uint32_t x = threadIdx.x; x < kLowSpan * kMaxRadix; x += kCountThreads)
            if (hc[x]) atomicAdd(&chunks[((x % kMaxRadix) << P.shift) + l0 + x / kMaxRadix], hc[x]);
}

// start[c] := the first index of chunk c, for c in [0, total], from the chunk counts held in
// start[0, total) (start[total] := their sum). Three launches over segments of kScanSeg counts:
// each segment's sum into seg[], the exclusive prefix of seg[] (one block), then each segment
// rewritten as its prefix plus the running sum of its counts.
constexpr uint32_t kScanSeg = 1024;  // counts per segment: 256 threads x 4
__global__ __launch_bounds__(256) void part_seg_sum_kernel(const uint32_t *__restrict__ start, uint32_t total,
                                                           uint32_t *__restrict__ seg)
{
    __shared__ uint32_t tmp[kScanWaves];
    const uint32_t c0 = blockIdx.x * kScanSeg + threadIdx.x * 4;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) sum += c0 + j < total ? start[c0 + j] : 0u;
    uint32_t all;
    (void)block_exclusive_scan(sum, tmp, &all);
    if (threadIdx.x == 0) seg[blockIdx.x] = all;
}

__global__ __launch_bounds__(1024) void part_seg_scan_kernel(uint32_t *__restrict__ seg, uint32_t nseg,
                                                             uint32_t *__restrict__ start, uint32_t total)
{
    __shared__ uint32_t tmp[kScanWaves];
    const uint32_t per = (nseg + blockDim.x - 1) / blockDim.x;
    const uint32_t s0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t j = 0; j < per; ++j)
        if (s0 + j < nseg) sum += seg[s0 + j];
    uint32_t all;
    uint32_t run = block_exclusive_scan(sum, tmp, &all);
    for (uint32_t j = 0; j < per; ++j) {
        if (s0 + j < nseg) {
            const uint32_t x = seg[s0 + j];
            seg[s0 + j] = run;
            run += x;
        }
    }
    if (threadIdx.x == 0) start[total] = all;
}

__global__ __launch_bounds__(256) void part_seg_write_kernel(uint32_t *__restrict__ start, uint32_t total,
                                                             const uint32_t *__restrict__ seg)
{
    __shared__ uint32_t tmp[kScanWaves];
    const uint32_t c0 = blockIdx.x * kScanSeg + threadIdx.x * 4;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        v[j] = c0 + j < total ? start[c0 + j] : 0u;
        sum += v[j];
    }
    uint32_t all;
    uint32_t run = seg[blockIdx.x] + block_exclusive_scan(sum, tmp, &all);
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        if (c0 + j < total) start[c0 + j] = run;
        run += v[j];
    }
}

// The lanes among vm (a ballot) whose digit equals this lane's d, a digit of `bits` bits (BITS if
// it is positive). Per bit: sb = 0 or ~0 (the lane's bit), and the mask keeps the lanes whose bit
// equals it, m &= ~(ballot ^ sb), on the two 32-bit halves.
template <int BITS>
__device__ inline uint64_t match_mask(uint64_t vm, uint32_t d, uint32_t bits)
{
    uint32_t mlo = (uint32_t)vm, mhi = (uint32_t)(vm >> 32);
    auto step = [&](uint32_t b) {
        const uint32_t sb = (uint32_t)((int32_t)(d << (31u - b)) >> 31);
        const uint64_t bb = __ballot(sb != 0u);  // lanes outside vm drop out through the AND
        mlo &= ~((uint32_t)bb ^ sb);
        mhi &= ~((uint32_t)(bb >> 32) ^ sb);
    };
    if constexpr (BITS > 0) {
#pragma unroll
        for (uint32_t b = 0; b < (uint32_t)BITS; ++b) step(b);
    } else {
        for (uint32_t b = 0; b < bits; ++b) step(b);
    }
    return ((uint64_t)mhi << 32) | mlo;
}

// One pass: the entries of each block's range (from src, chunk ids from cid) go to their places
// in ent_out (chunk ids to cid_out unless it is null), ordered by digit P. partial holds the
// blocks' exclusive prefixes per digit and total the digits' totals (part_colscan_kernel). The
// next tile's loads are issued before the current tile is ranked: a barrier here waits for LDS
// operations only, so they stay in flight while the tile is placed and written out.
// Ranking within a wave (BITS):
//   > 0  stably (by digit, then index), with one ballot per digit bit; BITS = P.bits;
//   0    the same with P.bits read at run time;
//   < 0  not stably: each entry's rank is what its wave's LDS counter of the digit returned
//        when the entry added itself, in whatever order the lanes' atomics were applied. The
//        first pass of a least-significant-digit partition may do this (its input order means
//        nothing); every later pass must keep the order of the previous digits.
// Where block b's entries lie and where its entries of each digit go (Layout):
//   begin(b), end(b, begin): its input range;
//   base(b, d, R, tmp): the global place of its first entry of digit d (every thread of the block
//   calls it, d = threadIdx.x, and uses the result for d < R; tmp: kScanWaves words of LDS);
//   n: the entry count (no place reaches it).
// PassLayout is a digit pass over the whole batch: fixed block ranges (PartGrid), places from
// the per-block digit counts after part_colscan_kernel and the digit totals.
struct PassLayout {
    PartGrid G;
    const uint32_t *partial, *total;
    uint64_t n;
    __device__ uint64_t begin(uint32_t b) const { return part_begin(G, b); }
    __device__ uint64_t end(uint32_t, uint64_t lo) const { return part_end(G, lo); }
    __device__ uint32_t base(uint32_t b, uint32_t d, uint32_t R, uint32_t *tmp) const
    {
        uint32_t all;
        const uint32_t dbase = block_exclusive_scan(d < R ? total[d] : 0u, tmp, &all);
        return d < R ? dbase + partial[(uint64_t)b * G.sub * R + d] : 0u;
    }
};

// A pass over a batch grouped by table (the group bulk fill): block b covers blocks[b] (one
// table's range) and partial holds the place of its first entry of each local chunk.
struct SegLayout {
    const GroupItem *blocks;
    const uint32_t *partial;
    uint64_t n;
    __device__ uint64_t begin(uint32_t b) const { return blocks[b].begin; }
    __device__ uint64_t end(uint32_t b, uint64_t lo) const { return lo + blocks[b].len; }
    __device__ uint32_t base(uint32_t b, uint32_t d, uint32_t R, uint32_t *) const
    {
        return d < R ? partial[(uint64_t)b * kGroupBulkChunks + d] : 0u;
    }
};

template <typename Entry, typename Src, typename CidT, int THREADS, int BITS, typename Layout = PassLayout>
__global__ __launch_bounds__(THREADS) void part_scatter_kernel(Src src, const CidT *__restrict__ cid, Layout Lay,
                                                               RadixPass P, Entry *__restrict__ ent_out,
                                                               CidT *__restrict__ cid_out)
{
    constexpr int kWaves = THREADS / kWave;
    constexpr uint32_t kTile = THREADS * kSortItems;
    static_assert(THREADS >= (int)kMaxRadix && THREADS % kWave == 0, "one thread per digit");
    __shared__ Entry buf[kTile];
    __shared__ CidT cbuf[kTile];
    __shared__ uint32_t wcnt[kWaves][kMaxRadix];  // per-wave digit counts, then wave offsets
    __shared__ uint32_t gbase[kMaxRadix];         // global place of the block's next entry of d
    __shared__ uint32_t gdel[kMaxRadix];          // gbase[d] minus the tile's first place of d
    __shared__ uint32_t loc[kMaxRadix];           // first place of digit d in the tile
    __shared__ uint32_t tmp[kScanWaves];
    const uint32_t R = 1u << P.bits;
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid / kWave;
    const uint64_t lo = Lay.begin(blockIdx.x);
    const uint64_t hi = Lay.end(blockIdx.x, lo);
    if (lo >= hi) return;  // block-uniform: nothing to place (a clipped block of a fill's later round)
    {
        const uint32_t b = Lay.base(blockIdx.x, tid, R, tmp);
        if (tid < R) gbase[tid] = b;
    }
    const uint64_t lanemask = lanemask_lt();
    // wave w holds tile entries w * kSortItems * 64 + u * 64 + lane: stable ranking follows
    // (u, lane) inside the wave, and the waves in order
    const uint32_t wbase = w * kSortItems * kWave;
    Entry e[kSortItems], en[kSortItems];
    uint32_t c[kSortItems], cn[kSortItems];
    auto load_tile = [&](uint64_t t0, Entry *ev, uint32_t *cv) {
#pragma unroll
        for (int u = 0; u < kSortItems; ++u) {
            const uint64_t i = t0 + wbase + u * kWave + lane;
            if (i < hi) {
                ev[u] = src.load(i);
                cv[u] = (uint32_t)cid[i];
            } else {
                cv[u] = 0u;
            }
        }
    };
    if (lo < hi) load_tile(lo, e, c);
    for (uint64_t t0 = lo; t0 < hi; t0 += kTile) {
        const uint32_t tn = hi - t0 < kTile ? (uint32_t)(hi - t0) : kTile;
        if (t0 + kTile < hi) load_tile(t0 + kTile, en, cn);
        for (uint32_t x = tid; x < kWaves * kMaxRadix; x += THREADS) (&wcnt[0][0])[x] = 0;
        __syncthreads();  // wcnt is zero; the previous tile's output is done with buf and gdel
        uint32_t r[kSortItems];
        if constexpr (BITS < 0) {
#pragma unroll
            for (int u = 0; u < kSortItems; ++u)
                r[u] = wbase + u * kWave + lane < tn ? atomicAdd(&wcnt[w][digit_of(c[u], P)], 1u) : 0u;
        } else {
            // the lanes of one digit in an item (their match mask) take consecutive ranks in lane
            // order after the digit's count so far
            const uint32_t bits = P.bits;
#pragma unroll
            for (int u = 0; u < kSortItems; ++u) {
                const bool valid = wbase + u * kWave + lane < tn;
                const uint32_t d = digit_of(c[u], P);
                const uint64_t m = match_mask<BITS>(__ballot(valid), d, bits);
                const uint32_t below = (uint32_t)__popcll(m & lanemask);
                const uint32_t base = valid ? wcnt[w][d] : 0u;
                r[u] = base + below;
                if (valid && below == 0) wcnt[w][d] = base + (uint32_t)__popcll(m);
            }
        }
        __syncthreads();
        // per digit: wave offsets (exclusive over the waves), the tile's first place, and the
        // shift from a tile place to its global place
        {
            uint32_t run = 0;
            if (tid < R) {
#pragma unroll
                for (int q = 0; q < kWaves; ++q) {
                    const uint32_t x = wcnt[q][tid];
                    wcnt[q][tid] = run;
                    run += x;
                }
            }
            uint32_t all;
            const uint32_t first = block_exclusive_scan(tid < R ? run : 0u, tmp, &all);
            if (tid < R) {
                loc[tid] = first;
                gdel[tid] = gbase[tid] - first;
                gbase[tid] += run;  // read again only after the next tile's offsets are made
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kSortItems; ++u) {
            if (wbase + u * kWave + lane < tn) {
                const uint32_t d = digit_of(c[u], P);
                const uint32_t pos = loc[d] + wcnt[w][d] + r[u];  // < tn by construction
                if (pos < kTile) {
                    buf[pos] = e[u];
                    cbuf[pos] = (CidT)c[u];
                }
            }
        }
        __syncthreads();
#pragma unroll 4
        for (uint32_t q = tid; q < tn; q += THREADS) {
            const uint32_t cq = (uint32_t)cbuf[q];
            const uint64_t g = (uint64_t)(uint32_t)(gdel[digit_of(cq, P)] + q);  // < n by construction
            if (g < Lay.n) {
                ent_out[g] = buf[q];
                if (cid_out) cid_out[g] = (CidT)cq;
            }
        }
#pragma unroll
        for (int u = 0; u < kSortItems; ++u) {
            e[u] = en[u];
            c[u] = cn[u];
        }
    }
}

}  // namespace

// The partition's launches after the pack, which has written the chunk ids of the n entries of
// src0 to cid[0] and the counts of the first digit (passes P[0..np)) of the G.blocks * G.sub
// counting blocks to partial. The first pass reads src0 and writes ent[first]; each later pass reads the entry
// buffer the previous one wrote and writes the other. Chunk ids go from cid[p & 1] to
// cid[(p + 1) & 1], except in the last pass. start (chunks + 1 words) receives the first index of
// every chunk. Returns the entry buffer the sorted entries end in.
template <typename Entry, typename Src, typename CidT, int THREADS>
hipError_t partition_passes(Src src0, int first, Entry *const ent[2], CidT *const cid[2], const PartGrid &G,
                            const RadixPass *P, uint32_t np, uint32_t *partial, uint32_t *total, uint32_t *start,
                            uint32_t chunks, hipStream_t st, int *out)
{
    int e = first;
    for (uint32_t p = 0; p < np; ++p) {
        const uint32_t R = 1u << P[p].bits;
        const bool last = p + 1 == np;
        if (last) {  // the chunk counts gather in start[0, chunks) (zeroed as 8-byte words)
            const hipError_t z = launch_zero(reinterpret_cast<unsigned long long *>(start), (chunks + 1) / 2, st);
            if (z != hipSuccess) return z;
        }
        // the first pass's digit counts come with the pack; the last count also counts chunks
        if (p > 0 || last)
            part_count_kernel<CidT><<<G.blocks * G.sub, kCountThreads, 0, st>>>(cid[p & 1], G, P[p], partial,
                                                                                 last ? start : nullptr);
        part_colscan_kernel<<<R, 256, 0, st>>>(partial, G.blocks * G.sub, R, total);
        CidT *const cout = last ? nullptr : cid[(p + 1) & 1];
        // the first pass ranks by atomics; the later ones stably, 8-bit digits unrolled
        const PassLayout lay{G, partial, total, G.n};
        if (p == 0)
            part_scatter_kernel<Entry, Src, CidT, THREADS, -1>
                <<<G.blocks, THREADS, 0, st>>>(src0, cid[0], lay, P[p], ent[e], cout);
        else if (P[p].bits == kMaxRadixBits)
            part_scatter_kernel<Entry, EntrySource<Entry>, CidT, THREADS, kMaxRadixBits><<<G.blocks, THREADS, 0, st>>>(
                EntrySource<Entry>{ent[1 - e]}, cid[p & 1], lay, P[p], ent[e], cout);
        else
            part_scatter_kernel<Entry, EntrySource<Entry>, CidT, THREADS, 0><<<G.blocks, THREADS, 0, st>>>(
                EntrySource<Entry>{ent[1 - e]}, cid[p & 1], lay, P[p], ent[e], cout);
        if (!last) e = 1 - e;
    }
    const uint32_t nseg = (chunks + kScanSeg - 1) / kScanSeg;  // <= kPartialWords: partial is free now
    part_seg_sum_kernel<<<nseg, 256, 0, st>>>(start, chunks, partial);
    part_seg_scan_kernel<<<1, 1024, 0, st>>>(partial, nseg, start, chunks);
    part_seg_write_kernel<<<nseg, 256, 0, st>>>(start, chunks, partial);
    *out = e;
    return hipGetLastError();
}

}  // namespace dyc
