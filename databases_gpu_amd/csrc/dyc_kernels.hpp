// dyc_kernels.hpp -- launchers for the DyCuckoo HIP kernels (internal to libdyc).
#pragma once

#include <stddef.h>

#include "dyc_device.hpp"

namespace dyc {

// Per-call device counters, zeroed before each launch that uses them. This is the base copy; the
// device block (kCounterBytes) holds it, padded to 256 B, followed by kTallyStripes stripes of the
// same 16 words, each 256 B apart. Atomics to one address are serialized at the memory-side atomic
// unit (about 12 ns each on MI355X), so per-wave tallies (kicks, deleted, ...) go to the stripe of
// the wave's slot; append indexes (pending, lost) stay in the base copy, where the device reads
// them back. The host copies the whole block and folds the stripes into the base (fold_tallies).
struct Counters {
    unsigned long long kicks;      // successful exchanges (insert)
    unsigned long long rejected;   // reserved keys in an insert batch
    unsigned long long pending;    // entries appended to the pending list (append index)
    unsigned long long lost;       // appends that found the pending list full
    unsigned long long deleted;    // erase hits
    unsigned long long retries;    // CAS attempts that lost a race
    unsigned long long misplaced;  // resize: entries outside their bucket pair (corrupt image)
    unsigned long long occupied;   // count kernel
    unsigned long long stalls;     // serving tick: dependency waits that hit their bound
    unsigned long long updated;    // upsert: keys whose stored value was replaced
    unsigned long long residuals;  // downsize: entries of the upper half that did not stay
    unsigned long long placed;     // insert with PendSink::count_placed: entries taken in
    unsigned long long max_chain;  // insert: the longest eviction chain (exchanges); folded by max
    unsigned long long pad[3];
};
constexpr int kMaxChainWord = 12;  // index of max_chain: stripes fold into it by max, not by sum
static_assert(offsetof(Counters, max_chain) == kMaxChainWord * sizeof(unsigned long long), "max_chain's word");
constexpr int kTallyStripes = 64;
constexpr int kTallyStrideWords = 32;  // 256 B between stripes
constexpr int kCounterWords = kTallyStrideWords * (1 + kTallyStripes);
constexpr size_t kCounterBytes = (size_t)kCounterWords * sizeof(unsigned long long);
static_assert(sizeof(Counters) <= kTallyStrideWords * sizeof(unsigned long long), "Counters exceeds a stripe");

// Add every stripe of a counter block (kCounterWords words, as copied from the device) into its
// base copy, in place. Call once per copy.
inline void fold_tallies(void *block)
{
    constexpr int kWords = (int)(sizeof(Counters) / sizeof(unsigned long long));
    unsigned long long *w = static_cast<unsigned long long *>(block);
    for (int s = 1; s <= kTallyStripes; ++s)
        for (int i = 0; i < kWords; ++i) {
            const unsigned long long x = w[kTallyStrideWords * s + i];
            w[i] = i == kMaxChainWord ? (w[i] > x ? w[i] : x) : w[i] + x;
        }
}
constexpr uint32_t kMaxRegions = 1024;

struct LaunchShape {
    int block_threads;  // multiple of 64
    int max_blocks;     // 0 = as many as the work needs
};

// Where an insert launch parks entries that exhaust the eviction bound, and what it reports.
struct PendSink {
    uint64_t *pend;              // pending list; the append index is Counters::pending
    uint32_t *tags;              // optional: the table index of each pending entry
    uint64_t cap;                // entries pend can hold
    uint32_t tag;                // table index of this launch (ignored by group launches)
    uint32_t count_placed;       // nonzero: add the entries this launch took in (reserved keys and
                                 // empty places excluded) to the tally Counters::placed
    unsigned long long *rejected_at;  // optional: += reserved keys this launch met (an atomic only
                                      // from a wave that met one; a keys insert of n took in n
                                      // minus this)
    uint32_t *chain;             // optional: chain[i] = exchanges of the chain input i started
                                 // (the reference's iterator_count, cuckoo.cu:349-351)
};

struct InsertPolicy {
    uint32_t kick_bound;    // iterations that follow the reference victim rule
    uint32_t rescue_kicks;  // further iterations with the rescue victim rule
    uint32_t max_retries;   // CAS races tolerated per entry before it is pended
};

struct Pool;
// find: vals/hit may be null. lanes in {4, 8, 16}. pool (optional): misses also look in the
// pool's entries of `table`.
hipError_t launch_find(const Desc &D, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, int lanes,
                       LaunchShape sh, hipStream_t st, const Pool *pool = nullptr, uint32_t table = 0);
// erase: deleted may be null; counts into C->deleted.
hipError_t launch_erase(const Desc &D, const uint32_t *keys, uint8_t *deleted, uint64_t n, int lanes, LaunchShape sh,
                        Counters *C, hipStream_t st, const Pool *pool = nullptr, uint32_t table = 0);
// insert keys/vals (start rule of cuckoo.cu:336-341) or packed entries with a fixed previous
// table `pre` (kNoTable = use the key-parity start rule).
// skip (optional): skip[i] != 0 leaves keys[i] out (an upsert already updated it in place).
hipError_t launch_insert(const Desc &D, const uint32_t *keys, const uint32_t *vals, uint64_t n, int lanes,
                         InsertPolicy pol, LaunchShape sh, Counters *C, PendSink sink, hipStream_t st,
                         const uint8_t *skip = nullptr);
// update in place: every copy of keys[i] in the first bucket (pos1, then pos2) that holds it
// takes the value vals[i] (with a pool: failing that, every copy in the pool); updated[i] = 1 if
// there was one; the count goes to C->updated.
hipError_t launch_update(const Desc &D, const uint32_t *keys, const uint32_t *vals, uint8_t *updated, uint64_t n,
                         int lanes, LaunchShape sh, Counters *C, hipStream_t st, const Pool *pool = nullptr);
// A zero entry is an empty place (an erased pending entry, a slot without a residual): skipped,
// and counted neither as rejected nor as placed. live (optional, a slot-addressed residual list's
// tile flags from launch_downsize): entries of a tile whose byte is 0 are empty and not read.
hipError_t launch_insert_entries(const Desc &D, const uint64_t *entries, uint64_t n, uint32_t pre, int lanes,
                                 InsertPolicy pol, LaunchShape sh, Counters *C, PendSink sink, hipStream_t st,
                                 const uint8_t *live = nullptr);
// ---- bulk insert by bucket fill (insert_order 3)
// A chunk is kChunkBuckets consecutive buckets of one subtable; chunk ids of subtable t start
// at base[t]; total = base[d] chunks, bits = bits of the largest chunk id.
constexpr uint32_t kChunkShift = 9;
constexpr uint32_t kChunkBuckets = 1u << kChunkShift;
struct Chunks {
    uint32_t base[kMaxSubtables + 1];
    uint32_t total, bits;
};
Chunks make_chunks(const Desc &D);
__host__ __device__ inline uint32_t chunk_base(const Chunks &Ch, uint32_t t)
{
    const uint32_t b0 = Ch.base[0], b1 = Ch.base[1], b2 = Ch.base[2], b3 = Ch.base[3], b4 = Ch.base[4],
                   b5 = Ch.base[5];
    uint32_t x = t == 1 ? b1 : b0;
    x = t == 2 ? b2 : x;
    x = t == 3 ? b3 : x;
    x = t == 4 ? b4 : x;
    return t == 5 ? b5 : x;
}
// A wide-key entry as the bulk fill of the wide table moves it.
struct WideEntry {
    uint64_t k, v;
};
// Workspace of a bulk fill's partition by chunk (dyc_part.hpp) of up to `cap` entries: the
// entries (8 bytes each, or a WideEntry) and their chunk ids (bulk_cid_bytes each) in two
// buffers each, the per-block digit counts (kPartialWords), the digit totals (256 words) and
// the chunk starts (chunks + 1 words).
constexpr uint64_t kPartialWords = 2048ull * 256;
constexpr uint64_t kMaxRadixWords = 256;
struct BulkSpace {
    void *ent[2];
    void *cid[2];
    uint32_t *partial;
    uint32_t *total;
    uint32_t *start;
};
size_t bulk_cid_bytes(const Chunks &Ch);
// Scatter blocks of a partition launch: two per CU on 256 CUs, of 4096-entry tiles. One per CU
// (8192-entry tiles, or 4096) and three per CU measured slower.
constexpr uint32_t kBulkSortBlocks = 512;
constexpr uint32_t kWideSortBlocks = 512;
// Group n entries by chunk (n < 2^32): round 1 from keys/vals for each key's fill_target bucket
// (keys != nullptr; reserved keys go to chunk 0), round 2 from the entries in ws.ent[in] for the
// other bucket of the pair. The sorted entries end in ws.ent[*out] (the round-2 input buffer is
// reused), and ws.start receives the first index of every chunk.
hipError_t launch_bulk_partition(const Desc &D, const Chunks &Ch, const uint32_t *keys, const uint32_t *vals, int in,
                                 uint64_t n, const BulkSpace &ws, uint32_t max_blocks, hipStream_t st, int *out);
// Fill free slots chunk by chunk from the partitioned entries (n < 2^32); start holds total + 1
// words. Entries that do not fit are appended to ovf (*ovf_n, zeroed by the caller, counts them;
// ovf_cap >= n). Placements go to the kick tally, reserved keys to the rejected tally. zeroed:
// every slot of the table is zero (a fresh or cleared table), so a dense round need not read it.
hipError_t launch_bulk_fill(const Desc &D, const Chunks &Ch, const uint32_t *start, const uint64_t *ent, uint64_t n,
                            uint64_t *ovf, uint64_t ovf_cap, unsigned long long *ovf_n, Counters *C, bool zeroed,
                            hipStream_t st);
// Order an insert batch by the region of its first bucket: out[0, n) receives the packed
// entries, region by region. counts and cursors are R words each (counts zeroed by the caller).
hipError_t launch_partition(const Desc &D, const Regions &M, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                            unsigned long long *counts, unsigned long long *cursors, uint64_t *out, hipStream_t st);
// ---- many tables in one launch
// A work item: input indexes [begin, begin + len) that all belong to table `table`.
struct GroupItem {
    uint64_t begin;
    uint32_t len;
    uint32_t table;
};

// Persistent launches over a directory of table descriptors: each wave takes items w, w + W, ...
// find: vals/hit may be null. erase: per-table deleted counts are added to table_counts.
// pool (optional): misses of an item of table i also look in the pool's entries tagged i.
// waves_per_eu: 6 or 8 builds the kernel for at least that many waves per SIMD (fewer
// registers); anything else leaves the register budget to the compiler.
hipError_t launch_group_find(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                             uint32_t *vals, uint8_t *hit, int lanes, int blocks, hipStream_t st,
                             const Pool *pool = nullptr, int waves_per_eu = 0);
hipError_t launch_group_erase(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                              uint8_t *deleted, unsigned long long *table_counts, int lanes, int blocks,
                              hipStream_t st, const Pool *pool = nullptr, int waves_per_eu = 0);
// insert: failures go to sink.pend with their table index in sink.tags; per-table inserted
// counts (placed or pended, reserved keys excluded) are added to table_counts.
// insert: an item's entries are keys/vals[i], or entries[i] (packed) when entries is not null.
hipError_t launch_group_insert(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                               const uint32_t *vals, InsertPolicy pol, int lanes, int blocks, Counters *C,
                               PendSink sink, unsigned long long *table_counts, hipStream_t st,
                               int waves_per_eu = 0, const uint64_t *entries = nullptr);

// ---- bulk fill of many tables (dyc_group.hip, insert_order 3)
// The batch is grouped by table already (host offsets), so one pass partitions each table's
// entries by its own chunks (at most kGroupBulkChunks per table, one 8-bit digit): a table's
// local chunk lc numbers its chunks subtable by subtable (chunks of 512 buckets, the last of a
// subtable shorter), and the call's chunks are numbered table by table (chunk0).
constexpr uint32_t kGroupBulkChunks = 256;
constexpr uint32_t kGroupBulkBlockKeys = 16384;  // entries per partition block (within one table)
struct BulkTable {
    uint64_t off;                  // its entries' first index in the batch
    uint32_t n;                    // its entries
    uint32_t table;                // its index in the group
    uint32_t first_block, blocks;  // its partition blocks (GroupItems, contiguous)
    uint32_t chunk0, chunks;       // its chunks in the call's numbering
};
static_assert(sizeof(BulkTable) == 32, "BulkTable is uploaded as a host array");
// A table of 64-bit keys and values: key lines laid out as a Desc, value lines parallel to them.
struct WideDesc {
    Desc D;                        // key arrays (D.sub), geometry and pair scheme
    uint64_t *val[kMaxSubtables];  // value arrays, slot for slot parallel to the key arrays
};
// One chunk of a group fill round with all that its workgroup reads before the entries, so the
// workgroup starts from a single load: a chain of chunk -> table -> directory loads, each
// waiting behind the round's store traffic, kept the workgroups of a whole round idle for about
// half their lifetime. Written by the scan.
struct alignas(16) GroupChunk {
    SubRef S;          // the chunk's subtable
    uint64_t ovf_off;  // the table's overflow places, ovf[ovf_off, ovf_off + ovf_cap)
    uint32_t ovf_cap;
    uint32_t b0, len;  // buckets [b0, b0 + len) of the subtable
    uint32_t lo, cnt;  // its entries, ent[lo, lo + cnt)
    uint32_t j;        // the table's index in the call's BulkTable list (its overflow counter)
    uint32_t table;    // the table's index in the group (its live count)
};
// Per partition block (blocks[b]: one table's range): every key's local chunk into lcid, and the
// block's count of each local chunk into partial[b * 256 + lc]. Round 1 (entries null): keys,
// chunk of the fill_target bucket (reserved keys count as chunk 0, whose fill rejects them).
// Round 2: packed entries (a fill's overflow), chunk of the other bucket of the pair.
hipError_t launch_group_bulk_pack(const Desc *dir, const GroupItem *blocks, uint32_t n_blocks, const uint32_t *keys,
                                  const uint64_t *entries, uint8_t *lcid, uint32_t *partial, hipStream_t st);
// Per table: partial[b * 256 + lc] becomes the place of block b's first entry of chunk lc in the
// partitioned batch, and chunk c = chunk0 + lc of the call gets its record chunks[c] (subtable
// from dir[table], entries, overflow region of the table's BulkTable).
hipError_t launch_group_bulk_scan(const Desc *dir, const BulkTable *tabs, uint32_t n_tabs, uint32_t *partial,
                                  GroupChunk *chunks, hipStream_t st);
hipError_t launch_group_bulk_scan(const WideDesc *dir, const BulkTable *tabs, uint32_t n_tabs, uint32_t *partial,
                                  GroupChunk *chunks, hipStream_t st);
// The pass itself: keys/vals (or packed entries, when not null) of every block to ent[0, n) by
// local chunk (no stable order).
hipError_t launch_group_bulk_scatter(const GroupItem *blocks, uint32_t n_blocks, const uint32_t *keys,
                                     const uint32_t *vals, const uint64_t *entries, const uint8_t *lcid,
                                     const uint32_t *partial, uint64_t n, uint64_t *ent, hipStream_t st);
// One workgroup per chunk record of the call (dense: the chunk's image in LDS; sparse: stores
// into the free slots of the buckets that received entries). Entries that do not fit go to the
// table's region of ovf, counted in ovf_n[j] (zeroed by the caller). Placements are added to
// table_counts and to the kick tally, reserved keys to the rejected tally. zeroed (dense only):
// every slot of every table is zero.
// The next round's work from this round's overflow counts, on the device (no read-back): block b
// of blocks (table tabs[btab[b]]) clipped to its table's overflow, out[b] = its part of
// [off, off + ovf_n[j]); tabs_out (optional) = tabs with n = ovf_n[j]. Blocks past the overflow
// become empty.
hipError_t launch_group_bulk_clip(const GroupItem *blocks, const uint32_t *btab, uint32_t n_blocks,
                                  const BulkTable *tabs, uint32_t n_tabs, const unsigned long long *ovf_n,
                                  GroupItem *out, BulkTable *tabs_out, hipStream_t st);
hipError_t launch_group_bulk_fill(const GroupChunk *chunks, uint32_t n_chunks, const uint64_t *ent, uint64_t *ovf,
                                  unsigned long long *ovf_n, unsigned long long *table_counts, Counters *C, bool dense,
                                  bool zeroed, hipStream_t st);

// ---- serving ticks
// A work item of a serving tick: keys [0, len) of one request (pointers already offset to the
// item's first key) against table `table` of the directory.
struct ServeItem {
    const uint32_t *keys;
    const uint32_t *vals;  // insert: values
    uint32_t *out_vals;    // find: values out (may be null)
    uint8_t *flags;        // find: hit flags; erase: deleted flags (may be null)
    uint32_t len;
    uint32_t table;
};

// Entries that exhausted the eviction bound, shared by all tables of a server (tags = table)
// and searched by its finds and erases until the host re-places them. count is the append index
// (the `pending` word of the Counters its insert launches add to); entries beyond cap are lost.
// Entries a table holds outside its buckets (a serving loop's overflow pool, or a single table's
// pending list used as its stash). Finds, erases and updates that miss both buckets search it.
// count != nullptr: the live size is min(*count, cap), read on the device (the pool grows while a
// tick runs); otherwise it is n, known when the launch is made. tags == nullptr: every entry
// belongs to the table; otherwise only those tagged with the launch's table index.
struct Pool {
    uint64_t *pend;
    const uint32_t *tags;
    uint64_t cap;
    const unsigned long long *count;
    uint64_t n;
};

// An item of a whole-tick launch: op 0 insert, 1 find, 2 erase. Items of one table and one
// (round, op) form a stage; with `wait` set the item starts once its stage is ready (the
// table's previous stage in the tick has finished).
struct TickItem {
    ServeItem item;
    uint32_t op;
    uint32_t stage;
    uint32_t wait;
    uint32_t pad;
};

// Per-stage control of a whole-tick launch, written by the host before each launch: left[s] =
// items of stage s, next[s] = the table's next stage (~0u: none), ready[s * kReadyStride] =
// 1 if nothing precedes the stage (flags padded to 64 bytes so waiters of different stages
// poll different lines).
constexpr uint32_t kReadyStride = 16;
struct TickStages {
    uint32_t *left;
    const uint32_t *next;
    uint32_t *ready;
};

bool serve_tick_supported(int insert_lanes, int find_lanes);
// One launch for a whole tick (see serve_tick_kernel). cursor must be 0 at launch. spin_cap
// bounds each wait for a stage; a waiting wave backs off up to max_back x 256 cycles per poll.
hipError_t launch_serve_tick(const Desc *dir, const TickItem *items, uint32_t n_items, unsigned long long *cursor,
                             TickStages stages, InsertPolicy pol, int insert_lanes, int find_lanes, int blocks,
                             Counters *C, PendSink sink, Pool pool, unsigned long long *ins_total,
                             unsigned long long *del_total, uint32_t spin_cap, uint32_t max_back, hipStream_t st);
hipError_t launch_serve_insert(const Desc *dir, const ServeItem *items, uint32_t n_items, InsertPolicy pol, int lanes,
                               int blocks, Counters *C, PendSink sink, unsigned long long *ins_total, hipStream_t st);
hipError_t launch_serve_find(const Desc *dir, const ServeItem *items, uint32_t n_items, Pool pool, int lanes,
                             int blocks, hipStream_t st);
hipError_t launch_serve_erase(const Desc *dir, const ServeItem *items, uint32_t n_items, Pool pool,
                              unsigned long long *del_total, int lanes, int blocks, hipStream_t st);

// ---- device-resident serving loop
// One launch that keeps running and takes work items from a ring in host memory as the host
// publishes them. Host memory (coherent pinned, hipHostMallocCoherent): the ring, the control
// block, the per-request records and the completion queue. Device memory: the loop words, the
// per-table completion counts and the per-request progress words.
//
// A work item, 64 bytes, written by the host; stamp (the low 32 bits of the item's index + 1)
// is written last, and the control block's tail after the items it publishes.
struct LoopItem {
    ServeItem item;
    uint32_t op;       // 0 insert, 1 find, 2 erase
    uint32_t req;      // request slot (index into the records and progress words)
    uint64_t need;     // the item starts once tbl_done[item.table] >= need
    uint32_t n_items;  // items of the request
    uint32_t stamp;
};
static_assert(sizeof(LoopItem) == 64, "one 64-byte line per item");

// Host-written control: tail = items published; stop = 1 asks the launch to end, stop = 2 also
// cuts every wait for a table short (the host gave up on the launch).
struct LoopCtl {
    unsigned long long tail;
    unsigned long long pad0[7];
    unsigned long long stop;
    unsigned long long pad1[7];
};

// Written by the device for a finished request, before its completion queue entry: acc = keys
// taken in (insert), hits (find) or deletions (erase); pool = the overflow pool's append index
// when the request finished; stalls = items that gave up waiting for the table's earlier requests.
// The record of a slot is read only after the slot's completion queue entry.
struct LoopRec {
    unsigned long long acc;
    unsigned long long pool;
    uint32_t stalls;
    uint32_t pad;
    unsigned long long lost;  // entries lost so far (the pool was full)
};

// Device words of the loop, each on its own 128-byte line: item cursor, items published as the
// front wave last saw them, quit flag, completion queue tail, fault flag, abort flag (the host
// gave up: waits for a table's earlier requests end at once, as stalls).
constexpr int kLoopCursor = 0, kLoopAvail = 16, kLoopQuit = 32, kLoopCqTail = 48, kLoopFault = 64, kLoopAbort = 80;
constexpr int kLoopWords = 96;
// Per-request progress word: items done in the low kLoopItemBits bits, acc above them.
constexpr int kLoopItemBits = 24;

struct LoopArgs {
    const Desc *dir;
    const LoopItem *ring;  // host memory
    uint64_t ring_mask;    // ring entries - 1 (a power of two)
    const LoopCtl *ctl;    // host memory
    LoopRec *rec;          // host memory, one per request slot
    unsigned long long *cq;  // host memory: entry p = (p + 1) << 32 | request slot
    uint32_t req_mask;     // request slots - 1 (a power of two; also the completion queue's)
    uint32_t spin_cap;     // polls an item waits for its table's earlier requests
    unsigned long long *words;     // kLoopWords, device
    unsigned long long *tbl_done;  // requests finished per table, device
    unsigned long long *req_word;  // per request slot, device, 0 when free: items whose table work is done
    uint32_t *req_out;             // per request slot, device, 0 when free: items whose results are written
    uint32_t *req_stall;           // per request slot, device, 0 when free
    uint64_t idle_ticks;   // the launch ends after this long without new items (100 MHz clock)
    uint32_t max_back;     // a waiting wave sleeps up to max_back x 256 cycles between polls
    uint32_t pad;
};

// Launch the loop on st with `blocks` blocks of 256 threads (all of them resident at once).
// words[kLoopCursor] must equal words[kLoopAvail] and the quit flag must be 0 (loop_reset).
hipError_t launch_serve_loop(const LoopArgs &A, InsertPolicy pol, int insert_lanes, int find_lanes, int blocks,
                             Counters *C, PendSink sink, Pool pool, unsigned long long *ins_total,
                             unsigned long long *del_total, hipStream_t st);
// Before a launch: cursor = avail (items claimed past the published ones by the previous
// launch's waves were not worked on), quit, fault and abort = 0.
hipError_t launch_loop_reset(unsigned long long *words, hipStream_t st);
// Blocks of 256 threads the loop kernel can keep resident on this device (its occupancy).
int serve_loop_capacity(int insert_lanes, int find_lanes);

// upsize subtable t: old has old_nb buckets, dst has 2*old_nb and is written completely.
hipError_t launch_upsize(const uint64_t *old, uint32_t old_nb, uint64_t *dst, uint32_t t, Counters *C,
                         hipStream_t st);
// downsize compaction: old has 2*new_nb buckets, dst new_nb (written completely). Residuals are
// slot-addressed in res (res_cap >= new_nb * 16 entries): live (res_tiles(new_nb) bytes, written
// completely) marks the tiles of kResTileEntries entries that hold any, and only those tiles of
// res are written (0 where a slot has none). Their number is added to C->residuals. Re-insert
// them with launch_insert_entries over new_nb * 16 entries with live.
constexpr uint32_t kResTileBuckets = 8;
constexpr uint64_t kResTileEntries = kResTileBuckets * kSlots;
inline uint64_t res_tiles(uint32_t new_nb) { return (new_nb + kResTileBuckets - 1) / kResTileBuckets; }
hipError_t launch_downsize(const uint64_t *old, uint32_t new_nb, uint64_t *dst, Counters *C, uint64_t *res,
                           uint64_t res_cap, uint8_t *live, hipStream_t st);
hipError_t launch_zero(unsigned long long *p, uint64_t words, hipStream_t st);
// Read a counter block back: one wave folds the stripes into the 16 counter words, writes them to
// post[1..16] (host-pinned), then seq to post[0], and clears the whole block for the next launch
// group. The host waits for post[0] == seq (await_post) instead of copying 16.6 KB and
// synchronising the stream: 7 us against 20 us per read on MI355X.
hipError_t launch_fold_post(Counters *C, unsigned long long *post, unsigned long long seq, hipStream_t st);
constexpr int kPostWords = 1 + (int)(sizeof(Counters) / sizeof(unsigned long long));
// Kernel copy of up to two segments of 8-byte words (either side may be host-pinned memory).
hipError_t launch_copy(const void *src0, void *dst0, uint64_t words0, const void *src1, void *dst1, uint64_t words1,
                       hipStream_t st);
// buckets the pos1-then-pos2 find order reads for keys -> C->kicks, pos1 hits -> C->occupied
hipError_t launch_probe_count(const Desc &D, const uint32_t *keys, uint64_t n, Counters *C, hipStream_t st);
// count non-empty slots of one subtable into C->occupied
hipError_t launch_count(const uint64_t *sub, uint64_t slots, Counters *C, hipStream_t st);

}  // namespace dyc
