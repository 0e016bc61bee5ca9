// dyc_table.hip -- host side of libdyc: the C ABI of include/dyc.h.
//
// Owns the four subtables, the pending list (entries in hand after an eviction chain hit its
// bound, or downsize residuals that could not be placed), per-call device counters, and the
// resize policy of the reference host API (GPU/dyhash/api.cpp).
#include "dyc_host.hpp"
#include "dyc_kernels.hpp"

#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

using namespace dyc;
using dyc::host::await_post;
using dyc::host::choose_down;
using dyc::host::choose_up;
using dyc::host::config_subtables;
using dyc::host::DeviceGuard;
using dyc::host::fail;
using dyc::host::initial_buckets;
using dyc::host::kMaxBuckets;
using dyc::host::CopyPool;
using dyc::host::Reclaimer;

namespace {
thread_local char g_err[512] = "";

// api.cpp:4-11 (int argument; 1 passes the 6k+-1 test, so nextPrime(0) = 1)
bool is_prime_ref(uint32_t n)
{
    int64_t num = (int32_t)n;
    if (num == 2 || num == 3) return true;
    if (num % 6 != 1 && num % 6 != 5) return false;
    for (int64_t i = 5; i * i <= num; i += 6)
        if (num % i == 0 || num % (i + 2) == 0) return false;
    return true;
}
}  // namespace

namespace dyc {
namespace host {

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *last_error() { return g_err; }

// api.cpp:47-51: s = nextPrime(size / (TABLE_NUM * 16)), then rounded up to an even bucket count
uint32_t initial_buckets(uint64_t hint, uint32_t d)
{
    uint32_t s = (uint32_t)std::min<uint64_t>(hint / ((uint64_t)d * kSlots), 0x7ffffff0u);
    while (!is_prime_ref(s)) ++s;
    return (s & 1u) ? s + 1u : s;
}

}  // namespace host
}  // namespace dyc

// Pinned and device buffers of the host-pointer entry points (two of each, for double
// buffering), and the copy streams that move them (allocated on first use).
struct DycHostStage;

// One insert dyc_insert_async left in flight: its own counter block, pinned read-back words and
// staging list for entries that exhaust the eviction bound, so that inserts on several streams
// run at the same time against one table (allocated on first use, kept by the table).
struct AsyncSlot {
    Counters *ctr = nullptr;
    unsigned long long *post = nullptr;
    unsigned long long seq = 0;
    uint64_t *pend = nullptr;
    uint64_t cap = 0;
    uint64_t n = 0;
    hipStream_t stream = nullptr;
    bool busy = false;
};
constexpr int kAsyncSlots = 8;

struct dyc_table {
    dyc_config cfg;
    int device = 0;
    Desc D{};
    uint64_t count = 0;
    uint64_t *pend = nullptr;  // pending entries [0, pend_n)
    uint64_t pend_cap = 0;
    uint64_t pend_n = 0;
    Counters *ctr = nullptr;    // device (kCounterBytes: base and stripes)
    Counters *hctr = nullptr;   // host copy of the folded counters
    unsigned long long *hpost = nullptr;  // pinned: the read-back words (launch_fold_post)
    unsigned long long post_seq = 0;
    bool ctr_clean = false;     // the device block is zero and no launch has used it since
    // every slot is zero: set when the table is created or cleared, dropped by every call that
    // may store an entry (a bulk fill into such a table need not read it)
    bool zeroed = false;
    uint64_t *part = nullptr;   // insert batch ordered by bucket region
    uint64_t part_cap = 0;
    unsigned long long *part_counts = nullptr;   // kMaxRegions counts, then kMaxRegions cursors
    // bulk fill (insert_order 3): the partition's entry and chunk-id buffers (bk_cid_bytes per
    // chunk id), its per-block counts and digit totals, the chunk starts and the overflow count
    // (device, and a pinned read-back)
    void *bk_cid[2] = {};
    uint64_t *bk_ent[2] = {};
    uint64_t bk_cap = 0;
    size_t bk_cid_bytes = 0;
    uint32_t *bk_start = nullptr;
    uint64_t bk_start_cap = 0;
    uint32_t *bk_partial = nullptr;  // kPartialWords, then 256 digit totals
    unsigned long long *bk_ovf = nullptr;
    unsigned long long *bk_hovf = nullptr;
    DycHostStage *stage = nullptr;
    uint8_t *upd = nullptr;  // upsert: per-key "updated in place" flags
    uint64_t upd_cap = 0;
    uint8_t *live = nullptr;  // downsize: tile flags of the residual list
    uint64_t live_cap = 0;
    Reclaimer reclaim;  // frees retired subtables off the calling thread
    AsyncSlot async[kAsyncSlots];
    uint64_t async_keys = 0;  // keys of the inserts in flight
    // Subtables created together (create, load, restore) share one allocation: the GPU's address
    // translation covers one large allocation with far fewer TLB entries than one allocation per
    // subtable, and random bucket access at 2^28 keys misses the per-CU TLB several times per
    // key otherwise. A subtable a resize replaces gets its own allocation; the slab is freed when
    // none of its subtables is still in use.
    uint8_t *slab = nullptr;
    uint32_t slab_live = 0;
    bool in_slab[kMaxSubtables] = {};
    dyc_stats st{};

    uint64_t capacity() const
    {
        uint64_t c = 0;
        for (uint32_t i = 0; i < D.d; ++i) c += (uint64_t)D.nb[i] * kSlots;
        return c;
    }
    LaunchShape shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.max_blocks}; }
    LaunchShape insert_shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.insert_blocks}; }
    InsertPolicy policy() const { return InsertPolicy{cfg.kick_bound, cfg.rescue_kicks, cfg.max_retries}; }
};

namespace {

// Pending list of a table from index `at` on (an insert launch appends there).
PendSink sink_at(const dyc_table *t, uint64_t at)
{
    return PendSink{t->pend + at, nullptr, t->pend_cap - at, 0, 0, nullptr};
}

// The pending list is the table's stash: finds, erases and updates that miss both buckets of a
// key search it. Returns the pool to pass to a launch, or nullptr when nothing is pending.
const Pool *stash_of(const dyc_table *t, Pool &p)
{
    p = Pool{t->pend, nullptr, t->pend_cap, nullptr, t->pend_n};
    return t->pend_n ? &p : nullptr;
}

// An erase leaves the stash entries it deleted as zero words: drop them (synchronous; the list
// is short whenever it is not empty).
int compact_pending(dyc_table *t, hipStream_t s)
{
    if (t->pend_n == 0) return DYC_OK;
    std::vector<uint64_t> h(t->pend_n);
    HIP_TRY(hipMemcpyAsync(h.data(), t->pend, t->pend_n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint64_t m = 0;
    for (uint64_t i = 0; i < h.size(); ++i)
        if (h[i]) h[m++] = h[i];
    if (m != t->pend_n) {
        if (m) HIP_TRY(hipMemcpyAsync(t->pend, h.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        t->pend_n = m;
    }
    t->st.pending = t->pend_n;
    return DYC_OK;
}

// Callers zero the counters before a launch group and read them after it; a read clears the
// device block, so the zero is launched only when something else left the block in use.
int zero_counters(dyc_table *t, hipStream_t s)
{
    if (!t->ctr_clean) HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(t->ctr), kCounterWords, s));
    t->ctr_clean = false;  // the launch group that follows uses it
    return DYC_OK;
}

// Fold, post and clear the counters (launch_fold_post) and wait for them. Every kernel queued on
// s before has finished when this returns.
int read_counters(dyc_table *t, hipStream_t s)
{
    const unsigned long long seq = ++t->post_seq;
    HIP_TRY(launch_fold_post(t->ctr, t->hpost, seq, s));
    HIP_TRY(await_post(t->hpost, seq, s));
    memcpy(t->hctr, const_cast<const unsigned long long *>(t->hpost) + 1, sizeof(Counters));
    t->ctr_clean = true;
    return DYC_OK;
}

constexpr size_t kSlabAlign = 2u << 20;  // subtable regions of a slab start 2 MiB apart

// Release a subtable's storage: its own allocation, or its region of the slab (the slab itself
// once its last subtable is released). The free itself runs on the reclaimer's thread, after
// every kernel on the device has finished.
int release_block(dyc_table *t, uint64_t *p, bool in_slab)
{
    if (!p) return DYC_OK;
    if (in_slab) {
        if (--t->slab_live == 0) {
            t->reclaim.defer(t->device, t->slab);
            t->slab = nullptr;
        }
        return DYC_OK;
    }
    t->reclaim.defer(t->device, p);
    return DYC_OK;
}

// Free the bulk-fill and region-ordering workspaces (grown on demand, kept between calls); the
// next call that needs them allocates them again. The caller has synchronised the table's stream.
void release_workspace(dyc_table *t)
{
    for (int b = 0; b < 2; ++b) {
        if (t->bk_cid[b]) (void)hipFree(t->bk_cid[b]);
        if (t->bk_ent[b]) (void)hipFree(t->bk_ent[b]);
        t->bk_cid[b] = nullptr;
        t->bk_ent[b] = nullptr;
    }
    t->bk_cap = 0;
    t->bk_cid_bytes = 0;
    if (t->bk_start) (void)hipFree(t->bk_start);
    t->bk_start = nullptr;
    t->bk_start_cap = 0;
    if (t->part) (void)hipFree(t->part);
    t->part = nullptr;
    t->part_cap = 0;
}

// hipMalloc for a resize; when it fails, wait for the frees still queued, release the idle
// insert workspaces (up to 20-24 B per key of the largest bulk-filled batch) and try once more.
hipError_t resize_malloc(dyc_table *t, void **p, size_t bytes, hipStream_t s)
{
    if (hipMalloc(p, bytes) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    t->reclaim.drain();
    if (hipStreamSynchronize(s) == hipSuccess) release_workspace(t);
    (void)hipGetLastError();
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

int release_sub(dyc_table *t, uint32_t i)
{
    uint64_t *p = t->D.sub[i];
    const bool in_slab = t->in_slab[i];
    t->D.sub[i] = nullptr;
    t->in_slab[i] = false;
    return release_block(t, p, in_slab);
}

// One allocation for d subtables of nb[i] buckets; *base receives it and sub[i] the regions.
int alloc_slab(uint32_t d, const uint32_t *nb, uint8_t **base, uint64_t **sub)
{
    size_t off[kMaxSubtables + 1] = {0};
    for (uint32_t i = 0; i < d; ++i)
        off[i + 1] = off[i] + (((size_t)nb[i] * kBucketBytes + kSlabAlign - 1) & ~(kSlabAlign - 1));
    if (hipMalloc(base, off[d]) != hipSuccess) {
        *base = nullptr;
        (void)hipGetLastError();
        return fail(DYC_ERR_OOM, "hipMalloc of %zu bytes for the table failed", off[d]);
    }
    for (uint32_t i = 0; i < d; ++i) sub[i] = reinterpret_cast<uint64_t *>(*base + off[i]);
    return DYC_OK;
}

// Replace every subtable of t by the regions of a new slab (the old storage is released).
int install_slab(dyc_table *t, uint8_t *base, uint64_t *const *sub, const uint32_t *nb)
{
    t->zeroed = false;  // new memory: only the caller knows what it holds
    for (uint32_t i = 0; i < t->D.d; ++i) {
        const int rc = release_sub(t, i);
        if (rc) return rc;
    }
    t->slab = base;
    t->slab_live = t->D.d;
    for (uint32_t i = 0; i < t->D.d; ++i) {
        desc_set(t->D, (int)i, sub[i], nb[i]);
        t->in_slab[i] = true;
    }
    return DYC_OK;
}

// Make room for `extra` more pending entries after the current ones.
int reserve_pending(dyc_table *t, uint64_t extra, hipStream_t s)
{
    const uint64_t need = t->pend_n + extra;
    if (need <= t->pend_cap) return DYC_OK;
    const uint64_t cap = std::max<uint64_t>(need, t->pend_cap * 2);
    uint64_t *np = nullptr;
    if (hipMalloc(&np, cap * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DYC_ERR_OOM, "hipMalloc of the pending list (%llu entries) failed", (unsigned long long)cap);
    }
    if (t->pend_n) HIP_TRY(hipMemcpyAsync(np, t->pend, t->pend_n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (t->pend) HIP_TRY(hipFree(t->pend));
    t->pend = np;
    t->pend_cap = cap;
    return DYC_OK;
}

// Absorb the counters of an insert-like launch that appended to pend + pend_n.
int absorb_insert_counters(dyc_table *t, uint64_t room)
{
    const Counters &c = *t->hctr;
    t->st.kicks += c.kicks;
    t->st.last_kicks += c.kicks;
    t->st.cas_retries += c.retries;
    t->st.kick_failures += c.pending;
    t->st.lost += c.lost;
    t->st.last_max_chain = std::max<uint32_t>(t->st.last_max_chain, (uint32_t)std::min<uint64_t>(c.max_chain, ~0u));
    t->pend_n += std::min<uint64_t>(c.pending, room);
    if (c.lost) return fail(DYC_ERR_LOST, "pending list overflow: %llu entries lost", (unsigned long long)c.lost);
    return DYC_OK;
}

int do_resize_up(dyc_table *t, hipStream_t s)
{
    t->zeroed = false;
    const uint32_t sub = choose_up(t->D.nb, t->D.d);
    const uint32_t old_nb = t->D.nb[sub];
    if ((uint64_t)old_nb * 2 >= kMaxBuckets) return fail(DYC_ERR_INVALID, "subtable %u cannot grow past 2^29 buckets", sub);
    uint64_t *dst = nullptr;
    const size_t bytes = (size_t)old_nb * 2 * kBucketBytes;
    if (resize_malloc(t, reinterpret_cast<void **>(&dst), bytes, s) != hipSuccess)
        return fail(DYC_ERR_OOM, "hipMalloc of %zu bytes for upsize failed", bytes);
    int rc = zero_counters(t, s);
    if (rc) return rc;
    HIP_TRY(launch_upsize(t->D.sub[sub], old_nb, dst, sub, t->ctr, s));
    rc = read_counters(t, s);
    if (rc) return rc;
    rc = release_sub(t, sub);
    if (rc) return rc;
    desc_set(t->D, (int)sub, dst, old_nb * 2);
    t->st.resizes_up++;
    if (t->hctr->misplaced)
        return fail(DYC_ERR_CORRUPT, "upsize found %llu entries outside their bucket",
                    (unsigned long long)t->hctr->misplaced);
    return DYC_OK;
}

// After an operation that found `before` entries pending and left more: re-insert everything on
// the pending list, upsizing before each round (the reference would have dropped these entries,
// cuckoo.cu:422-431). Entries that never fit (copies of one key beyond its two buckets' 32 slots)
// stay in the stash without growing the table on every later call. DYC_ERR_PENDING when the
// list is still longer than `before` after max_fail_rounds rounds.
int place_pending(dyc_table *t, hipStream_t s, uint64_t before)
{
    t->zeroed = false;
    uint32_t rounds = 0;
    while (t->pend_n > before && rounds < t->cfg.max_fail_rounds) {
        int rc = do_resize_up(t, s);
        if (rc) return rc;
        // move the pending entries aside, then re-insert them appending failures to pend
        const uint64_t m = t->pend_n;
        uint64_t *src = nullptr;
        if (hipMalloc(&src, m * sizeof(uint64_t)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(DYC_ERR_OOM, "hipMalloc for pending re-insert failed");
        }
        HIP_TRY(hipMemcpyAsync(src, t->pend, m * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        t->pend_n = 0;
        rc = zero_counters(t, s);
        if (!rc) {
            hipError_t e = launch_insert_entries(t->D, src, m, kNoTable, (int)t->cfg.insert_lanes, t->policy(),
                                                 t->insert_shape(), t->ctr, sink_at(t, 0), s);
            if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "pending re-insert launch: %s", hipGetErrorString(e));
        }
        if (!rc) rc = read_counters(t, s);
        (void)hipFree(src);
        if (rc) return rc;
        rc = absorb_insert_counters(t, t->pend_cap);
        if (rc) return rc;
        ++rounds;
    }
    t->st.pending = t->pend_n;
    if (t->pend_n > before)
        return fail(DYC_ERR_PENDING, "%llu entries still pending after %u upsize rounds (%llu before the call)",
                    (unsigned long long)t->pend_n, rounds, (unsigned long long)before);
    return DYC_OK;
}

int do_resize_down(dyc_table *t, hipStream_t s)
{
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    const uint32_t sub = choose_down(t->D.nb, t->D.d);
    const uint32_t old_nb = t->D.nb[sub];
    if (old_nb & 1u)
        return fail(DYC_ERR_ODD_BUCKETS, "downsize of subtable %u refused: %u buckets is odd", sub, old_nb);
    const uint32_t new_nb = old_nb / 2;
    uint64_t *dst = nullptr;
    if (resize_malloc(t, reinterpret_cast<void **>(&dst), (size_t)new_nb * kBucketBytes, s) != hipSuccess)
        return fail(DYC_ERR_OOM, "hipMalloc for downsize failed");
    if (res_tiles(new_nb) > t->live_cap) {  // the residual list's tile flags, grown as needed
        if (t->live) (void)hipFree(t->live);
        t->live = nullptr;
        t->live_cap = 0;
        if (hipMalloc(&t->live, res_tiles(new_nb)) != hipSuccess) {
            (void)hipGetLastError();
            t->live = nullptr;
            (void)hipFree(dst);
            return fail(DYC_ERR_OOM, "hipMalloc for downsize failed");
        }
        t->live_cap = res_tiles(new_nb);
    }
    // The residuals go to the old subtable's upper half, slot-addressed: a group writes pair g's
    // residuals there only after reading both of its buckets, and no other group reads them. The
    // old subtable is released once the residuals are re-inserted.
    uint64_t *old = t->D.sub[sub];
    const bool old_in_slab = t->in_slab[sub];
    uint64_t *res = old + (uint64_t)new_nb * kSlots;
    const uint64_t res_n = (uint64_t)new_nb * kSlots;
    int rc = zero_counters(t, s);
    if (!rc) {
        hipError_t e = launch_downsize(old, new_nb, dst, t->ctr, res, res_n, t->live, s);
        if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "downsize launch: %s", hipGetErrorString(e));
    }
    if (!rc) rc = read_counters(t, s);
    if (rc) {
        // a launch refused changes nothing; a read that fails means the stream faulted, and the
        // compaction may already have rewritten the old upper half: the table is then lost
        // with the device context
        (void)hipFree(dst);
        return rc;
    }
    const uint64_t nres = t->hctr->residuals;
    t->D.sub[sub] = nullptr;  // the table now owns dst; old is released below
    t->in_slab[sub] = false;
    desc_set(t->D, (int)sub, dst, new_nb);
    t->st.resizes_down++;
    // residuals continue from the shrunk subtable into the other one of their pair
    // (warp_insert(..., num_table_to_resize), cuckoo.cu:719), after the compaction completed
    if (nres) {
        rc = reserve_pending(t, std::max<uint64_t>(nres, 1024), s);
        if (!rc) rc = zero_counters(t, s);
        if (!rc) {
            // the slot-addressed list: its live tiles, zeros skipped
            hipError_t e = launch_insert_entries(t->D, res, res_n, sub, (int)t->cfg.insert_lanes, t->policy(),
                                                 t->insert_shape(), t->ctr, sink_at(t, t->pend_n), s, t->live);
            if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "residual insert launch: %s", hipGetErrorString(e));
        }
        if (!rc) rc = read_counters(t, s);
        if (!rc) rc = absorb_insert_counters(t, t->pend_cap - t->pend_n);
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        (void)release_block(t, old, old_in_slab);
        return rc;
    }
    rc = release_block(t, old, old_in_slab);  // freed once the device is idle
    if (rc) return rc;
    if (t->pend_n > pend_before) return place_pending(t, s, pend_before);
    return DYC_OK;
}

// Number of bucket regions to order an insert batch by (insert_order 2), or 0 to insert in input
// order. Ordering makes the resident waves place one region_mib slice of the table at a time;
// at 2^28 keys and load factor 0.90 on MI355X it measured slower than input order (25-35 ms
// against 17.7 ms), so it is opt-in.
uint32_t insert_regions(const dyc_table *t, uint64_t n)
{
    if (t->cfg.insert_order != 2 || n == 0) return 0;
    const uint64_t table_bytes = t->capacity() * 8;
    const uint64_t region = (uint64_t)t->cfg.region_mib << 20;
    uint64_t R = (table_bytes + region - 1) / region;
    if (R > kMaxRegions) R = kMaxRegions;
    return (uint32_t)std::max<uint64_t>(R, 1);
}

int reserve_partition(dyc_table *t, uint64_t n, hipStream_t s)
{
    if (!t->part_counts) {
        if (hipMalloc(&t->part_counts, 2 * kMaxRegions * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(DYC_ERR_OOM, "hipMalloc of the region counters failed");
        }
    }
    if (n <= t->part_cap) return DYC_OK;
    if (t->part) {
        HIP_TRY(hipStreamSynchronize(s));  // the table's work is ordered on s
        HIP_TRY(hipFree(t->part));
        t->part = nullptr;
        t->part_cap = 0;
    }
    if (hipMalloc(&t->part, n * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DYC_ERR_OOM, "hipMalloc of the region-ordered batch (%llu entries) failed", (unsigned long long)n);
    }
    t->part_cap = n;
    return DYC_OK;
}

// Bulk fill (insert_order 3) takes batches of at least two keys per bucket: below that the
// fill's pass over every chunk's free-slot masks costs more than the atomics it saves.
bool use_bulk_fill(const dyc_table *t, uint64_t n)
{
    uint64_t buckets = 0;
    for (uint32_t i = 0; i < t->D.d; ++i) buckets += t->D.nb[i];
    return t->cfg.insert_order == 3 && n >= 2 * buckets && n < (1ull << 31);
}

int bulk_alloc(void **p, size_t bytes, const char *what)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(DYC_ERR_OOM, "hipMalloc of the bulk-fill %s (%llu bytes) failed", what,
                    (unsigned long long)bytes);
    }
    return DYC_OK;
}

// Grow-only workspace of the bulk fill for n keys over the chunks Ch: two buffers of n entries
// and n chunk ids (2 or 4 bytes), the partition's counts and the chunk starts.
int reserve_bulk(dyc_table *t, uint64_t n, const Chunks &Ch, hipStream_t s)
{
    const size_t cb = bulk_cid_bytes(Ch);
    const bool grow = n > t->bk_cap || cb > t->bk_cid_bytes || Ch.total + 1ull > t->bk_start_cap;
    // a buffer being replaced may still be in use by the table's earlier work, ordered on s
    if (grow) HIP_TRY(hipStreamSynchronize(s));
    int rc = DYC_OK;
    if (n > t->bk_cap || cb > t->bk_cid_bytes) {
        for (int b = 0; b < 2; ++b) {
            if (t->bk_cid[b]) HIP_TRY(hipFree(t->bk_cid[b]));
            if (t->bk_ent[b]) HIP_TRY(hipFree(t->bk_ent[b]));
            t->bk_cid[b] = nullptr;
            t->bk_ent[b] = nullptr;
        }
        t->bk_cap = 0;
        t->bk_cid_bytes = 0;
        for (int b = 0; b < 2 && !rc; ++b) {
            rc = bulk_alloc(&t->bk_cid[b], n * cb, "chunk ids");
            if (!rc) rc = bulk_alloc((void **)&t->bk_ent[b], n * sizeof(uint64_t), "entries");
        }
        if (rc) return rc;
        t->bk_cap = n;
        t->bk_cid_bytes = cb;
    }
    if (Ch.total + 1ull > t->bk_start_cap) {
        if (t->bk_start) HIP_TRY(hipFree(t->bk_start));
        t->bk_start = nullptr;
        t->bk_start_cap = 0;
        rc = bulk_alloc((void **)&t->bk_start, (Ch.total + 1ull) * sizeof(uint32_t), "chunk starts");
        if (rc) return rc;
        t->bk_start_cap = Ch.total + 1ull;
    }
    if (!t->bk_partial) {
        rc = bulk_alloc((void **)&t->bk_partial, (kPartialWords + kMaxRadixWords) * sizeof(uint32_t), "partition counts");
        if (rc) return rc;
    }
    if (!t->bk_ovf) {
        rc = bulk_alloc((void **)&t->bk_ovf, 256, "overflow count");
        if (rc) return rc;
    }
    if (!t->bk_hovf && hipHostMalloc((void **)&t->bk_hovf, 64, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        t->bk_hovf = nullptr;
        return fail(DYC_ERR_OOM, "hipHostMalloc of the bulk-fill read-back word failed");
    }
    return DYC_OK;
}

BulkSpace bulk_space(const dyc_table *t)
{
    return BulkSpace{{t->bk_ent[0], t->bk_ent[1]}, {t->bk_cid[0], t->bk_cid[1]}, t->bk_partial,
                     t->bk_partial + kPartialWords, t->bk_start};
}

// One fill round over n entries: round 1 from keys/vals (keys != nullptr), round 2 from the
// entries in bk_ent[in]. Partition them by chunk, fill, and read back how many entries overflowed
// into bk_ent[*out] (the buffer the partition did not end in).
int bulk_round(dyc_table *t, const Chunks &Ch, const uint32_t *keys, const uint32_t *vals, int in, uint64_t n,
               bool zeroed, hipStream_t s, int *out, uint64_t *overflow)
{
    int r = 0;
    HIP_TRY(launch_bulk_partition(t->D, Ch, keys, vals, in, n, bulk_space(t), kBulkSortBlocks, s, &r));
    HIP_TRY(launch_zero(t->bk_ovf, 1, s));
    HIP_TRY(launch_bulk_fill(t->D, Ch, t->bk_start, t->bk_ent[r], n, t->bk_ent[1 - r], t->bk_cap, t->bk_ovf, t->ctr,
                             zeroed, s));
    HIP_TRY(hipMemcpyAsync(t->bk_hovf, t->bk_ovf, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *overflow = *t->bk_hovf;
    *out = 1 - r;
    if (*overflow > n) return fail(DYC_ERR_CORRUPT, "bulk fill overflow count %llu exceeds the batch (%llu)",
                                   (unsigned long long)*overflow, (unsigned long long)n);
    return DYC_OK;
}

// insert_order 3: place keys by bucket fill in their fill_target bucket (pos1, the bucket a find
// reads first, wherever equal subtable shares allow), then the overflow in the other bucket of
// their pair, then the rest by the eviction insert. Counters are zeroed here
// and read by the caller.
int bulk_insert(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, bool was_zeroed, hipStream_t s,
                uint64_t *left)
{
    *left = 0;
    const Chunks Ch = make_chunks(t->D);  // the workspace was reserved for it
    int rc = zero_counters(t, s);
    if (rc) return rc;
    int ov = 0, ov2 = 0;
    uint64_t m1 = 0, m2 = 0;
    // an empty table (count 0, nothing pending) that no call has stored into since it was
    // created or cleared holds zeros only
    const bool zeroed = was_zeroed && t->count == 0 && t->pend_n == 0;
    rc = bulk_round(t, Ch, keys, vals, 0, n, zeroed, s, &ov, &m1);
    if (rc || m1 == 0) return rc;
    rc = bulk_round(t, Ch, nullptr, nullptr, ov, m1, false, s, &ov2, &m2);  // the overflow, in its other table
    if (rc || m2 == 0) return rc;
    *left = m2;
    HIP_TRY(launch_insert_entries(t->D, t->bk_ent[ov2], m2, kNoTable, (int)t->cfg.insert_lanes, t->policy(),
                                  t->insert_shape(), t->ctr, sink_at(t, t->pend_n), s));
    return DYC_OK;
}

// Complete the inserts dyc_insert_async left in flight: wait for each one's counters, absorb them
// as dyc_insert does, move its failed entries to the pending list, and re-place those (upsize and
// re-insert) as dyc_insert would. Every other call on the table does this first.
int settle(dyc_table *t)
{
    bool any = false;
    for (const AsyncSlot &a : t->async) any |= a.busy;
    if (!any) return DYC_OK;
    const uint64_t before = t->pend_n;
    int rc = DYC_OK;
    hipStream_t last = nullptr;
    for (AsyncSlot &a : t->async) {
        if (!a.busy) continue;
        a.busy = false;
        const hipError_t e = await_post(a.post, a.seq, a.stream);
        if (e != hipSuccess) {
            if (!rc) rc = fail(DYC_ERR_HIP, "an insert in flight on a stream failed: %s", hipGetErrorString(e));
            continue;
        }
        Counters c;
        memcpy(&c, const_cast<const unsigned long long *>(a.post) + 1, sizeof c);
        t->st.rejected += c.rejected;
        t->count += a.n - c.rejected;
        t->st.kicks += c.kicks;
        t->st.last_kicks += c.kicks;
        t->st.cas_retries += c.retries;
        t->st.kick_failures += c.pending;
        t->st.lost += c.lost;
        t->st.last_max_chain = std::max<uint32_t>(t->st.last_max_chain, (uint32_t)std::min<uint64_t>(c.max_chain, ~0u));
        if (c.lost && !rc) rc = fail(DYC_ERR_LOST, "pending staging overflow: %llu entries lost", (unsigned long long)c.lost);
        const uint64_t m = std::min<uint64_t>(c.pending, a.cap);
        if (m) {
            int r = reserve_pending(t, m, a.stream);
            if (!r && (hipMemcpyAsync(t->pend + t->pend_n, a.pend, m * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                      a.stream) != hipSuccess ||
                       hipStreamSynchronize(a.stream) != hipSuccess))
                r = fail(DYC_ERR_HIP, "moving failed entries of an insert in flight to the pending list failed");
            if (r) {
                if (!rc) rc = r;
                continue;
            }
            t->pend_n += m;
        }
        last = a.stream;
    }
    t->async_keys = 0;
    t->st.pending = t->pend_n;
    if (!rc && t->pend_n > before) rc = place_pending(t, last, before);
    return rc;
}

#define SETTLE(t)                  \
    do {                           \
        const int rc_ = settle(t); \
        if (rc_) return rc_;       \
    } while (0)

// Counter block, read-back words and a staging list for n entries in slot a (grown on demand;
// the slot is idle).
int async_reserve(AsyncSlot &a, uint64_t n)
{
    if (!a.ctr && hipMalloc(&a.ctr, kCounterBytes) != hipSuccess) {
        (void)hipGetLastError();
        a.ctr = nullptr;
        return fail(DYC_ERR_OOM, "hipMalloc of an async counter block failed");
    }
    if (!a.post) {
        if (hipHostMalloc(reinterpret_cast<void **>(&a.post), kPostWords * sizeof(unsigned long long),
                          hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            a.post = nullptr;
            return fail(DYC_ERR_OOM, "hipHostMalloc of async read-back words failed");
        }
        memset(a.post, 0, kPostWords * sizeof(unsigned long long));
    }
    if (n > a.cap) {
        if (a.pend) (void)hipFree(a.pend);
        a.pend = nullptr;
        a.cap = 0;
        if (hipMalloc(&a.pend, n * sizeof(uint64_t)) != hipSuccess) {
            (void)hipGetLastError();
            a.pend = nullptr;
            return fail(DYC_ERR_OOM, "hipMalloc of an async staging list (%llu entries) failed", (unsigned long long)n);
        }
        a.cap = n;
    }
    return DYC_OK;
}

void async_free(dyc_table *t)
{
    for (AsyncSlot &a : t->async) {
        if (a.ctr) (void)hipFree(a.ctr);
        if (a.post) (void)hipHostFree(a.post);
        if (a.pend) (void)hipFree(a.pend);
        a = AsyncSlot{};
    }
}

// The counter block of dyc_insert_noalloc: a library-static device array instead of a table's
// own block, so that the call allocates nothing. Each device holds its own copy until the
// process ends. In host code `scratch` names a host-side stand-in, not a device address: kernels
// and copies take the address hipGetSymbolAddress gives for the device (noalloc_counters).
static __device__ __align__(256) unsigned long long scratch[kCounterWords];
static_assert(kCounterBytes == DYC_NOALLOC_SCRATCH_BYTES, "DYC_NOALLOC_SCRATCH_BYTES is the counter block's size");

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
Counters *g_scratch[kMaxScratchDevices] = {};

// The device address of `scratch` on `device` (the current device), looked up once per device.
int noalloc_counters(int device, Counters **out)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(DYC_ERR_INVALID, "device %d is beyond the %d devices dyc_insert_noalloc serves", device,
                    kMaxScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(DYC_ERR_HIP, "hipGetSymbolAddress of the static counter block failed: %s",
                        hipGetErrorString(e));
        }
        g_scratch[device] = static_cast<Counters *>(p);
    }
    *out = g_scratch[device];
    return DYC_OK;
}

bool valid_lanes(uint32_t l) { return l == 4 || l == 8 || l == 16; }
bool valid_insert_lanes(uint32_t l) { return l == 1 || l == 2 || valid_lanes(l); }

int check_config(const dyc_config &cfg)
{
    if (!valid_lanes(cfg.find_lanes) || !valid_insert_lanes(cfg.insert_lanes))
        return fail(DYC_ERR_INVALID, "lanes per key must be 4, 8 or 16 (insert also 1 or 2)");
    if (cfg.block_threads == 0 || cfg.block_threads % 64 || cfg.block_threads > 256)
        return fail(DYC_ERR_INVALID, "block_threads must be a multiple of 64 in [64, 256]");
    if (!(cfg.upper_bound > 0.0 && cfg.upper_bound <= 1.0) || !(cfg.lower_bound >= 0.0) ||
        cfg.lower_bound >= cfg.upper_bound)
        return fail(DYC_ERR_INVALID, "need 0 <= lower_bound < upper_bound <= 1");
    if (cfg.insert_order > 3 || cfg.region_mib == 0)
        return fail(DYC_ERR_INVALID, "insert_order must be 0..3 and region_mib > 0");
    const uint32_t d = config_subtables(cfg);
    if (d < (uint32_t)kMinSubtables || d > (uint32_t)kMaxSubtables || cfg.reserved0 != 0)
        return fail(DYC_ERR_INVALID, "subtables must be in [%d, %d] (reserved0 = 0)", kMinSubtables, kMaxSubtables);
    return DYC_OK;
}

}  // namespace

struct DycHostStage {
    CopyPool *pool = nullptr;           // staging copies
    uint32_t *hk[2] = {}, *hv[2] = {};  // pinned: keys; values in (insert) or out (find)
    uint8_t *hf[2] = {};                // pinned: flags out
    uint32_t *dk[2] = {}, *dv[2] = {};
    uint8_t *df[2] = {};
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t in_ready[2] = {}, done[2] = {}, out_ready[2] = {};
};

namespace {

constexpr uint64_t kStageKeys = 1ull << 21;  // keys per pipeline chunk

void free_stage(dyc_table *t)
{
    DycHostStage *S = t->stage;
    if (!S) return;
    for (int b = 0; b < 2; ++b) {
        if (S->hk[b]) (void)hipHostFree(S->hk[b]);
        if (S->hv[b]) (void)hipHostFree(S->hv[b]);
        if (S->hf[b]) (void)hipHostFree(S->hf[b]);
        if (S->dk[b]) (void)hipFree(S->dk[b]);
        if (S->dv[b]) (void)hipFree(S->dv[b]);
        if (S->df[b]) (void)hipFree(S->df[b]);
        if (S->in_ready[b]) (void)hipEventDestroy(S->in_ready[b]);
        if (S->done[b]) (void)hipEventDestroy(S->done[b]);
        if (S->out_ready[b]) (void)hipEventDestroy(S->out_ready[b]);
    }
    if (S->h2d) (void)hipStreamDestroy(S->h2d);
    if (S->d2h) (void)hipStreamDestroy(S->d2h);
    delete S->pool;
    delete S;
    t->stage = nullptr;
}

int get_stage(dyc_table *t, DycHostStage **out)
{
    if (!t->stage) {
        DycHostStage *S = new DycHostStage();
        t->stage = S;
        S->pool = new CopyPool(std::min(8u, std::max(1u, std::thread::hardware_concurrency())) - 1);
        bool ok = hipStreamCreateWithFlags(&S->h2d, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&S->d2h, hipStreamNonBlocking) == hipSuccess;
        for (int b = 0; b < 2 && ok; ++b) {
            ok = hipHostMalloc((void **)&S->hk[b], kStageKeys * 4, hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void **)&S->hv[b], kStageKeys * 4, hipHostMallocDefault) == hipSuccess &&
                 hipHostMalloc((void **)&S->hf[b], kStageKeys, hipHostMallocDefault) == hipSuccess &&
                 hipMalloc(&S->dk[b], kStageKeys * 4) == hipSuccess && hipMalloc(&S->dv[b], kStageKeys * 4) == hipSuccess &&
                 hipMalloc(&S->df[b], kStageKeys) == hipSuccess &&
                 hipEventCreateWithFlags(&S->in_ready[b], hipEventDisableTiming) == hipSuccess &&
                 hipEventCreateWithFlags(&S->done[b], hipEventDisableTiming) == hipSuccess &&
                 hipEventCreateWithFlags(&S->out_ready[b], hipEventDisableTiming) == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            free_stage(t);
            return fail(DYC_ERR_OOM, "staging buffers for host-pointer calls could not be allocated");
        }
    }
    *out = t->stage;
    return DYC_OK;
}

// ---- checkpoint file: a 128-byte header, the d subtable images, then the pending entries,
// all little-endian; FNV-1a (64-bit, over 8-byte words) guards each part and the header.
struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t subtables;
    uint32_t buckets[kMaxSubtables];  // entries at and past `subtables` are 0
    uint64_t count;
    uint64_t pending;
    uint64_t checksum[kMaxSubtables];  // subtable images (0 past `subtables`)
    uint64_t pending_check;            // the pending list
    uint64_t header_check;             // over every byte before this field
    uint8_t reserved[8];
};
static_assert(sizeof(CkptHeader) == 128, "checkpoint header is 128 bytes");
constexpr char kCkptMagic[8] = {'D', 'Y', 'C', 'K', 'P', 'T', '0', '1'};
constexpr uint32_t kCkptVersion = 2;
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;
constexpr uint64_t kCkptChunk = 1ull << 23;  // 8-byte words per staged chunk (64 MiB)

uint64_t fnv_words(uint64_t h, const uint64_t *w, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) h = (h ^ w[i]) * kFnvPrime;
    return h;
}

uint64_t fnv_bytes(const void *p, size_t n)
{
    uint64_t h = kFnvBasis;
    for (size_t i = 0; i < n; ++i) h = (h ^ static_cast<const uint8_t *>(p)[i]) * kFnvPrime;
    return h;
}

struct PinnedPair {
    uint64_t *buf[2] = {nullptr, nullptr};
    ~PinnedPair()
    {
        for (auto *b : buf)
            if (b) (void)hipHostFree(b);
    }
    bool alloc(size_t bytes)
    {
        return hipHostMalloc((void **)&buf[0], bytes, hipHostMallocDefault) == hipSuccess &&
               hipHostMalloc((void **)&buf[1], bytes, hipHostMallocDefault) == hipSuccess;
    }
};

// Stream `words` device words to the file through two pinned buffers (copy of chunk i+1
// overlaps the write of chunk i); returns the FNV-1a of the words.
int dump_to_file(FILE *f, const uint64_t *dev, uint64_t words, PinnedPair &pp, uint64_t *hash, hipStream_t s)
{
    uint64_t h = kFnvBasis;
    hipEvent_t ev[2] = {nullptr, nullptr};
    HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int rc = DYC_OK;
    uint64_t off = 0, prev_n = 0;
    int b = 0, prev_b = 0;
    while (rc == DYC_OK) {
        const uint64_t m = std::min(kCkptChunk, words - off);
        if (m && (hipMemcpyAsync(pp.buf[b], dev + off, m * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipEventRecord(ev[b], s) != hipSuccess))
            rc = fail(DYC_ERR_HIP, "checkpoint copy failed");
        if (prev_n && rc == DYC_OK) {  // write the previous chunk while this one copies
            if (hipEventSynchronize(ev[prev_b]) != hipSuccess) {
                rc = fail(DYC_ERR_HIP, "checkpoint copy failed");
                break;
            }
            h = fnv_words(h, pp.buf[prev_b], prev_n);
            if (fwrite(pp.buf[prev_b], 8, prev_n, f) != prev_n) rc = fail(DYC_ERR_INVALID, "checkpoint write failed");
        }
        if (!m) break;
        prev_n = m;
        prev_b = b;
        b ^= 1;
        off += m;
    }
    if (hipStreamSynchronize(s) != hipSuccess && rc == DYC_OK) rc = fail(DYC_ERR_HIP, "checkpoint copy failed");
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    *hash = h;
    return rc;
}

// Read `words` words from the file into device memory through two pinned buffers; returns
// their FNV-1a.
int load_from_file(FILE *f, uint64_t *dev, uint64_t words, PinnedPair &pp, uint64_t *hash, hipStream_t s)
{
    uint64_t h = kFnvBasis;
    hipEvent_t ev[2] = {nullptr, nullptr};
    HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool used[2] = {false, false};
    int rc = DYC_OK;
    int b = 0;
    for (uint64_t off = 0; off < words && rc == DYC_OK; off += kCkptChunk, b ^= 1) {
        const uint64_t m = std::min(kCkptChunk, words - off);
        if (used[b] && hipEventSynchronize(ev[b]) != hipSuccess) rc = fail(DYC_ERR_HIP, "restore copy failed");
        if (rc == DYC_OK && fread(pp.buf[b], 8, m, f) != m) rc = fail(DYC_ERR_CORRUPT, "checkpoint file is truncated");
        if (rc != DYC_OK) break;
        h = fnv_words(h, pp.buf[b], m);
        if (hipMemcpyAsync(dev + off, pp.buf[b], m * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipEventRecord(ev[b], s) != hipSuccess)
            rc = fail(DYC_ERR_HIP, "restore copy failed");
        used[b] = true;
    }
    if (hipStreamSynchronize(s) != hipSuccess && rc == DYC_OK) rc = fail(DYC_ERR_HIP, "restore copy failed");
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    *hash = h;
    return rc;
}

}  // namespace

extern "C" {

int dyc_abi_version(void) { return DYC_ABI_VERSION; }

const char *dyc_last_error(void) { return dyc::host::last_error(); }

int dyc_config_init(dyc_config *cfg)
{
    if (!cfg) return fail(DYC_ERR_INVALID, "null config");
    memset(cfg, 0, sizeof *cfg);
    cfg->kick_bound = kDefaultKickBound;
    cfg->rescue_kicks = 128;
    cfg->max_retries = 64;
    cfg->max_fail_rounds = 8;
    cfg->lower_bound = 0.50;
    cfg->upper_bound = 0.90;
    cfg->auto_resize = 0;
    cfg->find_lanes = 8;
    cfg->insert_lanes = 1;
    cfg->block_threads = 256;
    cfg->max_blocks = 0;
    cfg->min_capacity = 5033164ull;  // TABLE_NUM * 1024 * 1024 * 1.2 (api.cpp:496)
    cfg->insert_order = 0;
    cfg->region_mib = 64;
    cfg->subtables = kDefaultSubtables;
    return DYC_OK;
}

int dyc_create(const dyc_config *cfg_in, uint64_t capacity_hint, dyc_table **out)
{
    if (!out) return fail(DYC_ERR_INVALID, "null output handle");
    *out = nullptr;
    dyc_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        dyc_config_init(&cfg);
    int rc0 = check_config(cfg);
    if (rc0) return rc0;
    cfg.subtables = config_subtables(cfg);
    const uint32_t nb = initial_buckets(capacity_hint, cfg.subtables);
    if (nb >= kMaxBuckets) return fail(DYC_ERR_INVALID, "capacity hint too large");

    dyc_table *t = new dyc_table();
    t->cfg = cfg;
    desc_init(t->D, cfg.subtables);
    HIP_TRY(hipGetDevice(&t->device));
    int rc = DYC_OK;
    if (hipMalloc(&t->ctr, kCounterBytes) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&t->hctr), sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&t->hpost), kPostWords * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(DYC_ERR_OOM, "counter allocation failed");
    } else {
        memset(t->hpost, 0, kPostWords * sizeof(unsigned long long));  // no sequence number yet
    }
    if (rc == DYC_OK) {
        const uint32_t nbs[kMaxSubtables] = {nb, nb, nb, nb, nb, nb};
        uint8_t *base = nullptr;
        uint64_t *sub[kMaxSubtables] = {};
        rc = alloc_slab(t->D.d, nbs, &base, sub);
        if (rc == DYC_OK) rc = install_slab(t, base, sub, nbs);
        for (uint32_t i = 0; i < t->D.d && rc == DYC_OK; ++i)
            if (launch_zero(reinterpret_cast<unsigned long long *>(sub[i]), (uint64_t)nb * kSlots, nullptr) != hipSuccess)
                rc = fail(DYC_ERR_HIP, "zeroing the table failed");
    }
    if (rc == DYC_OK) rc = reserve_pending(t, DYC_NOALLOC_MAX_KEYS, nullptr);
    if (rc == DYC_OK && hipStreamSynchronize(nullptr) != hipSuccess) rc = fail(DYC_ERR_HIP, "create sync failed");
    t->zeroed = rc == DYC_OK;
    if (rc != DYC_OK) {
        char keep[sizeof g_err];
        memcpy(keep, g_err, sizeof keep);
        dyc_destroy(t);
        memcpy(g_err, keep, sizeof keep);
        return rc;
    }
    *out = t;
    return DYC_OK;
}

int dyc_destroy(dyc_table *t)
{
    if (!t) return DYC_OK;
    DeviceGuard dg(t->device);
    (void)settle(t);
    (void)hipDeviceSynchronize();
    async_free(t);
    for (uint32_t i = 0; i < t->D.d; ++i) (void)release_sub(t, i);
    t->reclaim.drain();  // every subtable is freed when this returns
    if (t->pend) (void)hipFree(t->pend);
    if (t->ctr) (void)hipFree(t->ctr);
    if (t->part) (void)hipFree(t->part);
    if (t->part_counts) (void)hipFree(t->part_counts);
    for (int b = 0; b < 2; ++b) {
        if (t->bk_cid[b]) (void)hipFree(t->bk_cid[b]);
        if (t->bk_ent[b]) (void)hipFree(t->bk_ent[b]);
    }
    if (t->bk_start) (void)hipFree(t->bk_start);
    if (t->bk_partial) (void)hipFree(t->bk_partial);
    if (t->bk_ovf) (void)hipFree(t->bk_ovf);
    if (t->bk_hovf) (void)hipHostFree(t->bk_hovf);
    if (t->hctr) (void)hipHostFree(t->hctr);
    if (t->hpost) (void)hipHostFree(t->hpost);
    if (t->upd) (void)hipFree(t->upd);
    if (t->live) (void)hipFree(t->live);
    free_stage(t);
    delete t;
    return DYC_OK;
}

int dyc_insert(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream)
{
    return dyc_insert_counted(t, keys, vals, n, nullptr, stream);
}

int dyc_insert_counted(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, uint32_t *iter_count,
                       dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    SETTLE(t);
    bool zeroed = t->zeroed;  // every slot zero before this call (dropped by any resize below)
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = t->st.last_max_chain = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    int rc;
    if (t->cfg.auto_resize) {
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = do_resize_up(t, s);
            if (rc) return rc;
            zeroed = false;
        }
    }
    rc = reserve_pending(t, std::max<uint64_t>(65536, n / 16), s);
    if (rc) return rc;
    const uint64_t room = t->pend_cap - t->pend_n;
    // per-key counts follow input order, which the bulk fill and the region partition do not keep
    const uint32_t R = iter_count ? 0 : insert_regions(t, n);
    bool bulk = !iter_count && use_bulk_fill(t, n);
    if (bulk) {
        const dyc::host::ErrorKeep keep;
        rc = reserve_bulk(t, n, make_chunks(t->D), s);
        if (rc == DYC_ERR_OOM) {  // no room for the workspace: the eviction insert needs none
            bulk = false;
            keep.restore();  // the call goes on, so it leaves no failure message
        } else if (rc) {
            return rc;
        }
    }
    uint64_t left = 0;  // entries the bulk fill left to the eviction insert
    if (bulk) {
        rc = bulk_insert(t, keys, vals, n, zeroed, s, &left);
        if (rc) return rc;
    } else if (R) {
        rc = reserve_partition(t, n, s);
        if (rc) return rc;
        HIP_TRY(launch_zero(t->part_counts, kMaxRegions, s));
        HIP_TRY(launch_partition(t->D, make_regions(t->D, R), keys, vals, n, t->part_counts,
                                 t->part_counts + kMaxRegions, t->part, s));
        rc = zero_counters(t, s);
        if (rc) return rc;
        HIP_TRY(launch_insert_entries(t->D, t->part, n, kNoTable, (int)t->cfg.insert_lanes, t->policy(),
                                      t->insert_shape(), t->ctr, sink_at(t, t->pend_n), s));
    } else {
        rc = zero_counters(t, s);
        if (rc) return rc;
        PendSink sink = sink_at(t, t->pend_n);
        sink.chain = iter_count;
        HIP_TRY(launch_insert(t->D, keys, vals, n, (int)t->cfg.insert_lanes, t->policy(), t->insert_shape(), t->ctr,
                              sink, s));
    }
    rc = read_counters(t, s);
    if (rc) return rc;
    const uint64_t rejected = t->hctr->rejected;
    t->st.rejected += rejected;
    t->count += n - rejected;
    if (bulk) t->st.last_filled = n - rejected - left;
    rc = absorb_insert_counters(t, room);
    t->st.pending = t->pend_n;
    if (rc) return rc;
    if (t->pend_n > pend_before) return place_pending(t, s, pend_before);
    return DYC_OK;
}

int dyc_insert_async(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    t->zeroed = false;
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return DYC_OK;
    int rc;
    if (t->cfg.auto_resize &&
        (double)(t->count + t->async_keys + n) > t->cfg.upper_bound * (double)t->capacity()) {
        // a resize replaces subtables the inserts in flight write: complete them first
        rc = settle(t);
        if (rc) return rc;
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = do_resize_up(t, s);
            if (rc) return rc;
        }
    }
    AsyncSlot *a = nullptr;
    for (AsyncSlot &x : t->async)
        if (!x.busy) {
            a = &x;
            break;
        }
    if (!a) {  // every slot in flight: complete them
        rc = settle(t);
        if (rc) return rc;
        a = &t->async[0];
    }
    rc = async_reserve(*a, n);
    if (rc) return rc;
    if (t->async_keys == 0) t->st.last_kicks = t->st.last_max_chain = t->st.last_filled = 0;
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(a->ctr), kCounterWords, s));
    HIP_TRY(launch_insert(t->D, keys, vals, n, (int)t->cfg.insert_lanes, t->policy(), t->insert_shape(), a->ctr,
                          PendSink{a->pend, nullptr, a->cap, 0, 0, nullptr}, s));
    HIP_TRY(launch_fold_post(a->ctr, a->post, ++a->seq, s));
    a->busy = true;
    a->n = n;
    a->stream = s;
    t->async_keys += n;
    return DYC_OK;
}

int dyc_sync(dyc_table *t)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    return settle(t);
}

int dyc_insert_noalloc(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    if (n > DYC_NOALLOC_MAX_KEYS)
        return fail(DYC_ERR_INVALID, "dyc_insert_noalloc takes at most %u keys per call (got %llu)", DYC_NOALLOC_MAX_KEYS,
                    (unsigned long long)n);
    for (const AsyncSlot &a : t->async)
        if (a.busy)
            return fail(DYC_ERR_INVALID, "dyc_insert_noalloc: inserts of dyc_insert_async are in flight (dyc_sync "
                                         "completes them, and may allocate)");
    const uint64_t room = t->pend_cap - t->pend_n;
    if (n > room)
        return fail(DYC_ERR_INVALID,
                    "dyc_insert_noalloc: %llu keys, but the pending list has room for %llu entries and cannot grow "
                    "without allocating (dyc_insert re-places pending entries and grows it)",
                    (unsigned long long)n, (unsigned long long)room);
    DeviceGuard dg(t->device);
    t->zeroed = false;
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = t->st.last_max_chain = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    Counters *C = nullptr;
    int rc = noalloc_counters(t->device, &C);
    if (rc) return rc;
    // the block keeps the last call's contents (a call that failed may have left it dirty)
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(C), kCounterWords, s));
    HIP_TRY(launch_insert(t->D, keys, vals, n, (int)t->cfg.insert_lanes, t->policy(), t->insert_shape(), C,
                          sink_at(t, t->pend_n), s));
    // read back as read_counters does, from the static block (the fold also clears it)
    const unsigned long long seq = ++t->post_seq;
    HIP_TRY(launch_fold_post(C, t->hpost, seq, s));
    HIP_TRY(await_post(t->hpost, seq, s));
    memcpy(t->hctr, const_cast<const unsigned long long *>(t->hpost) + 1, sizeof(Counters));
    const uint64_t rejected = t->hctr->rejected;
    t->st.rejected += rejected;
    t->count += n - rejected;
    rc = absorb_insert_counters(t, room);
    t->st.pending = t->pend_n;
    return rc;
}

int dyc_upsert(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, uint8_t *updated,
               dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    SETTLE(t);
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = t->st.last_max_chain = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    int rc;
    if (t->cfg.auto_resize) {  // sized as if every key were new (api.cpp:360-369)
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = do_resize_up(t, s);
            if (rc) return rc;
        }
    }
    uint8_t *flags = updated;
    if (!flags) {
        if (n > t->upd_cap) {
            HIP_TRY(hipStreamSynchronize(s));
            if (t->upd) HIP_TRY(hipFree(t->upd));
            t->upd = nullptr;
            t->upd_cap = 0;
            if (hipMalloc(&t->upd, n) != hipSuccess) {
                (void)hipGetLastError();
                return fail(DYC_ERR_OOM, "hipMalloc of %llu upsert flags failed", (unsigned long long)n);
            }
            t->upd_cap = n;
        }
        flags = t->upd;
    }
    rc = reserve_pending(t, std::max<uint64_t>(65536, n / 16), s);
    if (rc) return rc;
    const uint64_t room = t->pend_cap - t->pend_n;
    rc = zero_counters(t, s);
    if (rc) return rc;
    Pool sp;
    HIP_TRY(launch_update(t->D, keys, vals, flags, n, (int)t->cfg.find_lanes, t->shape(), t->ctr, s, stash_of(t, sp)));
    HIP_TRY(launch_insert(t->D, keys, vals, n, (int)t->cfg.insert_lanes, t->policy(), t->insert_shape(), t->ctr,
                          sink_at(t, t->pend_n), s, flags));
    rc = read_counters(t, s);
    if (rc) return rc;
    const uint64_t rejected = t->hctr->rejected, upd = t->hctr->updated;
    t->st.rejected += rejected;
    t->count += n - rejected - upd;
    rc = absorb_insert_counters(t, room);
    t->st.pending = t->pend_n;
    if (rc) return rc;
    if (t->pend_n > pend_before) return place_pending(t, s, pend_before);
    return DYC_OK;
}

int dyc_find(dyc_table *t, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(t->device);
    SETTLE(t);
    Pool sp;
    HIP_TRY(launch_find(t->D, keys, vals, hit, n, (int)t->cfg.find_lanes, t->shape(), (hipStream_t)stream,
                        stash_of(t, sp)));
    return DYC_OK;
}

int dyc_erase(dyc_table *t, const uint32_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return DYC_OK;
    int rc = zero_counters(t, s);
    if (rc) return rc;
    Pool sp;
    HIP_TRY(launch_erase(t->D, keys, deleted, n, (int)t->cfg.find_lanes, t->shape(), t->ctr, s, stash_of(t, sp)));
    rc = read_counters(t, s);
    if (rc) return rc;
    t->count -= std::min<uint64_t>(t->count, t->hctr->deleted);
    rc = compact_pending(t, s);
    if (rc) return rc;
    if (t->cfg.auto_resize && (double)t->count < t->cfg.lower_bound * (double)t->capacity() &&
        t->capacity() > t->cfg.min_capacity) {
        rc = do_resize_down(t, s);
        if (rc == DYC_ERR_ODD_BUCKETS) rc = DYC_OK;  // policy: skip, the table stays larger
        if (rc) return rc;
    }
    return DYC_OK;
}

int dyc_resize_up(dyc_table *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    SETTLE(t);
    return do_resize_up(t, (hipStream_t)stream);
}

int dyc_resize_down(dyc_table *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    SETTLE(t);
    return do_resize_down(t, (hipStream_t)stream);
}

int dyc_get_config(const dyc_table *t, dyc_config *out)
{
    if (!t || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = t->cfg;
    return DYC_OK;
}

int dyc_set_config(dyc_table *t, const dyc_config *cfg)
{
    if (!t || !cfg) return fail(DYC_ERR_INVALID, "null argument");
    int rc = check_config(*cfg);
    if (rc) return rc;
    if (config_subtables(*cfg) != t->D.d)
        return fail(DYC_ERR_INVALID, "subtables is fixed at create (%u)", t->D.d);
    t->cfg = *cfg;
    t->cfg.subtables = t->D.d;
    return DYC_OK;
}

int dyc_get_stats(const dyc_table *t, dyc_stats *out)
{
    if (!t || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = t->st;
    out->count = t->count;
    out->capacity = t->capacity();
    for (uint32_t i = 0; i < t->D.d; ++i) out->buckets[i] = t->D.nb[i];
    out->pending = t->pend_n;
    out->subtables = t->D.d;
    return DYC_OK;
}

int dyc_count_occupied(dyc_table *t, uint64_t *occupied, dyc_stream stream)
{
    if (!t || !occupied) return fail(DYC_ERR_INVALID, "null argument");
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    int rc = zero_counters(t, s);
    if (rc) return rc;
    for (uint32_t i = 0; i < t->D.d; ++i) HIP_TRY(launch_count(t->D.sub[i], (uint64_t)t->D.nb[i] * kSlots, t->ctr, s));
    rc = read_counters(t, s);
    if (rc) return rc;
    *occupied = t->hctr->occupied;
    return DYC_OK;
}

int dyc_clear(dyc_table *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    for (uint32_t i = 0; i < t->D.d; ++i)
        HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(t->D.sub[i]), (uint64_t)t->D.nb[i] * kSlots, s));
    t->count = 0;
    t->pend_n = 0;
    t->st.pending = 0;
    t->zeroed = true;
    return DYC_OK;
}

int dyc_probe_count(dyc_table *t, const uint32_t *keys, uint64_t n, uint64_t *probes, uint64_t *pos1_hits,
                    dyc_stream stream)
{
    if (!t || (n && !keys)) return fail(DYC_ERR_INVALID, "null argument");
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    int rc = zero_counters(t, s);
    if (rc) return rc;
    HIP_TRY(launch_probe_count(t->D, keys, n, t->ctr, s));
    rc = read_counters(t, s);
    if (rc) return rc;
    if (probes) *probes = t->hctr->kicks;
    if (pos1_hits) *pos1_hits = t->hctr->occupied;
    return DYC_OK;
}

int dyc_dump(dyc_table *t, uint32_t sub, uint64_t *host_slots, uint64_t n_slots, dyc_stream stream)
{
    if (!t || !host_slots || sub >= t->D.d) return fail(DYC_ERR_INVALID, "bad argument");
    const uint64_t slots = (uint64_t)t->D.nb[sub] * kSlots;
    if (n_slots < slots) return fail(DYC_ERR_INVALID, "buffer holds %llu slots, subtable has %llu",
                                     (unsigned long long)n_slots, (unsigned long long)slots);
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(host_slots, t->D.sub[sub], slots * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DYC_OK;
}

int dyc_load(dyc_table *t, const uint32_t *nb, const uint64_t *const *images, dyc_stream stream)
{
    if (!t || !nb || !images) return fail(DYC_ERR_INVALID, "null argument");
    for (uint32_t i = 0; i < t->D.d; ++i)
        if (nb[i] == 0 || nb[i] >= kMaxBuckets || !images[i]) return fail(DYC_ERR_INVALID, "bad subtable %d", i);
    DeviceGuard dg(t->device);
    SETTLE(t);
    t->zeroed = false;
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    uint8_t *base = nullptr;
    uint64_t *sub[kMaxSubtables] = {};
    int rc = alloc_slab(t->D.d, nb, &base, sub);
    if (rc) return rc;
    uint64_t count = 0;
    for (uint32_t i = 0; i < t->D.d; ++i) {
        const size_t bytes = (size_t)nb[i] * kBucketBytes;
        if (hipMemcpyAsync(sub[i], images[i], bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            (void)hipFree(base);
            return fail(DYC_ERR_HIP, "copying subtable %u to the device failed", i);
        }
        for (uint64_t j = 0; j < (uint64_t)nb[i] * kSlots; ++j) count += images[i][j] != 0;
    }
    rc = install_slab(t, base, sub, nb);
    if (rc) return rc;
    t->count = count;
    t->pend_n = 0;
    t->st.pending = 0;
    return DYC_OK;
}

int dyc_pending(dyc_table *t, uint64_t *host_entries, uint64_t cap, uint64_t *n_out, dyc_stream stream)
{
    if (!t || !n_out) return fail(DYC_ERR_INVALID, "null argument");
    {
        DeviceGuard dg(t->device);
        SETTLE(t);
    }
    *n_out = t->pend_n;
    const uint64_t m = std::min(cap, t->pend_n);
    if (m && host_entries) {
        DeviceGuard dg(t->device);
        HIP_TRY(hipMemcpyAsync(host_entries, t->pend, m * sizeof(uint64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    }
    return DYC_OK;
}

// ---- host-pointer entry points: chunks of kStageKeys keys move host -> pinned (threads) ->
// device on a copy stream while the previous chunk's kernel runs on `stream`; results move back
// on a second copy stream. Resizes and pending entries are handled once per call, as in the
// device-pointer calls.
int dyc_insert_host(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    SETTLE(t);
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = t->st.last_max_chain = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    int rc;
    if (t->cfg.auto_resize) {
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = do_resize_up(t, s);
            if (rc) return rc;
        }
    }
    rc = reserve_pending(t, std::max<uint64_t>(65536, n / 16), s);
    if (rc) return rc;
    const uint64_t room = t->pend_cap - t->pend_n;
    DycHostStage *S = nullptr;
    rc = get_stage(t, &S);
    if (rc) return rc;
    rc = zero_counters(t, s);
    if (rc) return rc;
    bool used[2] = {false, false};
    uint64_t off = 0;
    for (int c = 0; off < n; ++c) {
        const int b = c & 1;
        const uint64_t m = std::min(kStageKeys, n - off);
        if (used[b]) HIP_TRY(hipEventSynchronize(S->done[b]));  // chunk c - 2 no longer reads buffer b
        S->pool->copy(S->hk[b], keys + off, m * 4);
        S->pool->copy(S->hv[b], vals + off, m * 4);
        HIP_TRY(hipMemcpyAsync(S->dk[b], S->hk[b], m * 4, hipMemcpyHostToDevice, S->h2d));
        HIP_TRY(hipMemcpyAsync(S->dv[b], S->hv[b], m * 4, hipMemcpyHostToDevice, S->h2d));
        HIP_TRY(hipEventRecord(S->in_ready[b], S->h2d));
        HIP_TRY(hipStreamWaitEvent(s, S->in_ready[b], 0));
        HIP_TRY(launch_insert(t->D, S->dk[b], S->dv[b], m, (int)t->cfg.insert_lanes, t->policy(), t->insert_shape(),
                              t->ctr, sink_at(t, t->pend_n), s));
        HIP_TRY(hipEventRecord(S->done[b], s));
        used[b] = true;
        off += m;
    }
    rc = read_counters(t, s);
    if (rc) return rc;
    const uint64_t rejected = t->hctr->rejected;
    t->st.rejected += rejected;
    t->count += n - rejected;
    rc = absorb_insert_counters(t, room);
    t->st.pending = t->pend_n;
    if (rc) return rc;
    if (t->pend_n > pend_before) return place_pending(t, s, pend_before);
    return DYC_OK;
}

namespace {
// find (ERASE false) or erase over host arrays; see dyc_insert_host.
int probe_host(dyc_table *t, const uint32_t *keys, uint32_t *vals, uint8_t *flags, uint64_t n, hipStream_t s,
               bool erase)
{
    DycHostStage *S = nullptr;
    int rc = get_stage(t, &S);
    if (rc) return rc;
    bool used[2] = {false, false};
    uint64_t prev_off[2] = {0, 0}, prev_m[2] = {0, 0};
    auto drain = [&](int b) -> int {  // results of the chunk that used buffer b, to the caller
        if (!used[b]) return DYC_OK;
        HIP_TRY(hipEventSynchronize(S->out_ready[b]));
        if (vals && !erase) S->pool->copy(vals + prev_off[b], S->hv[b], prev_m[b] * 4);
        if (flags) S->pool->copy(flags + prev_off[b], S->hf[b], prev_m[b]);
        used[b] = false;
        return DYC_OK;
    };
    uint64_t off = 0;
    for (int c = 0; off < n; ++c) {
        const int b = c & 1;
        const uint64_t m = std::min(kStageKeys, n - off);
        rc = drain(b);
        if (rc) return rc;
        S->pool->copy(S->hk[b], keys + off, m * 4);
        HIP_TRY(hipMemcpyAsync(S->dk[b], S->hk[b], m * 4, hipMemcpyHostToDevice, S->h2d));
        HIP_TRY(hipEventRecord(S->in_ready[b], S->h2d));
        HIP_TRY(hipStreamWaitEvent(s, S->in_ready[b], 0));
        Pool sp;
        if (erase)
            HIP_TRY(launch_erase(t->D, S->dk[b], flags ? S->df[b] : nullptr, m, (int)t->cfg.find_lanes, t->shape(),
                                 t->ctr, s, stash_of(t, sp)));
        else
            HIP_TRY(launch_find(t->D, S->dk[b], vals ? S->dv[b] : nullptr, flags ? S->df[b] : nullptr, m,
                                (int)t->cfg.find_lanes, t->shape(), s, stash_of(t, sp)));
        HIP_TRY(hipEventRecord(S->done[b], s));
        HIP_TRY(hipStreamWaitEvent(S->d2h, S->done[b], 0));
        if (vals && !erase) HIP_TRY(hipMemcpyAsync(S->hv[b], S->dv[b], m * 4, hipMemcpyDeviceToHost, S->d2h));
        if (flags) HIP_TRY(hipMemcpyAsync(S->hf[b], S->df[b], m, hipMemcpyDeviceToHost, S->d2h));
        HIP_TRY(hipEventRecord(S->out_ready[b], S->d2h));
        used[b] = true;
        prev_off[b] = off;
        prev_m[b] = m;
        off += m;
    }
    for (int b = 0; b < 2; ++b) {
        rc = drain(b);
        if (rc) return rc;
    }
    return DYC_OK;
}
}  // namespace

int dyc_find_host(dyc_table *t, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    if (n == 0) return DYC_OK;
    DeviceGuard dg(t->device);
    SETTLE(t);
    return probe_host(t, keys, vals, hit, n, (hipStream_t)stream, false);
}

int dyc_erase_host(dyc_table *t, const uint32_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    if (n == 0) return DYC_OK;
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    int rc = zero_counters(t, s);
    if (rc) return rc;
    rc = probe_host(t, keys, nullptr, deleted, n, s, true);
    if (rc) return rc;
    rc = read_counters(t, s);
    if (rc) return rc;
    t->count -= std::min<uint64_t>(t->count, t->hctr->deleted);
    rc = compact_pending(t, s);
    if (rc) return rc;
    if (t->cfg.auto_resize && (double)t->count < t->cfg.lower_bound * (double)t->capacity() &&
        t->capacity() > t->cfg.min_capacity) {
        rc = do_resize_down(t, s);
        if (rc == DYC_ERR_ODD_BUCKETS) rc = DYC_OK;
        if (rc) return rc;
    }
    return DYC_OK;
}

int dyc_checkpoint_write(dyc_table *t, const char *path, dyc_stream stream)
{
    if (!t || !path) return fail(DYC_ERR_INVALID, "null argument");
    DeviceGuard dg(t->device);
    SETTLE(t);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    FILE *f = fopen(path, "wb");
    if (!f) return fail(DYC_ERR_INVALID, "cannot open %s for writing", path);
    CkptHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.subtables = t->D.d;
    for (uint32_t i = 0; i < t->D.d; ++i) h.buckets[i] = t->D.nb[i];
    h.count = t->count;
    h.pending = t->pend_n;
    int rc = fwrite(&h, sizeof h, 1, f) == 1 ? DYC_OK : fail(DYC_ERR_INVALID, "checkpoint write failed");
    PinnedPair pp;
    if (rc == DYC_OK && !pp.alloc(kCkptChunk * 8)) rc = fail(DYC_ERR_OOM, "pinned staging for the checkpoint");
    for (uint32_t i = 0; i < t->D.d && rc == DYC_OK; ++i)
        rc = dump_to_file(f, t->D.sub[i], (uint64_t)t->D.nb[i] * kSlots, pp, &h.checksum[i], s);
    if (rc == DYC_OK) rc = dump_to_file(f, t->pend, t->pend_n, pp, &h.pending_check, s);
    h.header_check = fnv_bytes(&h, offsetof(CkptHeader, header_check));
    if (rc == DYC_OK && (fseek(f, 0, SEEK_SET) != 0 || fwrite(&h, sizeof h, 1, f) != 1))
        rc = fail(DYC_ERR_INVALID, "checkpoint header write failed");
    if (fclose(f) != 0 && rc == DYC_OK) rc = fail(DYC_ERR_INVALID, "checkpoint close failed");
    return rc;
}

int dyc_checkpoint_read(const char *path, const dyc_config *cfg, dyc_table **out)
{
    if (!path || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(DYC_ERR_INVALID, "cannot open %s", path);
    CkptHeader h;
    int rc = fread(&h, sizeof h, 1, f) == 1 ? DYC_OK : fail(DYC_ERR_CORRUPT, "checkpoint header is truncated");
    if (rc == DYC_OK && (memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion ||
                         h.subtables < (uint32_t)kMinSubtables || h.subtables > (uint32_t)kMaxSubtables))
        rc = fail(DYC_ERR_CORRUPT, "not a DyCuckoo checkpoint (magic/version/subtables)");
    if (rc == DYC_OK && h.header_check != fnv_bytes(&h, offsetof(CkptHeader, header_check)))
        rc = fail(DYC_ERR_CORRUPT, "checkpoint header checksum mismatch");
    uint64_t slots = 0;
    for (uint32_t i = 0; i < h.subtables && rc == DYC_OK; ++i) {
        if (h.buckets[i] == 0 || h.buckets[i] >= kMaxBuckets) rc = fail(DYC_ERR_CORRUPT, "bad bucket count");
        slots += (uint64_t)h.buckets[i] * kSlots;
    }
    if (rc == DYC_OK && (fseek(f, 0, SEEK_END) != 0 ||
                         (uint64_t)ftell(f) != sizeof h + (slots + h.pending) * 8 || fseek(f, sizeof h, SEEK_SET) != 0))
        rc = fail(DYC_ERR_CORRUPT, "checkpoint size does not match its header");
    dyc_table *t = nullptr;
    dyc_config c;
    if (cfg)
        c = *cfg;
    else
        dyc_config_init(&c);
    c.subtables = h.subtables;
    if (rc == DYC_OK) rc = dyc_create(&c, 64, &t);
    PinnedPair pp;
    if (rc == DYC_OK && !pp.alloc(kCkptChunk * 8)) rc = fail(DYC_ERR_OOM, "pinned staging for the restore");
    uint8_t *base = nullptr;
    uint64_t *sub[kMaxSubtables] = {};
    if (rc == DYC_OK) rc = alloc_slab(h.subtables, h.buckets, &base, sub);
    if (rc == DYC_OK) rc = install_slab(t, base, sub, h.buckets);  // the table now owns the slab
    for (uint32_t i = 0; i < h.subtables && rc == DYC_OK; ++i) {
        uint64_t hash = 0;
        rc = load_from_file(f, sub[i], (uint64_t)h.buckets[i] * kSlots, pp, &hash, nullptr);
        if (rc == DYC_OK && hash != h.checksum[i]) rc = fail(DYC_ERR_CORRUPT, "subtable %u checksum mismatch", i);
    }
    if (rc == DYC_OK) rc = reserve_pending(t, h.pending, nullptr);
    if (rc == DYC_OK) {
        uint64_t hash = 0;
        rc = load_from_file(f, t->pend, h.pending, pp, &hash, nullptr);
        if (rc == DYC_OK && hash != h.pending_check) rc = fail(DYC_ERR_CORRUPT, "pending list checksum mismatch");
    }
    fclose(f);
    if (rc == DYC_OK) {
        t->pend_n = h.pending;
        t->st.pending = h.pending;
        t->count = h.count;
        uint64_t occ = 0;
        rc = dyc_count_occupied(t, &occ, nullptr);
        if (rc == DYC_OK && occ + h.pending != h.count)
            rc = fail(DYC_ERR_CORRUPT, "checkpoint count %llu does not match its %llu stored entries",
                      (unsigned long long)h.count, (unsigned long long)(occ + h.pending));
    }
    if (rc != DYC_OK) {
        if (t) {
            char keep[sizeof g_err];
            memcpy(keep, g_err, sizeof keep);
            dyc_destroy(t);
            memcpy(g_err, keep, sizeof keep);
        }
        return rc;
    }
    *out = t;
    return DYC_OK;
}

}  // extern "C"
