// dyc_group.hip -- many DyCuckoo tables served together: the dyc_group_* half of include/dyc.h.
//
// All subtables of all tables live in one device arena. A device directory holds one Desc per
// table. A batch (keys segmented per table by host offsets) is cut into work items of at most
// item_keys keys of one table, uploaded, and served by one persistent launch; the per-table
// variants launch the single-table kernels once per table for comparison.
#include "dyc_bulk_layout.hpp"
#include "dyc_host.hpp"
#include "dyc_kernels.hpp"

#include <string.h>

#include <algorithm>
#include <vector>

using namespace dyc;
using dyc::host::BulkLayout;
using dyc::host::items_by_xcd;
using dyc::host::kXcdLayoutKeys;
using dyc::host::DeviceGuard;
using dyc::host::fail;

struct dyc_group {
    dyc_config cfg;
    dyc_group_launch launch;
    int device = 0;
    uint32_t n = 0;
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<Desc> hdir;
    Desc *ddir = nullptr;
    std::vector<uint64_t> count;  // per table: placed + pending
    Counters *ctr = nullptr;
    Counters *hctr = nullptr;
    unsigned long long *tcounts = nullptr;   // per-table counts of the last insert/erase
    unsigned long long *htcounts = nullptr;  // pinned mirror
    uint64_t *pend = nullptr;
    uint32_t *pend_tag = nullptr;
    uint64_t pend_cap = 0, pend_n = 0;
    GroupItem *ditems = nullptr;
    GroupItem *hitems = nullptr;  // pinned staging
    size_t items_cap = 0;
    hipEvent_t items_free = nullptr;  // the last upload from hitems has completed
    dyc_stats st{};
    // every slot of every table is zero: set at create and clear, dropped by every insert
    bool zeroed = false;
    // bulk fill (cfg.insert_order 3), grown on demand and kept by the group. Shared by the eager
    // calls and the plans (a plan refuses to run once one of these moved): partitioned entries and
    // the overflow (8 B per key each), local chunk ids (1 B per key), per-block chunk counts,
    // per-chunk start / count / table, and the overflow counts of both rounds (2 per table). The
    // eager calls' own block and table arrays (device, with pinned staging) and the read-back of
    // the last round's overflow counts.
    uint64_t *bk_ent = nullptr, *bk_ovf = nullptr;
    uint8_t *bk_lcid = nullptr;
    uint64_t bk_cap = 0;
    uint32_t *bk_partial = nullptr;
    uint64_t bk_partial_cap = 0;
    GroupChunk *bk_chunks = nullptr;  // chunk records of a round
    uint64_t bk_chunk_cap = 0;
    unsigned long long *bk_ovfn = nullptr, *bk_hovfn = nullptr;
    uint32_t bk_tabs_cap = 0;
    GroupItem *bk_blk = nullptr, *bk_hblk = nullptr;  // device: 3 per block (rounds 1, 2, eviction)
    uint32_t *bk_btab = nullptr, *bk_hbtab = nullptr;
    uint64_t bk_blk_cap = 0;
    BulkTable *bk_tabs = nullptr, *bk_htabs = nullptr;  // device: 2 per table (rounds 1, 2)

    InsertPolicy policy() const { return InsertPolicy{launch.kick_bound, cfg.rescue_kicks, cfg.max_retries}; }
    // The pending list is the group's stash: entries tagged with their table, searched by finds
    // and erases of that table after both buckets. nullptr when nothing is pending.
    const Pool *stash(Pool &p) const
    {
        p = Pool{pend, pend_tag, pend_cap, nullptr, pend_n};
        return pend_n ? &p : nullptr;
    }
};

namespace {

constexpr uint32_t kDefaultItemKeys = 4096;

// An erase leaves the stash entries it deleted as zero words: drop them with their tags
// (synchronous; the list is short whenever it is not empty).
int compact_stash(dyc_group *g, hipStream_t s)
{
    if (g->pend_n == 0) return DYC_OK;
    std::vector<uint64_t> e(g->pend_n);
    std::vector<uint32_t> tg(g->pend_n);
    HIP_TRY(hipMemcpyAsync(e.data(), g->pend, g->pend_n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(tg.data(), g->pend_tag, g->pend_n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint64_t m = 0;
    for (uint64_t i = 0; i < e.size(); ++i)
        if (e[i]) {
            e[m] = e[i];
            tg[m] = tg[i];
            ++m;
        }
    if (m != g->pend_n) {
        if (m) {
            HIP_TRY(hipMemcpyAsync(g->pend, e.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(g->pend_tag, tg.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        g->pend_n = m;
    }
    g->st.pending = g->pend_n;
    return DYC_OK;
}

int check_launch(const dyc_group_launch &l)
{
    const bool find_ok = l.find_lanes == 4 || l.find_lanes == 8 || l.find_lanes == 16;
    const bool ins_ok = l.insert_lanes == 1 || l.insert_lanes == 2 || l.insert_lanes == 4 || l.insert_lanes == 8 ||
                        l.insert_lanes == 16;
    if (!find_ok || !ins_ok) return fail(DYC_ERR_INVALID, "group lanes: find 4/8/16, insert 1/2/4/8/16");
    auto wpe_ok = [](uint32_t w) { return w == 0 || w == 1 || w == 6 || w == 8; };
    if (!wpe_ok(l.insert_waves_per_eu) || !wpe_ok(l.find_waves_per_eu))
        return fail(DYC_ERR_INVALID, "group waves per EU: 0 (compiler's choice), 6 or 8");
    return DYC_OK;
}

int check_offsets(const dyc_group *g, const uint64_t *offsets)
{
    if (!offsets) return fail(DYC_ERR_INVALID, "null offsets");
    if (offsets[0] != 0) return fail(DYC_ERR_INVALID, "offsets[0] must be 0");
    for (uint32_t i = 0; i < g->n; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(DYC_ERR_INVALID, "offsets decrease at table %u", i);
    return DYC_OK;
}

// Cut segments (begin, len, table) into items of at most `per` indexes and upload them; returns
// the item count in *n_items.
// Room for `need` items in the item buffers (host and device), once the last upload is done.
int reserve_items(dyc_group *g, uint64_t need)
{
    if (need > 0xffffffffull) return fail(DYC_ERR_INVALID, "batch too large for 32-bit item indexes");
    HIP_TRY(hipEventSynchronize(g->items_free));  // hitems is free to rewrite
    if (need > g->items_cap) {
        const size_t cap = std::max<size_t>(need, g->items_cap * 2);
        if (g->ditems) HIP_TRY(hipFree(g->ditems));
        if (g->hitems) HIP_TRY(hipHostFree(g->hitems));
        g->ditems = nullptr;
        g->hitems = nullptr;
        g->items_cap = 0;
        if (hipMalloc(&g->ditems, cap * sizeof(GroupItem)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&g->hitems), cap * sizeof(GroupItem), hipHostMallocDefault) !=
                hipSuccess) {
            (void)hipGetLastError();
            return fail(DYC_ERR_OOM, "work item buffers (%zu items) failed", cap);
        }
        g->items_cap = cap;
    }
    return DYC_OK;
}

int upload_item_list(dyc_group *g, const std::vector<GroupItem> &items, hipStream_t s, uint32_t *n_items)
{
    const int rc = reserve_items(g, items.size());
    if (rc) return rc;
    const uint32_t m = (uint32_t)items.size();
    std::copy(items.begin(), items.end(), g->hitems);
    if (m) HIP_TRY(hipMemcpyAsync(g->ditems, g->hitems, m * sizeof(GroupItem), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(g->items_free, s));
    *n_items = m;
    return DYC_OK;
}

int upload_segments(dyc_group *g, const std::vector<GroupItem> &segs, uint64_t per, hipStream_t s, uint32_t *n_items)
{
    uint64_t need = 0;
    for (const GroupItem &x : segs) need += (x.len + per - 1) / per;
    const int rc = reserve_items(g, need);
    if (rc) return rc;
    uint32_t m = 0;
    for (const GroupItem &x : segs)
        for (uint64_t b = x.begin; b < x.begin + x.len; b += per)
            g->hitems[m++] = GroupItem{b, (uint32_t)std::min<uint64_t>(per, x.begin + x.len - b), x.table};
    if (m) HIP_TRY(hipMemcpyAsync(g->ditems, g->hitems, m * sizeof(GroupItem), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(g->items_free, s));
    *n_items = m;
    return DYC_OK;
}

// Cut the batch into items (at most item_keys keys of one table) and upload them.
int upload_items(dyc_group *g, const uint64_t *offsets, hipStream_t s, uint32_t *n_items)
{
    const uint64_t per = g->launch.item_keys ? g->launch.item_keys : kDefaultItemKeys;
    std::vector<GroupItem> segs;
    segs.reserve(g->n);
    for (uint32_t i = 0; i < g->n; ++i) {
        const uint64_t len = offsets[i + 1] - offsets[i];
        if (len > 0xffffffffull) return fail(DYC_ERR_INVALID, "table %u's batch is too large for 32-bit lengths", i);
        if (len) segs.push_back(GroupItem{offsets[i], (uint32_t)len, i});
    }
    return upload_segments(g, segs, per, s, n_items);
}

int upload_items_by_xcd(dyc_group *g, const uint64_t *offsets, hipStream_t s, uint32_t *n_items)
{
    const uint64_t per = g->launch.item_keys ? g->launch.item_keys : kDefaultItemKeys;
    for (uint32_t i = 0; i < g->n; ++i)
        if (offsets[i + 1] - offsets[i] > 0xffffffffull)
            return fail(DYC_ERR_INVALID, "table %u's batch is too large for 32-bit lengths", i);
    return upload_item_list(g, items_by_xcd(offsets, g->n, per), s, n_items);
}

int zero_and_run_prologue(dyc_group *g, hipStream_t s)
{
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(g->ctr), kCounterWords, s));
    HIP_TRY(launch_zero(g->tcounts, g->n, s));
    return DYC_OK;
}

int read_back(dyc_group *g, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(g->hctr, g->ctr, kCounterBytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(g->htcounts, g->tcounts, g->n * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    fold_tallies(g->hctr);
    return DYC_OK;
}

int reserve_pending(dyc_group *g, uint64_t extra, hipStream_t s)
{
    const uint64_t need = g->pend_n + extra;
    if (need <= g->pend_cap) return DYC_OK;
    const uint64_t cap = std::max<uint64_t>(need, g->pend_cap * 2);
    uint64_t *np = nullptr;
    uint32_t *nt = nullptr;
    if (hipMalloc(&np, cap * sizeof(uint64_t)) != hipSuccess || hipMalloc(&nt, cap * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        if (np) (void)hipFree(np);
        return fail(DYC_ERR_OOM, "group pending list (%llu entries) failed", (unsigned long long)cap);
    }
    if (g->pend_n) {
        HIP_TRY(hipMemcpyAsync(np, g->pend, g->pend_n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(nt, g->pend_tag, g->pend_n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (g->pend) HIP_TRY(hipFree(g->pend));
    if (g->pend_tag) HIP_TRY(hipFree(g->pend_tag));
    g->pend = np;
    g->pend_tag = nt;
    g->pend_cap = cap;
    return DYC_OK;
}

// After an insert launch: absorb counters and per-table counts.
int finish_insert(dyc_group *g, uint64_t room)
{
    const Counters &c = *g->hctr;
    g->st.kicks += c.kicks;
    g->st.last_kicks = c.kicks;
    g->st.cas_retries += c.retries;
    g->st.kick_failures += c.pending;
    g->st.rejected += c.rejected;
    g->st.lost += c.lost;
    g->pend_n += std::min<uint64_t>(c.pending, room);
    g->st.pending = g->pend_n;
    for (uint32_t i = 0; i < g->n; ++i) g->count[i] += g->htcounts[i];
    if (c.lost) return fail(DYC_ERR_LOST, "group pending list overflow: %llu entries lost", (unsigned long long)c.lost);
    if (c.pending)
        return fail(DYC_ERR_PENDING, "%llu entries exhausted the eviction bound (group tables do not resize)",
                    (unsigned long long)c.pending);
    return DYC_OK;
}

// ------------------------------------------------------------------------------ bulk fill
int bulk_alloc(void **p, size_t bytes, const char *what)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(DYC_ERR_OOM, "hipMalloc of the group bulk-fill %s (%zu bytes) failed", what, bytes);
    }
    return DYC_OK;
}

// Grow-only workspace of a bulk-filled batch: n keys, `blocks` partition blocks, `chunks` chunks
// and `tabs` participating tables. The caller has synchronized the stream (nothing in flight uses
// a buffer being replaced).
int reserve_group_bulk(dyc_group *g, uint64_t n, uint64_t blocks, uint64_t chunks, uint32_t tabs)
{
    int rc = DYC_OK;
    auto regrow = [&](void **p, size_t bytes, const char *what) {
        if (rc) return;
        if (*p) (void)hipFree(*p);
        rc = bulk_alloc(p, bytes, what);
    };
    auto regrow_pinned = [&](void **p, size_t bytes) {
        if (rc) return;
        if (*p) (void)hipHostFree(*p);
        *p = nullptr;
        if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            rc = fail(DYC_ERR_OOM, "pinned group bulk-fill staging (%zu bytes) failed", bytes);
        }
    };
    if (n > g->bk_cap) {
        g->bk_cap = 0;
        regrow((void **)&g->bk_ent, n * 8, "entries");
        regrow((void **)&g->bk_ovf, n * 8, "overflow list");
        regrow((void **)&g->bk_lcid, n, "chunk ids");
        if (rc) return rc;
        g->bk_cap = n;
    }
    if (blocks * kGroupBulkChunks > g->bk_partial_cap) {
        g->bk_partial_cap = 0;
        regrow((void **)&g->bk_partial, blocks * kGroupBulkChunks * sizeof(uint32_t), "block counts");
        if (rc) return rc;
        g->bk_partial_cap = blocks * kGroupBulkChunks;
    }
    if (blocks > g->bk_blk_cap) {
        g->bk_blk_cap = 0;
        regrow((void **)&g->bk_blk, 3 * blocks * sizeof(GroupItem), "blocks");
        regrow((void **)&g->bk_btab, blocks * sizeof(uint32_t), "block tables");
        regrow_pinned((void **)&g->bk_hblk, blocks * sizeof(GroupItem));
        regrow_pinned((void **)&g->bk_hbtab, blocks * sizeof(uint32_t));
        if (rc) return rc;
        g->bk_blk_cap = blocks;
    }
    if (chunks > g->bk_chunk_cap) {
        g->bk_chunk_cap = 0;
        regrow((void **)&g->bk_chunks, chunks * sizeof(GroupChunk), "chunk records");
        if (rc) return rc;
        g->bk_chunk_cap = chunks;
    }
    if (tabs > g->bk_tabs_cap) {
        g->bk_tabs_cap = 0;
        regrow((void **)&g->bk_tabs, 2 * tabs * sizeof(BulkTable), "table list");
        regrow((void **)&g->bk_ovfn, 2 * tabs * sizeof(unsigned long long), "overflow counts");
        regrow_pinned((void **)&g->bk_htabs, tabs * sizeof(BulkTable));
        regrow_pinned((void **)&g->bk_hovfn, tabs * sizeof(unsigned long long));
        if (rc) return rc;
        g->bk_tabs_cap = tabs;
    }
    return DYC_OK;
}

// Device arrays of one laid-out batch: blocks of round 1, 2 and the eviction insert (nb each),
// each block's table index, and the tables of rounds 1 and 2 (j each).
struct BulkArrays {
    GroupItem *blk;
    uint32_t *btab;
    BulkTable *tabs;
    uint32_t nb, j, chunks;
};

// The whole bulk fill of a laid-out batch, launched on s with no read-back in between, so a
// graph can hold it: round 1 (keys/vals -> entries by each table's chunks -> chunk fill), round 2
// (each table's overflow, for the other bucket of its pair), then the eviction insert of what
// overflows that. Each round's next work comes from the overflow counts on the device (clip).
// Counters and per-table counts accumulate (the caller zeroes them); the last round's overflow
// counts are left in bk_ovfn + j. zeroed: every slot of every table is zero (round 1 then does
// not read the tables).
hipError_t enqueue_group_bulk(dyc_group *g, const BulkArrays &A, const uint32_t *keys, const uint32_t *vals,
                              uint64_t total, bool zeroed, PendSink sink, hipStream_t s)
{
    GroupItem *const blk1 = A.blk, *const blk2 = A.blk + A.nb, *const items = A.blk + 2 * A.nb;
    BulkTable *const tabs1 = A.tabs, *const tabs2 = A.tabs + A.j;
    unsigned long long *const ovf1 = g->bk_ovfn, *const ovf2 = g->bk_ovfn + A.j;
    hipError_t e;
    if ((e = launch_group_bulk_pack(g->ddir, blk1, A.nb, keys, nullptr, g->bk_lcid, g->bk_partial, s)) != hipSuccess ||
        (e = launch_group_bulk_scan(g->ddir, tabs1, A.j, g->bk_partial, g->bk_chunks, s)) != hipSuccess ||
        (e = launch_group_bulk_scatter(blk1, A.nb, keys, vals, nullptr, g->bk_lcid, g->bk_partial, total, g->bk_ent,
                                       s)) != hipSuccess ||
        (e = launch_zero(ovf1, 2 * (uint64_t)A.j, s)) != hipSuccess ||
        (e = launch_group_bulk_fill(g->bk_chunks, A.chunks, g->bk_ent, g->bk_ovf, ovf1, g->tcounts, g->ctr, true,
                                    zeroed, s)) != hipSuccess ||
        (e = launch_group_bulk_clip(blk1, A.btab, A.nb, tabs1, A.j, ovf1, blk2, tabs2, s)) != hipSuccess ||
        (e = launch_group_bulk_pack(g->ddir, blk2, A.nb, nullptr, g->bk_ovf, g->bk_lcid, g->bk_partial, s)) !=
            hipSuccess ||
        (e = launch_group_bulk_scan(g->ddir, tabs2, A.j, g->bk_partial, g->bk_chunks, s)) != hipSuccess ||
        (e = launch_group_bulk_scatter(blk2, A.nb, nullptr, nullptr, g->bk_ovf, g->bk_lcid, g->bk_partial, total,
                                       g->bk_ent, s)) != hipSuccess ||
        (e = launch_group_bulk_fill(g->bk_chunks, A.chunks, g->bk_ent, g->bk_ovf, ovf2, g->tcounts, g->ctr, false,
                                    false, s)) != hipSuccess ||
        (e = launch_group_bulk_clip(blk1, A.btab, A.nb, tabs1, A.j, ovf2, items, nullptr, s)) != hipSuccess)
        return e;
    return launch_group_insert(g->ddir, items, A.nb, nullptr, nullptr, g->policy(), (int)g->launch.insert_lanes,
                               (int)g->launch.insert_blocks, g->ctr, sink, g->tcounts, s, 0, g->bk_ovf);
}

// Sizes the shared workspace for a layout (synchronizes s first); DYC_ERR_OOM leaves no message
// behind for a caller that falls back to the eviction insert.
int reserve_for(dyc_group *g, const BulkLayout &L, hipStream_t s)
{
    HIP_TRY(hipStreamSynchronize(s));  // buffers being replaced are no longer in use
    return reserve_group_bulk(g, L.total, L.blocks.size(), L.chunks, (uint32_t)L.tabs.size());
}

// insert_order 3 for a group: each table's batch is partitioned by its own chunks in one pass
// (the batch is grouped by table already) and every chunk is filled by one workgroup; what
// overflows its first bucket (fill_target) gets a second round for the other bucket of its pair
// (sparse: a few percent of the keys), and what
// overflows that goes through the group's eviction insert. Declines (*n_tabs 0, nothing
// launched) when bulk_layout does, or when the workspace cannot be allocated. Counters and
// per-table counts are zeroed here and read by the caller, who also finds the overflow counts of
// the last round (*n_tabs of them) in bk_hovfn after its wait.
int group_bulk_insert(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *offsets, bool zeroed,
                      hipStream_t s, uint64_t room, uint32_t *n_tabs)
{
    *n_tabs = 0;
    BulkLayout L;
    if (!dyc::host::bulk_layout(g->n, offsets, [g](uint32_t i) -> const Desc & { return g->hdir[i]; }, &L))
        return DYC_OK;
    {
        const dyc::host::ErrorKeep keep;
        const int rc = reserve_for(g, L, s);
        if (rc == DYC_ERR_OOM) {  // no room for the workspace: the eviction insert needs none
            keep.restore();
            return DYC_OK;
        }
        if (rc) return rc;
    }
    const uint32_t NB = (uint32_t)L.blocks.size(), J = (uint32_t)L.tabs.size();
    // the pinned staging is free: the previous call waited for its work before returning
    memcpy(g->bk_hblk, L.blocks.data(), NB * sizeof(GroupItem));
    memcpy(g->bk_hbtab, L.btab.data(), NB * sizeof(uint32_t));
    memcpy(g->bk_htabs, L.tabs.data(), J * sizeof(BulkTable));
    HIP_TRY(hipMemcpyAsync(g->bk_blk, g->bk_hblk, NB * sizeof(GroupItem), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(g->bk_btab, g->bk_hbtab, NB * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(g->bk_tabs, g->bk_htabs, J * sizeof(BulkTable), hipMemcpyHostToDevice, s));
    int rc = zero_and_run_prologue(g, s);
    if (rc) return rc;
    const BulkArrays A{g->bk_blk, g->bk_btab, g->bk_tabs, NB, J, L.chunks};
    HIP_TRY(enqueue_group_bulk(g, A, keys, vals, L.total, zeroed,
                               PendSink{g->pend + g->pend_n, g->pend_tag + g->pend_n, room, 0, 0, nullptr}, s));
    HIP_TRY(hipMemcpyAsync(g->bk_hovfn, g->bk_ovfn + J, J * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    *n_tabs = J;
    return DYC_OK;
}

// The group's shared bulk-fill buffers a plan's graph uses (any of them moving voids the plan).
std::vector<const void *> bulk_workspace(const dyc_group *g)
{
    return {g->bk_ent, g->bk_ovf, g->bk_lcid, g->bk_partial, g->bk_chunks, g->bk_ovfn};
}

}  // namespace

extern "C" {

int dyc_group_create(const dyc_config *cfg_in, uint32_t n_tables, const uint64_t *hints, dyc_group **out)
{
    if (!out || !hints || n_tables == 0) return fail(DYC_ERR_INVALID, "bad group arguments");
    *out = nullptr;
    dyc_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        dyc_config_init(&cfg);
    const uint32_t d = dyc::host::config_subtables(cfg);
    if (d < (uint32_t)kMinSubtables || d > (uint32_t)kMaxSubtables)
        return fail(DYC_ERR_INVALID, "subtables must be in [%d, %d]", kMinSubtables, kMaxSubtables);
    cfg.subtables = d;
    dyc_group *g = new dyc_group();
    g->cfg = cfg;
    g->launch = dyc_group_launch{cfg.find_lanes, cfg.insert_lanes, 0, 0, kDefaultItemKeys, cfg.kick_bound, 0, 0};
    int rc = check_launch(g->launch);
    if (rc) {
        delete g;
        return rc;
    }
    g->n = n_tables;
    g->hdir.resize(n_tables);
    g->count.assign(n_tables, 0);
    // arena layout: table i, subtable t at a 256-byte aligned offset
    std::vector<uint32_t> nbs(n_tables);
    size_t bytes = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        nbs[i] = dyc::host::initial_buckets(hints[i], d);
        if (nbs[i] >= dyc::host::kMaxBuckets) {
            delete g;
            return fail(DYC_ERR_INVALID, "table %u: capacity hint too large", i);
        }
        bytes += d * (((size_t)nbs[i] * kBucketBytes + 255) & ~(size_t)255);
    }
    auto cleanup = [&](int code) {
        char keep[512];
        strncpy(keep, dyc_last_error(), sizeof keep - 1);
        keep[sizeof keep - 1] = 0;
        dyc_group_destroy(g);
        fail(code, "%s", keep);
        return code;
    };
    HIP_TRY(hipGetDevice(&g->device));
    if (hipMalloc(&g->arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(DYC_ERR_OOM, "group arena of %zu bytes failed", bytes);
        return cleanup(DYC_ERR_OOM);
    }
    g->arena_bytes = bytes;
    size_t off = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        desc_init(g->hdir[i], d);
        for (uint32_t t = 0; t < d; ++t) {
            desc_set(g->hdir[i], (int)t, reinterpret_cast<uint64_t *>(g->arena + off), nbs[i]);
            off += ((size_t)nbs[i] * kBucketBytes + 255) & ~(size_t)255;
        }
    }
    if (hipMalloc(&g->ddir, n_tables * sizeof(Desc)) != hipSuccess || hipMalloc(&g->ctr, kCounterBytes) != hipSuccess ||
        hipMalloc(&g->tcounts, n_tables * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->hctr), kCounterBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->htcounts), n_tables * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&g->items_free, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        fail(DYC_ERR_OOM, "group bookkeeping allocation failed");
        return cleanup(DYC_ERR_OOM);
    }
    if (hipMemcpy(g->ddir, g->hdir.data(), n_tables * sizeof(Desc), hipMemcpyHostToDevice) != hipSuccess ||
        launch_zero(reinterpret_cast<unsigned long long *>(g->arena), bytes / 8, nullptr) != hipSuccess ||
        hipStreamSynchronize(nullptr) != hipSuccess) {
        fail(DYC_ERR_HIP, "group initialisation failed");
        return cleanup(DYC_ERR_HIP);
    }
    rc = reserve_pending(g, 65536, nullptr);
    if (rc) return cleanup(rc);
    g->zeroed = true;
    *out = g;
    return DYC_OK;
}

int dyc_group_destroy(dyc_group *g)
{
    if (!g) return DYC_OK;
    DeviceGuard dg(g->device);
    (void)hipDeviceSynchronize();
    if (g->arena) (void)hipFree(g->arena);
    if (g->ddir) (void)hipFree(g->ddir);
    if (g->ctr) (void)hipFree(g->ctr);
    if (g->tcounts) (void)hipFree(g->tcounts);
    if (g->hctr) (void)hipHostFree(g->hctr);
    if (g->htcounts) (void)hipHostFree(g->htcounts);
    if (g->pend) (void)hipFree(g->pend);
    if (g->pend_tag) (void)hipFree(g->pend_tag);
    if (g->ditems) (void)hipFree(g->ditems);
    if (g->hitems) (void)hipHostFree(g->hitems);
    if (g->items_free) (void)hipEventDestroy(g->items_free);
    for (void *p : {(void *)g->bk_ent, (void *)g->bk_ovf, (void *)g->bk_lcid, (void *)g->bk_partial,
                    (void *)g->bk_chunks, (void *)g->bk_tabs,
                    (void *)g->bk_ovfn, (void *)g->bk_blk, (void *)g->bk_btab})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)g->bk_htabs, (void *)g->bk_hovfn, (void *)g->bk_hblk, (void *)g->bk_hbtab})
        if (p) (void)hipHostFree(p);
    delete g;
    return DYC_OK;
}

int dyc_group_get_launch(const dyc_group *g, dyc_group_launch *out)
{
    if (!g || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = g->launch;
    return DYC_OK;
}

int dyc_group_set_launch(dyc_group *g, const dyc_group_launch *l)
{
    if (!g || !l) return fail(DYC_ERR_INVALID, "null argument");
    int rc = check_launch(*l);
    if (rc) return rc;
    g->launch = *l;
    return DYC_OK;
}

int dyc_group_insert(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *offsets,
                     dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = check_offsets(g, offsets);
    if (rc) return rc;
    const uint64_t total = offsets[g->n];
    if (total && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    const bool zeroed = g->zeroed;
    g->zeroed = false;
    g->st.last_filled = 0;
    rc = reserve_pending(g, std::max<uint64_t>(65536, total / 16), s);
    if (rc) return rc;
    const uint64_t room = g->pend_cap - g->pend_n;
    if (g->cfg.insert_order == 3) {
        uint32_t n_tabs = 0;
        rc = group_bulk_insert(g, keys, vals, offsets, zeroed, s, room, &n_tabs);
        if (rc) return rc;
        if (n_tabs) {
            rc = read_back(g, s);  // waits: the overflow counts are in bk_hovfn too
            if (rc) return rc;
            uint64_t left = 0;
            for (uint32_t j = 0; j < n_tabs; ++j) left += g->bk_hovfn[j];
            g->st.last_filled = total - g->hctr->rejected - left;
            return finish_insert(g, room);
        }
    }
    uint32_t n_items = 0;
    rc = upload_items(g, offsets, s, &n_items);
    if (rc) return rc;
    rc = zero_and_run_prologue(g, s);
    if (rc) return rc;
    HIP_TRY(launch_group_insert(g->ddir, g->ditems, n_items, keys, vals, g->policy(), (int)g->launch.insert_lanes,
                                (int)g->launch.insert_blocks, g->ctr,
                                PendSink{g->pend + g->pend_n, g->pend_tag + g->pend_n, room, 0, 0, nullptr}, g->tcounts,
                                s, (int)g->launch.insert_waves_per_eu));
    rc = read_back(g, s);
    if (rc) return rc;
    return finish_insert(g, room);
}

int dyc_group_find(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint32_t *vals, uint8_t *hit,
                   dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    uint32_t n_items = 0;
    // the layout by XCD pays where the batch spans more than the L2s hold; below that its host
    // work (a sort of the tables) would be most of a small call
    rc = offsets[g->n] >= kXcdLayoutKeys ? upload_items_by_xcd(g, offsets, s, &n_items)
                                          : upload_items(g, offsets, s, &n_items);
    if (rc) return rc;
    Pool sp;
    HIP_TRY(launch_group_find(g->ddir, g->ditems, n_items, keys, vals, hit, (int)g->launch.find_lanes,
                              (int)g->launch.find_blocks, s, g->stash(sp), (int)g->launch.find_waves_per_eu));
    return DYC_OK;
}

int dyc_group_erase(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint8_t *deleted, dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    uint32_t n_items = 0;
    rc = upload_items(g, offsets, s, &n_items);
    if (rc) return rc;
    rc = zero_and_run_prologue(g, s);
    if (rc) return rc;
    Pool sp;
    HIP_TRY(launch_group_erase(g->ddir, g->ditems, n_items, keys, deleted, g->tcounts, (int)g->launch.find_lanes,
                               (int)g->launch.find_blocks, s, g->stash(sp), (int)g->launch.find_waves_per_eu));
    rc = read_back(g, s);
    if (rc) return rc;
    for (uint32_t i = 0; i < g->n; ++i) g->count[i] -= std::min<uint64_t>(g->count[i], g->htcounts[i]);
    return compact_stash(g, s);
}

int dyc_group_insert_per_table(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *offsets,
                               dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    g->zeroed = false;
    g->st.last_filled = 0;
    rc = reserve_pending(g, std::max<uint64_t>(65536, offsets[g->n] / 16), s);
    if (rc) return rc;
    rc = zero_and_run_prologue(g, s);
    if (rc) return rc;
    const uint64_t room = g->pend_cap - g->pend_n;
    const LaunchShape sh{256, (int)g->launch.insert_blocks};
    // one launch per table; all share the counters, so pended entries append one after another
    for (uint32_t i = 0; i < g->n; ++i) {
        const uint64_t m = offsets[i + 1] - offsets[i];
        if (!m) continue;
        HIP_TRY(launch_insert(g->hdir[i], keys + offsets[i], vals + offsets[i], m, (int)g->launch.insert_lanes,
                              g->policy(), sh, g->ctr,
                              PendSink{g->pend + g->pend_n, g->pend_tag + g->pend_n, room, i, 0, g->tcounts + i}, s));
    }
    rc = read_back(g, s);
    if (rc) return rc;
    // tcounts held each table's reserved keys; a table took in its batch less those
    for (uint32_t i = 0; i < g->n; ++i) g->htcounts[i] = (offsets[i + 1] - offsets[i]) - g->htcounts[i];
    return finish_insert(g, room);
}

int dyc_group_find_per_table(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint32_t *vals, uint8_t *hit,
                             dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = check_offsets(g, offsets);
    if (rc) return rc;
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    const LaunchShape sh{256, (int)g->launch.find_blocks};
    for (uint32_t i = 0; i < g->n; ++i) {
        const uint64_t m = offsets[i + 1] - offsets[i];
        if (!m) continue;
        Pool sp;
        HIP_TRY(launch_find(g->hdir[i], keys + offsets[i], vals ? vals + offsets[i] : nullptr,
                            hit ? hit + offsets[i] : nullptr, m, (int)g->launch.find_lanes, sh, s, g->stash(sp), i));
    }
    return DYC_OK;
}

int dyc_group_table_stats(const dyc_group *g, uint32_t table, dyc_stats *out)
{
    if (!g || !out || table >= g->n) return fail(DYC_ERR_INVALID, "bad argument");
    *out = g->st;
    out->count = g->count[table];
    uint64_t cap = 0;
    for (uint32_t t = 0; t < g->hdir[table].d; ++t) {
        out->buckets[t] = g->hdir[table].nb[t];
        cap += (uint64_t)g->hdir[table].nb[t] * kSlots;
    }
    out->capacity = cap;
    out->subtables = g->hdir[table].d;
    return DYC_OK;
}

int dyc_group_dump(dyc_group *g, uint32_t table, uint32_t sub, uint64_t *host_slots, uint64_t n_slots,
                   dyc_stream stream)
{
    if (!g || table >= g->n || !host_slots || sub >= g->hdir[table].d) return fail(DYC_ERR_INVALID, "bad argument");
    const uint64_t slots = (uint64_t)g->hdir[table].nb[sub] * kSlots;
    if (n_slots < slots) return fail(DYC_ERR_INVALID, "buffer too small");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(host_slots, g->hdir[table].sub[sub], slots * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DYC_OK;
}

int dyc_group_pending(dyc_group *g, uint64_t *host_entries, uint32_t *host_tables, uint64_t cap, uint64_t *n_out,
                      dyc_stream stream)
{
    if (!g || !n_out) return fail(DYC_ERR_INVALID, "null argument");
    *n_out = g->pend_n;
    const uint64_t m = std::min(cap, g->pend_n);
    if (m) {
        DeviceGuard dg(g->device);
        hipStream_t s = (hipStream_t)stream;
        if (host_entries) HIP_TRY(hipMemcpyAsync(host_entries, g->pend, m * 8, hipMemcpyDeviceToHost, s));
        if (host_tables) HIP_TRY(hipMemcpyAsync(host_tables, g->pend_tag, m * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return DYC_OK;
}

// ------------------------------------------------------------------- captured steps
struct dyc_group_plan {
    dyc_group *g = nullptr;
    GroupItem *d_items = nullptr;  // insert items, then find items
    uint32_t n_ins = 0, n_find = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t room = 0;  // pending entries the insert may append
    uint64_t *pend = nullptr;  // the pending list the graph writes (a larger eager insert may move it)
    // insert_order 3: the batch's own bulk-fill arrays, the group's shared workspace the graph
    // uses (an eager insert that grows it moves these), and the last round's overflow counts
    // (pinned, written by the graph) for last_filled
    bool bulk = false;
    uint64_t total = 0;
    GroupItem *bk_blk = nullptr;
    uint32_t *bk_btab = nullptr;
    BulkTable *bk_tabs = nullptr;
    uint32_t bk_j = 0;
    unsigned long long *hovfn = nullptr;
    std::vector<const void *> shared;
};


int dyc_group_plan_destroy(dyc_group_plan *p)
{
    if (!p) return DYC_OK;
    DeviceGuard dg(p->g->device);
    (void)hipDeviceSynchronize();
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_items) (void)hipFree(p->d_items);
    for (void *q : {(void *)p->bk_blk, (void *)p->bk_btab, (void *)p->bk_tabs})
        if (q) (void)hipFree(q);
    if (p->hovfn) (void)hipHostFree(p->hovfn);
    delete p;
    return DYC_OK;
}

int dyc_group_plan_create(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *ins_offsets,
                          const uint32_t *lookups, const uint64_t *find_offsets, uint32_t *vals_out, uint8_t *hit,
                          dyc_group_plan **out)
{
    if (!g || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = check_offsets(g, ins_offsets);
    if (!rc) rc = check_offsets(g, find_offsets);
    if (rc) return rc;
    if (ins_offsets[g->n] && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    if (find_offsets[g->n] && !lookups) return fail(DYC_ERR_INVALID, "null lookups");
    DeviceGuard dg(g->device);
    rc = reserve_pending(g, std::max<uint64_t>(65536, ins_offsets[g->n] / 16), nullptr);
    if (rc) return rc;
    // the work items of both launches, cut and laid out as the eager calls do, uploaded once
    const uint64_t per = g->launch.item_keys ? g->launch.item_keys : kDefaultItemKeys;
    std::vector<GroupItem> items;
    for (uint32_t i = 0; i < g->n; ++i)
        for (uint64_t b = ins_offsets[i]; b < ins_offsets[i + 1]; b += per)
            items.push_back(GroupItem{b, (uint32_t)std::min<uint64_t>(per, ins_offsets[i + 1] - b), i});
    for (uint32_t i = 0; i < g->n; ++i)
        if (find_offsets[i + 1] - find_offsets[i] > 0xffffffffull)
            return fail(DYC_ERR_INVALID, "table %u's batch is too large for 32-bit lengths", i);
    const std::vector<GroupItem> finds = items_by_xcd(find_offsets, g->n, per);
    items.insert(items.end(), finds.begin(), finds.end());
    if (items.size() > 0xffffffffull) return fail(DYC_ERR_INVALID, "batch too large for 32-bit item indexes");
    dyc_group_plan *p = new dyc_group_plan();
    p->g = g;
    for (uint32_t i = 0; i < g->n; ++i) p->n_ins += (uint32_t)((ins_offsets[i + 1] - ins_offsets[i] + per - 1) / per);
    p->n_find = (uint32_t)(items.size() - p->n_ins);
    p->room = g->pend_cap;
    p->pend = g->pend;
    auto undo = [&](int code) {
        char keep[512];
        strncpy(keep, dyc_last_error(), sizeof keep - 1);
        keep[sizeof keep - 1] = 0;
        dyc_group_plan_destroy(p);
        fail(code, "%s", keep);
        return code;
    };
    if (!items.empty()) {
        if (hipMalloc(&p->d_items, items.size() * sizeof(GroupItem)) != hipSuccess) {
            (void)hipGetLastError();
            fail(DYC_ERR_OOM, "plan work items (%zu) failed", items.size());
            return undo(DYC_ERR_OOM);
        }
        if (hipMemcpy(p->d_items, items.data(), items.size() * sizeof(GroupItem), hipMemcpyHostToDevice) != hipSuccess) {
            fail(DYC_ERR_HIP, "plan work item upload failed");
            return undo(DYC_ERR_HIP);
        }
    }
    // insert_order 3: the bulk fill of this batch, with its own block and table arrays
    BulkLayout L;
    if (g->cfg.insert_order == 3 &&
        dyc::host::bulk_layout(g->n, ins_offsets, [g](uint32_t i) -> const Desc & { return g->hdir[i]; }, &L)) {
        const dyc::host::ErrorKeep keep;
        const int r = reserve_for(g, L, nullptr);
        if (r == DYC_ERR_OOM) {
            keep.restore();  // the plan holds the eviction insert instead
        } else if (r) {
            return undo(r);
        } else {
            const uint32_t NB = (uint32_t)L.blocks.size(), J = (uint32_t)L.tabs.size();
            p->bulk = true;
            p->total = L.total;
            p->bk_j = J;
            if (hipMalloc(&p->bk_blk, 3ull * NB * sizeof(GroupItem)) != hipSuccess ||
                hipMalloc(&p->bk_btab, NB * sizeof(uint32_t)) != hipSuccess ||
                hipMalloc(&p->bk_tabs, 2ull * J * sizeof(BulkTable)) != hipSuccess ||
                hipHostMalloc((void **)&p->hovfn, J * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                fail(DYC_ERR_OOM, "plan bulk-fill arrays failed");
                return undo(DYC_ERR_OOM);
            }
            if (hipMemcpy(p->bk_blk, L.blocks.data(), NB * sizeof(GroupItem), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(p->bk_btab, L.btab.data(), NB * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(p->bk_tabs, L.tabs.data(), J * sizeof(BulkTable), hipMemcpyHostToDevice) != hipSuccess) {
                fail(DYC_ERR_HIP, "plan bulk-fill upload failed");
                return undo(DYC_ERR_HIP);
            }
            p->shared = bulk_workspace(g);
        }
    }
    // capture: clear, counters to zero, insert, find (its stash is the pending list, whose size
    // the launch reads on the device), counters out to pinned memory. Zeroing is done by a fill
    // kernel, not a memset node.
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        fail(DYC_ERR_HIP, "capture stream");
        return undo(DYC_ERR_HIP);
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const Pool stash{g->pend, g->pend_tag, g->pend_cap, &g->ctr->pending, 0};
        hipError_t k = launch_zero(reinterpret_cast<unsigned long long *>(g->arena), g->arena_bytes / 8, cs);
        if (k == hipSuccess) k = launch_zero(reinterpret_cast<unsigned long long *>(g->ctr), kCounterWords, cs);
        if (k == hipSuccess) k = launch_zero(g->tcounts, g->n, cs);
        const PendSink sink{g->pend, g->pend_tag, g->pend_cap, 0, 0, nullptr};
        if (k == hipSuccess && p->bulk) {
            // the graph clears every table first, so round 1 need not read them
            const BulkArrays A{p->bk_blk, p->bk_btab, p->bk_tabs, (uint32_t)L.blocks.size(), p->bk_j, L.chunks};
            k = enqueue_group_bulk(g, A, keys, vals, p->total, true, sink, cs);
            if (k == hipSuccess) k = launch_copy(g->bk_ovfn + p->bk_j, p->hovfn, p->bk_j, nullptr, nullptr, 0, cs);
        } else if (k == hipSuccess) {
            k = launch_group_insert(g->ddir, p->d_items, p->n_ins, keys, vals, g->policy(), (int)g->launch.insert_lanes,
                                    (int)g->launch.insert_blocks, g->ctr, sink, g->tcounts, cs,
                                    (int)g->launch.insert_waves_per_eu);
        }
        if (k == hipSuccess)
            k = launch_group_find(g->ddir, p->d_items + p->n_ins, p->n_find, lookups, vals_out, hit,
                                  (int)g->launch.find_lanes, (int)g->launch.find_blocks, cs, &stash,
                                  (int)g->launch.find_waves_per_eu);
        if (k == hipSuccess)
            k = launch_copy(g->ctr, g->hctr, kCounterWords, g->tcounts, g->htcounts, g->n, cs);
        e = hipStreamEndCapture(cs, &p->graph);
        if (k != hipSuccess) e = k;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        fail(DYC_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        return undo(DYC_ERR_HIP);
    }
    *out = p;
    return DYC_OK;
}

int dyc_group_plan_run(dyc_group_plan *p, dyc_stream stream)
{
    if (!p) return fail(DYC_ERR_INVALID, "null plan");
    dyc_group *g = p->g;
    if (g->pend != p->pend || g->pend_cap != p->room)
        return fail(DYC_ERR_INVALID, "the group's pending list moved since the plan was made (a larger insert grew "
                                     "it): make the plan again");
    if (p->bulk && bulk_workspace(g) != p->shared)
        return fail(DYC_ERR_INVALID, "the group's bulk-fill workspace moved since the plan was made (a larger bulk "
                                     "insert grew it): make the plan again");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    g->zeroed = false;
    g->st.last_filled = 0;
    HIP_TRY(hipGraphLaunch(p->exec, s));
    HIP_TRY(hipStreamSynchronize(s));
    fold_tallies(g->hctr);
    // the step cleared every table: counts and the pending list start again from zero
    std::fill(g->count.begin(), g->count.end(), 0);
    g->pend_n = 0;
    if (p->bulk) {
        uint64_t left = 0;
        for (uint32_t j = 0; j < p->bk_j; ++j) left += p->hovfn[j];
        g->st.last_filled = p->total - g->hctr->rejected - left;
    }
    return finish_insert(g, p->room);
}

int dyc_group_clear(dyc_group *g, dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    DeviceGuard dg(g->device);
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(g->arena), g->arena_bytes / 8, (hipStream_t)stream));
    std::fill(g->count.begin(), g->count.end(), 0);
    g->pend_n = 0;
    g->st.pending = 0;
    g->zeroed = true;
    return DYC_OK;
}

}  // extern "C"
