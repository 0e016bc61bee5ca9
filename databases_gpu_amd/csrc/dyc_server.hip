// dyc_server.hip -- the serving loop: the dyc_server_* part of include/dyc.h.
//
// Iteration-level continuous batching on one stream the server owns:
//   * requests (one reference batch call each: insert, find or erase of one table) queue on the
//     host in arrival order;
//   * a "tick" admits every queued request that fits (slots, arena space), enqueues the upsizes
//     those inserts need, stages the tick's work items (and the table directory if a resize
//     changed it) with one copy kernel from pinned memory, and launches, round by round, one
//     fused insert, one fused find and one fused erase over the round's requests; a copy kernel
//     then writes the device counters to pinned memory and an event is recorded. A table's
//     requests are spread over rounds so that within a round they follow insert -> find ->
//     erase (at most one erase), which keeps every table's requests in arrival order;
//   * up to pipeline_depth ticks are in flight: the next tick is built and enqueued while the
//     previous one runs, so the device does not wait on the host between ticks;
//   * a tick retires when its event has fired: its requests complete, table counts are
//     refreshed from the counter snapshot, tables whose erases retired are checked for a
//     downsize, and entries that exhausted the eviction bound are re-placed.
// All table storage is carved from one arena (dyc_arena.hpp), so admission by free HBM is exact
// and no hipMalloc/hipFree runs while ticks are in flight.
//
// Device-resident mode (resident = 1, or replay mode 2): instead of ticks, one launch of the loop
// kernel (serve_loop_kernel) stays on the stream and takes work items from a ring in host memory.
// Submission admits a request at once: it is cut into items, each stamped with the count of the
// table's earlier requests it must wait for (a table's requests run as stages: a run of inserts,
// a run of finds, or one erase, each after the previous stage), and the ring's tail is advanced.
// The device appends finished requests to a completion queue in host memory, which poll reads.
// Resizes run on a second stream while the loop keeps serving other tables: a table is resized
// only when it has no request in flight, and its next requests are published once the resize
// and the directory update have finished. Re-placing the overflow pool needs every table quiet:
// the host stops publishing, waits for the loop to finish what it has, stops the launch, re-places
// the pool and starts the loop again.
#include "dyc_arena.hpp"
#include "dyc_host.hpp"
#include "dyc_kernels.hpp"

#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <unordered_map>
#include <vector>

using namespace dyc;
using dyc::host::choose_down;
using dyc::host::choose_up;
using dyc::host::DeviceGuard;
using dyc::host::fail;
using dyc::host::initial_buckets;
using dyc::host::kMaxBuckets;

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kTargetItems = 8192;  // work items per tick: one per resident wave (256 CUs x 32)
constexpr uint32_t kMinItemKeys = 64;
constexpr uint32_t kMaxItemKeys = 4096;
constexpr uint32_t kMaxRounds = 64;    // phase-ordered rounds per tick
constexpr uint32_t kChainItems = 1024;  // items per stage of a table with several stages
constexpr uint64_t kDirectItems = 256;  // ticks this small read their items from pinned memory
static_assert(sizeof(ServeItem) % 8 == 0 && sizeof(TickItem) % 8 == 0 && sizeof(Desc) % 8 == 0,
              "staged with 8-byte word copies");

struct SrvTable {
    Desc D{};
    uint64_t off[kMaxSubtables] = {};  // arena offsets of the subtables
    uint64_t count = 0;             // live entries as of the last retired tick
    uint64_t inflight_ins = 0;      // insert keys admitted in ticks not yet retired
    uint32_t inflight = 0;          // ticks in flight that touch the table
    bool check_down = false;        // an erase retired since the last downsize check
    uint64_t sync_del = 0;          // entries deleted by synchronous erases (not in del_total)
    uint64_t sync_ins = 0;          // entries taken in by synchronous inserts (not in ins_total)
    uint64_t resizes_up = 0, resizes_down = 0;

    uint64_t capacity() const
    {
        uint64_t c = 0;
        for (uint32_t i = 0; i < D.d; ++i) c += (uint64_t)D.nb[i] * kSlots;
        return c;
    }
};

struct Req {
    dyc_request r;
    int status = DYC_OK;
    int state = 0;  // 0 queued, 1 admitted, 2 done
};

struct TickTable {
    uint32_t table;
    uint64_t ins;  // insert keys admitted for the table in this tick
    bool erase;
};

struct Tick {
    ServeItem *h_items = nullptr;  // pinned staging: insert, find, erase items back to back
    ServeItem *d_items = nullptr;
    uint64_t items_cap = 0;
    // one-launch ticks: word 0 = the launch's item cursor (staged as 0), then TickItems
    unsigned long long *h_block = nullptr;  // pinned
    unsigned long long *d_block = nullptr;
    Desc *h_dir = nullptr;                 // pinned copy of the directory this tick uploads
    unsigned long long *h_snap = nullptr;  // pinned: Counters, ins_total[n], del_total[n]
    Counters *d_rctr = nullptr;            // downsize compaction counters
    hipEvent_t done = nullptr;
    bool busy = false;
    bool has_snap = false;  // the tick copies the device counters out (it may change counts)
    bool solo_erase = false;  // one erase through the plain kernel: its deletions are in Counters::deleted
    bool solo_insert = false;  // one insert through the plain kernel: its entries are in Counters::placed
    uint32_t solo_table = 0;
    uint64_t seq = 0;
    std::vector<uint64_t> reqs;
    std::vector<TickTable> tables;
};

// A request slot of the resident loop (index into the device's per-request words and the
// host-memory records).
struct LoopSlot {
    uint64_t ticket = 0;
    uint64_t pseq = 0;  // publish order
    uint64_t n = 0;
    uint32_t table = 0, op = 0;
};

constexpr uint64_t kLoopRingItems = 1ull << 18;  // 16 MiB of items in host memory
constexpr uint32_t kLoopSlots = 1u << 16;        // requests in flight at once

struct Loop {
    bool on = false;       // requests go through the loop
    bool running = false;  // a launch is on the stream and has not been seen to end
    hipStream_t st2 = nullptr;  // resizes and counter reads while the loop is on
    hipEvent_t ended = nullptr;  // recorded after each launch
    LoopCtl *ctl = nullptr;      // host memory (coherent)
    LoopItem *ring = nullptr;
    LoopRec *rec = nullptr;
    unsigned long long *cq = nullptr;
    unsigned long long *d_words = nullptr, *d_tbl_done = nullptr, *d_req_word = nullptr;
    uint32_t *d_req_out = nullptr, *d_req_stall = nullptr;
    int blocks = 0;
    uint64_t tail = 0;     // items published
    uint64_t cq_head = 0;  // completion entries read
    uint64_t pseq = 0;     // requests published
    std::vector<LoopSlot> slot;
    std::vector<uint32_t> free_slots;
    // requests in publish order from the oldest unfinished one: (first item, finished)
    uint64_t order_base = 0;
    std::deque<std::pair<uint64_t, bool>> order;
    std::vector<uint64_t> pub;    // per table: requests published through the loop
    std::vector<uint64_t> need;   // per table: the current stage's wait count
    std::vector<int8_t> last_op;  // per table: the current stage's op (-1: a new stage starts)
    bool want_resolve = false;
    uint64_t in_flight() const { return slot.size() - free_slots.size(); }
    uint64_t oldest_item() const { return order.empty() ? tail : order.front().first; }
};

}  // namespace

struct dyc_server {
    dyc_config cfg;
    dyc_server_config scfg;
    int device = 0;
    hipStream_t st = nullptr;
    uint32_t n = 0;
    uint8_t *base = nullptr;  // arena memory
    Arena arena;
    std::vector<SrvTable> t;
    Desc *d_dir = nullptr;
    bool dir_dirty = true;
    unsigned long long *d_stat = nullptr;  // Counters words, then ins_total[n], then del_total[n]
    uint64_t *d_pool = nullptr;
    uint32_t *d_pool_tag = nullptr;
    uint64_t *d_resolve = nullptr;  // staging for re-placing pool entries (pool_cap entries)
    uint64_t pool_settled = 0;      // pool entries left by the last re-placement
    uint64_t deleted_seen = 0;      // Counters::deleted as of the last synchronous erase
    uint64_t placed_seen = 0;       // Counters::placed as of the last synchronous insert
    uint64_t misplaced_seen = 0;
    uint64_t stalls_seen = 0;
    Tick ticks[2];
    uint64_t tick_seq = 0;
    std::vector<Req> reqs;
    std::deque<uint64_t> queue;                 // tickets waiting for admission, arrival order
    std::deque<std::pair<uint64_t, int>> done;  // finished, not yet handed out
    std::vector<uint32_t> down_list;            // tables with check_down set
    int error = DYC_OK;                         // sticky: lost entries or a corrupt resize
    dyc_server_stats ss{};
    // admission scratch, indexed by table
    std::vector<int8_t> phase;     // last op admitted for the table in the tick being built, -1 none
    std::vector<uint8_t> blocked;  // 1 order, 2 memory
    std::vector<uint8_t> in_tick;  // the tick being built works on the table (request or downsize)
    std::vector<uint8_t> seen;     // listed in touched
    std::vector<uint64_t> tick_ins;
    std::vector<uint64_t> mem_head;  // the request a memory block held back (valid when blocked == 2)
    std::vector<uint16_t> round;     // round of the table's last admitted request
    std::vector<uint8_t> had_erase;  // the tick erases from the table
    std::vector<int32_t> stage;      // one-launch ticks: the table's current stage key (round * 3 + op)
    std::vector<uint32_t> last_sid;  // one-launch ticks: the table's current stage id, ~0u none
    std::vector<uint32_t> st_left, st_next, st_ready;  // per stage of the tick being built
    std::vector<uint32_t> nstages;   // stages (round, op) of the table in the tick being built
    std::unordered_map<uint64_t, uint64_t> stage_keys;  // keys per (table, round, op)
    std::vector<uint32_t> touched;
    std::vector<std::pair<uint64_t, uint32_t>> admitted;  // (ticket, round)
    std::vector<ServeItem> items[kMaxRounds][3];
    Loop loop;

    Counters *pc() const { return reinterpret_cast<Counters *>(d_stat); }
    // stream of resizes, pool re-placement and counter reads: the serving stream, or the second
    // stream while the loop holds the serving stream
    hipStream_t ms() const { return loop.on ? loop.st2 : st; }
    unsigned long long *ins_total() const { return d_stat + kCounterWords; }
    unsigned long long *del_total() const { return d_stat + kCounterWords + n; }
    size_t stat_bytes() const { return (kCounterWords + 2 * (size_t)n) * sizeof(unsigned long long); }
    uint64_t *sub_ptr(uint64_t off) const { return reinterpret_cast<uint64_t *>(base + off); }
    InsertPolicy policy() const { return InsertPolicy{cfg.kick_bound, cfg.rescue_kicks, cfg.max_retries}; }
    LaunchShape shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.max_blocks}; }
    LaunchShape insert_shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.insert_blocks}; }
    Pool pool() const { return Pool{d_pool, d_pool_tag, scfg.pool_cap, &pc()->pending, 0}; }
    // count_placed: the launch adds the entries it takes in to Counters::placed (a synchronous
    // insert, whose table the host credits from that tally)
    PendSink sink(uint32_t tag, bool count_placed = false) const
    {
        return PendSink{d_pool, d_pool_tag, scfg.pool_cap, tag, count_placed ? 1u : 0u, nullptr};
    }
    uint32_t depth() const { return scfg.pipeline_depth; }
};

namespace {

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

int sticky(dyc_server *s, int rc)
{
    if (rc != DYC_OK && s->error == DYC_OK) s->error = rc;
    return rc;
}

void refresh_counts(dyc_server *s, const unsigned long long *snap)
{
    const unsigned long long *ins = snap + kCounterWords, *del = snap + kCounterWords + s->n;
    for (uint32_t i = 0; i < s->n; ++i) s->t[i].count = ins[i] + s->t[i].sync_ins - del[i] - s->t[i].sync_del;
}

// Double one subtable of table i (choose_up); *ok = false (nothing changed) if the arena has
// no room for the new subtable.
int upsize_now(dyc_server *s, uint32_t i, bool *ok)
{
    SrvTable &T = s->t[i];
    *ok = false;
    const uint32_t sub = choose_up(T.D.nb, T.D.d);
    const uint32_t old_nb = T.D.nb[sub];
    if ((uint64_t)old_nb * 2 >= kMaxBuckets) return DYC_OK;
    const uint64_t off = s->arena.alloc((uint64_t)old_nb * 2 * kBucketBytes);
    if (off == Arena::kNone) return DYC_OK;
    HIP_TRY(launch_upsize(T.D.sub[sub], old_nb, s->sub_ptr(off), sub, s->pc(), s->ms()));
    s->arena.release(T.off[sub]);
    T.off[sub] = off;
    desc_set(T.D, (int)sub, s->sub_ptr(off), old_nb * 2);
    s->dir_dirty = true;
    T.resizes_up++;
    s->ss.resizes_up++;
    *ok = true;
    return DYC_OK;
}

// Halve one subtable of table i (choose_down) and re-insert the compaction's residuals into the
// other subtable of their pair (the slot-addressed residual list is read by its live tiles, so
// the host needs no count). *done = false if refused (odd bucket count) or the arena has no room
// for the new subtable and the list's tile flags.
int downsize_now(dyc_server *s, uint32_t i, Counters *rctr, bool *done)
{
    SrvTable &T = s->t[i];
    *done = false;
    const uint32_t sub = choose_down(T.D.nb, T.D.d);
    const uint32_t old_nb = T.D.nb[sub];
    if ((old_nb & 1u) || old_nb < 2) return DYC_OK;
    const uint32_t new_nb = old_nb / 2;
    const uint64_t off = s->arena.alloc((uint64_t)new_nb * kBucketBytes);
    if (off == Arena::kNone) return DYC_OK;
    const uint64_t loff = s->arena.alloc(res_tiles(new_nb));  // the residual list's tile flags
    if (loff == Arena::kNone) {
        s->arena.release(off);
        return DYC_OK;
    }
    uint8_t *live = reinterpret_cast<uint8_t *>(s->sub_ptr(loff));
    // The residuals go to the old subtable's upper half, slot-addressed (a group writes pair g's
    // residuals there only after reading both of its buckets); the old region is released once
    // the re-insert that reads them is on the stream (the arena hands ranges out stream-ordered).
    uint64_t *old = T.D.sub[sub];
    uint64_t *res = old + (uint64_t)new_nb * kSlots;
    const uint64_t res_n = (uint64_t)new_nb * kSlots;
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(rctr), kCounterWords, s->ms()));
    HIP_TRY(launch_downsize(old, new_nb, s->sub_ptr(off), rctr, res, res_n, live, s->ms()));
    const uint64_t old_off = T.off[sub];
    T.off[sub] = off;
    desc_set(T.D, (int)sub, s->sub_ptr(off), new_nb);
    s->dir_dirty = true;
    // residuals continue from the shrunk subtable into the other one of their pair
    // (cuckoo.cu:719), after the compaction; failures go to the pool
    HIP_TRY(launch_insert_entries(T.D, res, res_n, sub, (int)s->cfg.insert_lanes, s->policy(), s->insert_shape(),
                                  s->pc(), s->sink(i), s->ms(), live));
    s->arena.release(old_off);
    s->arena.release(loff);
    T.resizes_down++;
    s->ss.resizes_down++;
    *done = true;
    return DYC_OK;
}

bool wants_down(const dyc_server *s, const SrvTable &T)
{
    const uint64_t cap = T.capacity();
    return (double)T.count < s->cfg.lower_bound * (double)cap && (double)cap > (double)s->cfg.min_capacity;
}

// dst is host-pinned (a tick's snapshot buffer)
int read_stats_sync(dyc_server *s, unsigned long long *dst)
{
    HIP_TRY(launch_copy(s->d_stat, dst, s->stat_bytes() / 8, nullptr, nullptr, 0, s->ms()));
    HIP_TRY(hipStreamSynchronize(s->ms()));
    fold_tallies(dst);
    return DYC_OK;
}

// Re-place the pool: up to max_fail_rounds rounds of (upsize each table that has pool entries,
// re-insert them). Runs with the stream drained; entries that still fail stay in the pool,
// where finds and erases keep seeing them.
int resolve_pool(dyc_server *s, unsigned long long *scratch)
{
    HIP_TRY(hipStreamSynchronize(s->ms()));
    std::vector<uint64_t> ent;
    std::vector<uint32_t> tag;
    for (uint32_t round = 0; round < s->cfg.max_fail_rounds; ++round) {
        int rc = read_stats_sync(s, scratch);
        if (rc) return rc;
        const Counters &c = *reinterpret_cast<const Counters *>(scratch);
        if (c.lost) return fail(DYC_ERR_LOST, "overflow pool full: %llu entries lost", (unsigned long long)c.lost);
        const uint64_t m = std::min<uint64_t>(c.pending, s->scfg.pool_cap);
        if (m == 0) {
            s->pool_settled = 0;
            return DYC_OK;
        }
        ent.resize(m);
        tag.resize(m);
        HIP_TRY(hipMemcpy(ent.data(), s->d_pool, m * sizeof(uint64_t), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(tag.data(), s->d_pool_tag, m * sizeof(uint32_t), hipMemcpyDeviceToHost));
        // empty the pool (its append index), then re-insert table by table
        HIP_TRY(launch_zero(&s->pc()->pending, 1, s->ms()));
        std::vector<uint64_t> order(m);
        for (uint64_t j = 0; j < m; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return tag[a] < tag[b]; });
        std::vector<uint64_t> packed;
        packed.reserve(m);
        struct Span {
            uint32_t table;
            uint64_t begin, len;
        };
        std::vector<Span> spans;
        for (uint64_t j : order) {
            if (ent[j] == 0 || tag[j] >= s->n) continue;  // erased while pooled
            if (spans.empty() || spans.back().table != tag[j]) spans.push_back(Span{tag[j], packed.size(), 0});
            packed.push_back(ent[j]);
            spans.back().len++;
        }
        if (!packed.empty())
            HIP_TRY(hipMemcpyAsync(s->d_resolve, packed.data(), packed.size() * sizeof(uint64_t),
                                   hipMemcpyHostToDevice, s->ms()));
        for (const Span &sp : spans) {
            bool ok = false;
            rc = upsize_now(s, sp.table, &ok);
            if (rc) return rc;
            HIP_TRY(launch_insert_entries(s->t[sp.table].D, s->d_resolve + sp.begin, sp.len, kNoTable,
                                          (int)s->cfg.insert_lanes, s->policy(), s->insert_shape(), s->pc(),
                                          s->sink(sp.table), s->ms()));
        }
        HIP_TRY(hipStreamSynchronize(s->ms()));
        s->ss.pool_rounds++;
    }
    int rc = read_stats_sync(s, scratch);
    if (rc) return rc;
    const Counters &c = *reinterpret_cast<const Counters *>(scratch);
    if (c.lost) return fail(DYC_ERR_LOST, "overflow pool full: %llu entries lost", (unsigned long long)c.lost);
    s->pool_settled = std::min<uint64_t>(c.pending, s->scfg.pool_cap);
    return DYC_OK;
}

void finish(dyc_server *s, uint64_t tk, int status)
{
    Req &R = s->reqs[tk];
    R.state = 2;
    R.status = status;
    s->done.emplace_back(tk, status);
    s->ss.completed++;
    if (status != DYC_OK) s->ss.failed++;
    s->ss.keys += R.r.n;
}

// Retire tick k (its event has fired).
int retire(dyc_server *s, Tick &k)
{
    int status = s->error;
    if (!k.has_snap) {  // finds only: no count, pool or resize changed
        for (const TickTable &tt : k.tables) s->t[tt.table].inflight--;
        k.busy = false;
        for (uint64_t tk : k.reqs) finish(s, tk, status);
        k.reqs.clear();
        k.tables.clear();
        return DYC_OK;
    }
    fold_tallies(k.h_snap);  // the tick's copy has landed (its event fired); folded once, here
    const Counters &c = *reinterpret_cast<const Counters *>(k.h_snap);
    if (k.solo_erase && c.deleted > s->deleted_seen) {
        s->t[k.solo_table].sync_del += c.deleted - s->deleted_seen;
        s->deleted_seen = c.deleted;
    }
    if (k.solo_insert && c.placed > s->placed_seen) {
        s->t[k.solo_table].sync_ins += c.placed - s->placed_seen;
        s->placed_seen = c.placed;
    }
    if (c.lost) status = sticky(s, fail(DYC_ERR_LOST, "overflow pool full: %llu entries lost", (unsigned long long)c.lost));
    if (c.misplaced > s->misplaced_seen) {
        s->misplaced_seen = c.misplaced;
        status = sticky(s, fail(DYC_ERR_CORRUPT, "a resize found entries outside their bucket pair"));
    }
    if (c.stalls > s->stalls_seen) {
        s->stalls_seen = c.stalls;
        status = sticky(s, fail(DYC_ERR_HIP, "a tick item waited past spin_cap for its table's earlier rounds"));
    }
    refresh_counts(s, k.h_snap);
    for (const TickTable &tt : k.tables) {
        SrvTable &T = s->t[tt.table];
        T.inflight--;
        T.inflight_ins -= tt.ins;
        if (tt.erase && !T.check_down) {
            T.check_down = true;
            s->down_list.push_back(tt.table);
        }
    }
    k.busy = false;
    if (status == DYC_OK && std::min<uint64_t>(c.pending, s->scfg.pool_cap) > s->pool_settled) {
        int rc = resolve_pool(s, k.h_snap);
        if (rc) status = sticky(s, rc);
        else refresh_counts(s, k.h_snap);
    }
    for (uint64_t tk : k.reqs) finish(s, tk, status);
    k.reqs.clear();
    k.tables.clear();
    return DYC_OK;
}

// Retire finished ticks, oldest first; with `wait`, block on the oldest one.
int retire_ready(dyc_server *s, bool wait)
{
    while (true) {
        Tick *oldest = nullptr;
        for (uint32_t i = 0; i < s->depth(); ++i)
            if (s->ticks[i].busy && (!oldest || s->ticks[i].seq < oldest->seq)) oldest = &s->ticks[i];
        if (!oldest) return DYC_OK;
        hipError_t e = wait ? hipEventSynchronize(oldest->done) : hipEventQuery(oldest->done);
        if (e == hipErrorNotReady) return DYC_OK;
        if (e != hipSuccess) return sticky(s, fail(DYC_ERR_HIP, "tick event: %s", hipGetErrorString(e)));
        int rc = retire(s, *oldest);
        if (rc) return rc;
        wait = false;
    }
}

void mark(dyc_server *s, uint32_t x)
{
    if (!s->seen[x]) {
        s->seen[x] = 1;
        s->touched.push_back(x);
    }
}

// Close a tick: counters out (if it may change them), completion event, bookkeeping.
int finish_launch(dyc_server *s, Tick &k, bool *launched)
{
    if (k.has_snap) HIP_TRY(launch_copy(s->d_stat, k.h_snap, s->stat_bytes() / 8, nullptr, nullptr, 0, s->st));
    HIP_TRY(hipEventRecord(k.done, s->st));
    for (const TickTable &tt : k.tables) {
        SrvTable &T = s->t[tt.table];
        T.inflight++;
        T.inflight_ins += tt.ins;
    }
    k.busy = true;
    k.seq = ++s->tick_seq;
    s->ss.ticks++;
    *launched = true;
    return DYC_OK;
}

// Admit queued requests into a free tick slot and enqueue it. *launched tells whether a tick
// went out; *stuck is set when nothing could be admitted for lack of arena space while no tick
// is in flight (nothing would ever free it).
int build_tick(dyc_server *s, bool *launched, bool *stuck)
{
    *launched = false;
    *stuck = false;
    Tick *slot = nullptr;
    bool any_busy = false;
    for (uint32_t i = 0; i < s->depth(); ++i) {
        if (!s->ticks[i].busy && !slot) slot = &s->ticks[i];
        any_busy |= s->ticks[i].busy;
    }
    if (!slot) return DYC_OK;
    Tick &k = *slot;
    k.reqs.clear();
    k.tables.clear();
    for (uint32_t x : s->touched) {
        s->phase[x] = -1;
        s->blocked[x] = 0;
        s->in_tick[x] = 0;
        s->seen[x] = 0;
        s->tick_ins[x] = 0;
        s->round[x] = 0;
        s->had_erase[x] = 0;
        s->stage[x] = -1;
        s->last_sid[x] = ~0u;
        s->nstages[x] = 0;
    }
    s->touched.clear();
    bool resized = false;

    // downsizes of idle tables whose erases retired (freed space is admissible at once)
    if (!s->down_list.empty()) {
        std::vector<uint32_t> keep;
        for (uint32_t x : s->down_list) {
            SrvTable &T = s->t[x];
            if (T.inflight) {
                keep.push_back(x);
                continue;
            }
            T.check_down = false;
            if (!wants_down(s, T)) continue;
            bool did = false;
            int rc = downsize_now(s, x, k.d_rctr, &did);
            if (rc) return sticky(s, rc);
            if (did) {
                resized = true;
                mark(s, x);
                s->in_tick[x] = 1;
            }
        }
        s->down_list.swap(keep);
    }

    // admission, arrival order. A table's requests go into rounds: within a round they follow
    // insert -> find -> erase with at most one erase; a request that would break that opens the
    // table's next round. Round r runs after round r - 1 (three launches per round), so each
    // table sees its requests in arrival order.
    uint64_t tick_keys = 0;
    uint32_t nreq = 0, rounds = 0;
    bool memory_blocked = false;
    s->admitted.clear();
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        const uint64_t tk = *it;
        Req &R = s->reqs[tk];
        const uint32_t x = R.r.table, op = R.r.op;
        if (s->blocked[x]) {
            ++it;
            continue;
        }
        if (nreq > 0 && (nreq >= s->scfg.tick_max_requests || tick_keys + R.r.n > s->scfg.tick_max_keys)) {
            s->ss.deferred_slots++;
            break;
        }
        const int ph = s->phase[x];
        uint32_t r = s->round[x];
        if (ph >= 0 && (ph > (int)op || ph == (int)DYC_OP_ERASE)) {
            if (r + 1 >= kMaxRounds) {
                mark(s, x);
                s->blocked[x] = 1;
                s->ss.deferred_order++;
                ++it;
                continue;
            }
            ++r;
        }
        if (op == DYC_OP_INSERT) {
            SrvTable &T = s->t[x];
            bool room = true;
            while ((double)(T.count + T.inflight_ins + s->tick_ins[x] + R.r.n) >
                   s->cfg.upper_bound * (double)T.capacity()) {
                bool ok = false;
                int rc = upsize_now(s, x, &ok);
                if (rc) return sticky(s, rc);
                if (!ok) {
                    room = false;
                    break;
                }
                resized = true;
            }
            if (!room) {
                mark(s, x);
                s->blocked[x] = 2;
                s->mem_head[x] = tk;
                s->ss.deferred_memory++;
                memory_blocked = true;
                ++it;
                continue;
            }
        }
        mark(s, x);
        s->in_tick[x] = 1;
        if (op == DYC_OP_INSERT) s->tick_ins[x] += R.r.n;
        if (op == DYC_OP_ERASE) s->had_erase[x] = 1;
        if (s->phase[x] != (int8_t)op || s->round[x] != r || s->nstages[x] == 0) s->nstages[x]++;
        s->phase[x] = (int8_t)op;
        s->round[x] = (uint16_t)r;
        rounds = std::max(rounds, r + 1);
        s->admitted.emplace_back(tk, r);
        k.reqs.push_back(tk);
        R.state = 1;
        tick_keys += R.r.n;
        ++nreq;
        it = s->queue.erase(it);
    }
    if (k.reqs.empty() && !resized) {
        *stuck = memory_blocked && !any_busy;
        return DYC_OK;
    }

    // work items, sized so a tick has about one per resident wave
    uint64_t per = s->scfg.item_keys;
    if (per == 0) {
        per = (tick_keys / kTargetItems + kMinItemKeys - 1) / kMinItemKeys * kMinItemKeys;
        per = std::min<uint64_t>(std::max<uint64_t>(per, kMinItemKeys), kMaxItemKeys);
    }
    const uint64_t room_items = k.items_cap > nreq ? k.items_cap - nreq : 1;
    per = std::max<uint64_t>(per, (tick_keys + room_items - 1) / room_items);
    per = std::min<uint64_t>(per, 0xffffffffull);
    // A table with several stages (round, op) in this tick is a dependency chain: cut its
    // stages into many small items (up to kChainItems per stage, at least one tile each) so
    // that each stage finishes in about one tile's latency; other tables take the tick's size.
    s->stage_keys.clear();
    for (const auto &[tk, r] : s->admitted) {
        const dyc_request &q = s->reqs[tk].r;
        s->stage_keys[((uint64_t)q.table << 16 | r) * 3 + q.op] += q.n;
    }
    for (uint32_t r = 0; r < rounds; ++r)
        for (auto &v : s->items[r]) v.clear();
    for (const auto &[tk, r] : s->admitted) {
        const dyc_request &q = s->reqs[tk].r;
        uint64_t step_keys = per;
        if (s->scfg.item_keys == 0 && s->nstages[q.table] > 1) {
            const uint64_t sk = s->stage_keys[((uint64_t)q.table << 16 | r) * 3 + q.op];
            step_keys = (sk / kChainItems + kMinItemKeys - 1) / kMinItemKeys * kMinItemKeys;
            step_keys = std::max<uint64_t>(step_keys, std::max<uint64_t>(kMinItemKeys,
                                                                          (tick_keys + room_items - 1) / room_items));
        }
        for (uint64_t o = 0; o < q.n; o += step_keys) {
            ServeItem it{};
            it.keys = q.keys + o;
            it.vals = q.op == DYC_OP_INSERT ? q.vals + o : nullptr;
            it.out_vals = (q.op == DYC_OP_FIND && q.out_vals) ? q.out_vals + o : nullptr;
            it.flags = (q.op != DYC_OP_INSERT && q.out_flags) ? q.out_flags + o : nullptr;
            it.len = (uint32_t)std::min<uint64_t>(step_keys, q.n - o);
            it.table = q.table;
            s->items[r][q.op].push_back(it);
        }
    }
    for (uint32_t x : s->touched)
        if (s->in_tick[x]) k.tables.push_back(TickTable{x, s->tick_ins[x], s->had_erase[x] != 0});
    const bool dir_up = s->dir_dirty;
    if (dir_up)
        for (uint32_t i = 0; i < s->n; ++i) k.h_dir[i] = s->t[i].D;
    s->dir_dirty = false;
    uint64_t n_items = 0;
    for (uint32_t r = 0; r < rounds; ++r)
        for (uint32_t op = 0; op < 3; ++op) n_items += s->items[r][op].size();
    if (n_items > k.items_cap) return sticky(s, fail(DYC_ERR_INVALID, "tick needs more work items than staged"));
    // Alone on an idle server (light load), a request goes out exactly as the synchronous path
    // launches it -- the plain single-table kernel -- without waiting for it.
    k.solo_erase = false;
    k.solo_insert = false;
    if (k.reqs.size() == 1) {
        const dyc_request &q = s->reqs[k.reqs.front()].r;
        const SrvTable &T = s->t[q.table];
        const Pool pool = s->pool();
        k.has_snap = q.op != DYC_OP_FIND;
        k.solo_erase = q.op == DYC_OP_ERASE;
        k.solo_insert = q.op == DYC_OP_INSERT;
        k.solo_table = q.table;
        if (dir_up) s->dir_dirty = true;  // the plain kernels take the descriptor by value
        if (q.op == DYC_OP_INSERT)
            HIP_TRY(launch_insert(T.D, q.keys, q.vals, q.n, (int)s->cfg.insert_lanes, s->policy(), s->insert_shape(),
                                  s->pc(), s->sink(q.table, true), s->st));
        else if (q.op == DYC_OP_FIND)
            HIP_TRY(launch_find(T.D, q.keys, q.out_vals, q.out_flags, q.n, (int)s->cfg.find_lanes, s->shape(), s->st,
                                &pool, q.table));
        else
            HIP_TRY(launch_erase(T.D, q.keys, q.out_flags, q.n, (int)s->cfg.find_lanes, s->shape(), s->pc(), s->st,
                                 &pool, q.table));
        s->ss.rounds += 1;
        return finish_launch(s, k, launched);
    }
    // one launch pays off when a table needs several rounds; a single round goes out as up to
    // three plain launches (lower latency when the load is light)
    const bool one = s->scfg.one_launch && rounds > 1 &&
                     serve_tick_supported((int)s->cfg.insert_lanes, (int)s->cfg.find_lanes);
    uint64_t writes = 0;  // insert and erase items
    for (uint32_t r = 0; r < rounds; ++r) writes += s->items[r][0].size() + s->items[r][2].size();
    k.has_snap = writes > 0 || resized;
    if (one) {
        // items in (round, op) order; a table's items of one (round, op) form a stage that waits
        // for the table's previous stage in the tick
        k.h_block[0] = 0;
        TickItem *ti = reinterpret_cast<TickItem *>(k.h_block + 1);
        std::vector<uint32_t> &left = s->st_left, &next = s->st_next;
        left.clear();
        next.clear();
        s->st_ready.clear();
        uint64_t at = 0;
        for (uint32_t r = 0; r < rounds; ++r)
            for (uint32_t op = 0; op < 3; ++op)
                for (const ServeItem &it : s->items[r][op]) {
                    const uint32_t x = it.table;
                    const int32_t key = (int32_t)(r * 3 + op);
                    if (s->stage[x] != key) {
                        const uint32_t sid = (uint32_t)left.size();
                        left.push_back(0);
                        next.push_back(~0u);
                        s->st_ready.push_back(s->last_sid[x] == ~0u ? 1u : 0u);
                        if (s->last_sid[x] != ~0u) next[s->last_sid[x]] = sid;
                        s->last_sid[x] = sid;
                        s->stage[x] = key;
                    }
                    const uint32_t sid = s->last_sid[x];
                    left[sid]++;
                    TickItem &d = ti[at++];
                    d.item = it;
                    d.op = op;
                    d.stage = sid;
                    d.wait = s->st_ready[sid] ? 0u : 1u;
                    d.pad = 0;
                }
        const uint64_t ns = left.size();
        const uint64_t w_items = 1 + n_items * sizeof(TickItem) / 8;
        const uint64_t w_half = (ns + 1) / 2;  // a u32 array of ns entries, in 8-byte words
        uint32_t *h_left = reinterpret_cast<uint32_t *>(k.h_block + w_items);
        uint32_t *h_next = reinterpret_cast<uint32_t *>(k.h_block + w_items + w_half);
        uint32_t *h_ready = reinterpret_cast<uint32_t *>(k.h_block + w_items + 2 * w_half);
        memcpy(h_left, left.data(), ns * 4);
        memcpy(h_next, next.data(), ns * 4);
        for (uint64_t i = 0; i < ns; ++i) h_ready[i * kReadyStride] = s->st_ready[i];
        const uint64_t words = w_items + 2 * w_half + ns * kReadyStride / 2;
        HIP_TRY(launch_copy(k.h_block, k.d_block, words, k.h_dir, s->d_dir, dir_up ? s->n * sizeof(Desc) / 8 : 0,
                            s->st));
        uint32_t *d_left = reinterpret_cast<uint32_t *>(k.d_block + w_items);
        TickStages stg{d_left, reinterpret_cast<const uint32_t *>(k.d_block + w_items + w_half),
                       reinterpret_cast<uint32_t *>(k.d_block + w_items + 2 * w_half)};
        HIP_TRY(launch_serve_tick(s->d_dir, reinterpret_cast<const TickItem *>(k.d_block + 1), (uint32_t)n_items,
                                  k.d_block, stg, s->policy(), (int)s->cfg.insert_lanes, (int)s->cfg.find_lanes, 0,
                                  s->pc(), s->sink(0), s->pool(), s->ins_total(), s->del_total(), s->scfg.spin_cap,
                                  s->scfg.poll_backoff, s->st));
    } else {
        uint64_t at = 0;
        for (uint32_t r = 0; r < rounds; ++r)
            for (uint32_t op = 0; op < 3; ++op) {
                const auto &v = s->items[r][op];
                memcpy(k.h_items + at, v.data(), v.size() * sizeof(ServeItem));
                at += v.size();
            }
        // a small tick's items are read where the host wrote them (pinned memory)
        const bool direct = n_items <= kDirectItems;
        const ServeItem *items = direct ? k.h_items : k.d_items;
        if (direct) {
            if (dir_up) HIP_TRY(launch_copy(k.h_dir, s->d_dir, s->n * sizeof(Desc) / 8, nullptr, nullptr, 0, s->st));
        } else {
            HIP_TRY(launch_copy(k.h_items, k.d_items, n_items * sizeof(ServeItem) / 8, k.h_dir, s->d_dir,
                                dir_up ? s->n * sizeof(Desc) / 8 : 0, s->st));
        }
        at = 0;
        for (uint32_t r = 0; r < rounds; ++r) {
            const uint64_t ni = s->items[r][0].size(), nf = s->items[r][1].size(), ne = s->items[r][2].size();
            HIP_TRY(launch_serve_insert(s->d_dir, items + at, (uint32_t)ni, s->policy(), (int)s->cfg.insert_lanes, 0,
                                        s->pc(), s->sink(0), s->ins_total(), s->st));
            at += ni;
            HIP_TRY(launch_serve_find(s->d_dir, items + at, (uint32_t)nf, s->pool(), (int)s->cfg.find_lanes, 0, s->st));
            at += nf;
            HIP_TRY(launch_serve_erase(s->d_dir, items + at, (uint32_t)ne, s->pool(), s->del_total(),
                                       (int)s->cfg.find_lanes, 0, s->st));
            at += ne;
        }
    }
    s->ss.rounds += rounds;
    return finish_launch(s, k, launched);
}

// One non-blocking step: retire what finished, then fill the free slots.
int step(dyc_server *s)
{
    int rc = retire_ready(s, false);
    if (rc) return rc;
    for (uint32_t i = 0; i < s->depth(); ++i) {
        if (s->queue.empty() && s->down_list.empty()) break;
        bool launched = false, stuck = false;
        rc = build_tick(s, &launched, &stuck);
        if (rc) return rc;
        if (stuck) {
            // nothing in flight can free arena space: fail each insert that waits for it (its
            // table's later requests stay queued and run in order)
            for (auto it = s->queue.begin(); it != s->queue.end();) {
                const uint32_t x = s->reqs[*it].r.table;
                if (s->blocked[x] == 2 && s->mem_head[x] == *it) {
                    fail(DYC_ERR_OOM, "table %u cannot grow within the server's HBM budget", x);
                    finish(s, *it, DYC_ERR_OOM);
                    it = s->queue.erase(it);
                } else {
                    ++it;
                }
            }
            continue;  // try again with the failed requests gone
        }
        if (!launched) break;
    }
    return DYC_OK;
}

bool idle(const dyc_server *s)
{
    for (uint32_t i = 0; i < s->depth(); ++i)
        if (s->ticks[i].busy) return false;
    return s->queue.empty();
}

int drain(dyc_server *s)
{
    while (true) {
        int rc = step(s);
        if (rc) return rc;
        if (idle(s)) break;
        rc = retire_ready(s, true);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    return DYC_OK;
}

int check_request(const dyc_server *s, const dyc_request &r)
{
    if (r.table >= s->n) return fail(DYC_ERR_INVALID, "table %u out of range (%u tables)", r.table, s->n);
    if (r.op > DYC_OP_ERASE) return fail(DYC_ERR_INVALID, "unknown op %u", r.op);
    if (r.n && !r.keys) return fail(DYC_ERR_INVALID, "null keys");
    if (r.op == DYC_OP_INSERT && r.n && !r.vals) return fail(DYC_ERR_INVALID, "insert needs values");
    return DYC_OK;
}

// The reference's calling pattern: one request alone, plain single-table kernels, then a
// stream synchronisation (api.cpp:185-300).
int run_sync(dyc_server *s, uint64_t tk)
{
    Req &R = s->reqs[tk];
    const dyc_request &r = R.r;
    SrvTable &T = s->t[r.table];
    unsigned long long *snap = s->ticks[0].h_snap;
    int status = DYC_OK;
    if (r.op == DYC_OP_INSERT) {
        while ((double)(T.count + r.n) > s->cfg.upper_bound * (double)T.capacity()) {
            bool ok = false;
            int rc = upsize_now(s, r.table, &ok);
            if (rc) return rc;
            if (!ok) {
                status = fail(DYC_ERR_OOM, "table %u cannot grow within the server's HBM budget", r.table);
                break;
            }
        }
        if (status == DYC_OK)
            HIP_TRY(launch_insert(T.D, r.keys, r.vals, r.n, (int)s->cfg.insert_lanes, s->policy(), s->insert_shape(),
                                  s->pc(), s->sink(r.table, true), s->st));
    } else if (r.op == DYC_OP_FIND) {
        // a find changes nothing: launch, wait, done
        const Pool pool = s->pool();
        HIP_TRY(launch_find(T.D, r.keys, r.out_vals, r.out_flags, r.n, (int)s->cfg.find_lanes, s->shape(), s->st,
                            &pool, r.table));
        HIP_TRY(hipStreamSynchronize(s->st));
        finish(s, tk, status);
        return DYC_OK;
    } else {
        const Pool pool = s->pool();
        HIP_TRY(launch_erase(T.D, r.keys, r.out_flags, r.n, (int)s->cfg.find_lanes, s->shape(), s->pc(), s->st,
                             &pool, r.table));
    }
    int rc = read_stats_sync(s, snap);
    if (rc) return rc;
    const Counters &c = *reinterpret_cast<const Counters *>(snap);
    if (c.deleted > s->deleted_seen) {
        T.sync_del += c.deleted - s->deleted_seen;
        s->deleted_seen = c.deleted;
    }
    if (c.placed > s->placed_seen) {
        T.sync_ins += c.placed - s->placed_seen;
        s->placed_seen = c.placed;
    }
    refresh_counts(s, snap);
    if (c.lost) status = sticky(s, fail(DYC_ERR_LOST, "overflow pool full"));
    if (status == DYC_OK && std::min<uint64_t>(c.pending, s->scfg.pool_cap) > s->pool_settled) {
        rc = resolve_pool(s, snap);
        if (rc) status = sticky(s, rc);
        else refresh_counts(s, snap);
    }
    if (status == DYC_OK && r.op == DYC_OP_ERASE && wants_down(s, T)) {
        bool did = false;
        rc = downsize_now(s, r.table, s->ticks[0].d_rctr, &did);
        if (rc) return rc;
        if (did) {
            rc = read_stats_sync(s, snap);
            if (rc) return rc;
            if (std::min<uint64_t>(reinterpret_cast<const Counters *>(snap)->pending, s->scfg.pool_cap) >
                s->pool_settled) {
                rc = resolve_pool(s, snap);
                if (rc) status = sticky(s, rc);
            }
        }
    }
    if (c.misplaced > s->misplaced_seen) {
        s->misplaced_seen = c.misplaced;
        status = sticky(s, fail(DYC_ERR_CORRUPT, "a resize found entries outside their bucket pair"));
    }
    finish(s, tk, status);
    return DYC_OK;
}

// ------------------------------------------------------------------ device-resident loop

constexpr double kLoopGiveUpS = 60.0;  // host-side bound on waiting for the loop to finish or stop

// Copy the directory to the device on the maintenance stream and wait for it.
int loop_upload_dir(dyc_server *s)
{
    Desc *h = s->ticks[0].h_dir;  // pinned; the ticks are idle while the loop is on
    for (uint32_t i = 0; i < s->n; ++i) h[i] = s->t[i].D;
    HIP_TRY(launch_copy(h, s->d_dir, s->n * sizeof(Desc) / 8, nullptr, nullptr, 0, s->ms()));
    HIP_TRY(hipStreamSynchronize(s->ms()));
    s->dir_dirty = false;
    return DYC_OK;
}

void loop_free(dyc_server *s);

// Everything the loop needs, allocated on first use; on any failure nothing is kept.
int loop_alloc(dyc_server *s)
{
    Loop &L = s->loop;
    if (L.ctl) return DYC_OK;
    const int cap = serve_loop_capacity((int)s->cfg.insert_lanes, (int)s->cfg.find_lanes);
    if (cap < 2) return fail(DYC_ERR_INVALID, "the resident loop needs insert_lanes 1, 2 or 4 and find_lanes 4, 8 or 16");
    auto undo = [&](int code) {
        loop_free(s);
        return code;
    };
    // default: half the resident capacity, so resizes on the second stream (and the caller's own
    // kernels) find room beside the loop instead of waiting for it to end
    const int blocks = s->scfg.resident_blocks ? std::min<int>((int)s->scfg.resident_blocks, cap) : cap / 2;
    const unsigned fl = hipHostMallocCoherent;
    const size_t dev_bytes = (kLoopWords + (size_t)s->n + kLoopSlots) * 8 + (size_t)kLoopSlots * 8;
    unsigned long long *d = nullptr;
    if (hipStreamCreateWithFlags(&L.st2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.ended, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&L.ring, kLoopRingItems * sizeof(LoopItem), fl) != hipSuccess ||
        hipHostMalloc((void **)&L.rec, kLoopSlots * sizeof(LoopRec), fl) != hipSuccess ||
        hipHostMalloc((void **)&L.cq, kLoopSlots * sizeof(unsigned long long), fl) != hipSuccess ||
        hipMalloc((void **)&d, dev_bytes) != hipSuccess) {
        (void)hipGetLastError();
        L.d_words = d;
        fail(DYC_ERR_OOM, "the resident loop's buffers (%zu bytes of device memory, about 19 MiB pinned) failed",
             dev_bytes);
        return undo(DYC_ERR_OOM);
    }
    L.d_words = d;
    L.d_tbl_done = d + kLoopWords;
    L.d_req_word = L.d_tbl_done + s->n;
    L.d_req_out = reinterpret_cast<uint32_t *>(L.d_req_word + kLoopSlots);
    L.d_req_stall = L.d_req_out + kLoopSlots;
    memset(L.ring, 0, kLoopRingItems * sizeof(LoopItem));
    memset(L.rec, 0, kLoopSlots * sizeof(LoopRec));
    memset(L.cq, 0, kLoopSlots * sizeof(unsigned long long));
    if (launch_zero(d, dev_bytes / 8, L.st2) != hipSuccess || hipStreamSynchronize(L.st2) != hipSuccess) {
        fail(DYC_ERR_HIP, "zeroing the resident loop's words failed");
        return undo(DYC_ERR_HIP);
    }
    // the control block last: its presence marks the loop as allocated
    if (hipHostMalloc((void **)&L.ctl, sizeof(LoopCtl), fl) != hipSuccess) {
        (void)hipGetLastError();
        L.ctl = nullptr;
        fail(DYC_ERR_OOM, "the resident loop's control block failed");
        return undo(DYC_ERR_OOM);
    }
    memset(L.ctl, 0, sizeof(LoopCtl));
    L.blocks = std::max(blocks, 2);
    L.slot.assign(kLoopSlots, LoopSlot{});
    L.free_slots.resize(kLoopSlots);
    for (uint32_t i = 0; i < kLoopSlots; ++i) L.free_slots[i] = kLoopSlots - 1 - i;
    L.pub.assign(s->n, 0);
    L.need.assign(s->n, 0);
    L.last_op.assign(s->n, -1);
    return DYC_OK;
}

void loop_free(dyc_server *s)
{
    Loop &L = s->loop;
    if (L.st2) (void)hipStreamSynchronize(L.st2);
    if (L.ctl) (void)hipHostFree(L.ctl);
    if (L.ring) (void)hipHostFree(L.ring);
    if (L.rec) (void)hipHostFree(L.rec);
    if (L.cq) (void)hipHostFree(L.cq);
    if (L.d_words) (void)hipFree(L.d_words);
    if (L.ended) (void)hipEventDestroy(L.ended);
    if (L.st2) (void)hipStreamDestroy(L.st2);
    L = Loop{};
}

int loop_launch(dyc_server *s)
{
    Loop &L = s->loop;
    __atomic_store_n(&L.ctl->stop, 0ull, __ATOMIC_RELEASE);
    HIP_TRY(launch_loop_reset(L.d_words, s->st));
    LoopArgs A{};
    A.dir = s->d_dir;
    A.ring = L.ring;
    A.ring_mask = kLoopRingItems - 1;
    A.ctl = L.ctl;
    A.rec = L.rec;
    A.cq = L.cq;
    A.req_mask = kLoopSlots - 1;
    A.spin_cap = s->scfg.spin_cap;
    A.words = L.d_words;
    A.tbl_done = L.d_tbl_done;
    A.req_word = L.d_req_word;
    A.req_out = L.d_req_out;
    A.req_stall = L.d_req_stall;
    A.idle_ticks = (uint64_t)std::max<uint32_t>(1, s->scfg.resident_idle_ms) * 100000ull;  // 100 MHz clock
    A.max_back = std::max<uint32_t>(1, s->scfg.poll_backoff);
    HIP_TRY(launch_serve_loop(A, s->policy(), (int)s->cfg.insert_lanes, (int)s->cfg.find_lanes, L.blocks, s->pc(),
                              s->sink(0), s->pool(), s->ins_total(), s->del_total(), s->st));
    HIP_TRY(hipEventRecord(L.ended, s->st));
    L.running = true;
    s->ss.loop_launches++;
    return DYC_OK;
}

// Notice a launch that has ended (idle time-out, stop or fault).
int loop_check_ended(dyc_server *s)
{
    Loop &L = s->loop;
    if (!L.running) return DYC_OK;
    const hipError_t e = hipEventQuery(L.ended);
    if (e == hipErrorNotReady) return DYC_OK;
    if (e != hipSuccess) return sticky(s, fail(DYC_ERR_HIP, "loop launch: %s", hipGetErrorString(e)));
    L.running = false;
    unsigned long long fault = 0;
    HIP_TRY(hipMemcpy(&fault, L.d_words + kLoopFault, sizeof fault, hipMemcpyDeviceToHost));
    if (fault) return sticky(s, fail(DYC_ERR_HIP, "the loop found a ring entry the host had not finished writing"));
    return DYC_OK;
}

// Read the completion queue: finish each request it names.
int loop_collect(dyc_server *s, uint32_t *got)
{
    Loop &L = s->loop;
    *got = 0;
    while (true) {
        const unsigned long long e = __atomic_load_n(L.cq + (L.cq_head & (kLoopSlots - 1)), __ATOMIC_ACQUIRE);
        if ((uint32_t)(e >> 32) != (uint32_t)(L.cq_head + 1)) return DYC_OK;
        L.cq_head++;
        const uint32_t si = (uint32_t)e;
        if (si >= kLoopSlots || L.slot[si].pseq < L.order_base)
            return sticky(s, fail(DYC_ERR_CORRUPT, "the loop's completion queue names slot %u", si));
        const LoopRec *r = L.rec + si;
        const unsigned long long acc = __atomic_load_n(&r->acc, __ATOMIC_ACQUIRE);
        const unsigned long long pool = __atomic_load_n(&r->pool, __ATOMIC_ACQUIRE);
        const unsigned long long lost = __atomic_load_n(&r->lost, __ATOMIC_ACQUIRE);
        const uint32_t stalls = __atomic_load_n(&r->stalls, __ATOMIC_ACQUIRE);
        const LoopSlot q = L.slot[si];
        SrvTable &T = s->t[q.table];
        if (q.op == DYC_OP_INSERT) {
            T.count += acc;
            T.inflight_ins -= q.n;
        } else if (q.op == DYC_OP_ERASE) {
            T.count -= std::min<uint64_t>(T.count, acc);
            if (!T.check_down) {
                T.check_down = true;
                s->down_list.push_back(q.table);
            }
        }
        T.inflight--;
        int status = s->error;
        if (lost) status = sticky(s, fail(DYC_ERR_LOST, "overflow pool full: %llu entries lost", lost));
        if (stalls)
            status = sticky(s, fail(DYC_ERR_HIP, "a loop item waited past spin_cap for its table's earlier requests"));
        if (std::min<uint64_t>(pool, s->scfg.pool_cap) > s->pool_settled) L.want_resolve = true;
        L.order[q.pseq - L.order_base].second = true;
        while (!L.order.empty() && L.order.front().second) {
            L.order.pop_front();
            L.order_base++;
        }
        L.free_slots.push_back(si);
        finish(s, q.ticket, status);
        ++*got;
    }
}

// Keys per work item of a request: one item per worker wave of the launch (each item costs a
// read of its descriptor from host memory, so a wave takes one item of the request rather than
// several small ones), and at most 65,536 items per request (the ring holds 2^18). Sizing items
// by the keys already in flight instead (larger items under load) cut the saturated rate of the
// 257-table serving trace from 3.0 to 0.4-0.7 G keys/s: a table's later stages wait for its
// earlier ones, and a large item keeps its stage open longer.
uint64_t loop_item_keys(const dyc_server *s, uint64_t n)
{
    uint64_t per = s->scfg.item_keys;
    if (per == 0) {
        const uint64_t waves = (uint64_t)s->loop.blocks * 4 - 1;
        per = ((n + waves - 1) / waves + kMinItemKeys - 1) / kMinItemKeys * kMinItemKeys;
        per = std::min<uint64_t>(std::max<uint64_t>(per, kMinItemKeys), kMaxItemKeys);
    }
    const uint64_t floor_per = ((n + 65535) / 65536 + kMinItemKeys - 1) / kMinItemKeys * kMinItemKeys;
    return std::min<uint64_t>(std::max(per, floor_per), 0xffffffc0ull);
}

void loop_clear_scratch(dyc_server *s)
{
    for (uint32_t x : s->touched) {
        s->blocked[x] = 0;
        s->seen[x] = 0;
    }
    s->touched.clear();
}

// Publish queued requests in arrival order. *stuck: nothing could be published for lack of
// arena space while nothing is in flight.
int loop_publish(dyc_server *s, bool *stuck)
{
    Loop &L = s->loop;
    *stuck = false;
    loop_clear_scratch(s);
    bool memory_blocked = false;
    uint64_t published = 0;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        const uint64_t tk = *it;
        Req &R = s->reqs[tk];
        const uint32_t x = R.r.table, op = R.r.op;
        if (s->blocked[x]) {
            ++it;
            continue;
        }
        if (L.free_slots.empty()) {
            s->ss.deferred_slots++;
            break;
        }
        const uint64_t per = loop_item_keys(s, R.r.n);
        const uint64_t n_items = std::max<uint64_t>(1, (R.r.n + per - 1) / per);
        if (L.tail + n_items - L.oldest_item() > kLoopRingItems) {
            s->ss.deferred_slots++;
            break;
        }
        SrvTable &T = s->t[x];
        if (op == DYC_OP_INSERT) {
            bool room = true, busy = false;
            while ((double)(T.count + T.inflight_ins + R.r.n) > s->cfg.upper_bound * (double)T.capacity()) {
                if (T.inflight) {  // a table resizes only with no request in flight
                    busy = true;
                    break;
                }
                bool ok = false;
                int rc = upsize_now(s, x, &ok);
                if (rc) return sticky(s, rc);
                if (!ok) {
                    room = false;
                    break;
                }
            }
            if (s->dir_dirty) {
                int rc = loop_upload_dir(s);
                if (rc) return sticky(s, rc);
            }
            if (busy || !room) {
                mark(s, x);
                s->blocked[x] = busy ? 1 : 2;
                if (busy) {
                    s->ss.deferred_order++;
                } else {
                    s->mem_head[x] = tk;
                    s->ss.deferred_memory++;
                    memory_blocked = true;
                }
                ++it;
                continue;
            }
        }
        // a table's requests run in stages: a run of inserts, a run of finds, or one erase
        if (L.last_op[x] != (int8_t)op || op == DYC_OP_ERASE) {
            L.need[x] = L.pub[x];
            L.last_op[x] = (int8_t)op;
        }
        L.pub[x]++;
        const uint32_t si = L.free_slots.back();
        L.free_slots.pop_back();
        LoopSlot &q = L.slot[si];
        q.ticket = tk;
        q.pseq = L.pseq++;
        q.n = R.r.n;
        q.table = x;
        q.op = op;
        L.order.emplace_back(L.tail, false);
        for (uint64_t j = 0; j < n_items; ++j) {
            const uint64_t o = j * per;
            LoopItem &d = L.ring[(L.tail + j) & (kLoopRingItems - 1)];
            d.item.keys = R.r.keys + o;
            d.item.vals = op == DYC_OP_INSERT ? R.r.vals + o : nullptr;
            d.item.out_vals = (op == DYC_OP_FIND && R.r.out_vals) ? R.r.out_vals + o : nullptr;
            d.item.flags = (op != DYC_OP_INSERT && R.r.out_flags) ? R.r.out_flags + o : nullptr;
            d.item.len = (uint32_t)std::min<uint64_t>(per, R.r.n - std::min(o, R.r.n));
            d.item.table = x;
            d.op = op;
            d.req = si;
            d.need = L.need[x];
            d.n_items = (uint32_t)n_items;
            __atomic_store_n(&d.stamp, (uint32_t)(L.tail + j + 1), __ATOMIC_RELEASE);  // last
        }
        L.tail += n_items;
        T.inflight++;
        if (op == DYC_OP_INSERT) T.inflight_ins += R.r.n;
        R.state = 1;
        s->ss.loop_items += n_items;
        s->ss.rounds++;
        it = s->queue.erase(it);
        ++published;
    }
    if (published) __atomic_store_n(&L.ctl->tail, L.tail, __ATOMIC_RELEASE);
    *stuck = !published && memory_blocked && L.in_flight() == 0;
    return DYC_OK;
}

// Wait until every published request has finished, then stop the launch. On time-out the host
// gives up: the launch is told to cut its waits short (its requests fail) and to end.
int loop_quiesce(dyc_server *s)
{
    Loop &L = s->loop;
    const Clock::time_point t0 = Clock::now();
    while (L.in_flight()) {
        uint32_t got = 0;
        int rc = loop_collect(s, &got);
        if (rc) return rc;
        if (got) continue;
        rc = loop_check_ended(s);
        if (rc) return rc;
        if (!L.running && L.in_flight()) {  // it ended on idle before it saw the last items
            rc = loop_launch(s);
            if (rc) return sticky(s, rc);
        }
        if (seconds_since(t0) > kLoopGiveUpS) {
            __atomic_store_n(&L.ctl->stop, 2ull, __ATOMIC_RELEASE);
            return sticky(s, fail(DYC_ERR_HIP, "the loop did not finish its requests in %.0f s", kLoopGiveUpS));
        }
    }
    if (L.running) {
        __atomic_store_n(&L.ctl->stop, 1ull, __ATOMIC_RELEASE);
        while (true) {
            const hipError_t e = hipEventQuery(L.ended);
            if (e == hipSuccess) break;
            if (e != hipErrorNotReady) return sticky(s, fail(DYC_ERR_HIP, "loop launch: %s", hipGetErrorString(e)));
            if (seconds_since(t0) > kLoopGiveUpS) {
                __atomic_store_n(&L.ctl->stop, 2ull, __ATOMIC_RELEASE);
                return sticky(s, fail(DYC_ERR_HIP, "the loop did not stop in %.0f s", kLoopGiveUpS));
            }
        }
        int rc = loop_check_ended(s);
        if (rc) return rc;
    }
    return DYC_OK;
}

// Re-place the overflow pool with the loop stopped (every table quiet), as a retiring tick does.
int loop_resolve(dyc_server *s)
{
    int rc = loop_quiesce(s);
    if (rc) return rc;
    s->ss.loop_pauses++;
    unsigned long long *snap = s->ticks[0].h_snap;
    rc = read_stats_sync(s, snap);
    if (rc) return sticky(s, rc);
    refresh_counts(s, snap);
    const Counters &c = *reinterpret_cast<const Counters *>(snap);
    if (c.lost) return sticky(s, fail(DYC_ERR_LOST, "overflow pool full: %llu entries lost", (unsigned long long)c.lost));
    if (std::min<uint64_t>(c.pending, s->scfg.pool_cap) > s->pool_settled) {
        rc = resolve_pool(s, snap);
        if (rc) return sticky(s, rc);
        refresh_counts(s, snap);
    }
    s->loop.want_resolve = false;
    if (s->dir_dirty) return loop_upload_dir(s);
    return DYC_OK;
}

// Downsize the tables whose erases finished and that have no request in flight.
int loop_downsizes(dyc_server *s)
{
    if (s->down_list.empty()) return DYC_OK;
    std::vector<uint32_t> keep;
    bool resized = false;
    for (uint32_t x : s->down_list) {
        SrvTable &T = s->t[x];
        if (T.inflight) {
            keep.push_back(x);
            continue;
        }
        T.check_down = false;
        if (!wants_down(s, T)) continue;
        bool did = false;
        int rc = downsize_now(s, x, s->ticks[0].d_rctr, &did);
        if (rc) return sticky(s, rc);
        resized |= did;
    }
    s->down_list.swap(keep);
    if (!resized) return DYC_OK;
    int rc = loop_upload_dir(s);
    if (rc) return sticky(s, rc);
    // residuals that found no place went to the pool
    unsigned long long *snap = s->ticks[0].h_snap;
    rc = read_stats_sync(s, snap);
    if (rc) return sticky(s, rc);
    const Counters &c = *reinterpret_cast<const Counters *>(snap);
    if (c.misplaced > s->misplaced_seen) {
        s->misplaced_seen = c.misplaced;
        return sticky(s, fail(DYC_ERR_CORRUPT, "a resize found entries outside their bucket pair"));
    }
    if (std::min<uint64_t>(c.pending, s->scfg.pool_cap) > s->pool_settled) s->loop.want_resolve = true;
    return DYC_OK;
}

// One non-blocking step of the loop mode: finish what completed, resize and re-place what is
// due, publish what is queued, and make sure a launch is there to take it.
int loop_step(dyc_server *s)
{
    Loop &L = s->loop;
    uint32_t got = 0;
    int rc = loop_collect(s, &got);
    if (rc) return rc;
    rc = loop_check_ended(s);
    if (rc) return rc;
    if (L.want_resolve && s->error == DYC_OK) {
        rc = loop_resolve(s);
        if (rc) return rc;
    }
    rc = loop_downsizes(s);
    if (rc) return rc;
    while (!s->queue.empty()) {
        bool stuck = false;
        rc = loop_publish(s, &stuck);
        if (rc) return rc;
        if (!stuck) break;
        // nothing in flight can free arena space: fail each insert that waits for it
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            const uint32_t x = s->reqs[*it].r.table;
            if (s->blocked[x] == 2 && s->mem_head[x] == *it) {
                fail(DYC_ERR_OOM, "table %u cannot grow within the server's HBM budget", x);
                finish(s, *it, DYC_ERR_OOM);
                it = s->queue.erase(it);
            } else {
                ++it;
            }
        }
    }
    if (!L.running && L.in_flight()) {
        rc = loop_launch(s);
        if (rc) return sticky(s, rc);
    }
    return DYC_OK;
}

// Switch between host-built ticks and the loop; both directions start from an idle server.
int set_loop(dyc_server *s, bool on)
{
    if (s->loop.on == on) return DYC_OK;
    if (on) {
        int rc = drain(s);  // ticks
        if (rc) return rc;
        rc = loop_alloc(s);
        if (rc) return rc;
        s->loop.on = true;
        std::fill(s->loop.last_op.begin(), s->loop.last_op.end(), (int8_t)-1);
        return loop_upload_dir(s);
    }
    int rc = loop_quiesce(s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->loop.st2));
    s->loop.on = false;
    s->dir_dirty = true;  // ticks upload the directory before their first launch
    return DYC_OK;
}

// Block until every submitted request has finished (the launch keeps running).
int loop_drain(dyc_server *s)
{
    Clock::time_point last = Clock::now();
    uint64_t seen = s->ss.completed;
    while (!s->queue.empty() || s->loop.in_flight()) {
        int rc = loop_step(s);
        if (rc) return rc;
        if (s->ss.completed != seen) {
            seen = s->ss.completed;
            last = Clock::now();
        } else if (seconds_since(last) > kLoopGiveUpS) {
            __atomic_store_n(&s->loop.ctl->stop, 2ull, __ATOMIC_RELEASE);
            return sticky(s, fail(DYC_ERR_HIP, "no request finished in %.0f s", kLoopGiveUpS));
        }
    }
    return DYC_OK;
}

int progress(dyc_server *s) { return s->loop.on ? loop_step(s) : step(s); }

int drain_any(dyc_server *s) { return s->loop.on ? loop_drain(s) : drain(s); }

void destroy(dyc_server *s)
{
    if (s->loop.on) {
        (void)loop_drain(s);
        (void)loop_quiesce(s);
    }
    if (s->st) (void)hipStreamSynchronize(s->st);
    loop_free(s);
    for (Tick &k : s->ticks) {
        if (k.h_items) (void)hipHostFree(k.h_items);
        if (k.d_items) (void)hipFree(k.d_items);
        if (k.h_dir) (void)hipHostFree(k.h_dir);
        if (k.h_snap) (void)hipHostFree(k.h_snap);
        if (k.h_block) (void)hipHostFree(k.h_block);
        if (k.d_block) (void)hipFree(k.d_block);
        if (k.d_rctr) (void)hipFree(k.d_rctr);
        if (k.done) (void)hipEventDestroy(k.done);
    }
    if (s->base) (void)hipFree(s->base);
    if (s->d_dir) (void)hipFree(s->d_dir);
    if (s->d_stat) (void)hipFree(s->d_stat);
    if (s->d_pool) (void)hipFree(s->d_pool);
    if (s->d_pool_tag) (void)hipFree(s->d_pool_tag);
    if (s->d_resolve) (void)hipFree(s->d_resolve);

    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

int create(dyc_server *s, uint32_t n_tables, const uint64_t *hints)
{
    s->n = n_tables;
    s->t.resize(n_tables);
    s->phase.assign(n_tables, -1);
    s->blocked.assign(n_tables, 0);
    s->in_tick.assign(n_tables, 0);
    s->seen.assign(n_tables, 0);
    s->tick_ins.assign(n_tables, 0);
    s->mem_head.assign(n_tables, 0);
    s->round.assign(n_tables, 0);
    s->had_erase.assign(n_tables, 0);
    s->stage.assign(n_tables, -1);
    s->last_sid.assign(n_tables, ~0u);
    s->nstages.assign(n_tables, 0);

    HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    uint64_t budget = s->scfg.hbm_budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        budget = fr / 2;
    }
    if (hipMalloc(&s->base, budget) != hipSuccess) {
        s->base = nullptr;
        (void)hipGetLastError();
        return fail(DYC_ERR_OOM, "hipMalloc of the %llu-byte table arena failed", (unsigned long long)budget);
    }
    s->arena.reset(budget);
    auto dev_alloc = [](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            return fail(DYC_ERR_OOM, "hipMalloc of %zu bytes failed", bytes);
        }
        return DYC_OK;
    };
    int rc = dev_alloc((void **)&s->d_dir, std::max<size_t>(1, n_tables) * sizeof(Desc));
    if (!rc) rc = dev_alloc((void **)&s->d_stat, s->stat_bytes());
    if (!rc) rc = dev_alloc((void **)&s->d_pool, s->scfg.pool_cap * sizeof(uint64_t));
    if (!rc) rc = dev_alloc((void **)&s->d_pool_tag, s->scfg.pool_cap * sizeof(uint32_t));
    if (!rc) rc = dev_alloc((void **)&s->d_resolve, s->scfg.pool_cap * sizeof(uint64_t));

    if (rc) return rc;
    HIP_TRY(launch_zero(s->d_stat, s->stat_bytes() / sizeof(unsigned long long), s->st));

    const uint64_t items_cap = s->scfg.tick_max_keys / kMinItemKeys + s->scfg.tick_max_requests + 1;
    for (uint32_t i = 0; i < 2; ++i) {
        Tick &k = s->ticks[i];
        k.items_cap = items_cap;
        HIP_TRY(hipHostMalloc((void **)&k.h_items, items_cap * sizeof(ServeItem), hipHostMallocDefault));
        rc = dev_alloc((void **)&k.d_items, items_cap * sizeof(ServeItem));
        if (rc) return rc;
        HIP_TRY(hipHostMalloc((void **)&k.h_dir, std::max<size_t>(1, n_tables) * sizeof(Desc), hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void **)&k.h_snap, s->stat_bytes(), hipHostMallocDefault));
        rc = dev_alloc((void **)&k.d_rctr, kCounterBytes);
        if (rc) return rc;
        // cursor, items, then per stage (at most one per item): left, next, padded ready flag
        const size_t block_bytes =
            (1 + items_cap * sizeof(TickItem) / 8 + 2 * ((items_cap + 1) / 2) + items_cap * kReadyStride / 2) * 8;
        HIP_TRY(hipHostMalloc((void **)&k.h_block, block_bytes, hipHostMallocDefault));
        rc = dev_alloc((void **)&k.d_block, block_bytes);
        if (rc) return rc;
        HIP_TRY(hipEventCreateWithFlags(&k.done, hipEventDisableTiming));
    }
    for (uint32_t i = 0; i < n_tables; ++i) {
        const uint32_t d = s->cfg.subtables;
        const uint32_t nb = initial_buckets(hints[i], d);
        if (nb == 0 || nb >= kMaxBuckets) return fail(DYC_ERR_INVALID, "table %u: capacity hint out of range", i);
        SrvTable &T = s->t[i];
        desc_init(T.D, d);
        for (int sub = 0; sub < (int)d; ++sub) {
            const uint64_t off = s->arena.alloc((uint64_t)nb * kBucketBytes);
            if (off == Arena::kNone)
                return fail(DYC_ERR_OOM, "table %u does not fit in the %llu-byte HBM budget", i,
                            (unsigned long long)budget);
            T.off[sub] = off;
            desc_set(T.D, sub, s->sub_ptr(off), nb);
            HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(s->sub_ptr(off)), (uint64_t)nb * kSlots, s->st));
        }
    }
    s->dir_dirty = true;
    HIP_TRY(hipStreamSynchronize(s->st));
    return DYC_OK;
}

}  // namespace

extern "C" {

int dyc_server_config_init(dyc_server_config *scfg)
{
    if (!scfg) return fail(DYC_ERR_INVALID, "null config");
    memset(scfg, 0, sizeof *scfg);
    scfg->hbm_budget = 0;
    scfg->tick_max_keys = 1ull << 23;
    scfg->tick_max_requests = 4096;
    scfg->pipeline_depth = 2;
    scfg->item_keys = 0;
    scfg->pool_cap = 65536;
    scfg->one_launch = 1;
    scfg->spin_cap = 1u << 24;
    scfg->poll_backoff = 16;
    scfg->resident = 0;
    scfg->resident_blocks = 0;
    scfg->resident_idle_ms = 100;
    return DYC_OK;
}

int dyc_server_create(const dyc_config *cfg, const dyc_server_config *scfg, uint32_t n_tables,
                      const uint64_t *capacity_hints, dyc_server **out)
{
    if (!out || !capacity_hints || n_tables == 0) return fail(DYC_ERR_INVALID, "null argument or no tables");
    *out = nullptr;
    dyc_config c;
    if (cfg) c = *cfg;
    else dyc_config_init(&c);
    dyc_server_config sc;
    if (scfg) sc = *scfg;
    else dyc_server_config_init(&sc);
    if (sc.pipeline_depth < 1 || sc.pipeline_depth > 2) return fail(DYC_ERR_INVALID, "pipeline_depth must be 1 or 2");
    if (sc.tick_max_keys == 0 || sc.tick_max_requests == 0 || sc.pool_cap == 0 || sc.spin_cap == 0)
        return fail(DYC_ERR_INVALID, "tick_max_keys, tick_max_requests, pool_cap and spin_cap must be positive");
    if (sc.item_keys && (sc.item_keys % 64)) return fail(DYC_ERR_INVALID, "item_keys must be a multiple of 64");
    const bool lanes_ok = (c.find_lanes == 4 || c.find_lanes == 8 || c.find_lanes == 16) &&
                          (c.insert_lanes == 1 || c.insert_lanes == 2 || c.insert_lanes == 4 ||
                           c.insert_lanes == 8 || c.insert_lanes == 16);
    if (!lanes_ok) return fail(DYC_ERR_INVALID, "lanes per key: find 4/8/16, insert 1/2/4/8/16");
    c.subtables = dyc::host::config_subtables(c);
    if (c.subtables < (uint32_t)kMinSubtables || c.subtables > (uint32_t)kMaxSubtables)
        return fail(DYC_ERR_INVALID, "subtables must be in [%d, %d]", kMinSubtables, kMaxSubtables);
    if (c.block_threads == 0 || c.block_threads % 64 || c.block_threads > 256)
        return fail(DYC_ERR_INVALID, "block_threads must be a multiple of 64 in [64, 256]");
    if (!(c.upper_bound > 0.0 && c.upper_bound <= 1.0) || !(c.lower_bound >= 0.0) || c.lower_bound >= c.upper_bound)
        return fail(DYC_ERR_INVALID, "need 0 <= lower_bound < upper_bound <= 1");
    if (sc.resident > 1) return fail(DYC_ERR_INVALID, "resident must be 0 or 1");
    if (sc.resident && !serve_tick_supported((int)c.insert_lanes, (int)c.find_lanes))
        return fail(DYC_ERR_INVALID, "the resident loop needs insert_lanes 1, 2 or 4");
    dyc_server *s = new dyc_server();
    s->cfg = c;
    s->scfg = sc;
    HIP_TRY(hipGetDevice(&s->device));
    int rc = create(s, n_tables, capacity_hints);
    if (!rc && sc.resident) rc = set_loop(s, true);
    if (rc) {
        destroy(s);
        return rc;
    }
    *out = s;
    return DYC_OK;
}

int dyc_server_destroy(dyc_server *s)
{
    if (!s) return DYC_OK;
    DeviceGuard g(s->device);
    destroy(s);
    return DYC_OK;
}

int dyc_server_submit(dyc_server *s, const dyc_request *req, uint64_t *ticket)
{
    if (!s || !req) return fail(DYC_ERR_INVALID, "null argument");
    int rc = check_request(s, *req);
    if (rc) return rc;
    Req R;
    R.r = *req;
    s->reqs.push_back(R);
    const uint64_t tk = s->reqs.size() - 1;
    s->queue.push_back(tk);
    s->ss.submitted++;
    if (ticket) *ticket = tk;
    return DYC_OK;
}

int dyc_server_poll(dyc_server *s, uint64_t *done_tickets, int *done_status, uint32_t cap, uint32_t *n_done)
{
    if (!s) return fail(DYC_ERR_INVALID, "null server");
    DeviceGuard g(s->device);
    int rc = progress(s);
    uint32_t k = 0;
    while (k < cap && !s->done.empty()) {
        if (done_tickets) done_tickets[k] = s->done.front().first;
        if (done_status) done_status[k] = s->done.front().second;
        s->done.pop_front();
        ++k;
    }
    if (n_done) *n_done = k;
    return rc;
}

int dyc_server_drain(dyc_server *s)
{
    if (!s) return fail(DYC_ERR_INVALID, "null server");
    DeviceGuard g(s->device);
    int rc = drain_any(s);
    if (rc) return rc;
    return s->error;
}

int dyc_server_replay(dyc_server *s, const dyc_request *reqs, uint64_t n, const double *arrival_s, uint32_t mode,
                      double *done_s, int *status)
{
    if (!s || (n && (!reqs || !arrival_s || !done_s))) return fail(DYC_ERR_INVALID, "null argument");
    if (mode > 2) return fail(DYC_ERR_INVALID, "mode must be 0 (ticks), 1 (synchronous) or 2 (resident loop)");
    if (mode == 2 && !serve_tick_supported((int)s->cfg.insert_lanes, (int)s->cfg.find_lanes))
        return fail(DYC_ERR_INVALID, "the resident loop needs insert_lanes 1, 2 or 4");
    for (uint64_t i = 0; i < n; ++i) {
        int rc = check_request(s, reqs[i]);
        if (rc) return rc;
        if (i && arrival_s[i] < arrival_s[i - 1]) return fail(DYC_ERR_INVALID, "arrival times must not decrease");
    }
    DeviceGuard g(s->device);
    int rc = drain_any(s);  // start from a quiet server
    if (rc) return rc;
    rc = set_loop(s, mode == 2);
    if (rc) return rc;
    s->done.clear();
    const uint64_t first = s->reqs.size();
    const Clock::time_point t0 = Clock::now();
    if (mode == 1) {
        for (uint64_t i = 0; i < n; ++i) {
            while (seconds_since(t0) < arrival_s[i]) {
            }
            Req R;
            R.r = reqs[i];
            s->reqs.push_back(R);
            s->ss.submitted++;
            rc = run_sync(s, first + i);
            if (rc) return sticky(s, rc);
            done_s[i] = seconds_since(t0);
            if (status) status[i] = s->reqs[first + i].status;
        }
        s->done.clear();
        return set_loop(s, s->scfg.resident != 0);
    }
    uint64_t next = 0, finished = 0;
    double last_done = 0.0;  // nothing finished or arrived for kLoopGiveUpS with work outstanding: give up
    while (finished < n) {
        double now = seconds_since(t0);
        if (next > finished && now - std::max(last_done, arrival_s[next - 1]) > kLoopGiveUpS) {
            if (s->loop.on) __atomic_store_n(&s->loop.ctl->stop, 2ull, __ATOMIC_RELEASE);
            return sticky(s, fail(DYC_ERR_HIP, "replay: no request finished in %.0f s", kLoopGiveUpS));
        }
        while (next < n && arrival_s[next] <= now) {
            Req R;
            R.r = reqs[next];
            s->reqs.push_back(R);
            s->queue.push_back(first + next);
            s->ss.submitted++;
            ++next;
        }
        rc = progress(s);
        if (rc) return sticky(s, rc);
        if (!s->done.empty()) {
            now = seconds_since(t0);
            while (!s->done.empty()) {
                const uint64_t tk = s->done.front().first;
                if (tk >= first && tk < first + n) {
                    done_s[tk - first] = now;
                    if (status) status[tk - first] = s->done.front().second;
                    ++finished;
                    last_done = now;
                }
                s->done.pop_front();
            }
        }
    }
    return set_loop(s, s->scfg.resident != 0);
}

int dyc_server_get_stats(const dyc_server *s, dyc_server_stats *out)
{
    if (!s || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = s->ss;
    out->pool_entries = s->pool_settled;
    out->arena_bytes = s->arena.size();
    out->arena_used = s->arena.used();
    out->arena_peak = s->arena.peak();
    out->queued = s->queue.size();
    uint64_t inflight = s->loop.on ? s->loop.in_flight() : 0;
    for (uint32_t i = 0; i < s->depth(); ++i)
        if (s->ticks[i].busy) inflight += s->ticks[i].reqs.size();
    out->in_flight = inflight;
    return DYC_OK;
}

int dyc_server_table_stats(const dyc_server *s, uint32_t table, dyc_stats *out)
{
    if (!s || !out) return fail(DYC_ERR_INVALID, "null argument");
    if (table >= s->n) return fail(DYC_ERR_INVALID, "table %u out of range", table);
    const SrvTable &T = s->t[table];
    memset(out, 0, sizeof *out);
    out->count = T.count;
    out->capacity = T.capacity();
    for (uint32_t i = 0; i < T.D.d; ++i) out->buckets[i] = T.D.nb[i];
    out->subtables = T.D.d;
    out->resizes_up = T.resizes_up;
    out->resizes_down = T.resizes_down;
    return DYC_OK;
}

int dyc_server_dump(dyc_server *s, uint32_t table, uint32_t sub, uint64_t *host_slots, uint64_t n_slots)
{
    if (!s || !host_slots) return fail(DYC_ERR_INVALID, "null argument");
    if (table >= s->n || sub >= s->t[table].D.d) return fail(DYC_ERR_INVALID, "table or subtable out of range");
    DeviceGuard g(s->device);
    int rc = drain_any(s);
    if (rc) return rc;
    const SrvTable &T = s->t[table];
    const uint64_t slots = (uint64_t)T.D.nb[sub] * kSlots;
    if (n_slots < slots) return fail(DYC_ERR_INVALID, "dump needs %llu slots", (unsigned long long)slots);
    HIP_TRY(hipMemcpy(host_slots, T.D.sub[sub], slots * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return DYC_OK;
}

}  // extern "C"
