// dyc_wide.hip -- DyCuckoo with 64-bit keys and 64-bit values.
//
// The paper's wide-key layout (tkde-revision/secs/dynamic.tex:35-37): for keys larger than
// 4 bytes a bucket holds fewer pairs, and keys and values are stored apart. A bucket is one
// 128-byte line of 16 u64 keys, plus the parallel 128-byte line of their 16 u64 values in the
// subtable's value array. A lookup reads the key line and, on a hit, 8 bytes of the value line.
//
// Placement is the narrow table's applied to fold(k) = lo32(k) ^ h5(hi32(k)): the same
// geometry, pair scheme of every subtable count, start rule, victim vote and rescue walk.
// fold(k) = k for keys below 2^32, so a wide table of such keys places them exactly where the
// narrow table does. Keys 0 (empty) and 2^64 - 1 (a slot being rewritten) are reserved.
//
// A 16-byte entry cannot move with one 64-bit CAS, so an exchange claims the slot instead
// (the paper locks the bucket, dynamic.tex:37-38):
//   1. CAS the slot's key from the value the lane read to BUSY (a lost race patches the view);
//   2. write the slot's value: an agent-scope atomic store for an empty slot, an exchange for a
//      victim's (its old value travels with the evicted key);
//   3. wait for that write to be performed, then publish the key with an agent-scope atomic
//      store (BUSY -> the new key).
// Every step is coherent at agent scope (memory-side atomics, write-through stores), so no cache
// maintenance is needed between them; a slot holding BUSY is never chosen as a victim by another
// lane, and no lane ever waits on one, so nothing can deadlock. Finds, erases and updates run in
// separate launches.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "dyc.h"
#include "dyc_bulk_layout.hpp"
#include "dyc_host.hpp"
#include "dyc_part.hpp"
#include "dyc_wave.hpp"

namespace dyc {

constexpr uint64_t kWideBusy = ~0ull;

__host__ __device__ inline uint32_t fold64(uint64_t k) { return (uint32_t)k ^ h5((uint32_t)(k >> 32)); }
__host__ __device__ inline bool wide_reserved(uint64_t k) { return k == 0ull || k == kWideBusy; }

// Where an insert parks entries that exhaust the eviction bound (append index Counters::pending).
// Groups: tags (optional) receives the table index of each parked entry (`tag` for a
// single-table launch, the item's table for a group launch), and rejected_at (optional) counts
// the launch's reserved keys.
struct WideSink {
    uint64_t *keys, *vals;
    uint64_t cap;
    uint32_t *tags = nullptr;
    uint32_t tag = 0;
    unsigned long long *rejected_at = nullptr;
};

// The pending entries searched after pos2 by finds, erases and updates (n = 0: none). tags
// (optional, groups): only entries tagged `table` belong to the table being probed.
struct WideStash {
    uint64_t *keys;
    uint64_t *vals;
    uint64_t n;
    const uint32_t *tags = nullptr;
    uint32_t table = 0;
};

namespace {

__device__ inline uint32_t wide_loc(uint32_t t, uint32_t bucket) { return (t << 29) | bucket; }
__device__ inline uint64_t *wide_keys_at(const WideDesc &W, uint32_t l)
{
    return selt(l >> 29, W.D.sub) + (uint64_t)(l & 0x1fffffffu) * kSlots;
}
__device__ inline uint64_t *wide_vals_at(const WideDesc &W, uint32_t l)
{
    return selt(l >> 29, W.val) + (uint64_t)(l & 0x1fffffffu) * kSlots;
}

__device__ inline void load2(const uint64_t *p, uint64_t (&e)[2])
{
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    e[0] = ((uint64_t)v.y << 32) | v.x;
    e[1] = ((uint64_t)v.w << 32) | v.z;
}

__device__ inline void store2(uint64_t *p, uint64_t a, uint64_t b)
{
    *reinterpret_cast<uint4 *>(p) = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
}

// --------------------------------------------------------------------- find / erase / update
// A wave owns 64 consecutive keys; groups of 8 lanes probe 8 keys each, lane s reading keys
// 2s, 2s+1 of a bucket (one 128-byte line per group). The first bucket (pos1, then pos2) that
// holds a key answers: find (MODE 0) takes the first matching slot's value (read by the key's
// own lane from the value line), erase (MODE 1) CASes every matching key of that bucket to 0,
// update (MODE 2) writes vals[i] to the value of every matching slot of that bucket. Keys
// neither bucket holds are then looked up in the stash (erase and update act on every copy
// there). Update: keys of one launch are distinct, so no two lanes write one value; the key word
// is untouched, and the value is an agent-scope atomic store.
// One wave probes keys [base, base + 64) clipped to n; returns this lane's flag (found, deleted
// or updated), also written to flags. All 64 lanes call it together.
template <int MODE>
__device__ inline uint32_t wide_probe_tile(const WideDesc &W, const uint64_t *__restrict__ keys,
                                           uint64_t *__restrict__ vals, uint8_t *__restrict__ flags, uint64_t base,
                                           uint64_t n, const WideStash &stash)
{
    constexpr bool ERASE = MODE == 1;
    constexpr bool UPDATE = MODE == 2;
    constexpr int G = 8;
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    const uint64_t i = base + lane;
    const bool inb = i < n;
    const uint64_t k = inb ? keys[i] : 0ull;
    const uint64_t nv = (UPDATE && inb) ? vals[i] : 0ull;  // update: the new value
    const uint32_t f = fold64(k);
    const uint32_t p = pair_of(W.D, f);
    const uint32_t t1 = pos1(W.D, p), t2 = pos2(W.D, p);
    const uint32_t loc1 = wide_loc(t1, bucket_of(W.D, f, t1));
    const uint32_t loc2 = wide_loc(t2, bucket_of(W.D, f, t2));

    uint64_t e[G][2];
    uint64_t kk_key[G];
    uint64_t kk_new[UPDATE ? G : 1];
    bool kk_valid[G];
    uint32_t kk_l[G];
    uint32_t my_flag = 0, my_loc = 0, my_slot = 0, done = 0;
#pragma unroll
    for (int phase = 0; phase < 2; ++phase) {
        const uint32_t lp = phase == 0 ? loc1 : loc2;
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            const int src = g * G + kk;
            if (phase == 0) {
                kk_key[kk] = shfl64(k, src);
                kk_valid[kk] = !wide_reserved(kk_key[kk]);
                if constexpr (UPDATE) kk_new[kk] = shfl64(nv, src);
            }
            kk_l[kk] = __shfl(lp, src);
            e[kk][0] = e[kk][1] = 0;
            if (kk_valid[kk] && !(done & (1u << kk))) load2(wide_keys_at(W, kk_l[kk]) + s * 2, e[kk]);
        }
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            if (done & (1u << kk)) continue;  // group-uniform
            const int first = e[kk][0] == kk_key[kk] ? 0 : (e[kk][1] == kk_key[kk] ? 1 : -1);
            const uint32_t gm = (uint32_t)((__ballot(first >= 0 && kk_valid[kk]) >> (g * G)) & 0xffu);
            if (!gm) continue;
            done |= 1u << kk;
            if constexpr (ERASE) {
                int ok = 0;
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (kk_valid[kk] && e[kk][j] == kk_key[kk]) {
                        uint64_t exp = e[kk][j];
                        ok |= cas_slot(wide_keys_at(W, kk_l[kk]) + s * 2 + j, exp, 0ull);
                    }
                const uint32_t okm = (uint32_t)((__ballot(ok) >> (g * G)) & 0xffu);
                if (kk == s) my_flag = okm != 0;
            } else if constexpr (UPDATE) {
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (kk_valid[kk] && e[kk][j] == kk_key[kk])
                        __hip_atomic_store(wide_vals_at(W, kk_l[kk]) + s * 2 + j, kk_new[UPDATE ? kk : 0],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (kk == s) my_flag = 1;
            } else {
                const int w = __builtin_ctz(gm);
                const int fj = __shfl(first, g * G + w);
                if (kk == s) {
                    my_flag = 1;
                    my_loc = kk_l[kk];
                    my_slot = (uint32_t)(w * 2 + fj);
                }
            }
        }
    }
    uint64_t my_val = 0;
    if (MODE == 0 && my_flag) my_val = wide_vals_at(W, my_loc)[my_slot];
    if (inb && !my_flag && stash.n && !wide_reserved(k)) {  // the stash counts as one more bucket
        for (uint64_t j = 0; j < stash.n; ++j) {
            const uint64_t sk = __hip_atomic_load(stash.keys + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (sk != k || (stash.tags && stash.tags[j] != stash.table)) continue;
            if constexpr (ERASE) {
                uint64_t exp = sk;
                my_flag |= (uint32_t)cas_slot(stash.keys + j, exp, 0ull);
            } else if constexpr (UPDATE) {
                __hip_atomic_store(stash.vals + j, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                my_flag = 1;
            } else {
                my_val = stash.vals[j];
                my_flag = 1;
                break;
            }
        }
    }
    if (inb) {
        if (MODE == 0 && vals) vals[i] = my_val;
        if (flags) flags[i] = (uint8_t)my_flag;
    }
    return my_flag;
}

template <int MODE>
__global__ __launch_bounds__(256) void wide_probe_kernel(WideDesc W, const uint64_t *__restrict__ keys,
                                                         uint64_t *__restrict__ vals, uint8_t *__restrict__ flags,
                                                         uint64_t n, Counters *__restrict__ C, WideStash stash)
{
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    unsigned long long ndel = 0;
    for (uint64_t base = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kWave; base < n;
         base += waves * kWave)
        ndel += wide_probe_tile<MODE>(W, keys, vals, flags, base, n, stash);
    if constexpr (MODE == 1) wave_tally(C, &C->deleted, ndel);
    if constexpr (MODE == 2) wave_tally(C, &C->updated, ndel);
}

// ---------------------------------------------------------------------------------- insert
// One lane per in-hand entry; a lane that finishes its chain takes the next entry of its
// grid-stride sequence, so it never waits behind a longer chain in its wave. The chain is the
// narrow table's (victim = first empty slot; else, for kick_bound iterations, the first slot
// whose resident's pair index differs from the table number; else slot iteration & 15; then
// the rescue walk), over fold(k), with the three-step exchange described at the top.
struct WideCarry {
    uint64_t key, val;
    uint32_t pre, kicks, retries;
    bool active, fresh;
    uint64_t cur[kSlots];  // the lane's view of the bucket it is about to write
};

__device__ inline void wide_give_up(WideCarry &c, const WideSink &sink, Counters *C, uint32_t tag = 0)
{
    const unsigned long long at = atomicAdd(&C->pending, 1ull);
    if (at < sink.cap) {
        sink.keys[at] = c.key;
        sink.vals[at] = c.val;
        if (sink.tags) sink.tags[at] = tag;
    } else {
        atomicAdd(&C->lost, 1ull);
    }
    c.active = false;
}

// Take entry (k, v) into the lane's hand; pr is its previous table (kNoTable for a fresh key).
__device__ inline void wide_take(WideCarry &c, const Desc &D, uint64_t k, uint64_t v, uint32_t pr)
{
    if (pr == kNoTable && (k & 1u)) pr = pos1(D, pair_of(D, fold64(k)));  // odd keys start in pos2
    c.key = k;
    c.val = v;
    c.pre = pr;
    c.kicks = 0;
    c.retries = 0;
    c.fresh = true;
    c.active = true;
}

// One iteration of an active lane's eviction chain (see the top of this file); finished chains
// add their exchanges to kicks.
__device__ inline void wide_step(WideCarry &c, const WideDesc &W, const InsertPolicy &P, Counters *C,
                                 const WideSink &sink, uint32_t tag, unsigned long long &kicks,
                                 unsigned long long &retries)
{
    const Desc &D = W.D;
    const uint32_t total_kicks = P.kick_bound + P.rescue_kicks;
    const uint32_t f = fold64(c.key);
    const uint32_t t = other_pos(D, pair_of(D, f), c.pre);
    const uint32_t b = bucket_of(D, f, t);
    uint64_t *kp = selt(t, D.sub) + (uint64_t)b * kSlots;
    uint64_t *vp = selt(t, W.val) + (uint64_t)b * kSlots;
    if (c.fresh) {
#pragma unroll
        for (int j = 0; j < kSlots; j += 2) {
            uint64_t e[2];
            load2(kp + j, e);
            c.cur[j] = e[0];
            c.cur[j + 1] = e[1];
        }
        c.fresh = false;
    }
    int slot = -1;
#pragma unroll
    for (int j = kSlots - 1; j >= 0; --j)
        if (c.cur[j] == 0ull) slot = j;
    if (slot < 0) {
        if (c.kicks < P.kick_bound) {
#pragma unroll
            for (int j = kSlots - 1; j >= 0; --j)
                if (c.cur[j] != kWideBusy && vote_victim(D, fold64(c.cur[j]), t)) slot = j;
            if (slot < 0) slot = (int)(c.kicks & 15u);
        } else {
            slot = (int)rescue_lane(f, c.kicks);
        }
    }
    uint64_t expected = 0;
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
        if (j == slot) expected = c.cur[j];
    // claim the slot; on a lost race (or a slot another lane is rewriting, refreshed with a
    // BUSY -> BUSY CAS, which changes nothing) the slot's current key replaces the view's
    uint64_t exp = expected;
    const bool claimed = expected != kWideBusy ? cas_slot(kp + slot, exp, kWideBusy) : false;
    if (expected == kWideBusy) (void)cas_slot(kp + slot, exp, kWideBusy);
    if (!claimed) {
#pragma unroll
        for (int j = 0; j < kSlots; ++j)
            if (j == slot) c.cur[j] = exp;
        ++retries;
        if (++c.retries > P.max_retries) {
            kicks += c.kicks;
            wide_give_up(c, sink, C, tag);
        }
        return;
    }
    // The claimed slot's value belongs to this lane until the key is published: an empty
    // slot's value is written with an atomic store, a victim's is exchanged (its value goes
    // on in hand). Agent-scope atomic stores cost less than exchanges (random 8-byte stores
    // 24 G/s against 18 G/s CAS on MI355X) and stay coherent with the CAS on the same slot.
    uint64_t old_val = 0;
    if (expected == 0ull)
        __hip_atomic_store(vp + slot, c.val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        old_val = __hip_atomic_exchange(vp + slot, c.val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // wait until the value write is performed before publishing the key; the memory clobber
    // also keeps the compiler from moving the publish above the wait
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store(kp + slot, c.key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ++c.kicks;
    c.fresh = true;
    if (expected == 0ull) {
        kicks += c.kicks;
        c.active = false;
    } else {
        c.key = expected;
        c.val = old_val;
        c.pre = t;
        if (c.kicks >= total_kicks) {
            kicks += c.kicks;
            wide_give_up(c, sink, C, tag);
        }
    }
}

// skip (optional): a keys insert leaves out key i where skip[i] != 0 (an upsert updated it in
// place); an entries insert reads it as the tile flags of a slot-addressed residual list.
template <bool FROM_ENTRIES>
__global__ __launch_bounds__(256) void wide_insert_kernel(WideDesc W, const uint64_t *__restrict__ keys,
                                                          const uint64_t *__restrict__ vals, uint64_t n,
                                                          uint32_t pre_fixed, InsertPolicy P, Counters *__restrict__ C,
                                                          WideSink sink, const uint8_t *__restrict__ skip)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t next = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long kicks = 0, rejected = 0, retries = 0;
    WideCarry c;
    c.active = false;
    for (;;) {
        while (!c.active && next < n) {  // take the next entry
            if (FROM_ENTRIES && skip && !skip[next / kResTileEntries]) {  // a residual tile with none
                next += stride;
                continue;
            }
            if (!FROM_ENTRIES && skip && skip[next]) {  // updated in place
                next += stride;
                continue;
            }
            const uint64_t k = keys[next];
            const uint64_t v = vals[next];
            next += stride;
            if (FROM_ENTRIES && k == 0) continue;  // an empty place in a residual or pending list
            if (wide_reserved(k)) {
                ++rejected;
                continue;
            }
            wide_take(c, W.D, k, v, FROM_ENTRIES ? pre_fixed : kNoTable);
        }
        if (!__any(c.active)) break;
        if (c.active) wide_step(c, W, P, C, sink, sink.tag, kicks, retries);
    }
    if (sink.rejected_at) {
        const unsigned long long rej = wave_sum(rejected);
        if (__lane_id() == 0 && rej) atomicAdd(sink.rejected_at, rej);
    }
    wave_tally(C, &C->kicks, kicks);
    wave_tally(C, &C->rejected, rejected);
    wave_tally(C, &C->retries, retries);
}

// ------------------------------------------------------------------------------ many tables
// The group forms: one persistent launch over work items of one table each (GroupItem), the
// table's descriptor read from a directory, the per-tile and per-chain code above unchanged.
// Pending entries carry their table in sink.tags / stash.tags.
template <int MODE>
__global__ __launch_bounds__(256) void wide_group_probe_kernel(const WideDesc *__restrict__ dir,
                                                               const GroupItem *__restrict__ items, uint32_t n_items,
                                                               const uint64_t *__restrict__ keys,
                                                               uint64_t *__restrict__ vals,
                                                               uint8_t *__restrict__ flags,
                                                               unsigned long long *__restrict__ table_counts,
                                                               WideStash stash)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t it = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
         it < n_items; it += waves) {
        const GroupItem item = items[it];
        if (item.len == 0) continue;  // wave-uniform: an empty slot of a layout by XCD
        const WideDesc W = dir[item.table];
        WideStash st = stash;
        st.table = item.table;
        const uint64_t end = item.begin + item.len;
        unsigned long long nd = 0;
        for (uint64_t base = item.begin; base < end; base += kWave)
            nd += wide_probe_tile<MODE>(W, keys, vals, flags, base, end, st);
        if constexpr (MODE == 1) {
            nd = wave_sum(nd);
            if (__lane_id() == 0 && nd) atomicAdd(&table_counts[item.table], nd);
        }
    }
}

// Lanes of a wave take the item's entries one each (lane-per-chain, as wide_insert_kernel); the
// wave moves to its next item when the item is used up and every lane's chain has ended.
__global__ __launch_bounds__(256) void wide_group_insert_kernel(const WideDesc *__restrict__ dir,
                                                                const GroupItem *__restrict__ items, uint32_t n_items,
                                                                const uint64_t *__restrict__ keys,
                                                                const uint64_t *__restrict__ vals, InsertPolicy P,
                                                                Counters *__restrict__ C, WideSink sink,
                                                                unsigned long long *__restrict__ table_counts)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    unsigned long long kicks = 0, rejected = 0, retries = 0;
    WideCarry c;
    c.active = false;
    for (uint32_t it = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
         it < n_items; it += waves) {
        const GroupItem item = items[it];
        if (item.len == 0) continue;  // wave-uniform: a block the bulk fill's rounds emptied
        const WideDesc W = dir[item.table];
        const uint64_t end = item.begin + item.len;
        uint64_t next = item.begin + __lane_id();
        unsigned long long rej = 0;
        for (;;) {
            while (!c.active && next < end) {
                const uint64_t k = keys[next];
                const uint64_t v = vals[next];
                next += kWave;
                if (wide_reserved(k)) {
                    ++rej;
                    continue;
                }
                wide_take(c, W.D, k, v, kNoTable);
            }
            if (!__any(c.active)) break;
            if (c.active) wide_step(c, W, P, C, sink, item.table, kicks, retries);
        }
        rejected += rej;
        rej = wave_sum(rej);
        if (__lane_id() == 0 && item.len > rej) atomicAdd(&table_counts[item.table], item.len - rej);
    }
    wave_tally(C, &C->kicks, kicks);
    wave_tally(C, &C->rejected, rejected);
    wave_tally(C, &C->retries, retries);
}

// ---------------------------------------------------------------------------------- resize
// Upsize (cuckoo.cu:600-627): key and value of slot i of old bucket j move to slot i of new
// bucket h % 2n, which is j or j + n; each 8-lane group writes both destination buckets of
// both arrays in full (entry or zero).
__global__ __launch_bounds__(256) void wide_upsize_kernel(const uint64_t *__restrict__ ok_, const uint64_t *__restrict__ ov,
                                                          uint32_t old_nb, uint64_t *__restrict__ dk,
                                                          uint64_t *__restrict__ dv, uint32_t t, uint32_t m,
                                                          uint32_t s1, uint32_t s2, Counters *__restrict__ C)
{
    constexpr int G = 8, NG = kWave / G;
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint32_t nnb = old_nb * 2u;
    unsigned long long bad = 0;
    for (uint64_t tile = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * NG; tile < old_nb;
         tile += waves * NG) {
        const uint64_t j = tile + g;
        if (j >= old_nb) continue;
        uint64_t k[2], v[2], lk[2] = {0, 0}, lv[2] = {0, 0}, hk[2] = {0, 0}, hv[2] = {0, 0};
        load2(ok_ + j * kSlots + s * 2, k);
        load2(ov + j * kSlots + s * 2, v);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!k[q]) continue;
            const uint32_t d = fast_mod(hash_t(fold64(k[q]), t), nnb, m, s1, s2);
            if (d == j) {
                lk[q] = k[q];
                lv[q] = v[q];
            } else if (d == j + old_nb) {
                hk[q] = k[q];
                hv[q] = v[q];
            } else {
                ++bad;
            }
        }
        store2(dk + j * kSlots + s * 2, lk[0], lk[1]);
        store2(dv + j * kSlots + s * 2, lv[0], lv[1]);
        store2(dk + (j + old_nb) * kSlots + s * 2, hk[0], hk[1]);
        store2(dv + (j + old_nb) * kSlots + s * 2, hv[0], hv[1]);
    }
    wave_tally(C, &C->misplaced, bad);
}

// Downsize compaction (cuckoo.cu:664-699): new bucket g keeps old bucket g slot for slot and
// fills its empty slots from old bucket g + n read back to front (slot i <- slot 15 - i); an
// entry of bucket g + n is a residual unless it is the one placed there (the reference compares
// the placed entry with it, cuckoo.cu:685-688), here key and value both. Residuals are
// slot-addressed (slot i of new bucket g's partner -> rk/rv[g * 16 + i], key 0 where none), as
// in the 32-bit table, so no wave reserves list space through a single append index; live[tile]
// flags the tiles of kResTileBuckets buckets that left any, and only those are written.
__global__ __launch_bounds__(256) void wide_downsize_kernel(const uint64_t *__restrict__ ok_,
                                                            const uint64_t *__restrict__ ov, uint32_t new_nb,
                                                            uint64_t *__restrict__ dk, uint64_t *__restrict__ dv,
                                                            Counters *__restrict__ C, uint64_t *__restrict__ rk,
                                                            uint64_t *__restrict__ rv, uint8_t *__restrict__ live)
{
    constexpr int G = 8, NG = kWave / G;
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint32_t slot = wave_slot();
    unsigned long long nres = 0;
    for (uint64_t tile = (uint64_t)slot * NG; tile < new_nb; tile += waves * NG) {
        const uint64_t j = tile + g;
        const bool ok = j < new_nb;
        uint64_t ek[2] = {0, 0}, ev[2] = {0, 0}, ck[2] = {0, 0}, cv[2] = {0, 0};
        if (ok) {
            load2(ok_ + j * kSlots + s * 2, ek);
            load2(ov + j * kSlots + s * 2, ev);
            uint64_t xk[2], xv[2];
            load2(ok_ + (j + new_nb) * kSlots + (7 - s) * 2, xk);
            load2(ov + (j + new_nb) * kSlots + (7 - s) * 2, xv);
            ck[0] = xk[1];  // slot 15 - 2s
            cv[0] = xv[1];
            ck[1] = xk[0];  // slot 14 - 2s
            cv[1] = xv[0];
        }
        uint64_t pk[2], pv[2], rk_[2], rv_[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            pk[q] = ek[q] ? ek[q] : ck[q];
            pv[q] = ek[q] ? ev[q] : cv[q];
            const bool res = ck[q] != 0 && !(pk[q] == ck[q] && pv[q] == cv[q]);
            rk_[q] = res ? ck[q] : 0;
            rv_[q] = res ? cv[q] : 0;
        }
        if (ok) {
            store2(dk + j * kSlots + s * 2, pk[0], pk[1]);
            store2(dv + j * kSlots + s * 2, pv[0], pv[1]);
        }
        const bool any = __ballot(rk_[0] != 0 || rk_[1] != 0) != 0;  // wave-uniform
        if (lane == 0) live[tile / NG] = any ? 1 : 0;
        if (any && ok) {  // slot-addressed residuals, key 0 where none
            store2(rk + j * kSlots + s * 2, rk_[0], rk_[1]);
            store2(rv + j * kSlots + s * 2, rv_[0], rv_[1]);
        }
        nres += (rk_[0] != 0) + (rk_[1] != 0);
    }
    wave_tally(C, &C->residuals, nres, slot);
}

constexpr uint64_t kWideInsertBlocks = 16384;

// ------------------------------------------------------------ bulk fill (insert_order 3)
// The narrow table's bulk fill (dyc_bulk.hip) over fold(k): (chunk id, entry) pairs sorted by
// chunk, one workgroup per chunk filling free slots, a second round in the other bucket, the
// eviction insert for the rest. The dense round keeps the chunk's key lines and value lines in
// LDS (2 x 64 KB), so one 1024-thread workgroup runs per CU.
// Chunk ids of the batch in input order and each block's counts of the partition's first digit
// (dyc_part.hpp), over fold(k): round 1 from the keys (fill_target; reserved keys go to chunk 0,
// whose fill rejects them), round 2 from overflow entries (the other bucket of the pair).
template <bool IN_ENTRIES, typename CidT>
__global__ __launch_bounds__(kCountThreads) void wide_bulk_pack_kernel(WideDesc W, Chunks Ch,
                                                                       const uint64_t *__restrict__ keys,
                                                                       const WideEntry *__restrict__ in, PartGrid G,
                                                                       RadixPass P, CidT *__restrict__ cid,
                                                                       uint32_t *__restrict__ partial)
{
    const Desc &D = W.D;
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
        uint64_t k[kCountItems];
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i < hi) {
                if constexpr (IN_ENTRIES) k[u] = in[i].k;
                else k[u] = keys[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i >= hi) continue;
            uint32_t c = 0;
            if (IN_ENTRIES || !wide_reserved(k[u])) {
                const uint32_t f = fold64(k[u]), p = pair_of(D, f);
                const uint32_t t = tsel[(((uint32_t)k[u] & 1u) << 4) | p];
                const uint4 q = sub_tab[t];
                c = q.w + (fast_mod_sh(hash_t(f, t), q.x, q.y, q.z) >> kChunkShift);
            }
            cid[i] = (CidT)c;
            atomicAdd(&h[digit_of(c, P)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kCountThreads) partial[(uint64_t)blockIdx.x * R + d] = h[d];
}

// The fill of one chunk of a wide table by the whole workgroup: buckets [b0, b0 + len) of
// subtable S (key lines S.keys, value lines S.vals), entries ent[lo, hi) (all targeting S); entries that do not fit go to
// ovf[0, ovf_cap) at the index *ovf_n hands out. Placements and rejected reserved keys are added
// to this lane's `placed` and `rejected`. Every thread of the workgroup calls it alike.
template <bool DENSE, int THREADS>
__device__ inline void wide_fill_chunk(const SubRef &S, uint32_t b0, uint32_t len,
                                       const WideEntry *__restrict__ ent, uint32_t lo, uint32_t hi,
                                       WideEntry *__restrict__ ovf, uint64_t ovf_cap,
                                       unsigned long long *__restrict__ ovf_n, bool zeroed,
                                       unsigned long long &placed, unsigned long long &rejected)
{
    constexpr uint32_t kImg = DENSE ? kChunkBuckets * kSlots : 2;
    __shared__ uint64_t kimg[kImg], vimg[kImg];  // DENSE: the chunk's key and value lines
    __shared__ uint32_t fm[kChunkBuckets];       // free-slot mask of each bucket
    __shared__ uint32_t cnt[kChunkBuckets];      // entries per bucket, then slots taken
    __shared__ uint32_t wred[THREADS / kWave];
    __shared__ unsigned long long obase;
    __shared__ uint32_t onext;
    uint64_t *const ksub = S.keys;
    uint64_t *const vsub = S.vals;
    const uint4 *const ksrc = reinterpret_cast<const uint4 *>(ksub + (uint64_t)b0 * kSlots);
    constexpr uint32_t kParts = kChunkBuckets * kSlots / 2;  // 16-byte parts of a chunk's lines
    if constexpr (DENSE) {
        const uint4 *const vsrc = reinterpret_cast<const uint4 *>(vsub + (uint64_t)b0 * kSlots);
        uint4 *const ki = reinterpret_cast<uint4 *>(kimg);
        uint4 *const vi = reinterpret_cast<uint4 *>(vimg);
        // a bucket's 8 key parts sit in 8 consecutive lanes, whose two ballot bits per part
        // make its free-slot mask (no strided LDS reads)
        static_assert(kParts % THREADS == 0, "every lane takes the same number of parts");
        const uint32_t lane = __lane_id();
        if (zeroed) {
            // every slot of the table is known to be zero (the caller's host state says so):
            // the chunk's lines start empty in LDS without reading them
            for (uint32_t q = threadIdx.x; q < kParts; q += THREADS) {
                ki[q] = make_uint4(0u, 0u, 0u, 0u);
                vi[q] = make_uint4(0u, 0u, 0u, 0u);
            }
            for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += THREADS) {
                fm[j] = j < len ? 0xffffu : 0u;  // past the end: full
                cnt[j] = 0;
            }
        }
        for (uint32_t q = threadIdx.x; q < kParts && !zeroed; q += THREADS) {
            const bool in = q < len * (kSlots / 2);
            const uint4 kv = in ? ksrc[q] : make_uint4(~0u, ~0u, ~0u, ~0u);  // past the end: full
            ki[q] = kv;
            vi[q] = in ? vsrc[q] : make_uint4(0u, 0u, 0u, 0u);
            const uint64_t m0 = __ballot((kv.x | kv.y) == 0u), m1 = __ballot((kv.z | kv.w) == 0u);
            if ((lane & 7u) == 0) {
                const uint32_t a = (uint32_t)(m0 >> lane) & 0xffu, b = (uint32_t)(m1 >> lane) & 0xffu;
                uint32_t m = 0;
#pragma unroll
                for (int p = 0; p < 8; ++p) m |= (((a >> p) & 1u) << (2 * p)) | (((b >> p) & 1u) << (2 * p + 1));
                fm[q / 8] = m;
                cnt[q / 8] = 0;
            }
        }
    } else {
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += blockDim.x) cnt[j] = 0;
    }
    if (threadIdx.x == 0) onext = 0;
    __syncthreads();
    // count: entries per bucket. A chunk with at most kBatch entries per thread keeps them in
    // registers for the placement, each with its bucket (9 bits) and its rank among the
    // chunk's entries of that bucket (what the count's atomic returned), or kReject / kStray;
    // a bucket's first popcount(free mask) ranks take its free slots. Larger chunks count
    // here and count again as they place.
    constexpr uint32_t kBatch = 8;
    constexpr uint32_t kReject = ~0u, kStray = ~0u - 1u;
    static_assert(kChunkBuckets <= 512 && kBatch * THREADS < (1u << 22), "rank and bucket fit 32 bits");
    const bool held = hi - lo <= kBatch * THREADS;  // block-uniform
    WideEntry hb[kBatch];
    uint32_t hp[kBatch];
    uint32_t stray = 0;
    if (held) {
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t i = lo + threadIdx.x + u * THREADS;
            hb[u] = i < hi ? ent[i] : WideEntry{0ull, 0ull};
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            uint32_t at = kReject;  // also past the range (never used)
            if (lo + threadIdx.x + u * THREADS < hi && !wide_reserved(hb[u].k)) {
                const uint32_t b = bucket_of_uniform(S, fold64(hb[u].k));
                if (b >= b0 && b - b0 < len) {
                    at = (atomicAdd(&cnt[b - b0], 1u) << 9) | (b - b0);
                } else {
                    at = kStray;
                    ++stray;
                }
            }
            hp[u] = at;
        }
    } else {
        for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kBatch * THREADS) {
            uint64_t kb[kBatch];
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                const uint32_t i = i0 + u * THREADS;
                kb[u] = i < hi ? ent[i].k : 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < kBatch; ++u) {
                if (i0 + u * THREADS >= hi || wide_reserved(kb[u])) continue;
                const uint32_t b = bucket_of_uniform(S, fold64(kb[u]));
                if (b >= b0 && b - b0 < len) atomicAdd(&cnt[b - b0], 1u);
                else ++stray;
            }
        }
    }
    __syncthreads();
    if constexpr (!DENSE) {
        // free masks of the buckets that received entries only (a bucket's 8 parts are read
        // by 8 consecutive lanes and folded by ballot; the others read as full)
        constexpr uint32_t kPer = kParts / THREADS;
        uint4 v[kPer];
#pragma unroll
        for (uint32_t u = 0; u < kPer; ++u) {
            const uint32_t q = threadIdx.x + u * THREADS;
            v[u] = q < len * (kSlots / 2) && cnt[q / 8] ? ksrc[q] : make_uint4(1u, 1u, 1u, 1u);
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
    for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += THREADS) {
        const uint32_t f = (uint32_t)__builtin_popcount(fm[j]);
        excess += cnt[j] > f ? cnt[j] - f : 0u;
    }
    excess = (uint32_t)wave_sum((unsigned long long)excess);
    if (__lane_id() == 0) wred[threadIdx.x / kWave] = excess;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < THREADS / kWave; ++w) tot += wred[w];
        obase = tot ? atomicAdd(ovf_n, (unsigned long long)tot) : 0ull;
    }
    if (!held)  // the placement counts again
        for (uint32_t j = threadIdx.x; j < kChunkBuckets; j += THREADS) cnt[j] = 0;
    __syncthreads();
    auto put = [&](uint32_t j, uint32_t slot, const WideEntry &e) {
        if constexpr (DENSE) {
            kimg[j * kSlots + slot] = e.k;
            vimg[j * kSlots + slot] = e.v;
        } else {
            ksub[(uint64_t)(b0 + j) * kSlots + slot] = e.k;
            vsub[(uint64_t)(b0 + j) * kSlots + slot] = e.v;
        }
    };
    auto overflow = [&](const WideEntry &e) {  // the counted excess reserved exactly these places
        const unsigned long long at = obase + atomicAdd(&onext, 1u);
        if (at < ovf_cap) ovf[at] = e;
    };
    if (held) {
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (lo + threadIdx.x + u * THREADS >= hi) continue;
            const uint32_t at = hp[u];
            if (at == kReject) {
                ++rejected;
                continue;
            }
            if (at != kStray) {
                const uint32_t j = at & 511u, r = at >> 9;
                const uint32_t m = fm[j];
                if (r < (uint32_t)__builtin_popcount(m)) {
                    put(j, nth_set_bit16(m, r), hb[u]);
                    ++placed;
                    continue;
                }
            }
            overflow(hb[u]);
        }
    }
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi && !held; i0 += kBatch * THREADS) {
        WideEntry eb[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            const uint32_t i = i0 + u * THREADS;
            eb[u] = i < hi ? ent[i] : WideEntry{0ull, 0ull};
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (i0 + u * THREADS >= hi) continue;
            const WideEntry e = eb[u];
            if (wide_reserved(e.k)) {
                ++rejected;
                continue;
            }
            const uint32_t b = bucket_of_uniform(S, fold64(e.k));
            bool fits = false;
            if (b >= b0 && b - b0 < len) {
                const uint32_t j = b - b0;
                const uint32_t r = atomicAdd(&cnt[j], 1u);
                const uint32_t m = fm[j];
                if (r < (uint32_t)__builtin_popcount(m)) {
                    put(j, nth_set_bit16(m, r), e);
                    fits = true;
                    ++placed;
                }
            }
            if (!fits) overflow(e);
        }
    }
    if constexpr (DENSE) {
        __syncthreads();
        uint4 *const kd = reinterpret_cast<uint4 *>(ksub + (uint64_t)b0 * kSlots);
        uint4 *const vd = reinterpret_cast<uint4 *>(vsub + (uint64_t)b0 * kSlots);
        const uint4 *const ki = reinterpret_cast<const uint4 *>(kimg);
        const uint4 *const vi = reinterpret_cast<const uint4 *>(vimg);
        for (uint32_t q = threadIdx.x; q < len * (kSlots / 2); q += THREADS) {
            kd[q] = ki[q];
            vd[q] = vi[q];
        }
    }
}

// One workgroup per chunk of one wide table (chunks numbered subtable by subtable, Ch); the
// entries of chunk c are ent[start[c], start[c + 1]).
template <bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void wide_bulk_fill_kernel(WideDesc W, Chunks Ch, const uint32_t *__restrict__ start,
                                                                 const WideEntry *__restrict__ ent, uint64_t n,
                                                                 WideEntry *__restrict__ ovf, uint64_t ovf_cap,
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
        for (uint32_t q = 1; q < W.D.d; ++q)
            if (Ch.base[q] <= c) t = q;
        const uint32_t b0 = (c - chunk_base(Ch, t)) << kChunkShift;
        const uint32_t len = min(kChunkBuckets, selt(t, W.D.nb) - b0);
        wide_fill_chunk<DENSE, THREADS>(sub_ref(W.D, t, selt(t, W.val)), b0, len, ent, lo, hi, ovf, ovf_cap, ovf_n,
                                        zeroed, placed, rejected);
    }
    wave_tally(C, &C->kicks, placed);
    wave_tally(C, &C->rejected, rejected);
}

// The last overflow list (entries) as the key and value arrays the eviction insert reads.
__global__ __launch_bounds__(256) void wide_split_kernel(const WideEntry *__restrict__ in, uint64_t n,
                                                         uint64_t *__restrict__ keys, uint64_t *__restrict__ vals)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const WideEntry e = in[i];
        keys[i] = e.k;
        vals[i] = e.v;
    }
}

}  // namespace

// ------------------------------------------------------------------ bulk fill of many tables
// The wide form of the group bulk fill (dyc_bulk.hip): keys fold to 32 bits for the pair and the
// bucket (fold64), and entries are 16-byte key/value pairs.

template <bool IN_ENTRIES>
__global__ __launch_bounds__(kCountThreads) void wide_group_bulk_pack_kernel(const WideDesc *__restrict__ dir,
                                                                             const GroupItem *__restrict__ blocks,
                                                                             const uint64_t *__restrict__ keys,
                                                                             const WideEntry *__restrict__ in,
                                                                             uint8_t *__restrict__ lcid,
                                                                             uint32_t *__restrict__ partial)
{
    __shared__ uint32_t h[kGroupBulkChunks];
    __shared__ uint4 sub_tab[kMaxSubtables];  // per subtable: buckets, divider, the table's first local chunk of it
    __shared__ uint32_t tsel[32];  // the round's subtable by key parity and pair (fill_round_sub)
    const GroupItem item = blocks[blockIdx.x];
    if (item.len == 0) {  // block-uniform: a clipped block of round 2 counts nothing
        for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads)
            partial[(uint64_t)blockIdx.x * kGroupBulkChunks + d] = 0;
        return;
    }
    const Desc D = dir[item.table].D;  // by value: loaded once, not per use through memory
    for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads) h[d] = 0;
    if (threadIdx.x < 32) tsel[threadIdx.x] = fill_round_sub(D, threadIdx.x, IN_ENTRIES);
    if (threadIdx.x < kMaxSubtables) {
        const uint32_t t = threadIdx.x;
        uint32_t first = 0;
        for (uint32_t q = 0; q < t && q < D.d; ++q) first += (selt(q, D.nb) + kChunkBuckets - 1) >> kChunkShift;
        sub_tab[t] = make_uint4(selt(t, D.nb), selt(t, D.m), selt(t, D.sh), first);
    }
    __syncthreads();
    const uint64_t lo = item.begin, hi = item.begin + item.len;
    for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (uint64_t)kCountThreads * kCountItems) {
        uint64_t k[kCountItems];
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if constexpr (IN_ENTRIES) k[u] = i < hi ? in[i].k : 0ull;
            else k[u] = i < hi ? keys[i] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kCountItems; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kCountThreads;
            if (i >= hi) continue;
            uint32_t lc = 0;
            if (IN_ENTRIES || !wide_reserved(k[u])) {
                const uint32_t f = fold64(k[u]), p = pair_of(D, f);
                const uint32_t t = tsel[(((uint32_t)k[u] & 1u) << 4) | p];
                const uint4 q = sub_tab[t];
                lc = q.w + (fast_mod_sh(hash_t(f, t), q.x, q.y, q.z) >> kChunkShift);
            }
            lcid[i] = (uint8_t)lc;  // < kGroupBulkChunks: the host admits tables of at most that many
            atomicAdd(&h[lc & (kGroupBulkChunks - 1)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kGroupBulkChunks; d += kCountThreads)
        partial[(uint64_t)blockIdx.x * kGroupBulkChunks + d] = h[d];
}

template <bool DENSE, int THREADS>
__global__ __launch_bounds__(THREADS) void wide_group_bulk_fill_kernel(const GroupChunk *__restrict__ chunks,
                                                                       const WideEntry *__restrict__ ent,
                                                                       WideEntry *__restrict__ ovf,
                                                                       unsigned long long *__restrict__ ovf_n,
                                                                       unsigned long long *__restrict__ table_counts,
                                                                       Counters *__restrict__ C, bool zeroed)
{
    const GroupChunk R = chunks[blockIdx.x];
    unsigned long long placed = 0, rejected = 0;
    if (R.cnt)  // block-uniform
        wide_fill_chunk<DENSE, THREADS>(R.S, R.b0, R.len, ent, R.lo, R.lo + R.cnt, ovf + R.ovf_off, R.ovf_cap,
                                        ovf_n + R.j, zeroed, placed, rejected);
    const unsigned long long wp = wave_sum(placed);
    if (__lane_id() == 0 && wp) atomicAdd(&table_counts[R.table], wp);
    wave_tally(C, &C->kicks, placed);
    wave_tally(C, &C->rejected, rejected);
}

// The overflow entries of every item as the key and value arrays the group insert reads (same
// indexes); one workgroup per item.
__global__ __launch_bounds__(256) void wide_items_split_kernel(const GroupItem *__restrict__ items,
                                                               const WideEntry *__restrict__ in,
                                                               uint64_t *__restrict__ keys, uint64_t *__restrict__ vals)
{
    const GroupItem it = items[blockIdx.x];
    for (uint64_t i = it.begin + threadIdx.x; i < it.begin + it.len; i += blockDim.x) {
        const WideEntry e = in[i];
        keys[i] = e.k;
        vals[i] = e.v;
    }
}

// ------------------------------------------------------------------------------ launchers
hipError_t launch_wide_probe(const WideDesc &W, const uint64_t *keys, uint64_t *vals, uint8_t *flags, uint64_t n,
                             bool erase, LaunchShape sh, Counters *C, const WideStash &stash, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh);
    if (erase)
        wide_probe_kernel<1><<<grid, sh.block_threads, 0, st>>>(W, keys, nullptr, flags, n, C, stash);
    else
        wide_probe_kernel<0><<<grid, sh.block_threads, 0, st>>>(W, keys, vals, flags, n, C, stash);
    return hipGetLastError();
}

// Writes vals[i] to every copy of keys[i] that pos1, pos2 (the first that holds it) and the stash
// hold; updated[i] = 1 where one was found. Keys must be distinct within the call.
hipError_t launch_wide_update(const WideDesc &W, const uint64_t *keys, const uint64_t *vals, uint8_t *updated,
                              uint64_t n, LaunchShape sh, Counters *C, const WideStash &stash, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh);
    wide_probe_kernel<2><<<grid, sh.block_threads, 0, st>>>(W, keys, const_cast<uint64_t *>(vals), updated, n, C,
                                                            stash);
    return hipGetLastError();
}

hipError_t launch_wide_insert(const WideDesc &W, const uint64_t *keys, const uint64_t *vals, uint64_t n,
                              uint32_t pre_fixed, bool from_entries, InsertPolicy pol, LaunchShape sh, Counters *C,
                              const WideSink &sink, hipStream_t st, const uint8_t *skip = nullptr)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh, kWideInsertBlocks);
    if (from_entries)
        wide_insert_kernel<true><<<grid, sh.block_threads, 0, st>>>(W, keys, vals, n, pre_fixed, pol, C, sink, skip);
    else
        wide_insert_kernel<false><<<grid, sh.block_threads, 0, st>>>(W, keys, vals, n, kNoTable, pol, C, sink,
                                                                     skip);
    return hipGetLastError();
}

constexpr int kWideGroupBlocks = 2048;

hipError_t launch_wide_group_probe(const WideDesc *dir, const GroupItem *items, uint32_t n_items, const uint64_t *keys,
                                   uint64_t *vals, uint8_t *flags, bool erase, unsigned long long *table_counts,
                                   int blocks, const WideStash &stash, hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    const int grid = blocks > 0 ? blocks : std::min<int>(kWideGroupBlocks, (int)((n_items + 3) / 4));
    if (erase)
        wide_group_probe_kernel<1><<<grid, 256, 0, st>>>(dir, items, n_items, keys, nullptr, flags, table_counts, stash);
    else
        wide_group_probe_kernel<0><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, flags, nullptr, stash);
    return hipGetLastError();
}

hipError_t launch_wide_group_insert(const WideDesc *dir, const GroupItem *items, uint32_t n_items,
                                    const uint64_t *keys, const uint64_t *vals, InsertPolicy pol, int blocks,
                                    Counters *C, const WideSink &sink, unsigned long long *table_counts,
                                    hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    const int grid = blocks > 0 ? blocks : std::min<int>(kWideGroupBlocks, (int)((n_items + 3) / 4));
    wide_group_insert_kernel<<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, pol, C, sink, table_counts);
    return hipGetLastError();
}

// Wide entries (16 bytes): 4096-entry tiles (256 threads), two blocks per CU.
constexpr int kWideSortThreads = 256;
constexpr uint32_t kWideSortTile = kWideSortThreads * kSortItems;

template <typename CidT>
hipError_t wide_partition(const WideDesc &W, const Chunks &Ch, const uint64_t *keys, const uint64_t *vals, int in,
                          uint64_t n, const BulkSpace &ws, uint32_t max_blocks, hipStream_t st, int *out)
{
    RadixPass P[kMaxRadixPasses];
    const uint32_t np = radix_passes(Ch.bits, P);
    const PartGrid G = part_grid(n, max_blocks, kWideSortTile);
    CidT *const cid[2] = {static_cast<CidT *>(ws.cid[0]), static_cast<CidT *>(ws.cid[1])};
    WideEntry *const ent[2] = {static_cast<WideEntry *>(ws.ent[0]), static_cast<WideEntry *>(ws.ent[1])};
    if (keys) {
        wide_bulk_pack_kernel<false, CidT><<<G.blocks * G.sub, kCountThreads, 0, st>>>(W, Ch, keys, nullptr, G, P[0], cid[0],
                                                                               ws.partial);
        return partition_passes<WideEntry, WideKeyValSource, CidT, kWideSortThreads>(
            WideKeyValSource{keys, vals}, 0, ent, cid, G, P, np, ws.partial, ws.total, ws.start, Ch.total, st, out);
    }
    wide_bulk_pack_kernel<true, CidT><<<G.blocks * G.sub, kCountThreads, 0, st>>>(W, Ch, nullptr, ent[in], G, P[0], cid[0],
                                                                          ws.partial);
    return partition_passes<WideEntry, EntrySource<WideEntry>, CidT, kWideSortThreads>(
        EntrySource<WideEntry>{ent[in]}, 1 - in, ent, cid, G, P, np, ws.partial, ws.total, ws.start, Ch.total, st,
        out);
}

// Group n wide entries by chunk, as launch_bulk_partition does for the 32-bit table (round 1
// from keys/vals, round 2 from ws.ent[in]); the sorted entries end in ws.ent[*out].
hipError_t launch_wide_partition(const WideDesc &W, const Chunks &Ch, const uint64_t *keys, const uint64_t *vals,
                                 int in, uint64_t n, const BulkSpace &ws, uint32_t max_blocks, hipStream_t st,
                                 int *out)
{
    *out = keys ? 0 : 1 - in;
    if (n == 0) return hipSuccess;
    if (n >= (1ull << 32) || Ch.total == 0 || Ch.bits > kMaxRadixPasses * kMaxRadixBits) return hipErrorInvalidValue;
    if (bulk_cid_bytes(Ch) == 2) return wide_partition<uint16_t>(W, Ch, keys, vals, in, n, ws, max_blocks, st, out);
    return wide_partition<uint32_t>(W, Ch, keys, vals, in, n, ws, max_blocks, st, out);
}

hipError_t launch_wide_bulk_fill(const WideDesc &W, const Chunks &Ch, const uint32_t *start, const WideEntry *ent,
                                 uint64_t n, WideEntry *ovf, uint64_t ovf_cap, unsigned long long *ovf_n, Counters *C,
                                 bool zeroed, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (n >= (1ull << 32) || Ch.total == 0) return hipErrorInvalidValue;
    uint64_t buckets = 0;
    for (uint32_t t = 0; t < W.D.d; ++t) buckets += W.D.nb[t];
    if (n >= 2 * buckets)
        wide_bulk_fill_kernel<true, 1024><<<Ch.total, 1024, 0, st>>>(W, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C, zeroed);
    else
        wide_bulk_fill_kernel<false, 256><<<Ch.total, 256, 0, st>>>(W, Ch, start, ent, n, ovf, ovf_cap, ovf_n, C, false);
    return hipGetLastError();
}

hipError_t launch_wide_split(const WideEntry *in, uint64_t n, uint64_t *keys, uint64_t *vals, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, LaunchShape{256, 0}, 16384);
    wide_split_kernel<<<grid, 256, 0, st>>>(in, n, keys, vals);
    return hipGetLastError();
}

hipError_t launch_wide_upsize(const uint64_t *ok_, const uint64_t *ov, uint32_t old_nb, uint64_t *dk, uint64_t *dv,
                              uint32_t t, Counters *C, hipStream_t st)
{
    const FastDiv f = make_fastdiv(old_nb * 2u);
    const int grid = grid_for(((uint64_t)old_nb + 7) / 8, LaunchShape{256, 0});
    wide_upsize_kernel<<<grid, 256, 0, st>>>(ok_, ov, old_nb, dk, dv, t, f.m, f.s1, f.s2, C);
    return hipGetLastError();
}

hipError_t launch_wide_downsize(const uint64_t *ok_, const uint64_t *ov, uint32_t new_nb, uint64_t *dk, uint64_t *dv,
                                Counters *C, uint64_t *rk, uint64_t *rv, uint64_t res_cap, uint8_t *live,
                                hipStream_t st)
{
    if (new_nb == 0) return hipSuccess;
    if (res_cap < (uint64_t)new_nb * kSlots || !live) return hipErrorInvalidValue;
    const int grid = grid_for(((uint64_t)new_nb + kResTileBuckets - 1) / kResTileBuckets, LaunchShape{256, 0});
    wide_downsize_kernel<<<grid, 256, 0, st>>>(ok_, ov, new_nb, dk, dv, C, rk, rv, live);
    return hipGetLastError();
}

}  // namespace dyc

// ------------------------------------------------------------------------------- host side
using namespace dyc;
using dyc::host::await_post;
using dyc::host::choose_down;
using dyc::host::choose_up;
using dyc::host::config_subtables;
using dyc::host::DeviceGuard;
using dyc::host::fail;
using dyc::host::initial_buckets;
using dyc::host::kMaxBuckets;
using dyc::host::Reclaimer;

struct dyc_wide {
    dyc_config cfg;
    int device = 0;
    WideDesc W{};
    // storage: subtable t's keys (nb * 128 B) then its values (nb * 128 B); the subtables a
    // table is created or loaded with share one allocation (the slab), a resized one has its own
    uint8_t *slab = nullptr;
    uint32_t slab_live = 0;
    uint8_t *own[kMaxSubtables] = {};
    uint64_t count = 0;
    uint64_t *pk = nullptr, *pv = nullptr;  // pending keys / values [0, pend_n): the stash
    uint64_t pend_cap = 0, pend_n = 0;
    Counters *ctr = nullptr, *hctr = nullptr;  // device block (kCounterBytes); host folded copy
    unsigned long long *hpost = nullptr;       // pinned: the read-back words (launch_fold_post)
    unsigned long long post_seq = 0;
    bool ctr_clean = false;  // the device block is zero and no launch has used it since
    // every key and value word is zero: set when the table is created or cleared, dropped by
    // every call that may store an entry (a bulk fill into such a table need not read it)
    bool zeroed = false;
    uint8_t *live = nullptr;  // downsize: tile flags of the residual list
    uint64_t live_cap = 0;
    uint8_t *upd = nullptr;  // upsert: per-key "updated in place" flags
    uint64_t upd_cap = 0;
    // bulk fill (insert_order 3): (chunk id, entry) pairs and their sort buffers, the chunk
    // starts, the radix sort's scratch and the overflow count (device, and a pinned read-back)
    void *bk_cid[2] = {};
    WideEntry *bk_ent[2] = {};
    uint64_t bk_cap = 0;
    size_t bk_cid_bytes = 0;
    uint32_t *bk_start = nullptr;
    uint64_t bk_start_cap = 0;
    uint32_t *bk_partial = nullptr;  // kPartialWords, then 256 digit totals
    unsigned long long *bk_ovf = nullptr;
    unsigned long long *bk_hovf = nullptr;
    Reclaimer reclaim;  // frees retired subtables off the calling thread
    dyc_stats st{};

    uint64_t capacity() const
    {
        uint64_t c = 0;
        for (uint32_t i = 0; i < W.D.d; ++i) c += (uint64_t)W.D.nb[i] * kSlots;
        return c;
    }
    LaunchShape shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.max_blocks}; }
    LaunchShape insert_shape() const { return LaunchShape{(int)cfg.block_threads, (int)cfg.insert_blocks}; }
    InsertPolicy policy() const { return InsertPolicy{cfg.kick_bound, cfg.rescue_kicks, cfg.max_retries}; }
    WideSink sink_at(uint64_t at) const { return WideSink{pk + at, pv + at, pend_cap - at, nullptr}; }
    WideStash stash() const { return WideStash{pk, pv, pend_n, nullptr, 0}; }
};

namespace {

constexpr size_t kWideAlign = 2u << 20;
size_t wide_bytes(uint32_t nb) { return (size_t)nb * kBucketBytes; }  // one array of a subtable

// Callers zero the counters before a launch group and read them after it; a read clears the
// device block, so the zero is launched only when something else left the block in use.
int wzero(dyc_wide *t, hipStream_t s)
{
    if (!t->ctr_clean) HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(t->ctr), kCounterWords, s));
    t->ctr_clean = false;
    return DYC_OK;
}

// Fold, post and clear the counters, and wait for them (every kernel queued on s before has
// finished when this returns).
int wread(dyc_wide *t, hipStream_t s)
{
    const unsigned long long seq = ++t->post_seq;
    HIP_TRY(launch_fold_post(t->ctr, t->hpost, seq, s));
    HIP_TRY(await_post(t->hpost, seq, s));
    memcpy(t->hctr, const_cast<const unsigned long long *>(t->hpost) + 1, sizeof(Counters));
    t->ctr_clean = true;
    return DYC_OK;
}

// Release a subtable's storage: its own allocation `own`, or (own null) its share of the slab.
// The free runs on the reclaimer's thread, after every kernel on the device has finished.
int wrelease_block(dyc_wide *t, uint8_t *own)
{
    if (own) {
        t->reclaim.defer(t->device, own);
    } else if (t->slab && --t->slab_live == 0) {
        t->reclaim.defer(t->device, t->slab);
        t->slab = nullptr;
    }
    return DYC_OK;
}

// Free the bulk-fill workspace (grown on demand, kept between calls); the next bulk fill
// allocates it again. The caller has synchronised the table's stream.
void wrelease_workspace(dyc_wide *t)
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
}

// hipMalloc for a resize; when it fails, wait for the frees still queued, release the idle
// bulk-fill workspace (up to 36-40 B per key of the largest bulk-filled batch) and try once more.
hipError_t wresize_malloc(dyc_wide *t, uint8_t **p, size_t bytes, hipStream_t s)
{
    if (hipMalloc(p, bytes) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    t->reclaim.drain();
    if (hipStreamSynchronize(s) == hipSuccess) wrelease_workspace(t);
    (void)hipGetLastError();
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

int wrelease(dyc_wide *t, uint32_t i)
{
    if (!t->W.D.sub[i]) return DYC_OK;
    uint8_t *own = t->own[i];
    t->own[i] = nullptr;
    t->W.D.sub[i] = nullptr;
    t->W.val[i] = nullptr;
    return wrelease_block(t, own);
}

void wset(dyc_wide *t, uint32_t i, uint8_t *base, uint32_t nb)
{
    desc_set(t->W.D, (int)i, reinterpret_cast<uint64_t *>(base), nb);
    t->W.val[i] = reinterpret_cast<uint64_t *>(base + wide_bytes(nb));
}

// One allocation for the key and value arrays of d subtables of nb[i] buckets; replaces every
// subtable of t (zeroed when `zero`).
int winstall_slab(dyc_wide *t, const uint32_t *nb, bool zero, hipStream_t s, uint8_t **region)
{
    t->zeroed = false;  // new memory: only the caller knows what it holds
    const uint32_t d = t->W.D.d;
    size_t off[kMaxSubtables + 1] = {0};
    for (uint32_t i = 0; i < d; ++i) off[i + 1] = off[i] + ((2 * wide_bytes(nb[i]) + kWideAlign - 1) & ~(kWideAlign - 1));
    uint8_t *base = nullptr;
    if (hipMalloc(&base, off[d]) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DYC_ERR_OOM, "hipMalloc of %zu bytes for the table failed", off[d]);
    }
    for (uint32_t i = 0; i < d; ++i) {
        const int rc = wrelease(t, i);
        if (rc) return rc;
    }
    t->slab = base;
    t->slab_live = d;
    for (uint32_t i = 0; i < d; ++i) {
        wset(t, i, base + off[i], nb[i]);
        if (region) region[i] = base + off[i];
        if (zero) HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(base + off[i]), 2 * (uint64_t)nb[i] * kSlots, s));
    }
    return DYC_OK;
}

int wreserve_pending(dyc_wide *t, uint64_t extra, hipStream_t s)
{
    if (t->pend_n + extra <= t->pend_cap) return DYC_OK;
    const uint64_t cap = std::max<uint64_t>(2 * t->pend_cap, t->pend_n + extra);
    uint64_t *nk = nullptr, *nv = nullptr;
    if (hipMalloc(&nk, cap * 8) != hipSuccess || hipMalloc(&nv, cap * 8) != hipSuccess) {
        (void)hipGetLastError();
        if (nk) (void)hipFree(nk);
        return fail(DYC_ERR_OOM, "hipMalloc for the pending list failed");
    }
    if (t->pend_n) {
        HIP_TRY(hipMemcpyAsync(nk, t->pk, t->pend_n * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(nv, t->pv, t->pend_n * 8, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (t->pk) HIP_TRY(hipFree(t->pk));
    if (t->pv) HIP_TRY(hipFree(t->pv));
    t->pk = nk;
    t->pv = nv;
    t->pend_cap = cap;
    return DYC_OK;
}

int wabsorb(dyc_wide *t, uint64_t room)
{
    const Counters &c = *t->hctr;
    t->st.kicks += c.kicks;
    t->st.last_kicks += c.kicks;
    t->st.cas_retries += c.retries;
    t->st.kick_failures += c.pending;
    t->st.lost += c.lost;
    t->pend_n += std::min<uint64_t>(c.pending, room);
    if (c.lost) return fail(DYC_ERR_LOST, "pending list overflow: %llu entries lost", (unsigned long long)c.lost);
    return DYC_OK;
}

int wresize_up(dyc_wide *t, hipStream_t s)
{
    t->zeroed = false;
    const uint32_t sub = choose_up(t->W.D.nb, t->W.D.d);
    const uint32_t old_nb = t->W.D.nb[sub];
    if ((uint64_t)old_nb * 2 >= kMaxBuckets) return fail(DYC_ERR_INVALID, "subtable %u cannot grow past 2^29 buckets", sub);
    uint8_t *dst = nullptr;
    if (wresize_malloc(t, &dst, 2 * wide_bytes(old_nb * 2), s) != hipSuccess)
        return fail(DYC_ERR_OOM, "hipMalloc for upsize failed");
    int rc = wzero(t, s);
    if (rc) return rc;
    HIP_TRY(launch_wide_upsize(t->W.D.sub[sub], t->W.val[sub], old_nb, reinterpret_cast<uint64_t *>(dst),
                               reinterpret_cast<uint64_t *>(dst + wide_bytes(old_nb * 2)), sub, t->ctr, s));
    rc = wread(t, s);
    if (rc) return rc;
    rc = wrelease(t, sub);
    if (rc) return rc;
    wset(t, sub, dst, old_nb * 2);
    t->own[sub] = dst;
    t->st.resizes_up++;
    if (t->hctr->misplaced)
        return fail(DYC_ERR_CORRUPT, "upsize found %llu entries outside their bucket",
                    (unsigned long long)t->hctr->misplaced);
    return DYC_OK;
}

// Re-insert the pending entries after an operation that found `before` pending and left more
// (upsizing before each round, at most max_fail_rounds rounds); see place_pending of the narrow
// table.
int wplace_pending(dyc_wide *t, hipStream_t s, uint64_t before)
{
    t->zeroed = false;
    uint32_t rounds = 0;
    while (t->pend_n > before && rounds < t->cfg.max_fail_rounds) {
        int rc = wresize_up(t, s);
        if (rc) return rc;
        const uint64_t m = t->pend_n;
        uint64_t *sk = nullptr, *sv = nullptr;
        if (hipMalloc(&sk, m * 8) != hipSuccess || hipMalloc(&sv, m * 8) != hipSuccess) {
            (void)hipGetLastError();
            if (sk) (void)hipFree(sk);
            return fail(DYC_ERR_OOM, "hipMalloc for pending re-insert failed");
        }
        HIP_TRY(hipMemcpyAsync(sk, t->pk, m * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(sv, t->pv, m * 8, hipMemcpyDeviceToDevice, s));
        t->pend_n = 0;
        rc = wzero(t, s);
        if (!rc) {
            const hipError_t e = launch_wide_insert(t->W, sk, sv, m, kNoTable, true, t->policy(), t->insert_shape(),
                                                    t->ctr, t->sink_at(0), s);
            if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "pending re-insert launch: %s", hipGetErrorString(e));
        }
        if (!rc) rc = wread(t, s);
        (void)hipFree(sk);
        (void)hipFree(sv);
        if (rc) return rc;
        rc = wabsorb(t, t->pend_cap);
        if (rc) return rc;
        ++rounds;
    }
    t->st.pending = t->pend_n;
    if (t->pend_n > before)
        return fail(DYC_ERR_PENDING, "%llu entries still pending after %u upsize rounds (%llu before the call)",
                    (unsigned long long)t->pend_n, rounds, (unsigned long long)before);
    return DYC_OK;
}

int wresize_down(dyc_wide *t, hipStream_t s)
{
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    const uint32_t sub = choose_down(t->W.D.nb, t->W.D.d);
    const uint32_t old_nb = t->W.D.nb[sub];
    if (old_nb & 1u)
        return fail(DYC_ERR_ODD_BUCKETS, "downsize of subtable %u refused: %u buckets is odd", sub, old_nb);
    const uint32_t new_nb = old_nb / 2;
    uint8_t *dst = nullptr;
    if (wresize_malloc(t, &dst, 2 * wide_bytes(new_nb), s) != hipSuccess)
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
    // The residuals go to the upper halves of the old subtable's key and value arrays,
    // slot-addressed: a group writes pair g's residuals there only after reading both of its
    // buckets, and no other group reads them. The old subtable is released once they are
    // re-inserted.
    uint64_t *ok_old = t->W.D.sub[sub], *ov_old = t->W.val[sub];
    uint8_t *old_own = t->own[sub];
    const uint64_t res_n = (uint64_t)new_nb * kSlots;
    uint64_t *rk = ok_old + res_n, *rv = ov_old + res_n;
    int rc = wzero(t, s);
    if (!rc) {
        const hipError_t e =
            launch_wide_downsize(ok_old, ov_old, new_nb, reinterpret_cast<uint64_t *>(dst),
                                 reinterpret_cast<uint64_t *>(dst + wide_bytes(new_nb)), t->ctr, rk, rv, res_n,
                                 t->live, s);
        if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "downsize launch: %s", hipGetErrorString(e));
    }
    if (!rc) rc = wread(t, s);
    if (rc) {
        (void)hipFree(dst);
        return rc;
    }
    const uint64_t nres = t->hctr->residuals;
    t->own[sub] = nullptr;  // detach the old storage; the table now owns dst
    wset(t, sub, dst, new_nb);
    t->own[sub] = dst;
    t->st.resizes_down++;
    // residuals continue from the shrunk subtable into the other one of their pair (cuckoo.cu:719)
    if (nres) {
        rc = wreserve_pending(t, std::max<uint64_t>(nres, 1024), s);
        if (!rc) rc = wzero(t, s);
        if (!rc) {
            // the slot-addressed list: its live tiles (key 0 marks an empty place)
            const hipError_t e = launch_wide_insert(t->W, rk, rv, res_n, sub, true, t->policy(), t->insert_shape(),
                                                    t->ctr, t->sink_at(t->pend_n), s, t->live);
            if (e != hipSuccess) rc = fail(DYC_ERR_HIP, "residual insert launch: %s", hipGetErrorString(e));
        }
        if (!rc) rc = wread(t, s);
        if (!rc) rc = wabsorb(t, t->pend_cap - t->pend_n);
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        (void)wrelease_block(t, old_own);
        return rc;
    }
    rc = wrelease_block(t, old_own);  // freed once the device is idle
    if (rc) return rc;
    if (t->pend_n > pend_before) return wplace_pending(t, s, pend_before);
    return DYC_OK;
}

// Erases leave the stash entries they deleted as zero keys: drop them (synchronous; the list
// is short whenever it is not empty).
int wcompact_pending(dyc_wide *t, hipStream_t s)
{
    if (t->pend_n == 0) return DYC_OK;
    std::vector<uint64_t> k(t->pend_n), v(t->pend_n);
    HIP_TRY(hipMemcpyAsync(k.data(), t->pk, t->pend_n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(v.data(), t->pv, t->pend_n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint64_t m = 0;
    for (uint64_t i = 0; i < k.size(); ++i)
        if (k[i]) {
            k[m] = k[i];
            v[m] = v[i];
            ++m;
        }
    if (m != t->pend_n) {
        if (m) {
            HIP_TRY(hipMemcpyAsync(t->pk, k.data(), m * 8, hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(t->pv, v.data(), m * 8, hipMemcpyHostToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        t->pend_n = m;
    }
    t->st.pending = t->pend_n;
    return DYC_OK;
}

// Bulk fill for batches of at least two keys per bucket (insert_order 3), as for 32-bit keys.
bool wuse_bulk(const dyc_wide *t, uint64_t n)
{
    uint64_t buckets = 0;
    for (uint32_t i = 0; i < t->W.D.d; ++i) buckets += t->W.D.nb[i];
    return t->cfg.insert_order == 3 && n >= 2 * buckets && n < (1ull << 31);
}

int walloc(void **p, size_t bytes, const char *what)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(DYC_ERR_OOM, "hipMalloc of the bulk-fill %s (%llu bytes) failed", what,
                    (unsigned long long)bytes);
    }
    return DYC_OK;
}

int wreserve_bulk(dyc_wide *t, uint64_t n, const Chunks &Ch, hipStream_t s)
{
    const size_t cb = bulk_cid_bytes(Ch);
    // a buffer being replaced may still be in use by the table's earlier work, ordered on s
    if (n > t->bk_cap || cb > t->bk_cid_bytes || Ch.total + 1ull > t->bk_start_cap) HIP_TRY(hipStreamSynchronize(s));
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
            rc = walloc(&t->bk_cid[b], n * cb, "chunk ids");
            if (!rc) rc = walloc((void **)&t->bk_ent[b], n * sizeof(WideEntry), "entries");
        }
        if (rc) return rc;
        t->bk_cap = n;
        t->bk_cid_bytes = cb;
    }
    if (Ch.total + 1ull > t->bk_start_cap) {
        if (t->bk_start) HIP_TRY(hipFree(t->bk_start));
        t->bk_start = nullptr;
        t->bk_start_cap = 0;
        rc = walloc((void **)&t->bk_start, (Ch.total + 1ull) * sizeof(uint32_t), "chunk starts");
        if (rc) return rc;
        t->bk_start_cap = Ch.total + 1ull;
    }
    if (!t->bk_partial &&
        (rc = walloc((void **)&t->bk_partial, (kPartialWords + kMaxRadixWords) * sizeof(uint32_t), "partition counts")))
        return rc;
    if (!t->bk_ovf && (rc = walloc((void **)&t->bk_ovf, 256, "overflow count"))) return rc;
    if (!t->bk_hovf && hipHostMalloc((void **)&t->bk_hovf, 64, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        t->bk_hovf = nullptr;
        return fail(DYC_ERR_OOM, "hipHostMalloc of the bulk-fill read-back word failed");
    }
    return DYC_OK;
}

// One fill round over n entries: round 1 from keys/vals (keys != nullptr), round 2 from the
// entries in bk_ent[in]. Partition them by chunk, fill, and read back how many entries overflowed
// into bk_ent[*out] (the buffer the partition did not end in).
int wbulk_round(dyc_wide *t, const Chunks &Ch, const uint64_t *keys, const uint64_t *vals, int in, uint64_t n,
                bool zeroed, hipStream_t s, int *out, uint64_t *overflow)
{
    const BulkSpace ws{{t->bk_ent[0], t->bk_ent[1]}, {t->bk_cid[0], t->bk_cid[1]}, t->bk_partial,
                       t->bk_partial + kPartialWords, t->bk_start};
    int r = 0;
    HIP_TRY(launch_wide_partition(t->W, Ch, keys, vals, in, n, ws, kWideSortBlocks, s, &r));
    HIP_TRY(launch_zero(t->bk_ovf, 1, s));
    HIP_TRY(launch_wide_bulk_fill(t->W, Ch, t->bk_start, t->bk_ent[r], n, t->bk_ent[1 - r], t->bk_cap, t->bk_ovf,
                                  t->ctr, zeroed, s));
    HIP_TRY(hipMemcpyAsync(t->bk_hovf, t->bk_ovf, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *overflow = *t->bk_hovf;
    *out = 1 - r;
    if (*overflow > n)
        return fail(DYC_ERR_CORRUPT, "bulk fill overflow count %llu exceeds the batch (%llu)",
                    (unsigned long long)*overflow, (unsigned long long)n);
    return DYC_OK;
}

// insert_order 3 for wide keys: fill each key's fill_target bucket, then the other bucket, then
// the eviction insert for the rest (the narrow table's bulk_insert over fold(k)).
int wbulk_insert(dyc_wide *t, const uint64_t *keys, const uint64_t *vals, uint64_t n, bool was_zeroed, hipStream_t s,
                 uint64_t *left)
{
    *left = 0;
    const Chunks Ch = make_chunks(t->W.D);  // the workspace was reserved for it
    int ov = 0, ov2 = 0;
    uint64_t m1 = 0, m2 = 0;
    // an empty table (count 0, nothing pending) that no call has stored into since it was
    // created or cleared holds zeros only
    const bool zeroed = was_zeroed && t->count == 0 && t->pend_n == 0;
    int rc = wbulk_round(t, Ch, keys, vals, 0, n, zeroed, s, &ov, &m1);
    if (rc || m1 == 0) return rc;
    rc = wbulk_round(t, Ch, nullptr, nullptr, ov, m1, false, s, &ov2, &m2);  // the overflow, in its other table
    if (rc || m2 == 0) return rc;
    *left = m2;
    // the other entry buffer (16 B per key of capacity) holds m2 keys and then m2 values
    uint64_t *const sk = reinterpret_cast<uint64_t *>(t->bk_ent[1 - ov2]);
    HIP_TRY(launch_wide_split(t->bk_ent[ov2], m2, sk, sk + m2, s));
    HIP_TRY(launch_wide_insert(t->W, sk, sk + m2, m2, kNoTable, true, t->policy(), t->insert_shape(), t->ctr,
                               t->sink_at(t->pend_n), s));
    return DYC_OK;
}

bool wide_check_config(const dyc_config &cfg)
{
    const uint32_t d = config_subtables(cfg);
    return cfg.block_threads && cfg.block_threads % 64 == 0 && cfg.block_threads <= 256 &&
           cfg.upper_bound > 0.0 && cfg.upper_bound <= 1.0 && cfg.lower_bound >= 0.0 &&
           cfg.lower_bound < cfg.upper_bound && d >= (uint32_t)kMinSubtables && d <= (uint32_t)kMaxSubtables;
}

}  // namespace

extern "C" {

int dyc_wide_create(const dyc_config *cfg_in, uint64_t capacity_hint, dyc_wide **out)
{
    if (!out) return fail(DYC_ERR_INVALID, "null output handle");
    *out = nullptr;
    dyc_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        dyc_config_init(&cfg);
    if (!wide_check_config(cfg))
        return fail(DYC_ERR_INVALID, "bad configuration (block_threads, fill bounds or subtables)");
    cfg.subtables = config_subtables(cfg);
    const uint32_t nb = initial_buckets(capacity_hint, cfg.subtables);
    if (nb >= kMaxBuckets) return fail(DYC_ERR_INVALID, "capacity hint too large");
    dyc_wide *t = new dyc_wide();
    t->cfg = cfg;
    desc_init(t->W.D, cfg.subtables);
    int rc = DYC_OK;
    if (hipGetDevice(&t->device) != hipSuccess || hipMalloc(&t->ctr, kCounterBytes) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&t->hctr), sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&t->hpost), kPostWords * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(DYC_ERR_OOM, "counter allocation failed");
    } else {
        memset(t->hpost, 0, kPostWords * sizeof(unsigned long long));  // no sequence number yet
    }
    const uint32_t nbs[kMaxSubtables] = {nb, nb, nb, nb, nb, nb};
    if (rc == DYC_OK) rc = winstall_slab(t, nbs, true, nullptr, nullptr);
    if (rc == DYC_OK) rc = wreserve_pending(t, 1024, nullptr);
    if (rc == DYC_OK && hipStreamSynchronize(nullptr) != hipSuccess) rc = fail(DYC_ERR_HIP, "create sync failed");
    t->zeroed = rc == DYC_OK;
    if (rc != DYC_OK) {
        char keep[512];
        strncpy(keep, dyc_last_error(), sizeof keep - 1);
        keep[sizeof keep - 1] = 0;
        dyc_wide_destroy(t);
        fail(rc, "%s", keep);
        return rc;
    }
    *out = t;
    return DYC_OK;
}

int dyc_wide_destroy(dyc_wide *t)
{
    if (!t) return DYC_OK;
    DeviceGuard dg(t->device);
    (void)hipDeviceSynchronize();
    for (uint32_t i = 0; i < t->W.D.d; ++i) (void)wrelease(t, i);
    t->reclaim.drain();  // every subtable is freed when this returns
    if (t->pk) (void)hipFree(t->pk);
    if (t->pv) (void)hipFree(t->pv);
    if (t->ctr) (void)hipFree(t->ctr);
    if (t->hctr) (void)hipHostFree(t->hctr);
    if (t->hpost) (void)hipHostFree(t->hpost);
    if (t->live) (void)hipFree(t->live);
    if (t->upd) (void)hipFree(t->upd);
    for (int b = 0; b < 2; ++b) {
        if (t->bk_cid[b]) (void)hipFree(t->bk_cid[b]);
        if (t->bk_ent[b]) (void)hipFree(t->bk_ent[b]);
    }
    if (t->bk_start) (void)hipFree(t->bk_start);
    if (t->bk_partial) (void)hipFree(t->bk_partial);
    if (t->bk_ovf) (void)hipFree(t->bk_ovf);
    if (t->bk_hovf) (void)hipHostFree(t->bk_hovf);
    delete t;
    return DYC_OK;
}

int dyc_wide_insert(dyc_wide *t, const uint64_t *keys, const uint64_t *vals, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    bool zeroed = t->zeroed;  // every word zero before this call (dropped by any resize below)
    t->zeroed = false;
    const uint64_t pend_before = t->pend_n;
    int rc;
    if (t->cfg.auto_resize) {
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = wresize_up(t, s);
            if (rc) return rc;
            zeroed = false;
        }
    }
    rc = wreserve_pending(t, std::max<uint64_t>(65536, n / 16), s);
    if (rc) return rc;
    const uint64_t room = t->pend_cap - t->pend_n;
    bool bulk = wuse_bulk(t, n);
    if (bulk) {
        const dyc::host::ErrorKeep keep;
        rc = wreserve_bulk(t, n, make_chunks(t->W.D), s);
        if (rc == DYC_ERR_OOM) {  // no room for the workspace: the eviction insert needs none
            bulk = false;
            keep.restore();  // the call goes on, so it leaves no failure message
        } else if (rc) {
            return rc;
        }
    }
    rc = wzero(t, s);
    if (rc) return rc;
    uint64_t left = 0;  // entries the bulk fill left to the eviction insert
    if (bulk) {
        rc = wbulk_insert(t, keys, vals, n, zeroed, s, &left);
        if (rc) return rc;
    } else {
        HIP_TRY(launch_wide_insert(t->W, keys, vals, n, kNoTable, false, t->policy(), t->insert_shape(), t->ctr,
                                   t->sink_at(t->pend_n), s));
    }
    rc = wread(t, s);
    if (rc) return rc;
    const uint64_t rejected = t->hctr->rejected;
    t->st.rejected += rejected;
    t->count += n - rejected;
    if (bulk) t->st.last_filled = n - rejected - left;
    rc = wabsorb(t, room);
    t->st.pending = t->pend_n;
    if (rc) return rc;
    if (t->pend_n > pend_before) return wplace_pending(t, s, pend_before);
    return DYC_OK;
}

int dyc_wide_upsert(dyc_wide *t, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint8_t *updated,
                    dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(t->device);
    t->zeroed = false;
    hipStream_t s = (hipStream_t)stream;
    t->st.last_kicks = 0;
    t->st.last_filled = 0;
    if (n == 0) return DYC_OK;
    const uint64_t pend_before = t->pend_n;
    int rc;
    if (t->cfg.auto_resize) {  // sized as if every key were new
        while ((double)(t->count + n) > t->cfg.upper_bound * (double)t->capacity()) {
            rc = wresize_up(t, s);
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
    rc = wreserve_pending(t, std::max<uint64_t>(65536, n / 16), s);
    if (rc) return rc;
    const uint64_t room = t->pend_cap - t->pend_n;
    rc = wzero(t, s);
    if (rc) return rc;
    HIP_TRY(launch_wide_update(t->W, keys, vals, flags, n, t->shape(), t->ctr, t->stash(), s));
    HIP_TRY(launch_wide_insert(t->W, keys, vals, n, kNoTable, false, t->policy(), t->insert_shape(), t->ctr,
                               t->sink_at(t->pend_n), s, flags));
    rc = wread(t, s);
    if (rc) return rc;
    const uint64_t rejected = t->hctr->rejected, upd = t->hctr->updated;
    t->st.rejected += rejected;
    t->count += n - rejected - upd;
    rc = wabsorb(t, room);
    t->st.pending = t->pend_n;
    if (rc) return rc;
    if (t->pend_n > pend_before) return wplace_pending(t, s, pend_before);
    return DYC_OK;
}

int dyc_wide_find(dyc_wide *t, const uint64_t *keys, uint64_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(t->device);
    HIP_TRY(launch_wide_probe(t->W, keys, vals, hit, n, false, t->shape(), t->ctr, t->stash(), (hipStream_t)stream));
    return DYC_OK;
}

int dyc_wide_erase(dyc_wide *t, const uint64_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    if (n && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(t->device);
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) return DYC_OK;
    int rc = wzero(t, s);
    if (rc) return rc;
    HIP_TRY(launch_wide_probe(t->W, keys, nullptr, deleted, n, true, t->shape(), t->ctr, t->stash(), s));
    rc = wread(t, s);
    if (rc) return rc;
    t->count -= std::min<uint64_t>(t->count, t->hctr->deleted);
    rc = wcompact_pending(t, s);
    if (rc) return rc;
    if (t->cfg.auto_resize && (double)t->count < t->cfg.lower_bound * (double)t->capacity() &&
        t->capacity() > t->cfg.min_capacity) {
        rc = wresize_down(t, s);
        if (rc == DYC_ERR_ODD_BUCKETS) rc = DYC_OK;  // policy: skip, the table stays larger
        if (rc) return rc;
    }
    return DYC_OK;
}

int dyc_wide_resize_up(dyc_wide *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    return wresize_up(t, (hipStream_t)stream);
}

int dyc_wide_resize_down(dyc_wide *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    return wresize_down(t, (hipStream_t)stream);
}

int dyc_wide_get_stats(const dyc_wide *t, dyc_stats *out)
{
    if (!t || !out) return fail(DYC_ERR_INVALID, "null argument");
    *out = t->st;
    out->count = t->count;
    out->capacity = t->capacity();
    for (uint32_t i = 0; i < kMaxSubtables; ++i) out->buckets[i] = i < t->W.D.d ? t->W.D.nb[i] : 0;
    out->pending = t->pend_n;
    out->subtables = t->W.D.d;
    return DYC_OK;
}

int dyc_wide_clear(dyc_wide *t, dyc_stream stream)
{
    if (!t) return fail(DYC_ERR_INVALID, "null table");
    DeviceGuard dg(t->device);
    hipStream_t s = (hipStream_t)stream;
    for (uint32_t i = 0; i < t->W.D.d; ++i) {
        HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(t->W.D.sub[i]), (uint64_t)t->W.D.nb[i] * kSlots, s));
        HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(t->W.val[i]), (uint64_t)t->W.D.nb[i] * kSlots, s));
    }
    t->count = 0;
    t->pend_n = 0;
    t->st.pending = 0;
    t->zeroed = true;
    return DYC_OK;
}

int dyc_wide_dump(dyc_wide *t, uint32_t sub, uint64_t *host_keys, uint64_t *host_vals, uint64_t n_slots,
                  dyc_stream stream)
{
    if (!t || sub >= t->W.D.d || (!host_keys && !host_vals)) return fail(DYC_ERR_INVALID, "bad argument");
    const uint64_t slots = (uint64_t)t->W.D.nb[sub] * kSlots;
    if (n_slots < slots) return fail(DYC_ERR_INVALID, "buffer holds %llu slots, subtable has %llu",
                                     (unsigned long long)n_slots, (unsigned long long)slots);
    DeviceGuard dg(t->device);
    hipStream_t s = (hipStream_t)stream;
    if (host_keys) HIP_TRY(hipMemcpyAsync(host_keys, t->W.D.sub[sub], slots * 8, hipMemcpyDeviceToHost, s));
    if (host_vals) HIP_TRY(hipMemcpyAsync(host_vals, t->W.val[sub], slots * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DYC_OK;
}

int dyc_wide_load(dyc_wide *t, const uint32_t *nb, const uint64_t *const *key_images,
                  const uint64_t *const *val_images, dyc_stream stream)
{
    if (!t || !nb || !key_images || !val_images) return fail(DYC_ERR_INVALID, "null argument");
    for (uint32_t i = 0; i < t->W.D.d; ++i)
        if (nb[i] == 0 || nb[i] >= kMaxBuckets || !key_images[i] || !val_images[i])
            return fail(DYC_ERR_INVALID, "bad subtable %u", i);
    DeviceGuard dg(t->device);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipStreamSynchronize(s));
    uint8_t *region[kMaxSubtables] = {};
    int rc = winstall_slab(t, nb, false, s, region);
    if (rc) return rc;
    uint64_t count = 0;
    for (uint32_t i = 0; i < t->W.D.d; ++i) {
        HIP_TRY(hipMemcpyAsync(region[i], key_images[i], wide_bytes(nb[i]), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(region[i] + wide_bytes(nb[i]), val_images[i], wide_bytes(nb[i]), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (uint64_t j = 0; j < (uint64_t)nb[i] * kSlots; ++j) count += key_images[i][j] != 0;
    }
    t->count = count;
    t->pend_n = 0;
    t->st.pending = 0;
    return DYC_OK;
}

int dyc_wide_pending(dyc_wide *t, uint64_t *host_keys, uint64_t *host_vals, uint64_t cap, uint64_t *n_out,
                     dyc_stream stream)
{
    if (!t || !n_out) return fail(DYC_ERR_INVALID, "null argument");
    *n_out = t->pend_n;
    const uint64_t m = std::min(cap, t->pend_n);
    if (m) {
        DeviceGuard dg(t->device);
        hipStream_t s = (hipStream_t)stream;
        if (host_keys) HIP_TRY(hipMemcpyAsync(host_keys, t->pk, m * 8, hipMemcpyDeviceToHost, s));
        if (host_vals) HIP_TRY(hipMemcpyAsync(host_vals, t->pv, m * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return DYC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------ wide-key groups
// Many wide tables in one arena behind a directory of WideDesc, served by one persistent launch
// per batch (the dyc_group_* design with 64-bit keys and values). Tables do not resize; entries
// that exhaust the eviction bound wait in a pending list tagged with their table, which that
// table's finds and erases also search.
struct dyc_wide_group {
    dyc_config cfg;
    int device = 0;
    uint32_t n = 0;
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<WideDesc> hdir;
    WideDesc *ddir = nullptr;
    std::vector<uint64_t> count;  // per table: placed + pending
    Counters *ctr = nullptr;
    Counters *hctr = nullptr;                // pinned, kCounterBytes
    unsigned long long *tcounts = nullptr;   // per-table counts of the last insert / erase
    unsigned long long *htcounts = nullptr;  // pinned mirror
    uint64_t *pk = nullptr, *pv = nullptr;
    uint32_t *ptag = nullptr;
    uint64_t pend_cap = 0, pend_n = 0;
    GroupItem *ditems = nullptr;
    GroupItem *hitems = nullptr;  // pinned staging
    size_t items_cap = 0;
    hipEvent_t items_free = nullptr;
    dyc_stats st{};
    // every key and value word of every table is zero: set at create and clear, dropped by every
    // insert
    bool zeroed = false;
    // bulk fill (cfg.insert_order 3), grown on demand: partitioned entries and the overflow (16 B
    // per key each; the entry buffer also holds the last overflow split into keys and values),
    // local chunk ids, per-block chunk counts, per-chunk start / count / table, overflow counts
    // of both rounds, and the block and table arrays (device, with pinned staging)
    WideEntry *bk_ent = nullptr, *bk_ovf = nullptr;
    uint8_t *bk_lcid = nullptr;
    uint64_t bk_cap = 0;
    uint32_t *bk_partial = nullptr;
    uint64_t bk_partial_cap = 0;
    GroupChunk *bk_chunks = nullptr;  // chunk records of a round
    uint64_t bk_chunk_cap = 0;
    unsigned long long *bk_ovfn = nullptr, *bk_hovfn = nullptr;
    GroupItem *bk_blk = nullptr, *bk_hblk = nullptr;
    uint32_t *bk_btab = nullptr, *bk_hbtab = nullptr;
    uint64_t bk_blk_cap = 0;
    BulkTable *bk_tabs = nullptr, *bk_htabs = nullptr;
    uint32_t bk_tabs_cap = 0;

    InsertPolicy policy() const { return InsertPolicy{cfg.kick_bound, cfg.rescue_kicks, cfg.max_retries}; }
    WideStash stash() const { return WideStash{pk, pv, pend_n, ptag, 0}; }
};

namespace {

constexpr uint32_t kWideGroupItemKeys = 4096;

int wg_check_offsets(const dyc_wide_group *g, const uint64_t *offsets)
{
    if (!offsets) return fail(DYC_ERR_INVALID, "null offsets");
    if (offsets[0] != 0) return fail(DYC_ERR_INVALID, "offsets[0] must be 0");
    for (uint32_t i = 0; i < g->n; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(DYC_ERR_INVALID, "offsets decrease at table %u", i);
    return DYC_OK;
}

// by_xcd: the find's layout (dyc::host::items_by_xcd: each table whole on one XCD, empty slots
// between), else the items in table order.
int wg_upload_items(dyc_wide_group *g, const uint64_t *offsets, hipStream_t s, uint32_t *n_items, bool by_xcd = false)
{
    const uint64_t per = kWideGroupItemKeys;
    std::vector<GroupItem> laid;
    if (by_xcd) laid = dyc::host::items_by_xcd(offsets, g->n, per);
    uint64_t need = laid.size();
    if (!by_xcd)
        for (uint32_t i = 0; i < g->n; ++i) need += (offsets[i + 1] - offsets[i] + per - 1) / per;
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
    uint32_t m = 0;
    if (by_xcd) {
        std::copy(laid.begin(), laid.end(), g->hitems);
        m = (uint32_t)laid.size();
    } else {
        for (uint32_t i = 0; i < g->n; ++i)
            for (uint64_t b = offsets[i]; b < offsets[i + 1]; b += per)
                g->hitems[m++] = GroupItem{b, (uint32_t)std::min<uint64_t>(per, offsets[i + 1] - b), i};
    }
    if (m) HIP_TRY(hipMemcpyAsync(g->ditems, g->hitems, m * sizeof(GroupItem), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(g->items_free, s));
    *n_items = m;
    return DYC_OK;
}

int wg_reserve_pending(dyc_wide_group *g, uint64_t extra, hipStream_t s)
{
    const uint64_t need = g->pend_n + extra;
    if (need <= g->pend_cap) return DYC_OK;
    const uint64_t cap = std::max<uint64_t>(need, g->pend_cap * 2);
    uint64_t *nk = nullptr, *nv = nullptr;
    uint32_t *nt = nullptr;
    if (hipMalloc(&nk, cap * 8) != hipSuccess || hipMalloc(&nv, cap * 8) != hipSuccess ||
        hipMalloc(&nt, cap * 4) != hipSuccess) {
        (void)hipGetLastError();
        if (nk) (void)hipFree(nk);
        if (nv) (void)hipFree(nv);
        return fail(DYC_ERR_OOM, "wide group pending list (%llu entries) failed", (unsigned long long)cap);
    }
    if (g->pend_n) {
        HIP_TRY(hipMemcpyAsync(nk, g->pk, g->pend_n * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(nv, g->pv, g->pend_n * 8, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(nt, g->ptag, g->pend_n * 4, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (g->pk) HIP_TRY(hipFree(g->pk));
    if (g->pv) HIP_TRY(hipFree(g->pv));
    if (g->ptag) HIP_TRY(hipFree(g->ptag));
    g->pk = nk;
    g->pv = nv;
    g->ptag = nt;
    g->pend_cap = cap;
    return DYC_OK;
}

int wg_prologue(dyc_wide_group *g, hipStream_t s)
{
    HIP_TRY(launch_zero(reinterpret_cast<unsigned long long *>(g->ctr), kCounterWords, s));
    HIP_TRY(launch_zero(g->tcounts, g->n, s));
    return DYC_OK;
}

int wg_read_back(dyc_wide_group *g, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(g->hctr, g->ctr, kCounterBytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(g->htcounts, g->tcounts, g->n * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    fold_tallies(g->hctr);
    return DYC_OK;
}

// An erase leaves the pending entries it deleted as zero keys: drop them with their values and
// tags (synchronous; the list is short whenever it is not empty).
int wg_compact_stash(dyc_wide_group *g, hipStream_t s)
{
    if (g->pend_n == 0) return DYC_OK;
    std::vector<uint64_t> k(g->pend_n), v(g->pend_n);
    std::vector<uint32_t> tg(g->pend_n);
    HIP_TRY(hipMemcpyAsync(k.data(), g->pk, g->pend_n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(v.data(), g->pv, g->pend_n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(tg.data(), g->ptag, g->pend_n * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    uint64_t m = 0;
    for (uint64_t i = 0; i < k.size(); ++i)
        if (k[i]) {
            k[m] = k[i];
            v[m] = v[i];
            tg[m] = tg[i];
            ++m;
        }
    if (m != g->pend_n) {
        if (m) {
            HIP_TRY(hipMemcpyAsync(g->pk, k.data(), m * 8, hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(g->pv, v.data(), m * 8, hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(g->ptag, tg.data(), m * 4, hipMemcpyHostToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        g->pend_n = m;
    }
    g->st.pending = g->pend_n;
    return DYC_OK;
}

// ------------------------------------------------------------------ bulk fill
int wg_bulk_alloc(void **p, size_t bytes, const char *what)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(DYC_ERR_OOM, "hipMalloc of the wide group bulk-fill %s (%zu bytes) failed", what, bytes);
    }
    return DYC_OK;
}

// Grow-only workspace of a laid-out batch (the caller has synchronized the stream).
int wg_reserve_bulk(dyc_wide_group *g, const dyc::host::BulkLayout &L)
{
    const uint64_t n = L.total, blocks = L.blocks.size(), chunks = L.chunks;
    const uint32_t tabs = (uint32_t)L.tabs.size();
    int rc = DYC_OK;
    auto regrow = [&](void **p, size_t bytes, const char *what) {
        if (rc) return;
        if (*p) (void)hipFree(*p);
        rc = wg_bulk_alloc(p, bytes, what);
    };
    auto regrow_pinned = [&](void **p, size_t bytes) {
        if (rc) return;
        if (*p) (void)hipHostFree(*p);
        *p = nullptr;
        if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            rc = fail(DYC_ERR_OOM, "pinned wide group bulk-fill staging (%zu bytes) failed", bytes);
        }
    };
    if (n > g->bk_cap) {
        g->bk_cap = 0;
        regrow((void **)&g->bk_ent, n * sizeof(WideEntry), "entries");
        regrow((void **)&g->bk_ovf, n * sizeof(WideEntry), "overflow list");
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

// The whole wide bulk fill of a laid-out batch on s, rounds chained on the device as in the
// 32-bit group fill (dyc_group.hip): round 1, round 2 for the other bucket, then the eviction
// insert of what overflows that (its entries split into key and value arrays first).
hipError_t wg_enqueue_bulk(dyc_wide_group *g, uint32_t NB, uint32_t J, uint32_t chunks, const uint64_t *keys,
                           const uint64_t *vals, uint64_t total, bool zeroed, const WideSink &sink, hipStream_t s)
{
    GroupItem *const blk1 = g->bk_blk, *const blk2 = g->bk_blk + NB, *const items = g->bk_blk + 2 * NB;
    BulkTable *const tabs1 = g->bk_tabs, *const tabs2 = g->bk_tabs + J;
    unsigned long long *const ovf1 = g->bk_ovfn, *const ovf2 = g->bk_ovfn + J;
    uint8_t *const no_cid = nullptr;
    const SegLayout lay1{blk1, g->bk_partial, total}, lay2{blk2, g->bk_partial, total};
    hipError_t e;
    wide_group_bulk_pack_kernel<false><<<NB, kCountThreads, 0, s>>>(g->ddir, blk1, keys, nullptr, g->bk_lcid,
                                                                    g->bk_partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_group_bulk_scan(g->ddir, tabs1, J, g->bk_partial, g->bk_chunks, s)) != hipSuccess)
        return e;
    part_scatter_kernel<WideEntry, WideKeyValSource, uint8_t, kWideSortThreads, -1, SegLayout>
        <<<NB, kWideSortThreads, 0, s>>>(WideKeyValSource{keys, vals}, g->bk_lcid, lay1, RadixPass{0, 8}, g->bk_ent,
                                         no_cid);
    if ((e = hipGetLastError()) != hipSuccess || (e = launch_zero(ovf1, 2 * (uint64_t)J, s)) != hipSuccess) return e;
    wide_group_bulk_fill_kernel<true, 1024><<<chunks, 1024, 0, s>>>(g->bk_chunks, g->bk_ent, g->bk_ovf, ovf1,
                                                                   g->tcounts, g->ctr, zeroed);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = launch_group_bulk_clip(blk1, g->bk_btab, NB, tabs1, J, ovf1, blk2, tabs2, s)) != hipSuccess)
        return e;
    wide_group_bulk_pack_kernel<true><<<NB, kCountThreads, 0, s>>>(g->ddir, blk2, nullptr, g->bk_ovf, g->bk_lcid,
                                                                   g->bk_partial);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = launch_group_bulk_scan(g->ddir, tabs2, J, g->bk_partial, g->bk_chunks, s)) != hipSuccess)
        return e;
    part_scatter_kernel<WideEntry, EntrySource<WideEntry>, uint8_t, kWideSortThreads, -1, SegLayout>
        <<<NB, kWideSortThreads, 0, s>>>(EntrySource<WideEntry>{g->bk_ovf}, g->bk_lcid, lay2, RadixPass{0, 8},
                                         g->bk_ent, no_cid);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    wide_group_bulk_fill_kernel<false, 256><<<chunks, 256, 0, s>>>(g->bk_chunks, g->bk_ent, g->bk_ovf, ovf2,
                                                                   g->tcounts, g->ctr, false);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = launch_group_bulk_clip(blk1, g->bk_btab, NB, tabs1, J, ovf2, items, nullptr, s)) != hipSuccess)
        return e;
    // the entry buffer (16 B per key) is free now: keys, then values, at the overflow's indexes
    uint64_t *const sk = reinterpret_cast<uint64_t *>(g->bk_ent);
    uint64_t *const sv = sk + total;
    wide_items_split_kernel<<<NB, 256, 0, s>>>(items, g->bk_ovf, sk, sv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_wide_group_insert(g->ddir, items, NB, sk, sv, g->policy(), (int)g->cfg.insert_blocks, g->ctr, sink,
                                    g->tcounts, s);
}

// insert_order 3 for a wide group (see group_bulk_insert in dyc_group.hip). Declines (*n_tabs 0)
// when the batch does not qualify or the workspace cannot be allocated; counters and per-table
// counts are zeroed here; the last round's overflow counts (*n_tabs) are in bk_hovfn after the
// caller's wait.
int wg_bulk_insert(dyc_wide_group *g, const uint64_t *keys, const uint64_t *vals, const uint64_t *offsets,
                   bool zeroed, hipStream_t s, uint64_t room, uint32_t *n_tabs)
{
    *n_tabs = 0;
    dyc::host::BulkLayout L;
    if (!dyc::host::bulk_layout(g->n, offsets, [g](uint32_t i) -> const Desc & { return g->hdir[i].D; }, &L))
        return DYC_OK;
    {
        HIP_TRY(hipStreamSynchronize(s));  // buffers being replaced are no longer in use
        const dyc::host::ErrorKeep keep;
        const int rc = wg_reserve_bulk(g, L);
        if (rc == DYC_ERR_OOM) {  // no room for the workspace: the eviction insert needs none
            keep.restore();
            return DYC_OK;
        }
        if (rc) return rc;
    }
    const uint32_t NB = (uint32_t)L.blocks.size(), J = (uint32_t)L.tabs.size();
    memcpy(g->bk_hblk, L.blocks.data(), NB * sizeof(GroupItem));
    memcpy(g->bk_hbtab, L.btab.data(), NB * sizeof(uint32_t));
    memcpy(g->bk_htabs, L.tabs.data(), J * sizeof(BulkTable));
    HIP_TRY(hipMemcpyAsync(g->bk_blk, g->bk_hblk, NB * sizeof(GroupItem), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(g->bk_btab, g->bk_hbtab, NB * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(g->bk_tabs, g->bk_htabs, J * sizeof(BulkTable), hipMemcpyHostToDevice, s));
    const int rc = wg_prologue(g, s);
    if (rc) return rc;
    HIP_TRY(wg_enqueue_bulk(g, NB, J, L.chunks, keys, vals, L.total, zeroed,
                            WideSink{g->pk + g->pend_n, g->pv + g->pend_n, room, g->ptag + g->pend_n}, s));
    HIP_TRY(hipMemcpyAsync(g->bk_hovfn, g->bk_ovfn + J, J * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    *n_tabs = J;
    return DYC_OK;
}

}  // namespace

extern "C" {

int dyc_wide_group_destroy(dyc_wide_group *g)
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
    if (g->pk) (void)hipFree(g->pk);
    if (g->pv) (void)hipFree(g->pv);
    if (g->ptag) (void)hipFree(g->ptag);
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

int dyc_wide_group_create(const dyc_config *cfg_in, uint32_t n_tables, const uint64_t *hints, dyc_wide_group **out)
{
    if (!out || !hints || n_tables == 0) return fail(DYC_ERR_INVALID, "bad group arguments");
    *out = nullptr;
    dyc_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        dyc_config_init(&cfg);
    if (!wide_check_config(cfg)) return fail(DYC_ERR_INVALID, "bad configuration (block_threads, fill bounds or subtables)");
    cfg.subtables = config_subtables(cfg);
    const uint32_t d = cfg.subtables;
    std::vector<uint32_t> nbs(n_tables);
    size_t bytes = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        nbs[i] = initial_buckets(hints[i], d);
        if (nbs[i] >= kMaxBuckets) return fail(DYC_ERR_INVALID, "table %u: capacity hint too large", i);
        bytes += (size_t)d * 2 * ((wide_bytes(nbs[i]) + 255) & ~(size_t)255);
    }
    dyc_wide_group *g = new dyc_wide_group();
    g->cfg = cfg;
    g->n = n_tables;
    g->hdir.resize(n_tables);
    g->count.assign(n_tables, 0);
    auto cleanup = [&](int code) {
        char keep[512];
        strncpy(keep, dyc_last_error(), sizeof keep - 1);
        keep[sizeof keep - 1] = 0;
        dyc_wide_group_destroy(g);
        fail(code, "%s", keep);
        return code;
    };
    HIP_TRY(hipGetDevice(&g->device));
    if (hipMalloc(&g->arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(DYC_ERR_OOM, "wide group arena of %zu bytes failed", bytes);
        return cleanup(DYC_ERR_OOM);
    }
    g->arena_bytes = bytes;
    size_t off = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        desc_init(g->hdir[i].D, d);
        for (uint32_t t = 0; t < d; ++t) {
            const size_t a = (wide_bytes(nbs[i]) + 255) & ~(size_t)255;
            desc_set(g->hdir[i].D, (int)t, reinterpret_cast<uint64_t *>(g->arena + off), nbs[i]);
            g->hdir[i].val[t] = reinterpret_cast<uint64_t *>(g->arena + off + a);
            off += 2 * a;
        }
    }
    if (hipMalloc(&g->ddir, n_tables * sizeof(WideDesc)) != hipSuccess || hipMalloc(&g->ctr, kCounterBytes) != hipSuccess ||
        hipMalloc(&g->tcounts, n_tables * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->hctr), kCounterBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&g->htcounts), n_tables * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&g->items_free, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        fail(DYC_ERR_OOM, "wide group bookkeeping allocation failed");
        return cleanup(DYC_ERR_OOM);
    }
    if (hipMemcpy(g->ddir, g->hdir.data(), n_tables * sizeof(WideDesc), hipMemcpyHostToDevice) != hipSuccess ||
        launch_zero(reinterpret_cast<unsigned long long *>(g->arena), bytes / 8, nullptr) != hipSuccess ||
        hipStreamSynchronize(nullptr) != hipSuccess) {
        fail(DYC_ERR_HIP, "wide group initialisation failed");
        return cleanup(DYC_ERR_HIP);
    }
    int rc = wg_reserve_pending(g, 65536, nullptr);
    if (rc) return cleanup(rc);
    g->zeroed = true;
    *out = g;
    return DYC_OK;
}

int dyc_wide_group_insert(dyc_wide_group *g, const uint64_t *keys, const uint64_t *vals, const uint64_t *offsets,
                          dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = wg_check_offsets(g, offsets);
    if (rc) return rc;
    const uint64_t total = offsets[g->n];
    if (total && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    const bool zeroed = g->zeroed;
    g->zeroed = false;
    g->st.last_filled = 0;
    rc = wg_reserve_pending(g, std::max<uint64_t>(65536, total / 16), s);
    if (rc) return rc;
    const uint64_t room = g->pend_cap - g->pend_n;
    uint32_t n_tabs = 0;
    if (g->cfg.insert_order == 3) {
        rc = wg_bulk_insert(g, keys, vals, offsets, zeroed, s, room, &n_tabs);
        if (rc) return rc;
    }
    if (!n_tabs) {
        uint32_t n_items = 0;
        rc = wg_upload_items(g, offsets, s, &n_items);
        if (rc) return rc;
        rc = wg_prologue(g, s);
        if (rc) return rc;
        HIP_TRY(launch_wide_group_insert(g->ddir, g->ditems, n_items, keys, vals, g->policy(),
                                         (int)g->cfg.insert_blocks, g->ctr,
                                         WideSink{g->pk + g->pend_n, g->pv + g->pend_n, room, g->ptag + g->pend_n},
                                         g->tcounts, s));
    }
    rc = wg_read_back(g, s);
    if (rc) return rc;
    const Counters &c = *g->hctr;
    if (n_tabs) {
        uint64_t left = 0;
        for (uint32_t j = 0; j < n_tabs; ++j) left += g->bk_hovfn[j];
        g->st.last_filled = total - c.rejected - left;
    }
    g->st.kicks += c.kicks;
    g->st.last_kicks = c.kicks;
    g->st.cas_retries += c.retries;
    g->st.kick_failures += c.pending;
    g->st.rejected += c.rejected;
    g->st.lost += c.lost;
    g->pend_n += std::min<uint64_t>(c.pending, room);
    g->st.pending = g->pend_n;
    for (uint32_t i = 0; i < g->n; ++i) g->count[i] += g->htcounts[i];
    if (c.lost) return fail(DYC_ERR_LOST, "wide group pending list overflow: %llu entries lost", (unsigned long long)c.lost);
    if (c.pending)
        return fail(DYC_ERR_PENDING, "%llu entries exhausted the eviction bound (group tables do not resize)",
                    (unsigned long long)c.pending);
    return DYC_OK;
}

int dyc_wide_group_find(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint64_t *vals, uint8_t *hit,
                        dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = wg_check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    uint32_t n_items = 0;
    rc = wg_upload_items(g, offsets, s, &n_items, offsets[g->n] >= dyc::host::kXcdLayoutKeys);
    if (rc) return rc;
    HIP_TRY(launch_wide_group_probe(g->ddir, g->ditems, n_items, keys, vals, hit, false, nullptr, (int)g->cfg.max_blocks,
                                    g->stash(), s));
    return DYC_OK;
}

int dyc_wide_group_erase(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint8_t *deleted,
                         dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = wg_check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && !keys) return fail(DYC_ERR_INVALID, "null keys");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    uint32_t n_items = 0;
    rc = wg_upload_items(g, offsets, s, &n_items);
    if (rc) return rc;
    rc = wg_prologue(g, s);
    if (rc) return rc;
    HIP_TRY(launch_wide_group_probe(g->ddir, g->ditems, n_items, keys, nullptr, deleted, true, g->tcounts,
                                    (int)g->cfg.max_blocks, g->stash(), s));
    rc = wg_read_back(g, s);
    if (rc) return rc;
    for (uint32_t i = 0; i < g->n; ++i) g->count[i] -= std::min<uint64_t>(g->count[i], g->htcounts[i]);
    return wg_compact_stash(g, s);
}

// The one-launch-per-table forms (the baseline the group launch is measured against): the
// single-table kernels once per table, on the group's own tables and pending list.
int dyc_wide_group_insert_per_table(dyc_wide_group *g, const uint64_t *keys, const uint64_t *vals,
                                    const uint64_t *offsets, dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = wg_check_offsets(g, offsets);
    if (rc) return rc;
    if (offsets[g->n] && (!keys || !vals)) return fail(DYC_ERR_INVALID, "null keys/vals");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    g->zeroed = false;
    g->st.last_filled = 0;
    rc = wg_reserve_pending(g, std::max<uint64_t>(65536, offsets[g->n] / 16), s);
    if (rc) return rc;
    rc = wg_prologue(g, s);
    if (rc) return rc;
    const uint64_t room = g->pend_cap - g->pend_n;
    const LaunchShape sh{256, (int)g->cfg.insert_blocks};
    // one launch per table; all share the counters, so pended entries append one after another,
    // each tagged with its launch's table; tcounts[i] receives table i's reserved keys
    for (uint32_t i = 0; i < g->n; ++i) {
        const uint64_t m = offsets[i + 1] - offsets[i];
        if (!m) continue;
        HIP_TRY(launch_wide_insert(g->hdir[i], keys + offsets[i], vals + offsets[i], m, kNoTable, false, g->policy(),
                                   sh, g->ctr,
                                   WideSink{g->pk + g->pend_n, g->pv + g->pend_n, room, g->ptag + g->pend_n, i,
                                            g->tcounts + i},
                                   s));
    }
    rc = wg_read_back(g, s);
    if (rc) return rc;
    const Counters &c = *g->hctr;
    g->st.kicks += c.kicks;
    g->st.last_kicks = c.kicks;
    g->st.cas_retries += c.retries;
    g->st.kick_failures += c.pending;
    g->st.rejected += c.rejected;
    g->st.lost += c.lost;
    g->pend_n += std::min<uint64_t>(c.pending, room);
    g->st.pending = g->pend_n;
    for (uint32_t i = 0; i < g->n; ++i) g->count[i] += (offsets[i + 1] - offsets[i]) - g->htcounts[i];
    if (c.lost) return fail(DYC_ERR_LOST, "wide group pending list overflow: %llu entries lost", (unsigned long long)c.lost);
    if (c.pending)
        return fail(DYC_ERR_PENDING, "%llu entries exhausted the eviction bound (group tables do not resize)",
                    (unsigned long long)c.pending);
    return DYC_OK;
}

int dyc_wide_group_find_per_table(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint64_t *vals,
                                  uint8_t *hit, dyc_stream stream)
{
    if (!g) return fail(DYC_ERR_INVALID, "null group");
    int rc = wg_check_offsets(g, offsets);
    if (rc) return rc;
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    const LaunchShape sh{256, (int)g->cfg.max_blocks};
    for (uint32_t i = 0; i < g->n; ++i) {
        const uint64_t m = offsets[i + 1] - offsets[i];
        if (!m) continue;
        WideStash st = g->stash();
        st.table = i;
        HIP_TRY(launch_wide_probe(g->hdir[i], keys + offsets[i], vals ? vals + offsets[i] : nullptr,
                                  hit ? hit + offsets[i] : nullptr, m, false, sh, g->ctr, st, s));
    }
    return DYC_OK;
}

int dyc_wide_group_table_stats(const dyc_wide_group *g, uint32_t table, dyc_stats *out)
{
    if (!g || !out || table >= g->n) return fail(DYC_ERR_INVALID, "bad argument");
    *out = g->st;
    out->count = g->count[table];
    uint64_t cap = 0;
    for (uint32_t t = 0; t < g->hdir[table].D.d; ++t) {
        out->buckets[t] = g->hdir[table].D.nb[t];
        cap += (uint64_t)g->hdir[table].D.nb[t] * kSlots;
    }
    out->capacity = cap;
    out->subtables = g->hdir[table].D.d;
    return DYC_OK;
}

int dyc_wide_group_dump(dyc_wide_group *g, uint32_t table, uint32_t sub, uint64_t *host_keys, uint64_t *host_vals,
                        uint64_t n_slots, dyc_stream stream)
{
    if (!g || table >= g->n || sub >= g->hdir[table].D.d || (!host_keys && !host_vals))
        return fail(DYC_ERR_INVALID, "bad argument");
    const uint64_t slots = (uint64_t)g->hdir[table].D.nb[sub] * kSlots;
    if (n_slots < slots) return fail(DYC_ERR_INVALID, "buffer too small");
    DeviceGuard dg(g->device);
    hipStream_t s = (hipStream_t)stream;
    if (host_keys) HIP_TRY(hipMemcpyAsync(host_keys, g->hdir[table].D.sub[sub], slots * 8, hipMemcpyDeviceToHost, s));
    if (host_vals) HIP_TRY(hipMemcpyAsync(host_vals, g->hdir[table].val[sub], slots * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DYC_OK;
}

int dyc_wide_group_pending(dyc_wide_group *g, uint64_t *host_keys, uint64_t *host_vals, uint32_t *host_tables,
                           uint64_t cap, uint64_t *n_out, dyc_stream stream)
{
    if (!g || !n_out) return fail(DYC_ERR_INVALID, "null argument");
    *n_out = g->pend_n;
    const uint64_t m = std::min(cap, g->pend_n);
    if (m) {
        DeviceGuard dg(g->device);
        hipStream_t s = (hipStream_t)stream;
        if (host_keys) HIP_TRY(hipMemcpyAsync(host_keys, g->pk, m * 8, hipMemcpyDeviceToHost, s));
        if (host_vals) HIP_TRY(hipMemcpyAsync(host_vals, g->pv, m * 8, hipMemcpyDeviceToHost, s));
        if (host_tables) HIP_TRY(hipMemcpyAsync(host_tables, g->ptag, m * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return DYC_OK;
}

int dyc_wide_group_clear(dyc_wide_group *g, dyc_stream stream)
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
