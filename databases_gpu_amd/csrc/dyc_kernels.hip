// dyc_kernels.hip -- gfx950 kernels of the DyCuckoo hot path.
//
// Wave64 mapping shared by find and erase ("probe" kernels):
//   * a wave owns a tile of 64 consecutive keys; lane i loads key i (one coalesced 256 B read)
//     and hashes it (pair, both subtables, both bucket addresses) -- no lane hashes twice;
//   * the wave splits into 64/G groups of G lanes; group g probes keys g*G .. g*G+G-1 of the
//     tile, each lane reading 128/G bytes of every bucket (G = 8: one dwordx4 per lane, so a
//     group reads a whole 128-byte bucket in one coalesced line). All G pos1 probes of a group
//     are issued before the first compare, so a lane keeps G bucket loads in flight;
//   * a 64-bit __ballot per probe gives every group its match mask; the group's answer is
//     broadcast to its lanes, and lane g*G+s keeps the answer of key g*G+s -- the lane that
//     loaded that key -- so results are stored with one coalesced write per tile.
// pos2 is read only for keys pos1 did not resolve, in the reference order (cuckoo.cu:482-497).
//
// Insert: each group of G lanes carries one entry through the eviction chain with 64-bit CAS
// on the victim slot. Groups are independent: a group that finishes takes its next key while
// the rest of the wave keeps kicking, so no lane idles behind a long chain.
#include "dyc_kernels.hpp"
#include "dyc_wave.hpp"

#include <algorithm>

namespace dyc {

namespace {


// A bucket location packed in 32 bits: subtable in the top 3 bits, bucket index below (bucket
// counts are capped at 2^29). Shuffling this instead of a 64-bit address halves the cross-lane
// traffic, and rebuilding the pointer from the descriptor keeps the loads in the global
// address space (a pointer rebuilt from a shuffled integer would become a flat access, whose
// completion the compiler can only track together with the shuffles).
__device__ inline uint32_t pack_loc(uint32_t t, uint32_t bucket) { return (t << 29) | bucket; }
__device__ inline uint64_t *loc_ptr(const Desc &D, uint32_t l)
{
    return selt(l >> 29, D.sub) + (uint64_t)(l & 0x1fffffffu) * kSlots;
}

// Plain (non-atomic) read of SPL consecutive slots, 16-byte vector loads where possible.
template <int SPL>
__device__ inline void load_slots(const uint64_t *p, uint64_t (&e)[SPL])
{
    if constexpr (SPL == 1) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        e[0] = ((uint64_t)v.y << 32) | v.x;
    } else {
#pragma unroll
        for (int j = 0; j < SPL; j += 2) {
            uint4 v = *reinterpret_cast<const uint4 *>(p + j);
            e[j] = ((uint64_t)v.y << 32) | v.x;
            e[j + 1] = ((uint64_t)v.w << 32) | v.z;
        }
    }
}

// Device-coherent read of SPL slots (agent-scope relaxed atomic loads, which skip lines another
// XCD's writes may have left stale in this XCD's L2). For reads that must see exchanges made
// earlier in the same launch by other waves.
template <int SPL>
__device__ inline void load_slots_coherent(const uint64_t *p, uint64_t (&e)[SPL])
{
#pragma unroll
    for (int j = 0; j < SPL; ++j) e[j] = __hip_atomic_load(p + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int SPL, bool COH>
__device__ inline void read_slots(const uint64_t *p, uint64_t (&e)[SPL])
{
    if constexpr (COH)
        load_slots_coherent<SPL>(p, e);
    else
        load_slots<SPL>(p, e);
}

// An input word (key or value). COH: read device-coherently, for a launch that outlives the
// caller's writes to its inputs (a buffer read earlier in the launch may be rewritten and
// handed in again, and this XCD's L2 may still hold the old line).
template <bool COH>
__device__ inline uint32_t load_in(const uint32_t *p)
{
    if constexpr (COH)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}

// A result word or flag. COH: a system-scope store, which leaves no dirty line in this XCD's
// L2, so a launch that keeps running can publish results by waiting for its stores to complete
// (publish_wait) instead of writing the whole L2 back.
template <bool COH, typename T>
__device__ inline void store_out(T *p, T v)
{
    if constexpr (COH)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        *p = v;
}

// Wait until this wave's earlier memory operations have completed; the compiler keeps memory
// accesses on their side of it. Used after system-scope stores and atomics only, which need no
// cache write-back to be seen by other agents.
__device__ inline void publish_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }


// ---------------------------------------------------------------- find / erase / update

// Erase and update write-back of one probe phase. Bit kk * SPL + j of mm says that this lane's slot
// j of key kk's bucket slice holds key kk; each such slot is CAS'd from what was read to `rep` (0
// for erase, the key with its new value for update). Returns bit kk set where one of this lane's
// CASes on key kk succeeded. Up to kCasInFlight CASes are issued before any result is compared, so
// a wave waits about once per phase instead of once per key of its groups; a lane with more matches
// (a duplicated key, or several of its group's keys in its slice) takes the rest one at a time.
constexpr int kCasInFlight = 3;

template <int G, int SPL, bool UPDATE>
__device__ inline void pick_match(const Desc &D, int b, const uint64_t (&e)[G][SPL], const uint32_t (&loc)[G],
                                  const uint32_t (&key)[G], const uint32_t *nv, int s, uint64_t *&ptr,
                                  uint64_t &exp, uint64_t &rep, uint32_t &kk_out)
{
    const int kk = b / SPL, j = b % SPL;
    uint64_t x = 0;
    uint32_t l = 0, k = 0, v = 0;
#pragma unroll
    for (int q = 0; q < G; ++q) {
        if (q == kk) {
            l = loc[q];
            k = key[q];
            if constexpr (UPDATE) v = nv[q];
        }
#pragma unroll
        for (int r = 0; r < SPL; ++r)
            if (q * SPL + r == b) x = e[q][r];
    }
    ptr = loc_ptr(D, l) + s * SPL + j;
    exp = x;
    rep = UPDATE ? pack(k, v) : 0ull;
    kk_out = (uint32_t)kk;
}

template <int G, int SPL, bool UPDATE>
__device__ inline uint32_t write_matches(const Desc &D, uint32_t mm, const uint64_t (&e)[G][SPL],
                                         const uint32_t (&loc)[G], const uint32_t (&key)[G], const uint32_t *nv,
                                         int s)
{
    uint64_t found[kCasInFlight];
    int bit[kCasInFlight];  // the match each in-flight CAS serves (-1: none)
#pragma unroll
    for (int q = 0; q < kCasInFlight; ++q) {
        found[q] = 0;
        bit[q] = -1;
        if (mm) {
            bit[q] = __builtin_ctz(mm);
            mm &= mm - 1u;
            uint64_t *ptr, x, rep;
            uint32_t kk;
            pick_match<G, SPL, UPDATE>(D, bit[q], e, loc, key, nv, s, ptr, x, rep, kk);
            found[q] = cas_old(ptr, x, rep);
        }
    }
    // The compiler would otherwise turn found == expected into the CAS's own success flag and
    // test it (waiting for the CAS) before the next one issues; an empty asm hides the relation.
#pragma unroll
    for (int q = 0; q < kCasInFlight; ++q) asm volatile("" : "+v"(found[q]));
    uint32_t ok = 0;
#pragma unroll
    for (int q = 0; q < kCasInFlight; ++q) {
        if (bit[q] < 0) continue;
        uint64_t x = 0;  // the expected value, selected again rather than held across the wait
#pragma unroll
        for (int kk = 0; kk < G; ++kk)
#pragma unroll
            for (int j = 0; j < SPL; ++j)
                if (kk * SPL + j == bit[q]) x = e[kk][j];
        if (found[q] == x) ok |= 1u << (bit[q] / SPL);
    }
    while (mm) {
        const int b = __builtin_ctz(mm);
        mm &= mm - 1u;
        uint64_t *ptr, x, rep;
        uint32_t kk;
        pick_match<G, SPL, UPDATE>(D, b, e, loc, key, nv, s, ptr, x, rep, kk);
        if (cas_old(ptr, x, rep) == x) ok |= 1u << kk;
    }
    return ok;
}
// One wave probes the 64 keys keys[base, base + 64) clipped to n, in table D. Returns this
// lane's flag for its own key (hit, deleted or updated); all 64 lanes must call it together.
// MODE 0 find: vals (optional) receives the value. MODE 1 erase: every copy of the key in the
// first bucket (pos1, then pos2) that holds it is CAS'd to empty. MODE 2 update: those copies
// are CAS'd to the new value vals[i] (an input). COH: read the buckets and the inputs
// device-coherently and write the results with system-scope stores (load_slots_coherent,
// load_in, store_out).
template <int G, int MODE, bool COH = false>
__device__ inline uint32_t probe_tile_m(const Desc &D, const uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                        uint8_t *__restrict__ flags, uint64_t base, uint64_t n)
{
    constexpr bool ERASE = MODE != 0;  // erase and update both rewrite the matching slots
    constexpr bool UPDATE = MODE == 2;
    constexpr int SPL = kSlots / G;  // slots each lane reads per bucket
    constexpr uint64_t GMASK = (G >= 64) ? ~0ull : ((1ull << G) - 1ull);
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    {
        const uint64_t i = base + lane;
        const bool inb = i < n;
        const uint32_t k = inb ? load_in<COH>(keys + i) : 0u;  // past the end: key 0, which is reserved
        const uint32_t nv = (UPDATE && inb) ? load_in<COH>(vals + i) : 0u;
        const uint32_t p = pair_of(D, k);
        const uint32_t t1 = pos1(D, p), t2 = pos2(D, p);
        const uint32_t loc1 = pack_loc(t1, bucket_of(D, k, t1));
        const uint32_t loc2 = pack_loc(t2, bucket_of(D, k, t2));

        uint64_t e[G][SPL];
        uint32_t kk_key[G];
        uint32_t kk_new[UPDATE ? G : 1];  // update: the new value of key kk
        bool kk_valid[G];
        uint32_t kk_loc[ERASE ? G : 1];  // erase writes back to the bucket it matched in
        // phase 1: pos1 buckets of the group's G keys, all loads in flight together
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            const int src = g * G + kk;
            kk_key[kk] = __shfl(k, src);
            if constexpr (UPDATE) kk_new[kk] = __shfl(nv, src);
            kk_valid[kk] = !reserved_key(kk_key[kk]);
            const uint32_t l = __shfl(loc1, src);
            if constexpr (ERASE) kk_loc[kk] = l;
#pragma unroll
            for (int j = 0; j < SPL; ++j) e[kk][j] = 0;
            if (kk_valid[kk]) read_slots<SPL, COH>(loc_ptr(D, l) + s * SPL, e[kk]);
        }

        uint32_t my_val = 0, my_flag = 0;
        uint32_t done = 0;  // bit kk: key kk resolved in pos1
        uint32_t mm = 0;    // erase/update: this lane's matching slots (write_matches)
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            int first = -1;
#pragma unroll
            for (int j = SPL - 1; j >= 0; --j)
                if (key_of(e[kk][j]) == kk_key[kk]) first = j;
            const uint32_t gm = (uint32_t)((__ballot(first >= 0 && kk_valid[kk]) >> (g * G)) & GMASK);
            if (gm) {
                done |= 1u << kk;
                if constexpr (ERASE) {
#pragma unroll
                    for (int j = 0; j < SPL; ++j)
                        if (kk_valid[kk] && key_of(e[kk][j]) == kk_key[kk]) mm |= 1u << (kk * SPL + j);
                } else {
                    const int w = __builtin_ctz(gm);
                    uint32_t v = 0;
#pragma unroll
                    for (int j = 0; j < SPL; ++j)
                        if (j == first) v = val_of(e[kk][j]);
                    v = __shfl(v, g * G + w);
                    if (kk == s) {
                        my_val = v;
                        my_flag = 1;
                    }
                }
            }
        }

        if constexpr (ERASE) {
            const uint32_t ok = write_matches<G, SPL, UPDATE>(D, mm, e, kk_loc, kk_key, kk_new, s);
#pragma unroll
            for (int kk = 0; kk < G; ++kk) {
                if (!(done & (1u << kk))) continue;  // group-uniform
                const uint32_t okm = (uint32_t)((__ballot((ok >> kk) & 1u) >> (g * G)) & GMASK);
                if (kk == s) my_flag = okm != 0;
            }
            mm = 0;
        }

        // phase 2: pos2 buckets of the keys pos1 did not resolve
        [[maybe_unused]] uint32_t done2 = 0;  // erase/update: bit kk: key kk found in pos2
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            const uint32_t l = __shfl(loc2, g * G + kk);
            if constexpr (ERASE) kk_loc[kk] = l;
            if (kk_valid[kk] && !(done & (1u << kk))) read_slots<SPL, COH>(loc_ptr(D, l) + s * SPL, e[kk]);
        }
        // Straight-line over every key (a key pos1 resolved is masked out, not skipped), so the
        // wait for the pos2 slots falls on every path before write_matches issues its CASes.
#pragma unroll
        for (int kk = 0; kk < G; ++kk) {
            const bool todo = kk_valid[kk] && !(done & (1u << kk));
            int first = -1;
#pragma unroll
            for (int j = SPL - 1; j >= 0; --j)
                if (key_of(e[kk][j]) == kk_key[kk]) first = j;
            const uint32_t gm = (uint32_t)((__ballot(first >= 0 && todo) >> (g * G)) & GMASK);
            if constexpr (ERASE) {
                done2 |= (gm != 0 ? 1u : 0u) << kk;
#pragma unroll
                for (int j = 0; j < SPL; ++j)
                    mm |= (todo && key_of(e[kk][j]) == kk_key[kk] ? 1u : 0u) << (kk * SPL + j);
            } else {
                if (!gm) continue;
                const int w = __builtin_ctz(gm);
                uint32_t v = 0;
#pragma unroll
                for (int j = 0; j < SPL; ++j)
                    if (j == first) v = val_of(e[kk][j]);
                v = __shfl(v, g * G + w);
                if (kk == s) {
                    my_val = v;
                    my_flag = 1;
                }
            }
        }

        if constexpr (ERASE) {
            const uint32_t ok = write_matches<G, SPL, UPDATE>(D, mm, e, kk_loc, kk_key, kk_new, s);
#pragma unroll
            for (int kk = 0; kk < G; ++kk) {
                if (!(done2 & (1u << kk))) continue;  // group-uniform
                const uint32_t okm = (uint32_t)((__ballot((ok >> kk) & 1u) >> (g * G)) & GMASK);
                if (kk == s) my_flag = okm != 0;
            }
        }

        if (inb) {
            if (!ERASE && vals) store_out<COH>(vals + i, my_val);
            if (flags) store_out<COH>(flags + i, (uint8_t)my_flag);
        }
        return my_flag;
    }
}

template <int G, bool ERASE, bool COH = false>
__device__ inline uint32_t probe_tile(const Desc &D, const uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                      uint8_t *__restrict__ flags, uint64_t base, uint64_t n)
{
    return probe_tile_m<G, ERASE ? 1 : 0, COH>(D, keys, vals, flags, base, n);
}

template <int MODE, bool COH = false>
__device__ inline uint32_t pool_probe_at(const Pool &pool, uint64_t np, const uint32_t *keys, uint32_t *vals,
                                         uint8_t *flags, uint64_t n, uint32_t table, uint64_t base, uint32_t f);

// Live entries of a pool (0 for Pool{}: a table without one).
__device__ inline uint64_t pool_size(const Pool &pool)
{
    if (!pool.count) return pool.n < pool.cap ? pool.n : pool.cap;
    const unsigned long long pc = __hip_atomic_load(pool.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return pc < pool.cap ? pc : pool.cap;
}

// Waves per SIMD the find / erase kernels are held to (the occupancy their register use gives
// with a four-subtable selection; the six-way one would otherwise cost a wave at G = 8). Erase
// keeps up to three CASes per lane in flight (write_matches), which costs G = 4 its sixth wave.
template <int G, bool ERASE>
constexpr int probe_waves()
{
    return G == 4 ? (ERASE ? 5 : 7) : (G == 8 ? (ERASE ? 5 : 7) : (ERASE ? 4 : 5));
}

// Misses also look in the pool's entries of `table` (a serving table's overflow pool, a single
// table's stash); an empty pool costs nothing.
template <int G, bool ERASE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(probe_waves<G, ERASE>()))) void probe_kernel(Desc D, const uint32_t *__restrict__ keys,
                                                    uint32_t *__restrict__ vals, uint8_t *__restrict__ flags,
                                                    uint64_t n, Counters *__restrict__ C, Pool pool, uint32_t table)
{
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t np = pool_size(pool);
    unsigned long long ndel = 0;
    const uint32_t slot = wave_slot();
    for (uint64_t base = (uint64_t)slot * kWave; base < n; base += waves * kWave) {
        uint32_t f = probe_tile<G, ERASE>(D, keys, vals, flags, base, n);
        if (np) f |= pool_probe_at<ERASE ? 1 : 0>(pool, np, keys, vals, flags, n, table, base, f);
        ndel += f;
    }
    if constexpr (ERASE) wave_tally(C, &C->deleted, ndel, slot);
}

// Upsert's first pass: rewrite the value of keys already stored (probe_tile_m mode 2).
template <int G>
__global__ __launch_bounds__(256) void update_kernel(Desc D, const uint32_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals, uint8_t *__restrict__ updated,
                                                     uint64_t n, Counters *__restrict__ C, Pool pool)
{
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t np = pool_size(pool);
    unsigned long long nu = 0;
    const uint32_t slot = wave_slot();
    for (uint64_t base = (uint64_t)slot * kWave; base < n; base += waves * kWave) {
        uint32_t f = probe_tile_m<G, 2>(D, keys, vals, updated, base, n);
        if (np) f |= pool_probe_at<2>(pool, np, keys, vals, updated, n, 0, base, f);
        nu += f;
    }
    wave_tally(C, &C->updated, nu, slot);
}

// ------------------------------------------------------------------------------- insert
// One group of G lanes per in-hand entry. Per iteration (cuckoo.cu:377-421):
//   t = the other subtable of the entry's pair relative to the previous one; read the bucket;
//   victim = first empty slot, else (first kick_bound iterations) the first slot whose resident
//   pair index differs from t (cuckoo.cu:398-400), else slot (iteration & 15); after kick_bound
//   iterations the victim is rescue_lane(). CAS the victim slot from the value read to the
//   entry: on success carry the old value on (done if it was empty).
// The bucket is read with plain vector loads. The view may be stale (another XCD's L2 may hold
// a newer line), but only the CAS decides: a stale view can make a CAS fail, never lose or
// duplicate an entry. A failed CAS returns the slot's current value, which replaces that slot in
// the view; the group then chooses again without re-reading, so every retry learns one slot's
// true value and at most max_retries races are spent on one entry before it is pended.
//
// Input is taken in chunks of kChunk indexes: wave w owns chunks w, w + W, w + 2W, ... and
// hands each chunk's indexes to its idle groups by ballot rank as they free up, so no group
// waits for the slowest chain of its wave. Resident waves sweep the input front to back
// together; when the batch was ordered by bucket region (launch_partition), the window they
// work on touches one region's buckets at a time.
// State of one group's in-hand entry (group-uniform except the view, which is per lane).
template <int SPL>
struct Carry {
    uint64_t entry;
    uint64_t idx;  // input index of the key that started the chain
    uint32_t pre, kicks, retries;
    uint64_t cur[SPL];  // this lane's SPL slots of the current bucket
    bool fresh;         // the view must be (re)read
    bool active;
};

struct InsertTally {
    unsigned long long kicks = 0, rejected = 0, retries = 0;
    unsigned long long empty = 0;  // zero entries skipped by a from-entries insert
    unsigned long long max_chain = 0;
};

// A chain ended after c.kicks exchanges (its entry placed or parked): tally it, and record it
// for the input index that started it.
template <int SPL>
__device__ inline void chain_end(const Carry<SPL> &c, const PendSink &sink, int s, InsertTally &tally)
{
    if (s == 0) {
        tally.kicks += c.kicks;
        tally.max_chain = tally.max_chain > c.kicks ? tally.max_chain : c.kicks;
        if (sink.chain) sink.chain[c.idx] = c.kicks;
    }
}

// Take entry e (previous table pr, kNoTable for a fresh key) into the group's hand.
template <int SPL>
__device__ inline void carry_start(const Desc &D, Carry<SPL> &c, uint64_t e, uint32_t pr, int s, InsertTally &tally,
                                   uint64_t idx = 0, uint32_t *chain = nullptr)
{
    const uint32_t k = key_of(e);
    if (reserved_key(k)) {
        if (s == 0) {
            ++tally.rejected;
            if (chain) chain[idx] = 0;
        }
        return;
    }
    if (pr == kNoTable && (k & 1u)) pr = pos1(D, pair_of(D, k));  // odd keys start in pos2 (cuckoo.cu:336-341)
    c.entry = e;
    c.idx = idx;
    c.pre = pr;
    c.kicks = 0;
    c.retries = 0;
    c.fresh = true;
    c.active = true;
}

// One eviction-chain iteration for an active group (see the comment above insert_kernel).
template <int G>
__device__ inline void kick_step(const Desc &D, Carry<kSlots / G> &c, const InsertPolicy &P, Counters *C,
                                 const PendSink &sink, uint32_t tag, InsertTally &tally)
{
    constexpr int SPL = kSlots / G;
    constexpr uint64_t GMASK = (G >= 64) ? ~0ull : ((1ull << G) - 1ull);
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    const uint32_t total_kicks = P.kick_bound + P.rescue_kicks;

    auto give_up = [&]() {  // park the in-hand entry on the pending list
        if (s == 0) {
            unsigned long long at = atomicAdd(&C->pending, 1ull);
            if (at < sink.cap) {
                sink.pend[at] = c.entry;
                if (sink.tags) sink.tags[at] = tag;
                // rare: publish the entry to every XCD (a serving tick's later finds read the pool
                // within the same launch)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            } else {
                atomicAdd(&C->lost, 1ull);
            }
        }
        c.active = false;
    };

    const uint32_t k = key_of(c.entry);
    const uint32_t t = other_pos(D, pair_of(D, k), c.pre);
    uint64_t *b = bucket_ptr(D, k, t);
    if (c.fresh) load_slots<SPL>(b + s * SPL, c.cur);
    c.fresh = false;

    // victim slot (group-uniform)
    int fe = -1;
#pragma unroll
    for (int j = SPL - 1; j >= 0; --j)
        if (key_of(c.cur[j]) == 0u) fe = j;
    uint32_t gm = (uint32_t)((__ballot(fe >= 0) >> (g * G)) & GMASK);
    int slot;
    if (gm) {
        const int w = __builtin_ctz(gm);
        slot = w * SPL + __shfl(fe, g * G + w);
    } else if (c.kicks < P.kick_bound) {
        int fo = -1;
#pragma unroll
        for (int j = SPL - 1; j >= 0; --j)
            if (vote_victim(D, key_of(c.cur[j]), t)) fo = j;
        gm = (uint32_t)((__ballot(fo >= 0) >> (g * G)) & GMASK);
        if (gm) {
            const int w = __builtin_ctz(gm);
            slot = w * SPL + __shfl(fo, g * G + w);
        } else {
            slot = (int)(c.kicks & 15u);
        }
    } else {
        slot = (int)rescue_lane(k, c.kicks);
    }
    const int owner = slot / SPL, sj = slot % SPL;
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < SPL; ++j)
        if (j == sj) mine = c.cur[j];
    uint64_t expected = shfl64(mine, g * G + owner);
    int won = 0;
    if (s == owner) {
        uint64_t exp = expected;
        won = cas_slot(b + slot, exp, c.entry);
        if (!won) {  // the slot's current value replaces the stale one in the view
#pragma unroll
            for (int j = 0; j < SPL; ++j)
                if (j == sj) c.cur[j] = exp;
        }
    }
    won = __shfl(won, g * G + owner);
    if (won) {
        ++c.kicks;
        c.fresh = true;
        if (expected == 0) {
            chain_end(c, sink, s, tally);
            c.active = false;
        } else {
            c.entry = expected;
            c.pre = t;
            if (c.kicks >= total_kicks) {
                chain_end(c, sink, s, tally);
                give_up();
            }
        }
    } else {
        if (s == 0) ++tally.retries;
        if (++c.retries > P.max_retries) {
            chain_end(c, sink, s, tally);
            give_up();
        }
    }
}

// Feed the wave's idle groups from input indexes [next, end), advancing next; returns when every
// group is busy or the range is used up (some groups may then still be idle).
template <int G, bool FROM_ENTRIES, bool COH = false>
__device__ inline void refill(const Desc &D, Carry<kSlots / G> &c, uint64_t &next, uint64_t end, const uint32_t *__restrict__ keys,
                              const uint32_t *__restrict__ vals, const uint64_t *__restrict__ entries,
                              uint32_t pre_fixed, InsertTally &tally, uint64_t &taken,
                              const uint8_t *__restrict__ skip = nullptr, uint32_t *chain = nullptr)
{
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    while (next < end) {
        const uint64_t idle = __ballot(!c.active && s == 0);  // bit g*G for every idle group
        if (!idle) return;
        const uint64_t avail = end - next;
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const uint32_t take = (uint32_t)(avail < nidle ? avail : nidle);
        const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << (g * G)) - 1ull));
        if (!c.active && rank < take) {
            const uint64_t idx = next + rank;
            if constexpr (FROM_ENTRIES) {
                const uint64_t e = entries[idx];
                if (e)
                    carry_start(D, c, e, pre_fixed, s, tally, idx, chain);
                else if (s == 0)
                    ++tally.empty;
            } else if (!skip || !skip[idx])
                carry_start(D, c, pack(load_in<COH>(keys + idx), load_in<COH>(vals + idx)), kNoTable, s, tally, idx,
                            chain);
            else if (chain && s == 0)
                chain[idx] = 0;
        }
        next += take;
        taken += take;
    }
}

// skip (optional): a keys insert leaves out key i where skip[i] != 0; an entries insert reads it
// as the tile flags of a slot-addressed residual list (entries of a tile flagged 0 are empty).
template <int G, bool FROM_ENTRIES>
__global__ __launch_bounds__(256) void insert_kernel(Desc D, const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ vals,
                                                     const uint64_t *__restrict__ entries, uint64_t n,
                                                     uint32_t pre_fixed, InsertPolicy P, Counters *__restrict__ C,
                                                     PendSink sink, const uint8_t *__restrict__ skip)
{
    constexpr int NG = kWave / G;
    constexpr uint64_t kChunk = 2 * NG;  // indexes a wave takes at a time
    static_assert(kResTileEntries % kChunk == 0, "a chunk lies inside one residual tile");
    Carry<kSlots / G> c;
    c.active = false;
    InsertTally tally;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    uint64_t next_chunk = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    uint64_t next = 0, end = 0;  // wave-uniform: the wave's current chunk
    uint64_t taken = 0;          // wave-uniform: input indexes this wave took
    while (true) {
        refill<G, FROM_ENTRIES>(D, c, next, end, keys, vals, entries, pre_fixed, tally, taken, skip, sink.chain);
        if (next == end) {
            // chunks w, w + W, w + 2W, ... (W = waves in the grid)
            const uint64_t base = next_chunk * kChunk;
            if (base < n) {
                next_chunk += waves;
                // an entries insert reads skip as the tile flags of a slot-addressed list; a chunk
                // lies inside one tile, and a tile flagged 0 holds no entry
                if constexpr (FROM_ENTRIES)
                    if (skip && !skip[base / kResTileEntries]) continue;
                next = base;
                end = base + kChunk < n ? base + kChunk : n;
                continue;
            }
        }
        if (!__ballot(c.active)) break;  // input exhausted and every group idle
        if (c.active) kick_step<G>(D, c, P, C, sink, sink.tag, tally);
    }
    if (sink.count_placed || sink.rejected_at) {
        const unsigned long long rej = wave_sum(tally.rejected);
        const unsigned long long skipped = rej + wave_sum(tally.empty);
        if (__lane_id() == 0) {
            if (sink.count_placed && taken > skipped) tally_add(C, &C->placed, taken - skipped, wave_slot());
            if (sink.rejected_at && rej) atomicAdd(sink.rejected_at, rej);
        }
    }
    wave_tally(C, &C->kicks, tally.kicks);
    wave_tally(C, &C->rejected, tally.rejected);
    wave_tally(C, &C->retries, tally.retries);
    wave_tally_max(C, &C->max_chain, tally.max_chain);
}

// -------------------------------------------------------------------------- partition
// Orders an insert batch by the region of each key's first bucket (first_region), so that the
// queue-driven insert kernel places one region's keys at a time. Three kernels: a histogram of
// regions (LDS counters per block), an exclusive scan into per-region write cursors, and a
// scatter. The scatter stages each tile of kTile keys in LDS sorted by region (a counting sort:
// LDS histogram, block scan, rank by LDS atomic), claims one global range per region present in
// the tile, and writes each region's run contiguously -- the only LDS use on the insert path is
// this staging of the batch's hashed keys.
constexpr int kPartThreads = 256;
constexpr int kPartPer = 8;
constexpr int kTile = kPartThreads * kPartPer;

__global__ __launch_bounds__(kPartThreads) void partition_count_kernel(Desc D, Regions M,
                                                                       const uint32_t *__restrict__ keys, uint64_t n,
                                                                       unsigned long long *__restrict__ counts)
{
    __shared__ unsigned int hist[kMaxRegions];
    for (uint32_t r = threadIdx.x; r < M.R; r += blockDim.x) hist[r] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&hist[first_region(D, M, keys[i])], 1u);
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < M.R; r += blockDim.x)
        if (hist[r]) atomicAdd(&counts[r], (unsigned long long)hist[r]);
}

__global__ void partition_scan_kernel(const unsigned long long *__restrict__ counts,
                                      unsigned long long *__restrict__ cursors, uint32_t R)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long acc = 0;
        for (uint32_t r = 0; r < R; ++r) {
            cursors[r] = acc;
            acc += counts[r];
        }
    }
}

__global__ __launch_bounds__(kPartThreads) void partition_scatter_kernel(Desc D, Regions M,
                                                                         const uint32_t *__restrict__ keys,
                                                                         const uint32_t *__restrict__ vals, uint64_t n,
                                                                         unsigned long long *__restrict__ cursors,
                                                                         uint64_t *__restrict__ out)
{
    __shared__ uint64_t stage[kTile];
    __shared__ unsigned short stage_r[kTile];
    __shared__ unsigned int hist[kMaxRegions];
    __shared__ unsigned int lbase[kMaxRegions];
    __shared__ unsigned long long gbase[kMaxRegions];
    __shared__ unsigned int wsum[kPartThreads / kWave];
    const uint32_t R = M.R;
    constexpr uint32_t kPerThread = (kMaxRegions + kPartThreads - 1) / kPartThreads;  // regions per thread in the scan

    for (uint64_t base = (uint64_t)blockIdx.x * kTile; base < n; base += (uint64_t)gridDim.x * kTile) {
        for (uint32_t r = threadIdx.x; r < R; r += kPartThreads) hist[r] = 0;
        __syncthreads();
        uint64_t e[kPartPer];
        uint32_t reg[kPartPer], rank[kPartPer];
#pragma unroll
        for (int j = 0; j < kPartPer; ++j) {
            const uint64_t i = base + (uint64_t)j * kPartThreads + threadIdx.x;
            reg[j] = 0xffffffffu;
            if (i < n) {
                const uint32_t k = keys[i];
                e[j] = pack(k, vals[i]);
                reg[j] = first_region(D, M, k);
                rank[j] = atomicAdd(&hist[reg[j]], 1u);
            }
        }
        __syncthreads();
        // block-wide exclusive scan of hist[0, R): each thread sums kPerThread consecutive
        // regions, the wave scans those sums with shuffles, the waves combine through LDS
        uint32_t mine[kPerThread];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPerThread; ++q) {
            const uint32_t r = threadIdx.x * kPerThread + q;
            mine[q] = r < R ? hist[r] : 0u;
            sum += mine[q];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t up = __shfl_up(incl, off);
            if (__lane_id() >= off) incl += up;
        }
        if (__lane_id() == kWave - 1) wsum[threadIdx.x / kWave] = incl;
        __syncthreads();
        uint32_t wbase = 0;
        for (uint32_t w = 0; w < threadIdx.x / kWave; ++w) wbase += wsum[w];
        uint32_t run = wbase + incl - sum;
#pragma unroll
        for (uint32_t q = 0; q < kPerThread; ++q) {
            const uint32_t r = threadIdx.x * kPerThread + q;
            if (r < R) {
                lbase[r] = run;
                if (mine[q]) gbase[r] = atomicAdd(&cursors[r], (unsigned long long)mine[q]);
            }
            run += mine[q];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPartPer; ++j)
            if (reg[j] != 0xffffffffu) {
                const uint32_t at = lbase[reg[j]] + rank[j];
                stage[at] = e[j];
                stage_r[at] = (unsigned short)reg[j];
            }
        __syncthreads();
        const uint64_t cnt = n - base < (uint64_t)kTile ? n - base : (uint64_t)kTile;
        for (uint32_t at = threadIdx.x; at < cnt; at += kPartThreads) {
            const uint32_t r = stage_r[at];
            out[gbase[r] + (at - lbase[r])] = stage[at];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------- upsize
// cuckoo.cu:600-627: slot i of old bucket j goes to slot i of new bucket h % 2n, which is j or
// j + n. Each group of 8 lanes writes both destination buckets in full (entry or zero), so the
// new subtable needs no separate clearing pass: 8 B read + 16 B written per old slot.
__global__ __launch_bounds__(256) void upsize_kernel(const uint64_t *__restrict__ old, uint32_t old_nb,
                                                     uint64_t *__restrict__ dst, uint32_t t, uint32_t m,
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
        uint64_t e[2];
        load_slots<2>(old + j * kSlots + s * 2, e);
        uint64_t lo[2], hi[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            lo[q] = hi[q] = 0;
            if (e[q]) {
                const uint32_t d = fast_mod(hash_t(key_of(e[q]), t), nnb, m, s1, s2);
                if (d == j)
                    lo[q] = e[q];
                else if (d == j + old_nb)
                    hi[q] = e[q];
                else
                    ++bad;
            }
        }
        *reinterpret_cast<uint4 *>(dst + j * kSlots + s * 2) =
            make_uint4((uint32_t)lo[0], (uint32_t)(lo[0] >> 32), (uint32_t)lo[1], (uint32_t)(lo[1] >> 32));
        *reinterpret_cast<uint4 *>(dst + (j + old_nb) * kSlots + s * 2) =
            make_uint4((uint32_t)hi[0], (uint32_t)(hi[0] >> 32), (uint32_t)hi[1], (uint32_t)(hi[1] >> 32));
    }
    wave_tally(C, &C->misplaced, bad);
}

// ----------------------------------------------------------------------------- downsize
// cuckoo.cu:664-699: new bucket g keeps old bucket g slot-for-slot and fills its empty slots
// from old bucket g+n read back to front (slot i <- slot 15-i). Lane s of a group holds slots
// 2s, 2s+1; their cross partners 15-2s, 14-2s are the two halves of lane 7-s's 16 bytes.
// An entry of bucket g+n that found no hole is a residual. res is slot-addressed: the residual
// from slot i of new bucket g's partner goes to res[g * 16 + i], 0 where there is none, so no
// wave reserves list space (an append index is one address, and atomics to one address are
// serialized: one per wave cost 0.9 ms on a 582,560-bucket downsize). A wave compacts a tile of
// kResTileBuckets buckets; live[tile] says whether it left any residual, and only such tiles
// write their kResTileEntries list entries, so a nearly full table writes almost no list. The
// caller re-inserts the list (launch_insert_entries with live) only after this kernel has
// finished (no race with the plain stores here). Counters::residuals receives their number.
__global__ __launch_bounds__(256) void downsize_kernel(const uint64_t *__restrict__ old, uint32_t new_nb,
                                                       uint64_t *__restrict__ dst, Counters *__restrict__ C,
                                                       uint64_t *__restrict__ res, uint8_t *__restrict__ live)
{
    static_assert(kWave / 8 == (int)kResTileBuckets, "a wave compacts one liveness tile");
    constexpr int G = 8, NG = kWave / G;
    const int lane = __lane_id();
    const int g = lane / G, s = lane % G;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint32_t slot = wave_slot();
    unsigned long long nres = 0;
    for (uint64_t tile = (uint64_t)slot * NG; tile < new_nb; tile += waves * NG) {
        const uint64_t j = tile + g;
        const bool ok = j < new_nb;
        uint64_t e[2] = {0, 0}, c[2] = {0, 0};
        if (ok) {
            load_slots<2>(old + j * kSlots + s * 2, e);
            uint64_t x[2];
            load_slots<2>(old + (j + new_nb) * kSlots + (7 - s) * 2, x);
            c[0] = x[1];  // slot 15-2s
            c[1] = x[0];  // slot 14-2s
        }
        uint64_t placed[2], r[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            placed[q] = e[q] ? e[q] : c[q];
            r[q] = (placed[q] != c[q] && c[q] != 0) ? c[q] : 0;
            nres += r[q] != 0;
        }
        if (ok)
            *reinterpret_cast<uint4 *>(dst + j * kSlots + s * 2) =
                make_uint4((uint32_t)placed[0], (uint32_t)(placed[0] >> 32), (uint32_t)placed[1],
                           (uint32_t)(placed[1] >> 32));
        const bool any = __ballot(r[0] != 0 || r[1] != 0) != 0;  // wave-uniform
        if (lane == 0) live[tile / NG] = any ? 1 : 0;
        if (any && ok)
            *reinterpret_cast<uint4 *>(res + j * kSlots + s * 2) =
                make_uint4((uint32_t)r[0], (uint32_t)(r[0] >> 32), (uint32_t)r[1], (uint32_t)(r[1] >> 32));
    }
    wave_tally(C, &C->residuals, nres, slot);
}

// Buckets the reference find order reads per key: 1 if pos1 holds the key, else 2 (0 for
// reserved keys). One thread per key; a diagnostic pass, not on the hot path.
__global__ __launch_bounds__(256) void probe_count_kernel(Desc D, const uint32_t *__restrict__ keys, uint64_t n,
                                                          Counters *__restrict__ C)
{
    unsigned long long probes = 0, pos1_hits = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t k = keys[i];
        if (reserved_key(k)) continue;
        const uint64_t *b = bucket_ptr(D, k, pos1(D, pair_of(D, k)));
        bool hit = false;
        for (int j = 0; j < kSlots; ++j) hit |= key_of(b[j]) == k;
        probes += hit ? 1 : 2;
        pos1_hits += hit;
    }
    wave_tally(C, &C->kicks, probes);
    wave_tally(C, &C->occupied, pos1_hits);
}

// ---------------------------------------------------------------------------- many tables
// One persistent launch serves a batch spread over many tables. The batch is cut into items of
// at most a few thousand keys of one table; wave w takes items w, w + W, w + 2W, ... and loads
// the item's table descriptor from the directory with scalar loads (the item index is
// wave-uniform), so the per-tile code is the single-table code with a different Desc.
__device__ inline uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// W: the least waves per SIMD the compiler must leave room for (1: its own choice); the group
// launches take it as an occupancy knob, chosen by measurement.
template <int G, bool ERASE, int W>
__global__ __launch_bounds__(256, W) void group_probe_kernel(const Desc *__restrict__ dir,
                                                          const GroupItem *__restrict__ items, uint32_t n_items,
                                                          const uint32_t *__restrict__ keys,
                                                          uint32_t *__restrict__ vals, uint8_t *__restrict__ flags,
                                                          unsigned long long *__restrict__ table_counts, Pool pool)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    const uint64_t np = pool_size(pool);
    for (uint32_t it = wave_uniform(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); it < n_items;
         it += waves) {
        const GroupItem item = items[it];
        if (item.len == 0) continue;  // wave-uniform: an empty slot of a layout by XCD
        const Desc D = dir[item.table];
        const uint64_t end = item.begin + item.len;
        unsigned long long nd = 0;
        for (uint64_t base = item.begin; base < end; base += kWave) {
            uint32_t f = probe_tile<G, ERASE>(D, keys, vals, flags, base, end);
            if (np) f |= pool_probe_at<ERASE ? 1 : 0>(pool, np, keys, vals, flags, end, item.table, base, f);
            nd += f;
        }
        if constexpr (ERASE) {
            nd = wave_sum(nd);
            if (__lane_id() == 0 && nd) atomicAdd(&table_counts[item.table], nd);
        }
    }
}

// Groups of one wave share an item: they are fed from it until it is used up, and the wave
// moves on only when all its groups have placed (or pended) their entries. FROM_ENTRIES: the
// items index packed entries (a bulk fill's overflow) instead of keys/vals.
template <int G, int W, bool FROM_ENTRIES = false>
__global__ __launch_bounds__(256, W) void group_insert_kernel(const Desc *__restrict__ dir,
                                                           const GroupItem *__restrict__ items, uint32_t n_items,
                                                           const uint32_t *__restrict__ keys,
                                                           const uint32_t *__restrict__ vals, InsertPolicy P,
                                                           Counters *__restrict__ C, PendSink sink,
                                                           unsigned long long *__restrict__ table_counts,
                                                           const uint64_t *__restrict__ entries = nullptr)
{
    Carry<kSlots / G> c;
    c.active = false;
    InsertTally tally;
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t it = wave_uniform(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); it < n_items;
         it += waves) {
        const GroupItem item = items[it];
        if (item.len == 0) continue;  // wave-uniform: a block the bulk fill's rounds emptied
        const Desc D = dir[item.table];
        uint64_t next = item.begin;
        const uint64_t end = item.begin + item.len;
        const unsigned long long rej0 = tally.rejected;
        uint64_t taken = 0;
        while (true) {
            refill<G, FROM_ENTRIES>(D, c, next, end, keys, vals, entries, kNoTable, tally, taken);
            if (!__ballot(c.active)) break;  // item used up and every group idle
            if (c.active) kick_step<G>(D, c, P, C, sink, item.table, tally);
        }
        const unsigned long long rej = wave_sum(tally.rejected - rej0);
        if (__lane_id() == 0 && item.len > rej) atomicAdd(&table_counts[item.table], item.len - rej);
    }
    wave_tally(C, &C->kicks, tally.kicks);
    wave_tally(C, &C->rejected, tally.rejected);
    wave_tally(C, &C->retries, tally.retries);
}

// ------------------------------------------------------------------------------ serving
// Items of a serving tick carry their request's own pointers; otherwise the tile code is the
// single-table code. Finds and erases that miss both buckets also look in the server's pool of
// entries that exhausted the eviction bound (normally empty: one scalar load per tile).

// The pool search of one tile's lanes whose key missed both buckets (f = 0). The pool counts as
// one more bucket after pos2. MODE 0 find: vals receives the first matching entry's value.
// MODE 1 erase: every copy is CAS'd to 0. MODE 2 update: every copy is CAS'd to the new value
// vals[i] (an input). Returns this lane's flag, also written to flags when set.
template <int MODE, bool COH>
__device__ inline uint32_t pool_probe_at(const Pool &pool, uint64_t np, const uint32_t *keys, uint32_t *vals,
                                         uint8_t *flags, uint64_t n, uint32_t table, uint64_t base, uint32_t f)
{
    const uint64_t i = base + __lane_id();
    if (i >= n || f) return 0;
    const uint32_t k = keys[i];
    if (reserved_key(k)) return 0;
    const uint32_t nv = MODE == 2 ? vals[i] : 0u;
    uint32_t got = 0;
    for (uint64_t j = 0; j < np; ++j) {
        const uint64_t e = __hip_atomic_load(pool.pend + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (key_of(e) != k ||
            (pool.tags && __hip_atomic_load(pool.tags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != table))
            continue;
        if constexpr (MODE != 0) {
            uint64_t exp = e;
            got |= (uint32_t)cas_slot(pool.pend + j, exp, MODE == 2 ? pack(k, nv) : 0ull);
        } else {
            if (vals) store_out<COH>(vals + i, val_of(e));
            got = 1;
            break;
        }
    }
    if (got && flags) store_out<COH>(flags + i, (uint8_t)1);
    return got;
}

template <bool ERASE, bool COH = false>
__device__ inline uint32_t pool_probe(const Pool &pool, uint64_t np, const ServeItem &item, uint64_t base, uint32_t f)
{
    return pool_probe_at<ERASE ? 1 : 0, COH>(pool, np, item.keys, item.out_vals, item.flags, item.len, item.table, base,
                                             f);
}

// One item of finds or erases; returns this lane's count of hits (erase: deletions).
template <int G, bool ERASE, bool COH = false>
__device__ inline unsigned long long serve_probe_item(const Desc &D, const ServeItem &item, const Pool &pool,
                                                      uint64_t np)
{
    unsigned long long nd = 0;
    for (uint64_t base = 0; base < item.len; base += kWave) {
        uint32_t f = probe_tile<G, ERASE, COH>(D, item.keys, ERASE ? nullptr : item.out_vals, item.flags, base,
                                               item.len);
        if (np) f |= pool_probe<ERASE, COH>(pool, np, item, base, f);
        nd += f;
    }
    return nd;
}

// One item of inserts; the wave's groups are all idle on return. COH: read the inputs
// device-coherently (load_in).
template <int G, bool COH = false>
__device__ inline unsigned long long serve_insert_item(const Desc &D, const ServeItem &item, Carry<kSlots / G> &c,
                                                       const InsertPolicy &P, Counters *C, const PendSink &sink,
                                                       InsertTally &tally)
{
    uint64_t next = 0;
    const uint64_t end = item.len;
    const unsigned long long rej0 = tally.rejected;
    uint64_t taken = 0;
    while (true) {
        refill<G, false, COH>(D, c, next, end, item.keys, item.vals, nullptr, kNoTable, tally, taken);
        if (!__ballot(c.active)) break;
        if (c.active) kick_step<G>(D, c, P, C, sink, item.table, tally);
    }
    const unsigned long long rej = wave_sum(tally.rejected - rej0);
    return item.len > rej ? item.len - rej : 0;  // entries the item took in (wave-uniform)
}

// A wave's running total for one table, added to that table's word when the wave moves on to
// another table and at the end. Items of one table come in runs, and an atomic per item put
// every item of a busy table on one address (atomics to one address are serialized).
struct WaveTotal {
    uint32_t table = ~0u;
    unsigned long long n = 0;
    __device__ void add(unsigned long long *totals, uint32_t t, unsigned long long v)  // v wave-uniform
    {
        if (t != table) {
            flush(totals);
            table = t;
        }
        n += v;
    }
    __device__ void flush(unsigned long long *totals)
    {
        if (__lane_id() == 0 && table != ~0u && n) atomicAdd(&totals[table], n);
        n = 0;
    }
};


template <int G, bool ERASE>
__global__ __launch_bounds__(256) void serve_probe_kernel(const Desc *__restrict__ dir,
                                                          const ServeItem *__restrict__ items, uint32_t n_items,
                                                          Pool pool, unsigned long long *__restrict__ del_total)
{
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    const uint64_t np = pool_size(pool);
    WaveTotal del;
    for (uint32_t it = wave_uniform(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); it < n_items;
         it += waves) {
        const ServeItem item = items[it];
        const Desc D = dir[item.table];
        unsigned long long nd = serve_probe_item<G, ERASE>(D, item, pool, np);
        if constexpr (ERASE) del.add(del_total, item.table, wave_sum(nd));
    }
    if constexpr (ERASE) del.flush(del_total);
}

template <int G>
__global__ __launch_bounds__(256) void serve_insert_kernel(const Desc *__restrict__ dir,
                                                           const ServeItem *__restrict__ items, uint32_t n_items,
                                                           InsertPolicy P, Counters *__restrict__ C, PendSink sink,
                                                           unsigned long long *__restrict__ ins_total)
{
    Carry<kSlots / G> c;
    c.active = false;
    InsertTally tally;
    WaveTotal ins;
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t it = wave_uniform(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); it < n_items;
         it += waves) {
        const ServeItem item = items[it];
        ins.add(ins_total, item.table, serve_insert_item<G>(dir[item.table], item, c, P, C, sink, tally));
    }
    ins.flush(ins_total);
    wave_tally(C, &C->kicks, tally.kicks);
    wave_tally(C, &C->rejected, tally.rejected);
    wave_tally(C, &C->retries, tally.retries);
}

// A whole serving tick in one launch. Waves claim items in list order from a counter; an item
// whose stage has a predecessor (the same table's previous round or phase) first waits for the
// stage's ready flag; when a stage's last item finishes it raises the flag of the table's next
// stage. No acquire/release fences: on this chip an agent-scope fence writes back or
// invalidates the whole L2 of the XCD. Table changes are 64-bit exchanges (device-coherent, and
// their results are consumed before the stage count drops), and finds and erases read buckets
// with coherent loads. Items wait only on items earlier in the list, which were claimed by
// running waves, so the earliest unfinished item never waits and the launch always drains. The
// wait is bounded all the same: past spin_cap polls the wave records a stall and goes on (the
// host reports the tick failed).
template <int GI, int GF>
__global__ __launch_bounds__(256) void serve_tick_kernel(const Desc *__restrict__ dir, const TickItem *__restrict__ items,
                                                         uint32_t n_items, unsigned long long *__restrict__ cursor,
                                                         TickStages stages, InsertPolicy P, Counters *__restrict__ C,
                                                         PendSink sink, Pool pool,
                                                         unsigned long long *__restrict__ ins_total,
                                                         unsigned long long *__restrict__ del_total, uint32_t spin_cap,
                                                         uint32_t max_back)
{
    Carry<kSlots / GI> c;
    c.active = false;
    InsertTally tally;
    WaveTotal ins, del;
    unsigned long long stalls = 0;
    while (true) {
        uint32_t it = 0;
        if (__lane_id() == 0) it = (uint32_t)atomicAdd(cursor, 1ull);
        it = wave_uniform(it);
        if (it >= n_items) break;
        const TickItem ti = items[it];
        if (ti.wait) {
            uint32_t ok = 1;
            if (__lane_id() == 0) {
                uint32_t spins = 0, back = 1;
                const uint32_t *flag = stages.ready + (uint64_t)ti.stage * kReadyStride;
                while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                    for (uint32_t b = 0; b < back; ++b) __builtin_amdgcn_s_sleep(4);
                    back = back < max_back ? back * 2 : max_back;
                    if (++spins > spin_cap) {
                        ok = 0;
                        break;
                    }
                }
            }
            if (!wave_uniform(ok)) ++stalls;
        }
        const Desc D = dir[ti.item.table];
        if (ti.op == 0) {
            ins.add(ins_total, ti.item.table, serve_insert_item<GI>(D, ti.item, c, P, C, sink, tally));
        } else {
            // an item that waited reads what earlier stages of this launch wrote: coherently;
            // a table's first stage reads what earlier launches wrote: through the caches
            unsigned long long nd;
            const uint64_t np = pool_size(pool);
            if (ti.op == 1)
                nd = ti.wait ? serve_probe_item<GF, false, true>(D, ti.item, pool, np)
                             : serve_probe_item<GF, false, false>(D, ti.item, pool, np);
            else
                nd = ti.wait ? serve_probe_item<GF, true, true>(D, ti.item, pool, np)
                             : serve_probe_item<GF, true, false>(D, ti.item, pool, np);
            if (ti.op == 2) del.add(del_total, ti.item.table, wave_sum(nd));
        }
        if (__lane_id() == 0) {
            const uint32_t before =
                __hip_atomic_fetch_sub(stages.left + ti.stage, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (before == 1) {
                const uint32_t nx = stages.next[ti.stage];
                if (nx != ~0u)
                    __hip_atomic_store(stages.ready + (uint64_t)nx * kReadyStride, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    ins.flush(ins_total);
    del.flush(del_total);
    wave_tally(C, &C->kicks, tally.kicks);
    wave_tally(C, &C->rejected, tally.rejected);
    wave_tally(C, &C->retries, tally.retries);
    if (__lane_id() == 0 && stalls) tally_add(C, &C->stalls, stalls, wave_slot());
}

// ----------------------------------------------------------- device-resident serving loop
// One launch serves requests as the host publishes them (LoopItem). Wave 0 of block 0 is the
// front wave: it polls the control block in host memory and copies the published item count
// into the device word avail. Every other wave claims item indexes from the device cursor in
// order and waits, polling device memory, until its index is published. An item of a table
// whose earlier requests have not finished waits until tbl_done[table] reaches its need. Items
// wait only on items with smaller indexes, which running waves have claimed (the grid is
// resident as a whole), so the earliest unfinished item never waits and every published item
// finishes. A finished item's table writes are atomics whose results it has used: it counts
// itself into its request's progress word at once, and the wave that counts the request's last
// item advances tbl_done. Its results are system-scope stores: it waits for them, then counts
// itself into the request's output count, and the wave that counts the last item writes the
// request's record and appends the request to the completion queue in host memory.
// Every wait is bounded. The front wave ends the launch on the host's stop flag or after
// idle_ticks without new items: it raises quit, and waves waiting for an index past the
// published ones leave. An item gives up waiting for its table after spin_cap polls, or at once
// once the host has given up on the launch (stop = 2 raises abort), and counts as a stall (the
// host fails its request). A ring entry whose stamp does not match after
// spin_cap polls raises the fault flag and quit.
// Tables, inputs and the directory are read device-coherently: the launch outlives the writes
// of resizes and of the caller, and this XCD's L2 may hold lines they made stale.

__device__ inline uint64_t ld_system(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ inline uint64_t ld_agent(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}

__device__ inline void sleep_units(uint32_t n)  // n x 256 cycles
{
    for (uint32_t b = 0; b < n; ++b) __builtin_amdgcn_s_sleep(4);
}

// The W words lanes 0..W-1 hold, as one wave-uniform T.
template <int W, typename T>
__device__ inline T wave_gather(uint64_t v)
{
    static_assert(sizeof(T) == W * 8, "T is W words");
    uint64_t w[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), j);
        w[j] = (uint64_t)hi << 32 | lo;
    }
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

__device__ inline Desc load_desc_coherent(const Desc *p)
{
    constexpr int W = sizeof(Desc) / 8;
    const unsigned long long *w = reinterpret_cast<const unsigned long long *>(p);
    const uint64_t v = __lane_id() < W ? ld_agent(w + __lane_id()) : 0;
    return wave_gather<W, Desc>(v);
}

__device__ inline void loop_front(const LoopArgs &A)
{
    unsigned long long *avail = A.words + kLoopAvail;
    uint64_t seen = uniform64(ld_agent(avail));
    uint64_t last = __builtin_amdgcn_s_memrealtime();
    while (true) {
        const uint64_t t = uniform64(ld_system(&A.ctl->tail));
        const uint64_t stop = uniform64(ld_system(&A.ctl->stop));
        if (stop) {
            if (stop == 2 && __lane_id() == 0)
                __hip_atomic_store(A.words + kLoopAbort, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (t != seen) {
            if (__lane_id() == 0) __hip_atomic_store(avail, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            seen = t;
            last = now;
            continue;
        }
        if (now - last > A.idle_ticks) break;
        __builtin_amdgcn_s_sleep(2);
    }
    if (__lane_id() == 0) __hip_atomic_store(A.words + kLoopQuit, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int GI, int GF>
__global__ __launch_bounds__(256) void serve_loop_kernel(LoopArgs A, InsertPolicy P, Counters *__restrict__ C,
                                                         PendSink sink, Pool pool,
                                                         unsigned long long *__restrict__ ins_total,
                                                         unsigned long long *__restrict__ del_total)
{
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        loop_front(A);
        return;
    }
    Carry<kSlots / GI> c;
    c.active = false;
    InsertTally tally;
    WaveTotal ins, del;
    const int lane = __lane_id();
    unsigned long long *cursor = A.words + kLoopCursor, *avail = A.words + kLoopAvail, *quit = A.words + kLoopQuit;
    while (true) {
        uint64_t idx = 0;
        if (lane == 0) idx = atomicAdd(cursor, 1ull);
        idx = uniform64(idx);
        // wait until the index is published, or the launch ends
        bool go = true;
        for (uint32_t back = 1;;) {
            if (idx < uniform64(ld_agent(avail))) break;
            // the front wave stores avail before quit (a release), and this reload is issued only
            // after quit was read, so it sees the last published count
            if (uniform64(ld_agent(quit))) {
                go = idx < uniform64(ld_agent(avail));
                break;
            }
            sleep_units(back);
            back = back < A.max_back ? back * 2 : A.max_back;
        }
        if (!go) break;
        // the item, 8 words from host memory; its stamp tells that the host finished writing it
        const unsigned long long *src = reinterpret_cast<const unsigned long long *>(A.ring + (idx & A.ring_mask));
        LoopItem it;
        bool ok = false;
        for (uint32_t spins = 0; spins <= A.spin_cap; ++spins) {
            it = wave_gather<8, LoopItem>(lane < 8 ? ld_system(src + lane) : 0ull);
            if (it.stamp == (uint32_t)(idx + 1)) {
                ok = true;
                break;
            }
            sleep_units(1);
        }
        if (!ok) {
            if (lane == 0) {
                __hip_atomic_store(A.words + kLoopFault, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(quit, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        const uint32_t table = it.item.table;
        // the descriptor can be read before the wait: a table is resized only while none of its
        // requests is in flight, and the host writes the directory before it publishes items
        const Desc D = load_desc_coherent(A.dir + table);
        uint32_t stalled = 0;
        if (it.need) {
            uint32_t back = 1, spins = 0;
            while (uniform64(ld_agent(A.tbl_done + table)) < it.need) {
                if (++spins > A.spin_cap || uniform64(ld_agent(A.words + kLoopAbort))) {
                    stalled = 1;
                    break;
                }
                sleep_units(back);
                back = back < A.max_back ? back * 2 : A.max_back;
            }
        }
        unsigned long long acc;
        if (it.op == 0) {
            acc = serve_insert_item<GI, true>(D, it.item, c, P, C, sink, tally);
            ins.add(ins_total, table, acc);
        } else {
            const uint64_t np = pool_size(pool);
            if (it.op == 1) {
                acc = wave_sum(serve_probe_item<GF, false, true>(D, it.item, pool, np));
            } else {
                acc = wave_sum(serve_probe_item<GF, true, true>(D, it.item, pool, np));
                del.add(del_total, table, acc);
            }
        }
        // Two counts per request. The table's work of the item is done here (its exchanges and
        // erasures are atomics whose results the wave has used; a parked entry was fenced), so
        // the item counts at once towards releasing the table's next requests. Its results are
        // system-scope stores, which it waits for before counting towards the host's completion.
        if (lane == 0) {
            const unsigned long long old = atomicAdd(A.req_word + it.req, (acc << kLoopItemBits) | 1ull);
            if ((old & ((1ull << kLoopItemBits) - 1)) + 1 == it.n_items)
                atomicAdd(A.tbl_done + table, 1ull);  // the table's next requests may start
            if (stalled) atomicAdd(A.req_stall + it.req, 1u);
        }
        publish_wait();
        if (lane == 0) {
            const uint32_t done = atomicAdd(A.req_out + it.req, 1u) + 1;
            if (done == it.n_items) {
                // every item counted into req_word before it counted here
                const unsigned long long total = ld_agent(A.req_word + it.req) >> kLoopItemBits;
                const uint32_t stalls = __hip_atomic_load(A.req_stall + it.req, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(A.req_word + it.req, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(A.req_out + it.req, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(A.req_stall + it.req, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                LoopRec *r = A.rec + it.req;
                __hip_atomic_store(&r->acc, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&r->pool, pool.count ? ld_agent(pool.count) : 0ull, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&r->lost, ld_agent(&C->lost), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&r->stalls, stalls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                publish_wait();  // the record is in host memory before the queue entry
                const unsigned long long pos = atomicAdd(A.words + kLoopCqTail, 1ull);
                __hip_atomic_store(A.cq + (pos & A.req_mask), ((pos + 1) << 32) | it.req, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
    ins.flush(ins_total);
    del.flush(del_total);
    wave_tally(C, &C->kicks, tally.kicks);
    wave_tally(C, &C->rejected, tally.rejected);
    wave_tally(C, &C->retries, tally.retries);
}

__global__ void loop_reset_kernel(unsigned long long *words)
{
    words[kLoopCursor] = words[kLoopAvail];
    words[kLoopQuit] = 0;
    words[kLoopFault] = 0;
    words[kLoopAbort] = 0;
}

// Copy between host-pinned and device memory with a kernel instead of a DMA transfer: a
// serving tick moves a few KB each way, where queueing a copy engine costs more than the
// transfer. Up to two (src, dst, words) segments per launch.
struct CopySeg {
    const unsigned long long *src;
    unsigned long long *dst;
    uint64_t words;
};

__global__ __launch_bounds__(256) void copy_kernel(CopySeg a, CopySeg b)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.words; i += stride) a.dst[i] = a.src[i];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < b.words; i += stride) b.dst[i] = b.src[i];
}

__global__ void zero_kernel(unsigned long long *p, uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(256) void count_kernel(const uint64_t *__restrict__ sub, uint64_t slots,
                                                    Counters *__restrict__ C)
{
    unsigned long long c = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += stride) c += sub[i] != 0;
    wave_tally(C, &C->occupied, c);
}

// Blocks for `work_waves` waves of work. A launch may not exceed 2^32 work-items in total, and
// every kernel here loops grid-stride, so the grid is capped (default_cap when the caller set
// no max_blocks); the rest of the work is taken by the resident waves as they finish.

// Insert groups take their next key when they finish one, so a grid of a few times the resident
// capacity (256 CUs x 2048 threads) keeps every CU busy without launching a block per key.
constexpr uint64_t kInsertDefaultBlocks = 65536;

// Eviction chains vary in length, so a grid whose groups each take one or two keys ends on a tail
// of late blocks: a batch of 1.6 M keys inserts 18% faster as 1024 resident blocks that take about
// eight keys per group than as one block per 256 keys. With no max_blocks from the caller, the
// default grid gives each group at least kInsertKeysPerGroup keys, never below kInsertFloorBlocks
// (enough to fill the device) unless the batch is smaller than that.
constexpr uint64_t kInsertKeysPerGroup = 8;
constexpr uint64_t kInsertFloorBlocks = 1024;

inline uint64_t insert_default_cap(uint64_t work_waves, LaunchShape sh)
{
    const uint64_t wpb = (uint64_t)sh.block_threads / kWave;
    const uint64_t work_blocks = (work_waves + wpb - 1) / wpb;
    uint64_t cap = work_blocks / kInsertKeysPerGroup;
    if (cap < kInsertFloorBlocks) cap = kInsertFloorBlocks;
    if (cap > kInsertDefaultBlocks) cap = kInsertDefaultBlocks;
    return cap;
}

}  // namespace

hipError_t launch_find(const Desc &D, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, int lanes,
                       LaunchShape sh, hipStream_t st, const Pool *pool, uint32_t table)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh);
    const Pool p = pool ? *pool : Pool{};
    switch (lanes) {
    case 4: probe_kernel<4, false><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, hit, n, nullptr, p, table); break;
    case 16: probe_kernel<16, false><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, hit, n, nullptr, p, table); break;
    default: probe_kernel<8, false><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, hit, n, nullptr, p, table); break;
    }
    return hipGetLastError();
}

hipError_t launch_erase(const Desc &D, const uint32_t *keys, uint8_t *deleted, uint64_t n, int lanes, LaunchShape sh,
                        Counters *C, hipStream_t st, const Pool *pool, uint32_t table)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh);
    const Pool p = pool ? *pool : Pool{};
    switch (lanes) {
    case 4: probe_kernel<4, true><<<grid, sh.block_threads, 0, st>>>(D, keys, nullptr, deleted, n, C, p, table); break;
    case 16: probe_kernel<16, true><<<grid, sh.block_threads, 0, st>>>(D, keys, nullptr, deleted, n, C, p, table); break;
    default: probe_kernel<8, true><<<grid, sh.block_threads, 0, st>>>(D, keys, nullptr, deleted, n, C, p, table); break;
    }
    return hipGetLastError();
}

template <bool FE>
static hipError_t insert_common(const Desc &D, const uint32_t *keys, const uint32_t *vals, const uint64_t *entries,
                                uint64_t n, uint32_t pre, int lanes,
                                InsertPolicy pol, LaunchShape sh, Counters *C, PendSink sink, hipStream_t st,
                                const uint8_t *skip = nullptr)
{
    if (n == 0) return hipSuccess;
    const int per_wave = kWave / lanes;  // entries in flight per wave
    const uint64_t work_waves = (n + per_wave - 1) / per_wave;
    const int grid = grid_for(work_waves, sh, insert_default_cap(work_waves, sh));
    switch (lanes) {
    case 8:
        insert_kernel<8, FE><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, entries, n, pre, pol, C,
                                                                      sink, skip);
        break;
    case 4:
        insert_kernel<4, FE><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, entries, n, pre, pol, C,
                                                                      sink, skip);
        break;
    case 2:
        insert_kernel<2, FE><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, entries, n, pre, pol, C,
                                                                      sink, skip);
        break;
    case 1:
        insert_kernel<1, FE><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, entries, n, pre, pol, C,
                                                                      sink, skip);
        break;
    default:
        insert_kernel<16, FE><<<grid, sh.block_threads, 0, st>>>(D, keys, vals, entries, n, pre, pol, C,
                                                                      sink, skip);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_insert(const Desc &D, const uint32_t *keys, const uint32_t *vals, uint64_t n, int lanes,
                         InsertPolicy pol, LaunchShape sh, Counters *C, PendSink sink, hipStream_t st,
                         const uint8_t *skip)
{
    return insert_common<false>(D, keys, vals, nullptr, n, kNoTable, lanes, pol, sh, C, sink, st, skip);
}

hipError_t launch_update(const Desc &D, const uint32_t *keys, const uint32_t *vals, uint8_t *updated, uint64_t n,
                         int lanes, LaunchShape sh, Counters *C, hipStream_t st, const Pool *pool)
{
    if (n == 0) return hipSuccess;
    const int grid = grid_for((n + kWave - 1) / kWave, sh);
    uint32_t *v = const_cast<uint32_t *>(vals);  // read-only in update mode
    const Pool p = pool ? *pool : Pool{};
    switch (lanes) {
    case 4: update_kernel<4><<<grid, sh.block_threads, 0, st>>>(D, keys, v, updated, n, C, p); break;
    case 16: update_kernel<16><<<grid, sh.block_threads, 0, st>>>(D, keys, v, updated, n, C, p); break;
    default: update_kernel<8><<<grid, sh.block_threads, 0, st>>>(D, keys, v, updated, n, C, p); break;
    }
    return hipGetLastError();
}

hipError_t launch_insert_entries(const Desc &D, const uint64_t *entries, uint64_t n, uint32_t pre, int lanes,
                                 InsertPolicy pol, LaunchShape sh, Counters *C, PendSink sink, hipStream_t st,
                                 const uint8_t *live)
{
    return insert_common<true>(D, nullptr, nullptr, entries, n, pre, lanes, pol, sh, C, sink, st, live);
}

hipError_t launch_partition(const Desc &D, const Regions &M, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                            unsigned long long *counts, unsigned long long *cursors, uint64_t *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (M.R == 0 || M.R > kMaxRegions) return hipErrorInvalidValue;
    uint64_t blocks = (n + kTile - 1) / kTile;
    if (blocks > 8192) blocks = 8192;
    partition_count_kernel<<<(int)blocks, kPartThreads, 0, st>>>(D, M, keys, n, counts);
    partition_scan_kernel<<<1, 64, 0, st>>>(counts, cursors, M.R);
    partition_scatter_kernel<<<(int)blocks, kPartThreads, 0, st>>>(D, M, keys, vals, n, cursors, out);
    return hipGetLastError();
}

hipError_t launch_upsize(const uint64_t *old, uint32_t old_nb, uint64_t *dst, uint32_t t, Counters *C,
                         hipStream_t st)
{
    if (old_nb == 0) return hipSuccess;
    const FastDiv f = make_fastdiv(old_nb * 2u);
    const int grid = grid_for((old_nb + 7) / 8, LaunchShape{256, 0});
    upsize_kernel<<<grid, 256, 0, st>>>(old, old_nb, dst, t, f.m, f.s1, f.s2, C);
    return hipGetLastError();
}

hipError_t launch_downsize(const uint64_t *old, uint32_t new_nb, uint64_t *dst, Counters *C, uint64_t *res,
                           uint64_t res_cap, uint8_t *live, hipStream_t st)
{
    if (new_nb == 0) return hipSuccess;
    if (res_cap < (uint64_t)new_nb * kSlots || !live) return hipErrorInvalidValue;
    const int grid = grid_for((new_nb + kResTileBuckets - 1) / kResTileBuckets, LaunchShape{256, 0});
    downsize_kernel<<<grid, 256, 0, st>>>(old, new_nb, dst, C, res, live);
    return hipGetLastError();
}

constexpr int kGroupDefaultBlocks = 2048;  // 8192 waves: 256 CUs x 32 resident waves

template <int G, bool ERASE>
static void group_probe_w(int wpe, int grid, hipStream_t st, const Desc *dir, const GroupItem *items, uint32_t n_items,
                          const uint32_t *keys, uint32_t *vals, uint8_t *flags, unsigned long long *table_counts,
                          const Pool &p)
{
    switch (wpe) {
    case 6:
        group_probe_kernel<G, ERASE, 6><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, flags, table_counts, p);
        break;
    case 8:
        group_probe_kernel<G, ERASE, 8><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, flags, table_counts, p);
        break;
    default:
        group_probe_kernel<G, ERASE, 1><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, flags, table_counts, p);
        break;
    }
}

hipError_t launch_group_find(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                             uint32_t *vals, uint8_t *hit, int lanes, int blocks, hipStream_t st, const Pool *pool,
                             int waves_per_eu)
{
    const Pool p = pool ? *pool : Pool{};
    if (n_items == 0) return hipSuccess;
    const int grid = blocks > 0 ? blocks : std::min<int>(kGroupDefaultBlocks, (int)((n_items + 3) / 4));
    switch (lanes) {
    case 4: group_probe_w<4, false>(waves_per_eu, grid, st, dir, items, n_items, keys, vals, hit, nullptr, p); break;
    case 16: group_probe_w<16, false>(waves_per_eu, grid, st, dir, items, n_items, keys, vals, hit, nullptr, p); break;
    default: group_probe_w<8, false>(waves_per_eu, grid, st, dir, items, n_items, keys, vals, hit, nullptr, p); break;
    }
    return hipGetLastError();
}

hipError_t launch_group_erase(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                              uint8_t *deleted, unsigned long long *table_counts, int lanes, int blocks,
                              hipStream_t st, const Pool *pool, int waves_per_eu)
{
    const Pool p = pool ? *pool : Pool{};
    if (n_items == 0) return hipSuccess;
    const int grid = blocks > 0 ? blocks : std::min<int>(kGroupDefaultBlocks, (int)((n_items + 3) / 4));
    switch (lanes) {
    case 4:
        group_probe_w<4, true>(waves_per_eu, grid, st, dir, items, n_items, keys, nullptr, deleted, table_counts, p);
        break;
    case 16:
        group_probe_w<16, true>(waves_per_eu, grid, st, dir, items, n_items, keys, nullptr, deleted, table_counts, p);
        break;
    default:
        group_probe_w<8, true>(waves_per_eu, grid, st, dir, items, n_items, keys, nullptr, deleted, table_counts, p);
        break;
    }
    return hipGetLastError();
}

template <int G>
static void group_insert_w(int wpe, int grid, hipStream_t st, const Desc *dir, const GroupItem *items,
                           uint32_t n_items, const uint32_t *keys, const uint32_t *vals, InsertPolicy pol, Counters *C,
                           PendSink sink, unsigned long long *table_counts, const uint64_t *entries)
{
    if (entries) {  // a bulk fill's overflow: the compiler's occupancy only
        group_insert_kernel<G, 1, true>
            <<<grid, 256, 0, st>>>(dir, items, n_items, nullptr, nullptr, pol, C, sink, table_counts, entries);
        return;
    }
    switch (wpe) {
    case 6:
        group_insert_kernel<G, 6><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, pol, C, sink, table_counts);
        break;
    case 8:
        group_insert_kernel<G, 8><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, pol, C, sink, table_counts);
        break;
    default:
        group_insert_kernel<G, 1><<<grid, 256, 0, st>>>(dir, items, n_items, keys, vals, pol, C, sink, table_counts);
        break;
    }
}

hipError_t launch_group_insert(const Desc *dir, const GroupItem *items, uint32_t n_items, const uint32_t *keys,
                               const uint32_t *vals, InsertPolicy pol, int lanes, int blocks, Counters *C,
                               PendSink sink, unsigned long long *table_counts, hipStream_t st, int waves_per_eu,
                               const uint64_t *entries)
{
    if (n_items == 0) return hipSuccess;
    const int grid = blocks > 0 ? blocks : std::min<int>(kGroupDefaultBlocks, (int)((n_items + 3) / 4));
    const int w = waves_per_eu;
    const uint64_t *const e = entries;
    switch (lanes) {
    case 2: group_insert_w<2>(w, grid, st, dir, items, n_items, keys, vals, pol, C, sink, table_counts, e); break;
    case 4: group_insert_w<4>(w, grid, st, dir, items, n_items, keys, vals, pol, C, sink, table_counts, e); break;
    case 8: group_insert_w<8>(w, grid, st, dir, items, n_items, keys, vals, pol, C, sink, table_counts, e); break;
    case 16: group_insert_w<16>(w, grid, st, dir, items, n_items, keys, vals, pol, C, sink, table_counts, e); break;
    default: group_insert_w<1>(w, grid, st, dir, items, n_items, keys, vals, pol, C, sink, table_counts, e); break;
    }
    return hipGetLastError();
}

hipError_t launch_copy(const void *src0, void *dst0, uint64_t words0, const void *src1, void *dst1, uint64_t words1,
                       hipStream_t st)
{
    const uint64_t words = std::max(words0, words1);
    if (words == 0) return hipSuccess;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 256) blocks = 256;
    copy_kernel<<<(int)blocks, 256, 0, st>>>(
        CopySeg{static_cast<const unsigned long long *>(src0), static_cast<unsigned long long *>(dst0), words0},
        CopySeg{static_cast<const unsigned long long *>(src1), static_cast<unsigned long long *>(dst1), words1});
    return hipGetLastError();
}

static int serve_grid(uint32_t n_items, int blocks)
{
    return blocks > 0 ? blocks : std::min<int>(kGroupDefaultBlocks, (int)((n_items + 3) / 4));
}

hipError_t launch_serve_find(const Desc *dir, const ServeItem *items, uint32_t n_items, Pool pool, int lanes,
                             int blocks, hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    const int grid = serve_grid(n_items, blocks);
    switch (lanes) {
    case 4: serve_probe_kernel<4, false><<<grid, 256, 0, st>>>(dir, items, n_items, pool, nullptr); break;
    case 16: serve_probe_kernel<16, false><<<grid, 256, 0, st>>>(dir, items, n_items, pool, nullptr); break;
    default: serve_probe_kernel<8, false><<<grid, 256, 0, st>>>(dir, items, n_items, pool, nullptr); break;
    }
    return hipGetLastError();
}

hipError_t launch_serve_erase(const Desc *dir, const ServeItem *items, uint32_t n_items, Pool pool,
                              unsigned long long *del_total, int lanes, int blocks, hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    const int grid = serve_grid(n_items, blocks);
    switch (lanes) {
    case 4: serve_probe_kernel<4, true><<<grid, 256, 0, st>>>(dir, items, n_items, pool, del_total); break;
    case 16: serve_probe_kernel<16, true><<<grid, 256, 0, st>>>(dir, items, n_items, pool, del_total); break;
    default: serve_probe_kernel<8, true><<<grid, 256, 0, st>>>(dir, items, n_items, pool, del_total); break;
    }
    return hipGetLastError();
}

hipError_t launch_serve_insert(const Desc *dir, const ServeItem *items, uint32_t n_items, InsertPolicy pol, int lanes,
                               int blocks, Counters *C, PendSink sink, unsigned long long *ins_total, hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    const int grid = serve_grid(n_items, blocks);
    switch (lanes) {
    case 2: serve_insert_kernel<2><<<grid, 256, 0, st>>>(dir, items, n_items, pol, C, sink, ins_total); break;
    case 4: serve_insert_kernel<4><<<grid, 256, 0, st>>>(dir, items, n_items, pol, C, sink, ins_total); break;
    case 8: serve_insert_kernel<8><<<grid, 256, 0, st>>>(dir, items, n_items, pol, C, sink, ins_total); break;
    case 16: serve_insert_kernel<16><<<grid, 256, 0, st>>>(dir, items, n_items, pol, C, sink, ins_total); break;
    default: serve_insert_kernel<1><<<grid, 256, 0, st>>>(dir, items, n_items, pol, C, sink, ins_total); break;
    }
    return hipGetLastError();
}

template <int GI>
static hipError_t tick_by_find_lanes(int grid, int find_lanes, const Desc *dir, const TickItem *items,
                                     uint32_t n_items, unsigned long long *cursor, TickStages done,
                                     InsertPolicy pol, Counters *C, PendSink sink, Pool pool,
                                     unsigned long long *ins_total, unsigned long long *del_total, uint32_t spin_cap,
                                     uint32_t max_back, hipStream_t st)
{
    switch (find_lanes) {
    case 4:
        serve_tick_kernel<GI, 4><<<grid, 256, 0, st>>>(dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                                       del_total, spin_cap, max_back);
        break;
    case 16:
        serve_tick_kernel<GI, 16><<<grid, 256, 0, st>>>(dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                                        del_total, spin_cap, max_back);
        break;
    default:
        serve_tick_kernel<GI, 8><<<grid, 256, 0, st>>>(dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                                       del_total, spin_cap, max_back);
        break;
    }
    return hipGetLastError();
}

bool serve_tick_supported(int insert_lanes, int find_lanes)
{
    return (insert_lanes == 1 || insert_lanes == 2 || insert_lanes == 4) &&
           (find_lanes == 4 || find_lanes == 8 || find_lanes == 16);
}

hipError_t launch_serve_tick(const Desc *dir, const TickItem *items, uint32_t n_items, unsigned long long *cursor,
                             TickStages done, InsertPolicy pol, int insert_lanes, int find_lanes, int blocks,
                             Counters *C, PendSink sink, Pool pool, unsigned long long *ins_total,
                             unsigned long long *del_total, uint32_t spin_cap, uint32_t max_back, hipStream_t st)
{
    if (n_items == 0) return hipSuccess;
    if (!serve_tick_supported(insert_lanes, find_lanes)) return hipErrorInvalidValue;
    if (max_back == 0) max_back = 1;
    const int grid = serve_grid(n_items, blocks);
    switch (insert_lanes) {
    case 2:
        return tick_by_find_lanes<2>(grid, find_lanes, dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                     del_total, spin_cap, max_back, st);
    case 4:
        return tick_by_find_lanes<4>(grid, find_lanes, dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                     del_total, spin_cap, max_back, st);
    default:
        return tick_by_find_lanes<1>(grid, find_lanes, dir, items, n_items, cursor, done, pol, C, sink, pool, ins_total,
                                     del_total, spin_cap, max_back, st);
    }
}

template <int GI>
static hipError_t loop_by_find_lanes(int grid, int find_lanes, const LoopArgs &A, InsertPolicy pol, Counters *C,
                                     PendSink sink, Pool pool, unsigned long long *ins_total,
                                     unsigned long long *del_total, hipStream_t st)
{
    switch (find_lanes) {
    case 4: serve_loop_kernel<GI, 4><<<grid, 256, 0, st>>>(A, pol, C, sink, pool, ins_total, del_total); break;
    case 16: serve_loop_kernel<GI, 16><<<grid, 256, 0, st>>>(A, pol, C, sink, pool, ins_total, del_total); break;
    default: serve_loop_kernel<GI, 8><<<grid, 256, 0, st>>>(A, pol, C, sink, pool, ins_total, del_total); break;
    }
    return hipGetLastError();
}

hipError_t launch_serve_loop(const LoopArgs &A, InsertPolicy pol, int insert_lanes, int find_lanes, int blocks,
                             Counters *C, PendSink sink, Pool pool, unsigned long long *ins_total,
                             unsigned long long *del_total, hipStream_t st)
{
    if (!serve_tick_supported(insert_lanes, find_lanes) || blocks < 2) return hipErrorInvalidValue;
    switch (insert_lanes) {
    case 2: return loop_by_find_lanes<2>(blocks, find_lanes, A, pol, C, sink, pool, ins_total, del_total, st);
    case 4: return loop_by_find_lanes<4>(blocks, find_lanes, A, pol, C, sink, pool, ins_total, del_total, st);
    default: return loop_by_find_lanes<1>(blocks, find_lanes, A, pol, C, sink, pool, ins_total, del_total, st);
    }
}

hipError_t launch_loop_reset(unsigned long long *words, hipStream_t st)
{
    loop_reset_kernel<<<1, 1, 0, st>>>(words);
    return hipGetLastError();
}

template <int GI, int GF>
static int loop_blocks_per_cu()
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, serve_loop_kernel<GI, GF>, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return per;
}

template <int GI>
static int loop_blocks_by_find_lanes(int find_lanes)
{
    switch (find_lanes) {
    case 4: return loop_blocks_per_cu<GI, 4>();
    case 16: return loop_blocks_per_cu<GI, 16>();
    default: return loop_blocks_per_cu<GI, 8>();
    }
}

int serve_loop_capacity(int insert_lanes, int find_lanes)
{
    if (!serve_tick_supported(insert_lanes, find_lanes)) return 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    int per = 0;
    switch (insert_lanes) {
    case 2: per = loop_blocks_by_find_lanes<2>(find_lanes); break;
    case 4: per = loop_blocks_by_find_lanes<4>(find_lanes); break;
    default: per = loop_blocks_by_find_lanes<1>(find_lanes); break;
    }
    return per * cus;
}

hipError_t launch_zero(unsigned long long *p, uint64_t words, hipStream_t st)
{
    if (words == 0) return hipSuccess;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    zero_kernel<<<(int)blocks, 256, 0, st>>>(p, words);
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void fold_post_kernel(unsigned long long *block, unsigned long long *post,
                                                       unsigned long long seq)
{
    constexpr int kWords = (int)(sizeof(Counters) / sizeof(unsigned long long));
    const int i = threadIdx.x;
    unsigned long long v = 0;
    if (i < kWords) {
        v = block[i];
        for (int s = 1; s <= kTallyStripes; ++s) {
            const unsigned long long x = block[kTallyStrideWords * s + i];
            v = i == kMaxChainWord ? (v > x ? v : x) : v + x;
        }
        post[1 + i] = v;  // stored once every load of the sum has returned
    }
    // the clears come after those stores in program order, so after every load of the block
    for (int w = i; w < kCounterWords; w += 64) block[w] = 0;
    __threadfence_system();
    if (i == 0) post[0] = seq;
}

hipError_t launch_fold_post(Counters *C, unsigned long long *post, unsigned long long seq, hipStream_t st)
{
    fold_post_kernel<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long *>(C), post, seq);
    return hipGetLastError();
}

hipError_t launch_probe_count(const Desc &D, const uint32_t *keys, uint64_t n, Counters *C, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    probe_count_kernel<<<(int)blocks, 256, 0, st>>>(D, keys, n, C);
    return hipGetLastError();
}

hipError_t launch_count(const uint64_t *sub, uint64_t slots, Counters *C, hipStream_t st)
{
    uint64_t blocks = (slots + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    count_kernel<<<(int)blocks, 256, 0, st>>>(sub, slots, C);
    return hipGetLastError();
}

}  // namespace dyc
