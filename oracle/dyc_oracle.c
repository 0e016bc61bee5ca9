/*
 * dyc_oracle.c -- sequential CPU restatement of DyCuckoo (GPU/dyhash). TEST INFRASTRUCTURE:
 * a checker for the HIP library, never part of it. See dyc_oracle.h for the contract.
 *
 * Every function cites the reference lines it restates (paths relative to the reference
 * repository root). Parity status of each part is recorded in the tests that pin it:
 *   hashes h1..h5     -- pinned by independent inverse functions (tests/test_oracle_hashes.py)
 *   bucket / pair     -- pinned by big-integer arithmetic and the exhaustive pair table;
 *                        the pair tables of d = 2, 3, 5, 6 by the property that they list
 *                        every pair of distinct subtables once, and by the reference's
 *                        comments that spell the pairs out
 *   sizing            -- pinned by a sieve (brute force)
 *   insert replay     -- pinned by hand-derived placements, the structural invariant and
 *                        set equality against a dict
 *   find / erase      -- pinned by a dict of the inserted set
 *   upsize / downsize -- pinned by the bucket-arithmetic identities h%2n in {h%n, h%n+n}
 *                        and set preservation
 */
#include "dyc_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define KEY_OF(e) ((uint32_t)((e) >> 32))
#define VAL_OF(e) ((uint32_t)((e)&0xffffffffu))
#define PACK(k, v) ((((uint64_t)(uint32_t)(k)) << 32) | (uint64_t)(uint32_t)(v))

static int reserved_key(uint32_t k) { return k == 0u || k == 0xffffffffu; }

/* ---------------------------------------------------------------- hashes */

/* cuckoo.cu:27-36 -- Thomas Wang's 32-bit shift hash */
uint32_t orc_h1(uint32_t k)
{
    k = ~k + (k << 15);
    k = k ^ (k >> 12);
    k = k + (k << 2);
    k = k ^ (k >> 4);
    k = k * 2057u;
    k = k ^ (k >> 16);
    return k;
}

/* cuckoo.cu:39-48 -- Robert Jenkins' 6-shift 32-bit hash */
uint32_t orc_h2(uint32_t a)
{
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

/* cuckoo.cu:50-53 */
uint32_t orc_h3(uint32_t k) { return ((k ^ 59064253u) + 72355969u) % ORC_PRIME; }

/* cuckoo.cu:56-64 -- Thomas Wang's 32-bit shift-multiply hash */
uint32_t orc_h4(uint32_t a)
{
    a = (a ^ 61u) ^ (a >> 16);
    a = a + (a << 3);
    a = a ^ (a >> 4);
    a = a * 0x27d4eb2du;
    a = a ^ (a >> 15);
    return a;
}

/* cuckoo.cu:110-120 -- Jenkins' 7-shift hash, used only for the pair choice */
uint32_t orc_h5(uint32_t a)
{
    a -= (a << 6);
    a ^= (a >> 17);
    a -= (a << 9);
    a ^= (a << 4);
    a -= (a << 3);
    a ^= (a << 10);
    a ^= (a >> 15);
    return a;
}

/* cuckoo.cu:80-89: the hash of subtable t, reduced modulo PRIME_uint (dy_hash.h:24) */
uint32_t orc_hash_t(uint32_t k, uint32_t t)
{
    switch (t % 4u) {
    case 0: return orc_h1(k) % ORC_PRIME;
    case 1: return orc_h2(k) % ORC_PRIME;
    case 2: return orc_h3(k) % ORC_PRIME;
    default: return orc_h4(k) % ORC_PRIME;
    }
}

/* cuckoo.cu:76-90 with get_table_length = Lsize >> 4 (cuckoo.cu:17-18) */
uint32_t orc_bucket(uint32_t k, uint32_t t, uint32_t nbuckets) { return orc_hash_t(k, t) % nbuckets; }

/* cuckoo.cu:190-193 */
uint32_t orc_pair(uint32_t k) { return orc_h5(k) % 6u; }

/* cuckoo.cu:197-204 */
uint32_t orc_pos1(uint32_t p)
{
    if (p & 1u) return 0u;
    if (p & 2u) return 2u;
    return 1u;
}

/* cuckoo.cu:206-213 */
uint32_t orc_pos2(uint32_t p)
{
    if (p & 1u) return (p >> 1) + 1u;
    if (p & 2u) return 3u;
    return (p >> 2) + 2u;
}

/* cuckoo.cu:310-312 */
uint32_t orc_get_pos(uint32_t p, uint32_t pre) { return pre == orc_pos1(p) ? orc_pos2(p) : orc_pos1(p); }

/* ---- the other subtable counts (diff_hash_num_pos_fun.cpp:14-188, the n-choose-2 variants
 * that cuckoo.cu:122-187 keeps commented out). Each d has its own choose_hash / pos1 / pos2;
 * d = 4 is the code above. */

/* number of pairs choose_hash ranges over: 2 for d = 2 (k & 1), else C(d, 2) */
uint32_t orc_pairs(uint32_t d) { return d == 2u ? 2u : d * (d - 1u) / 2u; }

/* choose_hash: d = 2 sig & 1 (:16-19); d = 3 h5 % 3 (:38-41); d = 4 h5 % 6 (cuckoo.cu:190-193);
 * d = 5 h5 % 10 (:120-123); d = 6 h5 % 15 (:153-156) */
uint32_t orc_pair_d(uint32_t k, uint32_t d)
{
    if (d == 2u) return k & 1u;
    if (d == 4u) return orc_pair(k);
    return orc_h5(k) % orc_pairs(d);
}

/* pos1 per d. d = 2: pos (:21-24). d = 3: the shift form, pos & 1 ? 1 : 0 (:43-48).
 * d = 5: 3 for pos 9, else the d = 4 rule (:126-135). d = 6: the switch of :158-172, whose
 * comma-separated case labels list the pairs (0,1)..(0,5), (1,2)..(1,5), (2,3)..(2,5), (3,4),
 * (3,5), (4,5) in order: pos1 is the block the pair index falls in. */
uint32_t orc_pos1_d(uint32_t p, uint32_t d)
{
    switch (d) {
    case 2: return p;
    case 3: return (p & 1u) ? 1u : 0u;
    case 4: return orc_pos1(p);
    case 5:
        if (p == 9u) return 3u;
        if (p & 1u) return 0u;
        if (p & 2u) return 2u;
        return 1u;
    default:
        if (p <= 4u) return 0u;
        if (p <= 8u) return 1u;
        if (p <= 11u) return 2u;
        if (p <= 13u) return 3u;
        return 4u;
    }
}

/* pos2 per d. d = 2: pos ^ 1 (:26-29). d = 3: pos & 1 ? 2 : (pos >> 1) + 1 (:50-55).
 * d = 5: 4 for pos 9; odd: (pos >> 1) + 1; pos & 2: (pos >> 2) + 3; else (pos >> 2) + 2
 * (:137-146). d = 6: pos + 1, pos - 3, pos - 6, pos - 8, 5 by block (:174-188). */
uint32_t orc_pos2_d(uint32_t p, uint32_t d)
{
    switch (d) {
    case 2: return p ^ 1u;
    case 3: return (p & 1u) ? 2u : (p >> 1) + 1u;
    case 4: return orc_pos2(p);
    case 5:
        if (p == 9u) return 4u;
        if (p & 1u) return (p >> 1) + 1u;
        if (p & 2u) return (p >> 2) + 3u;
        return (p >> 2) + 2u;
    default:
        if (p <= 4u) return p + 1u;
        if (p <= 8u) return p - 3u;
        if (p <= 11u) return p - 6u;
        if (p <= 13u) return p - 8u;
        return 5u;
    }
}

uint32_t orc_get_pos_d(uint32_t p, uint32_t pre, uint32_t d)
{
    return pre == orc_pos1_d(p, d) ? orc_pos2_d(p, d) : orc_pos1_d(p, d);
}

/* Victim vote over a full bucket (cuckoo.cu:396-401): the first slot whose resident's pair
 * index differs from the table number t. For d = 2 the vote is __ballot(true)
 * (diff_hash_num_pos_fun.cpp:3-6): every slot qualifies, so slot 0 is taken. */
static int vote_lane(const uint64_t *view, uint32_t t, uint32_t d)
{
    if (d == 2u) return 0;
    for (int i = 0; i < ORC_SLOTS; ++i)
        if (orc_pair_d(KEY_OF(view[i]), d) != t) return i;
    return -1;
}

void orc_hash_batch(const uint32_t *keys, uint64_t n, uint32_t which, uint32_t *out)
{
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t k = keys[i];
        switch (which) {
        case 1: out[i] = orc_h1(k); break;
        case 2: out[i] = orc_h2(k); break;
        case 3: out[i] = orc_h3(k); break;
        case 4: out[i] = orc_h4(k); break;
        case 5: out[i] = orc_h5(k); break;
        case 6: out[i] = orc_pair(k); break;
        default: out[i] = 0; break;
        }
    }
}

/* ---------------------------------------------------------------- sizing */

/* api.cpp:4-11. Note the reference reports 1 as prime (the 6k+-1 test admits it). */
int orc_is_prime(uint32_t n)
{
    int64_t num = (int32_t)n; /* the reference takes an int */
    if (num == 2 || num == 3) return 1;
    if (num % 6 != 1 && num % 6 != 5) return 0;
    for (int64_t i = 5; i * i <= num; i += 6)
        if (num % i == 0 || num % (i + 2) == 0) return 0;
    return 1;
}

/* api.cpp:13-22 */
uint32_t orc_next_prime(uint32_t n)
{
    while (!orc_is_prime(n)) ++n;
    return n;
}

/* api.cpp:47-51: s = nextPrime(size / (TABLE_NUM*BUCKET_SIZE)); buckets = s rounded up to even */
uint32_t orc_initial_buckets_d(uint64_t capacity_hint, uint32_t d)
{
    uint32_t s = (uint32_t)(capacity_hint / ((uint64_t)d * ORC_SLOTS));
    s = orc_next_prime(s);
    return (s & 1u) ? s + 1u : s;
}

uint32_t orc_initial_buckets(uint64_t capacity_hint) { return orc_initial_buckets_d(capacity_hint, ORC_SUBTABLES); }

/* ---------------------------------------------------------------- resize choice */

/* api.cpp:116-122 (loop bound TABLE_NUM): the first subtable that is half the size of its left
 * neighbour, else 0 */
uint32_t orc_choose_up_d(const uint32_t *nb, uint32_t d)
{
    for (uint32_t i = 0; i + 1 < d; ++i)
        if (nb[i] == 2u * nb[i + 1]) return i + 1;
    return 0;
}

/* api.cpp:147-153: the first subtable larger than its right neighbour, else the last */
uint32_t orc_choose_down_d(const uint32_t *nb, uint32_t d)
{
    for (uint32_t i = 0; i + 1 < d; ++i)
        if (nb[i] > nb[i + 1]) return i;
    return d - 1;
}

uint32_t orc_choose_up(const uint32_t nb[ORC_SUBTABLES]) { return orc_choose_up_d(nb, ORC_SUBTABLES); }
uint32_t orc_choose_down(const uint32_t nb[ORC_SUBTABLES]) { return orc_choose_down_d(nb, ORC_SUBTABLES); }

/* ---------------------------------------------------------------- insert */

/* Victim lane for the rescue phase that follows the reference's kick bound: a multiplicative
 * hash of the in-hand key and the iteration number (top 4 bits). Not in the reference, which
 * drops the entry at the kick bound (cuckoo.cu:422-431); the rescue turns its nearly
 * deterministic first-lane walk into a pseudo-random walk so that chains that cycle under
 * the reference rule terminate. */
uint32_t orc_rescue_lane(uint32_t key, uint32_t it)
{
    return (key * 0x9E3779B1u + it * 0x632BE5ABu) >> 28;
}

/* The eviction loop of warp_insert (cuckoo.cu:377-421), one key at a time:
 *   t      = the other table of the in-hand key's pair, relative to `pre`
 *   victim = first empty slot; else the vote of vote_lane (the reference compares a pair index
 *            with a table number, cuckoo.cu:398-400); else slot (iteration & 15)
 *   exchange, carry the old entry; done when the old entry is empty. */
int orc_insert_entry(orc_table *tb, uint64_t entry, uint32_t pre, uint32_t kick_bound,
                     uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry)
{
    uint32_t n_kicks = 0;
    const uint32_t total = kick_bound + rescue_kicks;
    for (uint32_t it = 0; it < total; ++it) {
        uint32_t k = KEY_OF(entry);
        uint32_t t = orc_get_pos_d(orc_pair_d(k, tb->d), pre, tb->d);
        uint64_t *b = tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS;
        int lane = -1;
        for (int i = 0; i < ORC_SLOTS; ++i)
            if (KEY_OF(b[i]) == 0u) { lane = i; break; }
        if (lane < 0 && it < kick_bound) {
            lane = vote_lane(b, t, tb->d);
            if (lane < 0) lane = (int)(it & 15u);
        }
        if (lane < 0) lane = (int)orc_rescue_lane(k, it);
        uint64_t old = b[lane];
        b[lane] = entry;
        entry = old;
        ++n_kicks;
        if (entry == 0) {
            if (kicks) *kicks = n_kicks;
            return 1;
        }
        pre = t;
    }
    if (kicks) *kicks = n_kicks;
    if (carry) *carry = entry;
    return 0;
}

/* cuckoo_insert (cuckoo.cu:336-348): odd keys start in pos2, even keys in pos1 */
int orc_insert_one(orc_table *tb, uint32_t key, uint32_t value, uint32_t kick_bound,
                   uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry)
{
    uint32_t pre = ORC_NO_TABLE;
    if (key & 1u) pre = orc_get_pos_d(orc_pair_d(key, tb->d), ORC_NO_TABLE, tb->d);
    return orc_insert_entry(tb, PACK(key, value), pre, kick_bound, rescue_kicks, kicks, carry);
}

uint64_t orc_insert(orc_table *tb, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                    uint32_t kick_bound, uint32_t rescue_kicks, uint32_t *kicks_out,
                    uint64_t *fail_out, uint64_t fail_cap, uint64_t *rejected)
{
    uint64_t nfail = 0, nrej = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t kicks = 0;
        uint64_t carry = 0;
        if (reserved_key(keys[i])) {
            ++nrej;
            if (kicks_out) kicks_out[i] = 0;
            continue;
        }
        if (!orc_insert_one(tb, keys[i], vals[i], kick_bound, rescue_kicks, &kicks, &carry)) {
            if (fail_out && nfail < fail_cap) fail_out[nfail] = carry;
            ++nfail;
        }
        if (kicks_out) kicks_out[i] = kicks;
    }
    if (rejected) *rejected = nrej;
    return nfail;
}

/* ---------------------------------------------------------------- find / erase */

static int first_match(const uint64_t *b, uint32_t k)
{
    for (int i = 0; i < ORC_SLOTS; ++i)
        if (KEY_OF(b[i]) == k) return i;
    return -1;
}

/* cuckoo.cu:478-497: pos1 bucket, then pos2 bucket; a miss leaves 0 (api.cpp:275) */
void orc_find(const orc_table *tb, const uint32_t *keys, uint64_t n, uint32_t *vals_out,
              uint8_t *hit_out, uint64_t *probes_out)
{
    uint64_t probes = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t k = keys[i], v = 0;
        uint8_t hit = 0;
        if (!reserved_key(k)) {
            uint32_t p = orc_pair_d(k, tb->d);
            uint32_t t = orc_pos1_d(p, tb->d);
            const uint64_t *b = tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS;
            int lane = first_match(b, k);
            ++probes;
            if (lane < 0) {
                t = orc_pos2_d(p, tb->d);
                b = tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS;
                lane = first_match(b, k);
                ++probes;
            }
            if (lane >= 0) {
                v = VAL_OF(b[lane]);
                hit = 1;
            }
        }
        if (vals_out) vals_out[i] = v;
        if (hit_out) hit_out[i] = hit;
    }
    if (probes_out) *probes_out = probes;
}

/* cuckoo.cu:541-563: clear every matching slot of the first bucket holding the key */
static int erase_in(uint64_t *b, uint32_t k)
{
    int any = 0;
    for (int i = 0; i < ORC_SLOTS; ++i)
        if (KEY_OF(b[i]) == k) {
            b[i] = 0;
            any = 1;
        }
    return any;
}

uint64_t orc_erase(orc_table *tb, const uint32_t *keys, uint64_t n, uint8_t *deleted_out)
{
    uint64_t ndel = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t k = keys[i];
        int del = 0;
        if (!reserved_key(k)) {
            uint32_t p = orc_pair_d(k, tb->d);
            uint32_t t = orc_pos1_d(p, tb->d);
            del = erase_in(tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS, k);
            if (!del) {
                t = orc_pos2_d(p, tb->d);
                del = erase_in(tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS, k);
            }
        }
        ndel += (uint64_t)del;
        if (deleted_out) deleted_out[i] = (uint8_t)del;
    }
    return ndel;
}

/* ---------------------------------------------------------------- resize */

/* cuckoo.cu:612-624: slot i of old bucket j moves to slot i of new bucket h % 2n */
void orc_upsize(const uint64_t *old_sub, uint32_t old_nb, uint64_t *new_sub, uint32_t t)
{
    uint32_t nnb = old_nb * 2u;
    memset(new_sub, 0, (size_t)nnb * ORC_SLOTS * sizeof(uint64_t));
    for (uint64_t j = 0; j < old_nb; ++j)
        for (int i = 0; i < ORC_SLOTS; ++i) {
            uint64_t e = old_sub[j * ORC_SLOTS + i];
            if (e != 0) new_sub[(uint64_t)orc_bucket(KEY_OF(e), t, nnb) * ORC_SLOTS + i] = e;
        }
}

/* cuckoo.cu:664-699: new bucket g takes old bucket g lane-for-lane and fills its holes from
 * old bucket g+n read back to front (lane i <- slot 15-i); entries of old bucket g+n that
 * found no hole are residuals. */
uint64_t orc_downsize_compact(const uint64_t *old_sub, uint32_t old_nb, uint64_t *new_sub,
                              uint64_t *res_out, uint64_t res_cap)
{
    if (old_nb & 1u) return UINT64_MAX;
    uint32_t n = old_nb / 2u;
    uint64_t nres = 0;
    for (uint64_t g = 0; g < n; ++g) {
        const uint64_t *lo = old_sub + g * ORC_SLOTS;
        const uint64_t *hi = old_sub + (g + n) * ORC_SLOTS;
        for (int i = 0; i < ORC_SLOTS; ++i) {
            uint64_t e = lo[i];
            uint64_t c = hi[ORC_SLOTS - 1 - i];
            if (e == 0) e = c;
            new_sub[g * ORC_SLOTS + i] = e;
            if (e != c && c != 0) {
                if (res_out && nres < res_cap) res_out[nres] = c;
                ++nres;
            }
        }
    }
    return nres;
}

/* ---------------------------------------------------------------- checks */

uint64_t orc_count(const orc_table *tb)
{
    uint64_t c = 0;
    for (uint32_t t = 0; t < tb->d; ++t)
        for (uint64_t s = 0; s < (uint64_t)tb->nb[t] * ORC_SLOTS; ++s) c += tb->sub[t][s] != 0;
    return c;
}

uint64_t orc_check_invariant(const orc_table *tb)
{
    uint64_t bad = 0;
    for (uint32_t t = 0; t < tb->d; ++t)
        for (uint64_t b = 0; b < tb->nb[t]; ++b)
            for (int i = 0; i < ORC_SLOTS; ++i) {
                uint64_t e = tb->sub[t][b * ORC_SLOTS + i];
                if (e == 0) continue;
                uint32_t k = KEY_OF(e), p = orc_pair_d(k, tb->d);
                if (reserved_key(k) || (orc_pos1_d(p, tb->d) != t && orc_pos2_d(p, tb->d) != t) ||
                    orc_bucket(k, t, tb->nb[t]) != b)
                    ++bad;
            }
    return bad;
}

/* ---------------------------------------------------------------- wide keys */

#define WIDE_BUSY (~(uint64_t)0)

static int wide_reserved(uint64_t k) { return k == 0 || k == WIDE_BUSY; }

/* fold of a 64-bit key onto the 32-bit hashes: lo ^ h5(hi); h5(0) = 0, so keys below 2^32
 * hash as the 32-bit table hashes them */
uint32_t orc_fold64(uint64_t k) { return (uint32_t)k ^ orc_h5((uint32_t)(k >> 32)); }

static uint64_t wide_at(const orc_wide *tb, uint32_t t, uint32_t fk)
{
    return (uint64_t)orc_bucket(fk, t, tb->nb[t]) * ORC_SLOTS;
}

/* warp_insert's loop (cuckoo.cu:377-421) over fold(k), key and value exchanged together */
int orc_wide_insert_entry(orc_wide *tb, uint64_t key, uint64_t val, uint32_t pre, uint32_t kick_bound,
                          uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry_key, uint64_t *carry_val)
{
    uint32_t n_kicks = 0;
    const uint32_t total = kick_bound + rescue_kicks;
    for (uint32_t it = 0; it < total; ++it) {
        const uint32_t f = orc_fold64(key);
        const uint32_t t = orc_get_pos_d(orc_pair_d(f, tb->d), pre, tb->d);
        const uint64_t at = wide_at(tb, t, f);
        uint64_t *bk = tb->keys[t] + at, *bv = tb->vals[t] + at;
        int lane = -1;
        for (int i = 0; i < ORC_SLOTS; ++i)
            if (bk[i] == 0) { lane = i; break; }
        if (lane < 0 && it < kick_bound) {
            if (tb->d == 2u) {
                lane = 0;
            } else {
                for (int i = 0; i < ORC_SLOTS; ++i)
                    if (orc_pair_d(orc_fold64(bk[i]), tb->d) != t) { lane = i; break; }
            }
            if (lane < 0) lane = (int)(it & 15u);
        }
        if (lane < 0) lane = (int)orc_rescue_lane(f, it);
        const uint64_t ok = bk[lane], ov = bv[lane];
        bk[lane] = key;
        bv[lane] = val;
        ++n_kicks;
        if (ok == 0) {
            if (kicks) *kicks = n_kicks;
            return 1;
        }
        key = ok;
        val = ov;
        pre = t;
    }
    if (kicks) *kicks = n_kicks;
    if (carry_key) *carry_key = key;
    if (carry_val) *carry_val = val;
    return 0;
}

uint64_t orc_wide_insert(orc_wide *tb, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint32_t kick_bound,
                         uint32_t rescue_kicks, uint32_t *kicks_out, uint64_t *fail_keys, uint64_t *fail_vals,
                         uint64_t fail_cap, uint64_t *rejected)
{
    uint64_t nfail = 0, nrej = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t kicks = 0;
        uint64_t ck = 0, cv = 0;
        if (wide_reserved(keys[i])) {
            ++nrej;
            if (kicks_out) kicks_out[i] = 0;
            continue;
        }
        uint32_t pre = ORC_NO_TABLE;  /* odd keys start in pos2 (cuckoo.cu:336-341) */
        if (keys[i] & 1u) pre = orc_get_pos_d(orc_pair_d(orc_fold64(keys[i]), tb->d), ORC_NO_TABLE, tb->d);
        if (!orc_wide_insert_entry(tb, keys[i], vals[i], pre, kick_bound, rescue_kicks, &kicks, &ck, &cv)) {
            if (nfail < fail_cap) {
                if (fail_keys) fail_keys[nfail] = ck;
                if (fail_vals) fail_vals[nfail] = cv;
            }
            ++nfail;
        }
        if (kicks_out) kicks_out[i] = kicks;
    }
    if (rejected) *rejected = nrej;
    return nfail;
}

/* cuckoo.cu:478-497 over fold(k): pos1 bucket, then pos2; a miss gives 0 */
void orc_wide_find(const orc_wide *tb, const uint64_t *keys, uint64_t n, uint64_t *vals_out, uint8_t *hit_out)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t k = keys[i];
        uint64_t v = 0;
        uint8_t hit = 0;
        if (!wide_reserved(k)) {
            const uint32_t f = orc_fold64(k), p = orc_pair_d(f, tb->d);
            for (int step = 0; step < 2 && !hit; ++step) {
                const uint32_t t = step == 0 ? orc_pos1_d(p, tb->d) : orc_pos2_d(p, tb->d);
                const uint64_t at = wide_at(tb, t, f);
                for (int j = 0; j < ORC_SLOTS; ++j)
                    if (tb->keys[t][at + j] == k) {
                        v = tb->vals[t][at + j];
                        hit = 1;
                        break;
                    }
            }
        }
        if (vals_out) vals_out[i] = v;
        if (hit_out) hit_out[i] = hit;
    }
}

uint64_t orc_wide_erase(orc_wide *tb, const uint64_t *keys, uint64_t n, uint8_t *deleted_out)
{
    uint64_t ndel = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t k = keys[i];
        int del = 0;
        if (!wide_reserved(k)) {
            const uint32_t f = orc_fold64(k), p = orc_pair_d(f, tb->d);
            for (int step = 0; step < 2 && !del; ++step) {
                const uint32_t t = step == 0 ? orc_pos1_d(p, tb->d) : orc_pos2_d(p, tb->d);
                const uint64_t at = wide_at(tb, t, f);
                for (int j = 0; j < ORC_SLOTS; ++j)
                    if (tb->keys[t][at + j] == k) {
                        tb->keys[t][at + j] = 0;
                        del = 1;
                    }
            }
        }
        ndel += (uint64_t)del;
        if (deleted_out) deleted_out[i] = (uint8_t)del;
    }
    return ndel;
}

/* cuckoo.cu:612-624: slot i of old bucket j moves to slot i of new bucket h % 2n (keys and
 * values); the new arrays are fully written */
void orc_wide_upsize(const uint64_t *old_keys, const uint64_t *old_vals, uint32_t old_nb, uint64_t *new_keys,
                     uint64_t *new_vals, uint32_t t)
{
    const uint32_t nnb = old_nb * 2u;
    memset(new_keys, 0, (size_t)nnb * ORC_SLOTS * sizeof(uint64_t));
    memset(new_vals, 0, (size_t)nnb * ORC_SLOTS * sizeof(uint64_t));
    for (uint64_t j = 0; j < old_nb; ++j)
        for (int i = 0; i < ORC_SLOTS; ++i) {
            const uint64_t k = old_keys[j * ORC_SLOTS + i];
            if (k == 0) continue;
            const uint64_t at = (uint64_t)orc_bucket(orc_fold64(k), t, nnb) * ORC_SLOTS + i;
            new_keys[at] = k;
            new_vals[at] = old_vals[j * ORC_SLOTS + i];
        }
}

/* cuckoo.cu:664-699: new bucket g keeps old bucket g slot for slot and fills its empty slots from
 * old bucket g+n back to front; an entry of bucket g+n is a residual unless it equals (key and
 * value) the one placed in its mirrored slot (the reference's comparison, cuckoo.cu:685-688) */
uint64_t orc_wide_downsize_compact(const uint64_t *old_keys, const uint64_t *old_vals, uint32_t old_nb,
                                   uint64_t *new_keys, uint64_t *new_vals, uint64_t *res_keys, uint64_t *res_vals,
                                   uint64_t res_cap)
{
    if (old_nb & 1u) return UINT64_MAX;
    const uint32_t n = old_nb / 2u;
    uint64_t nres = 0;
    for (uint64_t g = 0; g < n; ++g)
        for (int i = 0; i < ORC_SLOTS; ++i) {
            const uint64_t lo = g * ORC_SLOTS + i, hi = (g + n) * ORC_SLOTS + (ORC_SLOTS - 1 - i);
            uint64_t k = old_keys[lo], v = old_vals[lo];
            const uint64_t ck = old_keys[hi], cv = old_vals[hi];
            if (k == 0) {
                k = ck;
                v = cv;
            }
            new_keys[lo] = k;
            new_vals[lo] = v;
            if (ck != 0 && !(k == ck && v == cv)) {
                if (nres < res_cap) {
                    res_keys[nres] = ck;
                    res_vals[nres] = cv;
                }
                ++nres;
            }
        }
    return nres;
}

uint64_t orc_wide_check_invariant(const orc_wide *tb)
{
    uint64_t bad = 0;
    for (uint32_t t = 0; t < tb->d; ++t)
        for (uint64_t b = 0; b < tb->nb[t]; ++b)
            for (int i = 0; i < ORC_SLOTS; ++i) {
                const uint64_t k = tb->keys[t][b * ORC_SLOTS + i];
                if (k == 0) continue;
                const uint32_t f = orc_fold64(k), p = orc_pair_d(f, tb->d);
                if (wide_reserved(k) || (orc_pos1_d(p, tb->d) != t && orc_pos2_d(p, tb->d) != t) ||
                    orc_bucket(f, t, tb->nb[t]) != b)
                    ++bad;
            }
    return bad;
}

/* ---------------------------------------------------------------- CPU baseline */

/* A parallel loop over [0, n) in chunks taken from a shared counter, on `threads` POSIX threads
 * (the calling thread is one of them). Plain threads rather than OpenMP, so that the sanitizer
 * builds see every synchronisation (the joins and the atomics). */
typedef struct mt_loop {
    uint64_t n, chunk, next;
    void (*body)(struct mt_loop *, uint64_t lo, uint64_t hi);
    void *ctx;
    uint64_t sum0, sum1, sum2; /* per-loop totals, added atomically by each thread */
} mt_loop;

static void *mt_worker(void *arg)
{
    mt_loop *L = (mt_loop *)arg;
    for (;;) {
        const uint64_t lo = __atomic_fetch_add(&L->next, L->chunk, __ATOMIC_RELAXED);
        if (lo >= L->n) return 0;
        L->body(L, lo, lo + L->chunk < L->n ? lo + L->chunk : L->n);
    }
}

static void mt_run(mt_loop *L, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int i = 1; i < threads; ++i)
        if (pthread_create(&th[started], 0, mt_worker, L) == 0) ++started;
    mt_worker(L);
    for (int i = 0; i < started; ++i) pthread_join(th[i], 0);
}

typedef struct {
    const orc_table *tb;
    const uint32_t *keys;
    uint32_t *vals_out;
    uint8_t *hit_out;
} find_ctx;

static void find_body(mt_loop *L, uint64_t lo, uint64_t hi)
{
    const find_ctx *c = (const find_ctx *)L->ctx;
    orc_find(c->tb, c->keys + lo, hi - lo, c->vals_out ? c->vals_out + lo : 0, c->hit_out ? c->hit_out + lo : 0, 0);
}

void orc_find_mt(const orc_table *tb, const uint32_t *keys, uint64_t n, uint32_t *vals_out,
                 uint8_t *hit_out, int threads)
{
    find_ctx c = {tb, keys, vals_out, hit_out};
    mt_loop L = {n, 1u << 16, 0, find_body, &c, 0, 0, 0};
    mt_run(&L, threads);
}

/* Multithreaded insert: the same start table, hashes and victim rule as orc_insert_entry, with
 * the bucket read slot by slot (relaxed atomic loads) and each exchange a 64-bit CAS, keys
 * shared out in chunks of 4096 over POSIX threads (mt_run). A CAS that loses a race re-reads the bucket and repeats the step (at most
 * ORC_MT_RETRIES times per key). Entries still in hand after the bound, or after too many lost
 * races, are written to fail_out (up to fail_cap) and counted in the return value. Placement
 * depends on thread timing; the stored key->value set does not. Used only as the CPU timing
 * baseline and checked against the sequential insert by tests. */
#define ORC_MT_RETRIES 64u
typedef struct {
    orc_table *tb;
    const uint32_t *keys, *vals;
    uint32_t kick_bound, total;
    uint64_t *fail_out, fail_cap;
    uint64_t nfail;
} insert_ctx;

static void insert_body(mt_loop *L, uint64_t lo, uint64_t hi)
{
    insert_ctx *c = (insert_ctx *)L->ctx;
    orc_table *tb = c->tb;
    uint64_t kicks = 0, nrej = 0;
    for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t key = c->keys[i];
        if (reserved_key(key)) {
            ++nrej;
            continue;
        }
        uint32_t pre = ORC_NO_TABLE;
        if (key & 1u) pre = orc_get_pos_d(orc_pair_d(key, tb->d), ORC_NO_TABLE, tb->d);
        uint64_t entry = PACK(key, c->vals[i]);
        uint32_t it = 0, races = 0;
        int placed = 0;
        while (it < c->total && races <= ORC_MT_RETRIES) {
            const uint32_t k = KEY_OF(entry);
            const uint32_t t = orc_get_pos_d(orc_pair_d(k, tb->d), pre, tb->d);
            uint64_t *b = tb->sub[t] + (uint64_t)orc_bucket(k, t, tb->nb[t]) * ORC_SLOTS;
            uint64_t view[ORC_SLOTS];
            for (int j = 0; j < ORC_SLOTS; ++j) view[j] = __atomic_load_n(&b[j], __ATOMIC_RELAXED);
            int lane = -1;
            for (int j = 0; j < ORC_SLOTS; ++j)
                if (KEY_OF(view[j]) == 0u) { lane = j; break; }
            if (lane < 0 && it < c->kick_bound) {
                lane = vote_lane(view, t, tb->d);
                if (lane < 0) lane = (int)(it & 15u);
            }
            if (lane < 0) lane = (int)orc_rescue_lane(k, it);
            uint64_t old = view[lane];
            if (!__atomic_compare_exchange_n(&b[lane], &old, entry, 0, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED)) {
                ++races;
                continue;
            }
            ++it;
            entry = old;
            if (entry == 0) {
                placed = 1;
                break;
            }
            pre = t;
        }
        kicks += it;
        if (!placed) {
            uint64_t at = __atomic_fetch_add(&c->nfail, 1, __ATOMIC_RELAXED);
            if (c->fail_out && at < c->fail_cap) c->fail_out[at] = entry;
        }
    }
    __atomic_fetch_add(&L->sum0, kicks, __ATOMIC_RELAXED);
    __atomic_fetch_add(&L->sum1, nrej, __ATOMIC_RELAXED);
}

uint64_t orc_insert_mt(orc_table *tb, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                       uint32_t kick_bound, uint32_t rescue_kicks, uint64_t *fail_out, uint64_t fail_cap,
                       uint64_t *kicks_out, uint64_t *rejected_out, int threads)
{
    insert_ctx c = {tb, keys, vals, kick_bound, kick_bound + rescue_kicks, fail_out, fail_cap, 0};
    mt_loop L = {n, 4096, 0, insert_body, &c, 0, 0, 0};
    mt_run(&L, threads);
    if (kicks_out) *kicks_out = L.sum0;
    if (rejected_out) *rejected_out = L.sum1;
    return c.nfail;
}

int orc_max_threads(void)
{
    const char *e = getenv("OMP_NUM_THREADS");  /* the thread-count convention of this image */
    if (e && atoi(e) > 0) return atoi(e);
    const long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}
