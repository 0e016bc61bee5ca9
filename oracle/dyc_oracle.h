/*
 * dyc_oracle.h -- CPU restatement of DyCuckoo (GPU/dyhash) used ONLY as a test checker.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only to check or time against the HIP product path. Nothing
 * in databases_gpu_amd/ links or calls it.
 *
 * The oracle replays the reference semantics one key at a time (a sequential replay of the
 * 16-lane warp_insert / cuckoo_search / cuckoo_delete / resize kernels), on a caller-owned
 * table image whose byte layout is the layout the GPU library dumps:
 *   d subtables (4 by default); subtable t is nb[t] buckets x 16 slots of uint64 entries,
 *   entry = (uint64)key << 32 | value, empty slot = 0   (cuckoo.h:13-16, 30-34).
 *
 * Where the reference is ambiguous or buggy the oracle takes the documented decision:
 *   - values are packed unsigned (the reference sign-extends an int value, cuckoo.cu:338);
 *   - a kick-bound failure hands the in-hand entry back to the caller instead of dropping it
 *     (cuckoo.cu:422-431 drops it);
 *   - keys 0 and 0xFFFFFFFF are reserved: insert rejects them, find misses, erase skips;
 *   - downsize completes the whole compaction before any residual is re-inserted.
 */
#ifndef DYC_ORACLE_H
#define DYC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SUBTABLES 4            /* dy_hash.h:32 (TABLE_NUM), the default subtable count */
#define ORC_MAX_SUBTABLES 6
#define ORC_SLOTS 16
#define ORC_PRIME 294967291u      /* dy_hash.h:24 (PRIME_uint) */
#define ORC_KICK_BOUND 15         /* dy_hash.h:33 (MAX_ITER)   */
#define ORC_NO_TABLE 100u         /* cuckoo.cu:336 initial pre_table_no */

/* a caller-owned table image of d subtables (2..6): sub[t] points at nb[t]*16 uint64 slots */
typedef struct orc_table {
    uint64_t *sub[ORC_MAX_SUBTABLES];
    uint32_t nb[ORC_MAX_SUBTABLES];
    uint32_t d;
} orc_table;

/* ---- hashes (cuckoo.cu:27-64, 110-120) ---- */
uint32_t orc_h1(uint32_t k);
uint32_t orc_h2(uint32_t k);
uint32_t orc_h3(uint32_t k);
uint32_t orc_h4(uint32_t k);
uint32_t orc_h5(uint32_t k);
uint32_t orc_hash_t(uint32_t k, uint32_t t);                 /* h_{t%4}(k) % ORC_PRIME */
uint32_t orc_bucket(uint32_t k, uint32_t t, uint32_t nbuckets); /* cuckoo.cu:76-90 */
uint32_t orc_pair(uint32_t k);                               /* cuckoo.cu:190-193 */
uint32_t orc_pos1(uint32_t pair);                            /* cuckoo.cu:197-204 */
uint32_t orc_pos2(uint32_t pair);                            /* cuckoo.cu:206-213 */
uint32_t orc_get_pos(uint32_t pair, uint32_t pre);           /* cuckoo.cu:310-312 */
/* the same for d subtables (diff_hash_num_pos_fun.cpp:14-188); d = 4 gives the above */
uint32_t orc_pairs(uint32_t d);
uint32_t orc_pair_d(uint32_t k, uint32_t d);
uint32_t orc_pos1_d(uint32_t pair, uint32_t d);
uint32_t orc_pos2_d(uint32_t pair, uint32_t d);
uint32_t orc_get_pos_d(uint32_t pair, uint32_t pre, uint32_t d);

/* vectorised helpers for tests: out[i] = f(keys[i]) */
void orc_hash_batch(const uint32_t *keys, uint64_t n, uint32_t which, uint32_t *out);

/* ---- sizing (api.cpp:4-22, 47-56) ---- */
int orc_is_prime(uint32_t n);
uint32_t orc_next_prime(uint32_t n);
uint32_t orc_initial_buckets(uint64_t capacity_hint);  /* buckets per subtable */
uint32_t orc_initial_buckets_d(uint64_t capacity_hint, uint32_t d);

/* ---- resize choice (api.cpp:116-122, 147-153) ---- */
uint32_t orc_choose_up(const uint32_t nb[ORC_SUBTABLES]);
uint32_t orc_choose_down(const uint32_t nb[ORC_SUBTABLES]);
uint32_t orc_choose_up_d(const uint32_t *nb, uint32_t d);
uint32_t orc_choose_down_d(const uint32_t *nb, uint32_t d);

/* ---- operations ---- */
/* Victim lane used after the reference kick bound (rescue phase); see dyc_oracle.c. */
uint32_t orc_rescue_lane(uint32_t key, uint32_t it);

/* Insert one entry, replaying warp_insert (cuckoo.cu:336-348, 365-432) sequentially.
 * The first kick_bound iterations follow the reference victim rule; up to rescue_kicks more
 * iterations pick the victim with orc_rescue_lane (0 = the reference behaviour exactly).
 * Returns 1 on success, 0 on failure (*carry holds the entry left in hand).
 * *kicks receives the number of exchanges performed. */
int orc_insert_one(orc_table *tb, uint32_t key, uint32_t value, uint32_t kick_bound,
                   uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry);

/* Continue an eviction chain with an entry already in hand whose previous table is `pre`
 * (the residual path of cuckoo_resize_down, cuckoo.cu:719). Same return convention. */
int orc_insert_entry(orc_table *tb, uint64_t entry, uint32_t pre, uint32_t kick_bound,
                     uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry);

/* Batch insert in input order. kicks_out[i] (may be NULL) = exchanges for key i, 0 if rejected.
 * Entries left in hand after a failure are appended to fail_out (capacity fail_cap).
 * Returns the number of failures; *rejected (may be NULL) counts reserved keys. */
uint64_t orc_insert(orc_table *tb, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                    uint32_t kick_bound, uint32_t rescue_kicks, uint32_t *kicks_out,
                    uint64_t *fail_out, uint64_t fail_cap, uint64_t *rejected);

/* Find (cuckoo.cu:460-520): vals_out[i] = value or 0, hit_out[i] = 1/0 (either may be NULL).
 * probes_out (may be NULL) receives the number of buckets the pos1-then-pos2 order reads. */
void orc_find(const orc_table *tb, const uint32_t *keys, uint64_t n, uint32_t *vals_out,
              uint8_t *hit_out, uint64_t *probes_out);

/* Erase (cuckoo.cu:524-587): clears every matching slot of the first bucket (pos1, then pos2)
 * that holds the key. deleted_out[i] = 1 iff something was cleared. Returns #deleted. */
uint64_t orc_erase(orc_table *tb, const uint32_t *keys, uint64_t n, uint8_t *deleted_out);

/* Upsize subtable t (cuckoo.cu:600-627): old_sub has old_nb buckets, new_sub has 2*old_nb
 * buckets and is fully written (zeros included). */
void orc_upsize(const uint64_t *old_sub, uint32_t old_nb, uint64_t *new_sub, uint32_t t);

/* Downsize compaction of subtable t (cuckoo.cu:638-697): old_nb must be even; new_sub has
 * old_nb/2 buckets and is fully written. Residual entries are appended to res_out in
 * (new bucket, lane) order. Returns the number of residuals (or UINT64_MAX if old_nb is odd). */
uint64_t orc_downsize_compact(const uint64_t *old_sub, uint32_t old_nb, uint64_t *new_sub,
                              uint64_t *res_out, uint64_t res_cap);

/* Number of non-empty slots in the image. */
uint64_t orc_count(const orc_table *tb);

/* Structural invariant: every non-empty slot of subtable t, bucket b holds a key with
 * t in {pos1,pos2}(pair(key)) and b == bucket(key, t). Returns the number of violations. */
uint64_t orc_check_invariant(const orc_table *tb);

/* ---- wide keys: uint64 keys and values in parallel key / value arrays (tkde-revision/secs/
 * dynamic.tex:35-37). Placement is the 32-bit table's applied to fold(k) = lo ^ h5(hi); an
 * exchange swaps key and value together. Keys 0 and ~0 are reserved. ---- */
typedef struct orc_wide {
    uint64_t *keys[ORC_MAX_SUBTABLES];
    uint64_t *vals[ORC_MAX_SUBTABLES];
    uint32_t nb[ORC_MAX_SUBTABLES];
    uint32_t d;
} orc_wide;

uint32_t orc_fold64(uint64_t k);
int orc_wide_insert_entry(orc_wide *tb, uint64_t key, uint64_t val, uint32_t pre, uint32_t kick_bound,
                          uint32_t rescue_kicks, uint32_t *kicks, uint64_t *carry_key, uint64_t *carry_val);
/* Batch insert in input order; failed entries go to fail_keys/fail_vals. Returns #failures. */
uint64_t orc_wide_insert(orc_wide *tb, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint32_t kick_bound,
                         uint32_t rescue_kicks, uint32_t *kicks_out, uint64_t *fail_keys, uint64_t *fail_vals,
                         uint64_t fail_cap, uint64_t *rejected);
void orc_wide_find(const orc_wide *tb, const uint64_t *keys, uint64_t n, uint64_t *vals_out, uint8_t *hit_out);
/* Clears the keys of every matching slot of the first bucket holding the key (values stay). */
uint64_t orc_wide_erase(orc_wide *tb, const uint64_t *keys, uint64_t n, uint8_t *deleted_out);
void orc_wide_upsize(const uint64_t *old_keys, const uint64_t *old_vals, uint32_t old_nb, uint64_t *new_keys,
                     uint64_t *new_vals, uint32_t t);
uint64_t orc_wide_downsize_compact(const uint64_t *old_keys, const uint64_t *old_vals, uint32_t old_nb,
                                   uint64_t *new_keys, uint64_t *new_vals, uint64_t *res_keys, uint64_t *res_vals,
                                   uint64_t res_cap);
uint64_t orc_wide_check_invariant(const orc_wide *tb);

/* Multithreaded find over the same image (POSIX threads), used only as the CPU timing baseline. */
void orc_find_mt(const orc_table *tb, const uint32_t *keys, uint64_t n, uint32_t *vals_out,
                 uint8_t *hit_out, int threads);
/* Multithreaded insert (POSIX threads over key chunks, 64-bit CAS per exchange), used only as the CPU timing
 * baseline. Returns the number of entries left unplaced (written to fail_out up to fail_cap). */
uint64_t orc_insert_mt(orc_table *tb, const uint32_t *keys, const uint32_t *vals, uint64_t n,
                       uint32_t kick_bound, uint32_t rescue_kicks, uint64_t *fail_out, uint64_t fail_cap,
                       uint64_t *kicks_out, uint64_t *rejected_out, int threads);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
