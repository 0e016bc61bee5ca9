/*
 * dyc.h -- C ABI of the MI355X DyCuckoo library (libdyc.so).
 *
 * A dynamic cuckoo hash table of uint32 keys and uint32 values: d subtables (4 by default,
 * 2..6 through dyc_config.subtables) of 128-byte buckets (16 slots of (uint64)key << 32 |
 * value), two-layer pair hashing, bounded eviction, and single-subtable upsize/downsize -- the
 * data structure of GPU/dyhash in the reference.
 *
 * The reference exposes a C++ class (GPU/dyhash/api.h:9-55, hashAPI) over a device-pointer
 * layer (GPU/dyhash/cuckoo.h:100-114, gpu_lp_*). This header replaces both with an opaque
 * handle and plain pointers; include/dyc_hashapi.hpp rebuilds the hashAPI class on top of it.
 *
 * Conventions
 *   - Every function returns DYC_OK (0) or a DYC_ERR_* code; dyc_last_error() gives a message
 *     for the calling thread's last failure.
 *   - keys / vals / hit / deleted are DEVICE pointers (any memory the GPU can read or write:
 *     hipMalloc, managed or pinned host memory). The caller owns them; the library owns the
 *     table storage.
 *   - `stream` is a hipStream_t (NULL = the default stream). Work is stream-ordered. One table
 *     is driven from one host thread, and its calls run on one stream at a time (or on externally
 *     serialised streams), except dyc_insert_async, whose inserts run on several streams at once.
 *   - Keys 0 and 0xFFFFFFFF are reserved (0 marks an empty slot): insert skips them and counts
 *     them as rejected, find reports a miss, erase reports not-deleted.
 *   - Insert does not look for an existing copy of the key (reference semantics,
 *     dy_hash.h:42 update_on 0): the keys of one table should be unique.
 *   - An entry is never dropped. An eviction chain that exceeds kick_bound + rescue_kicks
 *     parks its entry on the table's pending list; a call that leaves the list longer than it
 *     found it upsizes and re-inserts pending entries (at most max_fail_rounds rounds). Entries still pending form the
 *     table's stash: find, erase and upsert search it for keys neither bucket holds (so copies
 *     of one key beyond its two buckets' 32 slots stay reachable).
 */
#ifndef DYC_H
#define DYC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DYC_ABI_VERSION 4

#define DYC_OK 0
#define DYC_ERR_INVALID 1        /* bad argument (null handle, bad config, size limit) */
#define DYC_ERR_HIP 2            /* a HIP runtime call failed */
#define DYC_ERR_OOM 3            /* device allocation failed */
#define DYC_ERR_PENDING 4        /* the call left more entries pending than it found, after
                                    max_fail_rounds upsizes (they stay in the stash) */
#define DYC_ERR_ODD_BUCKETS 5    /* downsize refused: the chosen subtable has an odd bucket count */
#define DYC_ERR_LOST 6           /* the pending list overflowed: entries were lost (table corrupt) */
#define DYC_ERR_CORRUPT 7        /* a resize found entries outside their bucket pair */

typedef struct dyc_table dyc_table;   /* opaque */
typedef void *dyc_stream;             /* hipStream_t */

/* Runtime configuration; the reference fixes these at compile time (dy_hash.h:15-73,
 * test_uilt.h:1-5). Fill with dyc_config_init() and override fields. */
typedef struct dyc_config {
    uint32_t kick_bound;       /* eviction iterations under the reference victim rule (MAX_ITER = 15) */
    uint32_t rescue_kicks;     /* further iterations with a pseudo-random victim (default 128) */
    uint32_t max_retries;      /* CAS races tolerated per entry before it is pended (default 64) */
    uint32_t max_fail_rounds;  /* upsize + re-insert rounds per dyc_insert call (default 8) */
    double lower_bound;        /* downsize when count < lower * capacity (default 0.50) */
    double upper_bound;        /* upsize while count + n > upper * capacity (default 0.90) */
    uint32_t auto_resize;      /* 1: apply the fill bounds in dyc_insert / dyc_erase (default 0) */
    uint32_t find_lanes;       /* lanes per key for find/erase: 4, 8 or 16 (default 8) */
    uint32_t insert_lanes;     /* lanes per key for insert: 1, 2, 4, 8 or 16 (default 1) */
    uint32_t block_threads;    /* threads per block, multiple of 64, <= 256 (default 256) */
    uint32_t max_blocks;       /* grid cap for find/erase, 0 = one wave per 64 keys up to 2^20 blocks */
    uint32_t insert_blocks;    /* grid of the insert kernel, 0 = default (65536 blocks) */
    uint64_t min_capacity;     /* no automatic downsize at or below this many slots
                                  (default 5033164 = TABLE_NUM*1024*1024*1.2, api.cpp:496) */
    uint32_t insert_order;     /* 0 or 1 = input order (default); 2 = order each batch by the bucket
                                  region of its keys' first bucket before inserting; 3 = bulk
                                  fill for batches of at least two keys per bucket: sort the batch
                                  by 512-bucket chunk of a first bucket (pos1, the one a find
                                  reads first, for as many keys as equal subtable shares allow),
                                  fill free slots chunk by chunk without atomics, the overflow
                                  likewise in its other bucket,
                                  then the eviction insert for the rest (workspace 24 B per key;
                                  wide tables 40 B). It reads two overflow counts back, so the
                                  call waits twice more on `stream` before its final wait */
    uint32_t region_mib;       /* table bytes per bucket region when ordering (default 64) */
    uint32_t subtables;        /* subtable count d, 2..6 (default 4 = TABLE_NUM, dy_hash.h:32); the
                                  pair scheme of each d is in diff_hash_num_pos_fun.cpp:14-188.
                                  Fixed at create: dyc_set_config must leave it unchanged */
    uint32_t reserved0;        /* must be 0 */
} dyc_config;

typedef struct dyc_stats {
    uint64_t count;             /* live entries: placed + pending */
    uint64_t capacity;          /* slots = 16 * sum(buckets) */
    uint32_t buckets[6];        /* buckets per subtable; entries at and past `subtables` are 0 */
    uint64_t pending;           /* entries owned but not placed */
    uint64_t kicks;             /* cumulative successful exchanges (insert + re-insert) */
    uint64_t last_kicks;        /* exchanges of the most recent dyc_insert */
    uint64_t kick_failures;     /* cumulative entries that exhausted the eviction bound */
    uint64_t cas_retries;       /* cumulative CAS races */
    uint64_t rejected;          /* cumulative reserved keys skipped by insert */
    uint64_t resizes_up;
    uint64_t resizes_down;
    uint64_t lost;              /* must stay 0 */
    uint32_t subtables;         /* subtable count d */
    uint32_t last_max_chain;    /* the longest eviction chain (exchanges) of the most recent insert call
                                   (32-bit tables; the maximum of the reference's iterator_count,
                                   api.cpp:219-243) */
    uint64_t last_filled;       /* entries the most recent insert call placed by bucket fill
                                   (insert_order 3); 0 when the eviction insert took the batch */
} dyc_stats;

/* Defaults (see field comments). */
int dyc_config_init(dyc_config *cfg);

/* Replaces hashAPI::hashAPI(TYPE size) (api.cpp:26-91): cfg->subtables subtables of
 * nextPrime(capacity_hint / (16 * subtables)) buckets each, rounded up to even (api.cpp:47-56),
 * zeroed. */
int dyc_create(const dyc_config *cfg, uint64_t capacity_hint, dyc_table **out);

/* Replaces hashAPI::~hashAPI (api.cpp:93-109). Frees everything the table owns. */
int dyc_destroy(dyc_table *t);

/* Replaces gpu_lp_insert (cuckoo.h:100-104, cuckoo.cu:858-877) and the insert half of
 * hashAPI::hash_insert / hash_insert_batch (api.cpp:184-265, 353-432).
 * With auto_resize, upsizes first while count + n > upper * capacity (api.cpp:360-369).
 * Waits once, at the end, for the work on `stream` (it reads the failure counters): when it
 * returns, every kernel queued on `stream` before and by the call has finished. */
int dyc_insert(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream);

/* Insert on a stream without waiting, so that inserts on several streams run at the same time
 * against one table: the reference's stream constructor and per-stream insert
 * (GPU/old_version/cuckoo_gtest/api.h:27-31, gpu_lp_insert(..., stream), cuckoo.cu:865-888) as
 * its HashStreamTEST.Four uses them (Main.cpp:72-113: one host thread, four streams, one insert
 * enqueued on each). Each call in flight has its own counter block and its own list for entries
 * that exhaust the eviction bound (up to 8 calls in flight; a ninth completes the others first).
 * keys/vals must stay valid until the call completes. dyc_sync(t) completes every insert in
 * flight: it waits for them, updates the count and counters, and re-places failed entries as
 * dyc_insert does (upsize + re-insert); every other call on the table does this first, so the
 * table is never observed half-way. Keys are placed with the eviction insert (insert_order is not
 * applied: the bulk fill's workspace belongs to the table). With auto_resize, a call that would
 * take the table past the upper bound completes the calls in flight and upsizes first. The keys of
 * calls in flight together should be unique, as for one dyc_insert. dyc_get_stats reports
 * completed calls only. */
int dyc_insert_async(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream);
int dyc_sync(dyc_table *t);

/* Insert without allocating. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Keys are placed as dyc_insert places them with the eviction insert
 * (insert_order 0; the bulk fill's workspace is never used), and the call makes no hipMalloc,
 * hipFree or other allocating call:
 *   - its counter block is a library-static device array of DYC_NOALLOC_SCRATCH_BYTES bytes (one
 *     copy per device, held until the process ends), not the table's own block;
 *   - nothing is resized (auto_resize is not applied here). Entries that exhaust the eviction
 *     bound join the table's stash (dyc_stats.pending), where find, erase and upsert see them;
 *     a later dyc_insert or dyc_resize_up + dyc_insert re-places them.
 * DYC_ERR_INVALID, with the table unchanged, while inserts of dyc_insert_async are in flight
 * (dyc_sync completes them, and may allocate), for n > DYC_NOALLOC_MAX_KEYS, or for n larger than the
 * free room of the table's pending list (dyc_create reserves DYC_NOALLOC_MAX_KEYS entries): the
 * list cannot grow here, and every entry must keep a place.
 * Every call on one device shares that device's counter block, which keeps the last call's
 * contents (each call clears it first): calls that overlap on one GPU -- from other host threads
 * or on other streams, for any table -- race, and the caller must serialise them. Waits once, as
 * dyc_insert. */
#define DYC_NOALLOC_MAX_KEYS 65536u
#define DYC_NOALLOC_SCRATCH_BYTES 16640u
int dyc_insert_noalloc(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream);

/* dyc_insert that also reports every key's eviction chain: iter_count[i] = the exchanges of the
 * chain keys[i] started (its placement counts as one; 0 for a reserved key, and for a chain that
 * parked on the pending list before its first exchange because its CAS retries ran out), the reference's
 * iterator_count (gpu_lp_insert(..., d_iter_count), cuckoo.h:100-104, written at cuckoo.cu:349-351
 * under ITERAOTOR_COUNT_FLAG; api.cpp:219-243 prints their maximum and mean). iter_count is a
 * device array of n words. Keys are placed with the eviction insert whatever insert_order says.
 * Re-placement of failed entries after an upsize is not counted again. Waits once, as dyc_insert. */
int dyc_insert_counted(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, uint32_t *iter_count,
                       dyc_stream stream);

/* Insert with update (the paper's insert, tkde-revision/secs/cuckoo.tex:127-129; the reference
 * code has it switched off, dy_hash.h:42 update_on 0). A key already stored gets vals[i] in every
 * copy held by the first bucket (pos1, then pos2) that holds it -- what dyc_find returns -- and
 * updated[i] = 1 (updated may be NULL); a key in neither bucket but in the stash gets it in every
 * stash copy. Every other key is inserted as by dyc_insert. Keys within one call must be
 * distinct. Waits once, as dyc_insert. */
int dyc_upsert(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, uint8_t *updated,
               dyc_stream stream);

/* Replaces gpu_lp_search (cuckoo.h:106-108, cuckoo.cu:881-899) / hashAPI::hash_search
 * (api.cpp:267-285). vals[i] = the value of keys[i] or 0 on a miss (what the reference's
 * pre-zeroed output shows); hit[i] = 1/0. Either output may be NULL. Asynchronous. */
int dyc_find(dyc_table *t, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream);

/* Replaces gpu_lp_delete (cuckoo.h:110-112, cuckoo.cu:901-916) / hashAPI::hash_delete
 * (api.cpp:289-303). Clears every copy of the key in the first bucket (pos1, then pos2) that
 * holds it; deleted[i] = 1 iff this call removed it (may be NULL; the reference's `ans` is
 * unused). With auto_resize, downsizes once when count < lower * capacity and capacity >
 * min_capacity (api.cpp:490-505). Waits once, as dyc_insert (it reads the delete count). */
int dyc_erase(dyc_table *t, const uint32_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream);

/* Replaces hashAPI::resize_up (api.cpp:112-141) + GPU_cuckoo_resize_up (cuckoo.cu:731-758):
 * doubles the subtable chosen by api.cpp:116-122. Synchronous. */
int dyc_resize_up(dyc_table *t, dyc_stream stream);

/* Replaces hashAPI::resize_low (api.cpp:143-173) + GPU_cuckoo_resize_down (cuckoo.cu:762-788):
 * halves the subtable chosen by api.cpp:147-153 and re-inserts the entries that no longer fit
 * into the other subtable of their pair. DYC_ERR_ODD_BUCKETS if that subtable's bucket count
 * is odd (the reference reads past the end there, cuckoo.cu:653,767-769). Synchronous. */
int dyc_resize_down(dyc_table *t, dyc_stream stream);

/* Host-pointer forms of insert / find / erase, for callers whose arrays live in host memory
 * (the reference's hashAPI takes host arrays and copies them, api.cpp:204-209,272-283). The
 * batch moves in chunks through library-owned pinned buffers: the copy of one chunk overlaps the
 * kernel of the previous one, and results stream back on a second copy stream. Same semantics
 * and resize policy as the device-pointer calls; synchronous (they return when the results are
 * in the caller's arrays). dyc_insert_host places each staged chunk with the eviction insert
 * whatever insert_order says (a chunk is too small for the bulk fill); the stored set is the
 * same. */
int dyc_insert_host(dyc_table *t, const uint32_t *keys, const uint32_t *vals, uint64_t n, dyc_stream stream);
int dyc_find_host(dyc_table *t, const uint32_t *keys, uint32_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream);
int dyc_erase_host(dyc_table *t, const uint32_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream);

/* Checkpoint a table to a file and restore it (no reference counterpart: the reference keeps
 * tables in device memory only). The file is a 128-byte header (magic "DYCKPT01", version 2,
 * subtable count d, buckets per subtable, live count, pending count, FNV-1a checksums of every
 * part and of the header), then the d subtable images and the pending entries as little-endian
 * u64 slots. Restore validates the header, the file size, every checksum and that the stored
 * entries match the count (DYC_ERR_CORRUPT otherwise), and creates a table with `cfg` (NULL =
 * defaults; its subtables field is replaced by the file's d) and the file's geometry. */
int dyc_checkpoint_write(dyc_table *t, const char *path, dyc_stream stream);
int dyc_checkpoint_read(const char *path, const dyc_config *cfg, dyc_table **out);

/* Read / replace the run-time configuration of a table (no reference counterpart: the
 * reference's configuration is compile-time, dy_hash.h:15-73). Sizes are not touched; every
 * field of dyc_config is otherwise taken, validated as in dyc_create. */
int dyc_get_config(const dyc_table *t, dyc_config *out);
int dyc_set_config(dyc_table *t, const dyc_config *cfg);

/* Counters and geometry (no reference counterpart beyond the occupancy print of
 * computer_table, cuckoo.cu:818-848). Does not touch the device. */
int dyc_get_stats(const dyc_table *t, dyc_stats *out);

/* Occupied slots counted on the device (computer_table, cuckoo.cu:818-848). Synchronous. */
int dyc_count_occupied(dyc_table *t, uint64_t *occupied, dyc_stream stream);

/* Empty the table (every slot zeroed, count and pending reset); sizes are kept. Asynchronous.
 * No reference counterpart (the reference allocates a new table, api.cpp:58-74). */
int dyc_clear(dyc_table *t, dyc_stream stream);

/* Diagnostic, not on the hot path: the number of buckets the find order (pos1, then pos2 on a
 * miss, cuckoo.cu:482-497) reads for keys[0..n), and how many keys pos1 resolves. The
 * algorithmic bytes of a find are 8 + 128 * probes / n (+1 with hit flags). Synchronous. */
int dyc_probe_count(dyc_table *t, const uint32_t *keys, uint64_t n, uint64_t *probes, uint64_t *pos1_hits,
                    dyc_stream stream);

/* Copy subtable `sub` (buckets[sub] * 16 uint64 slots) into host memory. Synchronous. */
int dyc_dump(dyc_table *t, uint32_t sub, uint64_t *host_slots, uint64_t n_slots, dyc_stream stream);

/* Replace the table contents with host images: nb[s] buckets per subtable, images[s] holding
 * nb[s] * 16 slots, for s < the table's subtable count. Pending entries are discarded and count
 * is recomputed. Synchronous. */
int dyc_load(dyc_table *t, const uint32_t *nb, const uint64_t *const *images, dyc_stream stream);

/* Copy up to `cap` pending entries to host memory; *n_out = number pending. */
int dyc_pending(dyc_table *t, uint64_t *host_entries, uint64_t cap, uint64_t *n_out, dyc_stream stream);

/* Message of the calling thread's last error ("" if none). */
const char *dyc_last_error(void);

/* DYC_ABI_VERSION of the loaded library. */
int dyc_abi_version(void);

/* ---------------------------------------------------------------------------------------
 * Many tables, one launch. No reference counterpart: the reference serves one table per
 * process (cuckoo.cu:16). A group holds n_tables independent DyCuckoo tables (same layout,
 * hashes and eviction rule as above) in one device arena. A batch is a CSR-style segmentation
 * of device arrays: the keys of table i are keys[offsets[i] .. offsets[i+1]) (offsets is a HOST
 * array of n_tables + 1 entries, offsets[0] = 0). One persistent kernel serves the whole batch
 * (work items of up to item_keys keys of one table); the *_per_table variants launch the
 * single-table kernels once per non-empty table instead -- the plain per-batch launch the
 * persistent form is measured against. Group tables do not resize: size them for their keys.
 * --------------------------------------------------------------------------------------- */
typedef struct dyc_group dyc_group;   /* opaque */

/* Launch shape of a group's kernels; dyc_group_tune() picks these by measurement. */
typedef struct dyc_group_launch {
    uint32_t find_lanes;     /* 4, 8 or 16 */
    uint32_t insert_lanes;   /* 1, 2, 4, 8 or 16 */
    uint32_t find_blocks;    /* persistent grid for find/erase (0 = default) */
    uint32_t insert_blocks;  /* persistent grid for insert (0 = default) */
    uint32_t item_keys;      /* keys per work item (0 = default 4096) */
    uint32_t kick_bound;     /* eviction iterations under the reference victim rule */
    uint32_t insert_waves_per_eu;  /* occupancy: 6 or 8 = the insert kernel built for at least that
                                      many waves per SIMD (fewer registers); 0 = the compiler's own */
    uint32_t find_waves_per_eu;    /* the same for find and erase */
} dyc_group_launch;

/* capacity_hints: host array of n_tables slot-count hints (sized per table as dyc_create). */
int dyc_group_create(const dyc_config *cfg, uint32_t n_tables, const uint64_t *capacity_hints, dyc_group **out);
int dyc_group_destroy(dyc_group *g);
int dyc_group_get_launch(const dyc_group *g, dyc_group_launch *out);
int dyc_group_set_launch(dyc_group *g, const dyc_group_launch *launch);

/* Insert / find / erase over all tables of a batch. Insert and erase synchronise `stream`
 * once (per-table counts); find is asynchronous. Insert entries that exhaust the eviction
 * bound stay pending in the group, tagged with their table (DYC_ERR_PENDING; dyc_group_pending
 * lists them); they are the group's stash, which finds and erases of that table search after
 * both buckets. */
int dyc_group_insert(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *offsets,
                     dyc_stream stream);
int dyc_group_find(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint32_t *vals, uint8_t *hit,
                   dyc_stream stream);
int dyc_group_erase(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint8_t *deleted,
                    dyc_stream stream);
int dyc_group_insert_per_table(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *offsets,
                               dyc_stream stream);
int dyc_group_find_per_table(dyc_group *g, const uint32_t *keys, const uint64_t *offsets, uint32_t *vals,
                             uint8_t *hit, dyc_stream stream);

/* Per-table geometry and counts (count, capacity, buckets; kick/retry figures are group-wide). */
int dyc_group_table_stats(const dyc_group *g, uint32_t table, dyc_stats *out);
/* Copy subtable `sub` of `table` into host memory (buckets[sub] * 16 slots). Synchronous. */
int dyc_group_dump(dyc_group *g, uint32_t table, uint32_t sub, uint64_t *host_slots, uint64_t n_slots,
                   dyc_stream stream);
/* Pending entries and the table each belongs to (host arrays of `cap`); *n_out = number pending. */
int dyc_group_pending(dyc_group *g, uint64_t *host_entries, uint32_t *host_tables, uint64_t cap, uint64_t *n_out,
                      dyc_stream stream);
/* Empty every table of the group. Asynchronous. */
int dyc_group_clear(dyc_group *g, dyc_stream stream);

/* A group step captured once into a HIP graph and replayed with one launch: clear every table,
 * insert keys/vals (segmented by ins_offsets, as dyc_group_insert), then find lookups (segmented
 * by find_offsets, as dyc_group_find) into vals_out / hit (optional). The pointers and offsets
 * are fixed when the plan is made; the buffers' contents may change between runs. Each run
 * equals dyc_group_clear + dyc_group_insert + dyc_group_find with the launch shape the group had
 * when the plan was made, and returns what that insert would (DYC_ERR_PENDING when entries
 * exhausted the eviction bound; the find of the same run searched them). It waits once, at the
 * end. No counterpart in the reference, whose calls launch and synchronise one by one
 * (cuckoo.cu:858-916); for callers that replay the same step shape many times (a per-batch
 * build-and-probe join over fixed partition buffers). A plan keeps work-item buffers of its own;
 * destroy plans before their group. */
typedef struct dyc_group_plan dyc_group_plan;  /* opaque */
int dyc_group_plan_create(dyc_group *g, const uint32_t *keys, const uint32_t *vals, const uint64_t *ins_offsets,
                          const uint32_t *lookups, const uint64_t *find_offsets, uint32_t *vals_out, uint8_t *hit,
                          dyc_group_plan **out);
int dyc_group_plan_run(dyc_group_plan *p, dyc_stream stream);
int dyc_group_plan_destroy(dyc_group_plan *p);

/* ---------------------------------------------------------------------------------------
 * Serving loop: many resizable tables behind one request queue (no reference counterpart; the
 * reference serves one synchronous batch per call, api.cpp:185-300). Requests are the
 * reference's batch calls -- hash_insert / hash_search / hash_delete of one table -- submitted
 * asynchronously. The server admits queued requests into "ticks" (continuous batching): each
 * tick is one fused insert launch, one fused find launch and one fused erase launch over every
 * admitted request, and the next tick is built and enqueued while the previous one runs.
 *
 * Admission, per tick, in arrival order:
 *   - slots: at most tick_max_requests requests and tick_max_keys keys (a larger request is
 *     admitted alone);
 *   - order: a table's requests run in arrival order. A tick runs in rounds (one fused insert,
 *     find and erase launch each); within a round a table's requests follow insert -> find ->
 *     erase with at most one erase, and a request that would break this goes to the table's next
 *     round (at most 64 rounds per tick; past that it waits for a later tick);
 *   - memory: all table storage lives in one arena of hbm_budget bytes. An insert that needs
 *     its table grown (count + n > upper_bound * capacity) is admitted only once the upsizes fit
 *     in the arena's free space; otherwise it (and its table's later requests) waits.
 * Resizes run in-line: upsizes are enqueued ahead of the tick that needs them; a table whose
 * erases bring it below lower_bound * capacity (and above min_capacity) is downsized once it
 * has no request in flight. Entries that exhaust the eviction bound go to a server-wide pool
 * that finds and erases also search, and are re-placed (upsize + re-insert) when their tick
 * completes, so every request's results equal those of running the requests one at a time in
 * arrival order.
 *
 * keys / vals / out_vals / out_flags are device pointers that must stay valid (and inputs
 * unchanged) until the request completes. One server is driven from one host thread; its work
 * runs on a stream the server owns. */
typedef struct dyc_server dyc_server;  /* opaque */

#define DYC_OP_INSERT 0u  /* hash_insert: vals required; out_* unused */
#define DYC_OP_FIND 1u    /* hash_search: out_vals (0 on a miss) and out_flags (hit) optional */
#define DYC_OP_ERASE 2u   /* hash_delete: out_flags (deleted) optional */

typedef struct dyc_request {
    uint32_t table;
    uint32_t op;
    const uint32_t *keys;
    const uint32_t *vals;
    uint32_t *out_vals;
    uint8_t *out_flags;
    uint64_t n;
} dyc_request;

typedef struct dyc_server_config {
    uint64_t hbm_budget;         /* arena bytes for all tables (0 = half of the free device memory) */
    uint64_t tick_max_keys;      /* keys per tick over all requests (default 2^23) */
    uint32_t tick_max_requests;  /* requests per tick (default 4096) */
    uint32_t pipeline_depth;     /* ticks in flight at once, 1 or 2 (default 2) */
    uint32_t item_keys;          /* keys per work item; 0 = chosen per tick from its size */
    uint32_t pool_cap;           /* entries the overflow pool holds (default 65536) */
    uint32_t one_launch;         /* 1 (default): a tick is one launch whose items wait on their table's
                                    earlier rounds; 0: three launches per round. One launch needs
                                    insert_lanes 1, 2 or 4 (otherwise rounds are launched) */
    uint32_t spin_cap;           /* polls an item waits for its table before the tick is failed
                                    (default 2^24) */
    uint32_t poll_backoff;       /* a waiting wave sleeps up to this many x 256 cycles between polls
                                    (default 16) */
    uint32_t resident;           /* 1: submit / poll / drain go through the device-resident loop: one
                                    launch stays on the GPU and takes requests from a ring in host
                                    memory as they are submitted (no launch per tick). Needs
                                    insert_lanes 1, 2 or 4. Default 0: host-built ticks */
    uint32_t resident_blocks;    /* blocks of 256 threads of the loop launch (0: half of what the GPU
                                    keeps resident at once, leaving room for resizes beside it; never
                                    more than that capacity) */
    uint32_t resident_idle_ms;   /* the loop launch ends after this long without new requests; the
                                    next submission starts it again (default 100) */
    uint32_t pad;
} dyc_server_config;

typedef struct dyc_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t ticks;
    uint64_t rounds;             /* phase-ordered rounds over all ticks (three launches each at most) */
    uint64_t keys;               /* keys of completed requests */
    uint64_t deferred_order;     /* admission passes that held a request back for table order */
    uint64_t deferred_memory;    /* admission passes that held a request back for arena space */
    uint64_t deferred_slots;     /* admission passes that ended with requests left for lack of slots */
    uint64_t resizes_up, resizes_down;
    uint64_t pool_entries;       /* entries in the overflow pool now */
    uint64_t pool_rounds;        /* re-placement rounds run */
    uint64_t arena_bytes, arena_used, arena_peak;
    uint64_t queued, in_flight;  /* requests waiting for admission / admitted, not completed */
    uint64_t loop_launches;      /* launches of the device-resident loop */
    uint64_t loop_items;         /* work items published to it */
    uint64_t loop_pauses;        /* times it was stopped to re-place the overflow pool */
} dyc_server_stats;

int dyc_server_config_init(dyc_server_config *scfg);
/* n_tables tables, table i sized for capacity_hints[i] slots (reference sizing). cfg gives
 * the eviction rule, fill bounds, min_capacity and lanes; auto_resize is implied. */
int dyc_server_create(const dyc_config *cfg, const dyc_server_config *scfg, uint32_t n_tables,
                      const uint64_t *capacity_hints, dyc_server **out);
int dyc_server_destroy(dyc_server *s);
/* Queue one request; *ticket numbers requests 0, 1, 2, ... in submission order. */
int dyc_server_submit(dyc_server *s, const dyc_request *req, uint64_t *ticket);
/* Make progress without blocking: retire finished ticks, admit and enqueue new ones. Writes up
 * to cap finished tickets (and their status, DYC_OK or an error code) and sets *n_done. */
int dyc_server_poll(dyc_server *s, uint64_t *done_tickets, int *done_status, uint32_t cap, uint32_t *n_done);
/* Block until every submitted request has completed. */
int dyc_server_drain(dyc_server *s);
/* Replay a request trace in real time: request i is submitted at arrival_s[i] seconds after the
 * call starts (non-decreasing). mode 0: continuous batching with host-built ticks (submit + poll).
 * mode 2: continuous batching through the device-resident loop (see resident). mode 1: synchronous
 * per-request launches -- each request alone, in arrival order, with the plain single-table
 * kernels and a stream synchronisation, the reference's calling pattern. done_s[i] receives
 * the completion time (seconds since the call started) and status[i] the request status. */
int dyc_server_replay(dyc_server *s, const dyc_request *reqs, uint64_t n, const double *arrival_s, uint32_t mode,
                      double *done_s, int *status);
int dyc_server_get_stats(const dyc_server *s, dyc_server_stats *out);
int dyc_server_table_stats(const dyc_server *s, uint32_t table, dyc_stats *out);
/* Copy subtable `sub` of `table` into host memory (buckets[sub] * 16 slots); drains first. */
int dyc_server_dump(dyc_server *s, uint32_t table, uint32_t sub, uint64_t *host_slots, uint64_t n_slots);

/* ---------------------------------------------------------------------------------------
 * Wide keys: a DyCuckoo table of uint64 keys and uint64 values. No counterpart in the reference
 * code (32-bit keys only); the paper's layout for keys larger than 4 bytes
 * (tkde-revision/secs/dynamic.tex:35-37): a bucket is a 128-byte line of 16 keys plus the
 * parallel line of their 16 values. Placement is the 32-bit table's (geometry, subtable count,
 * pair scheme, victim rule, resize choice) applied to fold(k) = lo32(k) ^ h5(hi32(k)), so keys
 * below 2^32 sit exactly where the 32-bit table puts them. Keys 0 and 2^64 - 1 are reserved.
 * Same conventions as the 32-bit calls above (device pointers, status codes, stream order,
 * insert and erase synchronise once, entries are never dropped and pending ones form a stash
 * that find and erase search); dyc_config fields that apply: kick_bound, rescue_kicks,
 * max_retries, max_fail_rounds, fill bounds, auto_resize, min_capacity, block_threads,
 * max_blocks, insert_blocks, subtables. Insert takes one lane per key. */
typedef struct dyc_wide dyc_wide;   /* opaque */

int dyc_wide_create(const dyc_config *cfg, uint64_t capacity_hint, dyc_wide **out);
int dyc_wide_destroy(dyc_wide *t);
int dyc_wide_insert(dyc_wide *t, const uint64_t *keys, const uint64_t *vals, uint64_t n, dyc_stream stream);
/* Insert with update for wide keys, with dyc_upsert's meaning (the paper's insert,
 * tkde-revision/secs/cuckoo.tex:127-129): a stored key gets vals[i] in every copy held by the
 * first bucket (pos1, then pos2) that holds it, or else in every stash copy, and updated[i] = 1
 * (updated may be NULL); every other key is inserted. Keys within one call must be distinct.
 * Waits once, as dyc_wide_insert. */
int dyc_wide_upsert(dyc_wide *t, const uint64_t *keys, const uint64_t *vals, uint64_t n, uint8_t *updated,
                    dyc_stream stream);
/* vals[i] = the value of keys[i] or 0, hit[i] = 1/0 (either may be NULL). Asynchronous. */
int dyc_wide_find(dyc_wide *t, const uint64_t *keys, uint64_t *vals, uint8_t *hit, uint64_t n, dyc_stream stream);
int dyc_wide_erase(dyc_wide *t, const uint64_t *keys, uint8_t *deleted, uint64_t n, dyc_stream stream);
int dyc_wide_resize_up(dyc_wide *t, dyc_stream stream);
int dyc_wide_resize_down(dyc_wide *t, dyc_stream stream);
int dyc_wide_get_stats(const dyc_wide *t, dyc_stats *out);
int dyc_wide_clear(dyc_wide *t, dyc_stream stream);
/* Subtable `sub`'s key and value arrays (buckets[sub] * 16 words each; either may be NULL). */
int dyc_wide_dump(dyc_wide *t, uint32_t sub, uint64_t *host_keys, uint64_t *host_vals, uint64_t n_slots,
                  dyc_stream stream);
/* Replace the contents with host images (per subtable: nb[s] buckets, keys and values). */
int dyc_wide_load(dyc_wide *t, const uint32_t *nb, const uint64_t *const *key_images,
                  const uint64_t *const *val_images, dyc_stream stream);
int dyc_wide_pending(dyc_wide *t, uint64_t *host_keys, uint64_t *host_vals, uint64_t cap, uint64_t *n_out,
                     dyc_stream stream);

/* ---------------------------------------------------------------------------------------
 * Many wide-key tables in one launch: the dyc_group_* calls with uint64 keys and values
 * (tables of the dyc_wide_* layout). No counterpart in the reference (one table per process,
 * 32-bit keys). Keys of table i are keys[offsets[i] .. offsets[i+1]) (host offsets, n_tables + 1
 * entries). Tables have fixed sizes; entries that exhaust the eviction bound stay pending,
 * tagged with their table, and that table's finds and erases search them (the insert returns
 * DYC_ERR_PENDING). cfg gives the eviction rule, subtable count, insert_blocks (insert grid) and
 * max_blocks (find / erase grid; 0 = default). Keys 0 and 2^64 - 1 are reserved. */
typedef struct dyc_wide_group dyc_wide_group; /* opaque */
int dyc_wide_group_create(const dyc_config *cfg, uint32_t n_tables, const uint64_t *capacity_hints,
                          dyc_wide_group **out);
int dyc_wide_group_destroy(dyc_wide_group *g);
int dyc_wide_group_insert(dyc_wide_group *g, const uint64_t *keys, const uint64_t *vals, const uint64_t *offsets,
                          dyc_stream stream);
int dyc_wide_group_find(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint64_t *vals,
                        uint8_t *hit, dyc_stream stream);
int dyc_wide_group_erase(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint8_t *deleted,
                         dyc_stream stream);
/* The same batches with one launch of the single-table kernels per table (the baseline). */
int dyc_wide_group_insert_per_table(dyc_wide_group *g, const uint64_t *keys, const uint64_t *vals,
                                    const uint64_t *offsets, dyc_stream stream);
int dyc_wide_group_find_per_table(dyc_wide_group *g, const uint64_t *keys, const uint64_t *offsets, uint64_t *vals,
                                  uint8_t *hit, dyc_stream stream);
int dyc_wide_group_table_stats(const dyc_wide_group *g, uint32_t table, dyc_stats *out);
int dyc_wide_group_dump(dyc_wide_group *g, uint32_t table, uint32_t sub, uint64_t *host_keys, uint64_t *host_vals,
                        uint64_t n_slots, dyc_stream stream);
int dyc_wide_group_pending(dyc_wide_group *g, uint64_t *host_keys, uint64_t *host_vals, uint32_t *host_tables,
                           uint64_t cap, uint64_t *n_out, dyc_stream stream);
int dyc_wide_group_clear(dyc_wide_group *g, dyc_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* DYC_H */
