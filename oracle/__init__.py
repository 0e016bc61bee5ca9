"""CPU oracle for DyCuckoo -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker (or as the timed CPU baseline). The product
package ``databases_gpu_amd`` never imports it.

``liboracle.so`` (built from ``oracle/dyc_oracle.c`` by ``oracle/Makefile``) restates the
reference kernels one key at a time. ``OracleTable`` adds the host-side policy of
``GPU/dyhash/api.cpp`` (sizing ``api.cpp:47-56``, resize choice ``api.cpp:112-173``, fill
bounds ``api.cpp:175-181,353-369,490-505``) with the decisions documented in
``dyc_oracle.h``: unsigned packing, no dropped entries, reserved keys 0/0xFFFFFFFF,
true-occupancy bookkeeping.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SUBTABLES = 4          # default subtable count (dy_hash.h:32 TABLE_NUM)
MAX_SUBTABLES = 6
SUBTABLE_COUNTS = (2, 3, 4, 5, 6)
SLOTS = 16
PRIME = 294967291
KICK_BOUND = 15
RESCUE_KICKS = 0
NO_TABLE = 100
RESERVED = (0, 0xFFFFFFFF)


class _OrcWide(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_void_p * MAX_SUBTABLES), ("vals", ctypes.c_void_p * MAX_SUBTABLES),
                ("nb", ctypes.c_uint32 * MAX_SUBTABLES), ("d", ctypes.c_uint32)]


class _OrcTable(ctypes.Structure):
    _fields_ = [("sub", ctypes.c_void_p * MAX_SUBTABLES), ("nb", ctypes.c_uint32 * MAX_SUBTABLES),
                ("d", ctypes.c_uint32)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
        for name in ("orc_h1", "orc_h2", "orc_h3", "orc_h4", "orc_h5", "orc_pair", "orc_pos1",
                     "orc_pos2", "orc_next_prime"):
            getattr(L, name).restype = u32
            getattr(L, name).argtypes = [u32]
        L.orc_is_prime.restype = ctypes.c_int
        L.orc_is_prime.argtypes = [u32]
        L.orc_hash_t.restype = u32
        L.orc_hash_t.argtypes = [u32, u32]
        L.orc_get_pos.restype = u32
        L.orc_get_pos.argtypes = [u32, u32]
        L.orc_pairs.restype = u32
        L.orc_pairs.argtypes = [u32]
        for name in ("orc_pair_d", "orc_pos1_d", "orc_pos2_d"):
            getattr(L, name).restype = u32
            getattr(L, name).argtypes = [u32, u32]
        L.orc_get_pos_d.restype = u32
        L.orc_get_pos_d.argtypes = [u32, u32, u32]
        L.orc_initial_buckets_d.restype = u32
        L.orc_initial_buckets_d.argtypes = [u64, u32]
        L.orc_choose_up_d.restype = u32
        L.orc_choose_up_d.argtypes = [vp, u32]
        L.orc_choose_down_d.restype = u32
        L.orc_choose_down_d.argtypes = [vp, u32]
        L.orc_bucket.restype = u32
        L.orc_bucket.argtypes = [u32, u32, u32]
        L.orc_initial_buckets.restype = u32
        L.orc_initial_buckets.argtypes = [u64]
        L.orc_rescue_lane.restype = u32
        L.orc_rescue_lane.argtypes = [u32, u32]
        L.orc_choose_up.restype = u32
        L.orc_choose_up.argtypes = [vp]
        L.orc_choose_down.restype = u32
        L.orc_choose_down.argtypes = [vp]
        L.orc_hash_batch.restype = None
        L.orc_hash_batch.argtypes = [vp, u64, u32, vp]
        L.orc_insert_one.restype = ctypes.c_int
        L.orc_insert_one.argtypes = [vp, u32, u32, u32, u32, vp, vp]
        L.orc_insert_entry.restype = ctypes.c_int
        L.orc_insert_entry.argtypes = [vp, u64, u32, u32, u32, vp, vp]
        L.orc_insert.restype = u64
        L.orc_insert.argtypes = [vp, vp, vp, u64, u32, u32, vp, vp, u64, vp]
        L.orc_find.restype = None
        L.orc_find.argtypes = [vp, vp, u64, vp, vp, vp]
        L.orc_find_mt.restype = None
        L.orc_find_mt.argtypes = [vp, vp, u64, vp, vp, ctypes.c_int]
        L.orc_insert_mt.restype = u64
        L.orc_insert_mt.argtypes = [vp, vp, vp, u64, u32, u32, vp, u64, vp, vp, ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_max_threads.argtypes = []
        L.orc_erase.restype = u64
        L.orc_erase.argtypes = [vp, vp, u64, vp]
        L.orc_upsize.restype = None
        L.orc_upsize.argtypes = [vp, u32, vp, u32]
        L.orc_downsize_compact.restype = u64
        L.orc_downsize_compact.argtypes = [vp, u32, vp, vp, u64]
        L.orc_count.restype = u64
        L.orc_count.argtypes = [vp]
        L.orc_check_invariant.restype = u64
        L.orc_check_invariant.argtypes = [vp]
        L.orc_fold64.restype = u32
        L.orc_fold64.argtypes = [u64]
        L.orc_wide_insert_entry.restype = ctypes.c_int
        L.orc_wide_insert_entry.argtypes = [vp, u64, u64, u32, u32, u32, vp, vp, vp]
        L.orc_wide_insert.restype = u64
        L.orc_wide_insert.argtypes = [vp, vp, vp, u64, u32, u32, vp, vp, vp, u64, vp]
        L.orc_wide_find.restype = None
        L.orc_wide_find.argtypes = [vp, vp, u64, vp, vp]
        L.orc_wide_erase.restype = u64
        L.orc_wide_erase.argtypes = [vp, vp, u64, vp]
        L.orc_wide_upsize.restype = None
        L.orc_wide_upsize.argtypes = [vp, vp, u32, vp, vp, u32]
        L.orc_wide_downsize_compact.restype = u64
        L.orc_wide_downsize_compact.argtypes = [vp, vp, u32, vp, vp, vp, vp, u64]
        L.orc_wide_check_invariant.restype = u64
        L.orc_wide_check_invariant.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- scalar / batch hashes

def hash_batch(keys, which: int) -> np.ndarray:
    """which: 1..5 = h1..h5 (cuckoo.cu:27-64,110-120), 6 = pair index (cuckoo.cu:190-193)."""
    k = _u32(keys)
    out = np.empty_like(k)
    lib().orc_hash_batch(_p(k), k.size, which, _p(out))
    return out


def initial_buckets(capacity_hint: int, subtables: int = SUBTABLES) -> int:
    return int(lib().orc_initial_buckets_d(capacity_hint, subtables))


def pair_table(subtables: int):
    """[(pos1, pos2)] of every pair index of a `subtables`-subtable table."""
    L = lib()
    return [(int(L.orc_pos1_d(p, subtables)), int(L.orc_pos2_d(p, subtables)))
            for p in range(int(L.orc_pairs(subtables)))]


def pair_of(keys, subtables: int = SUBTABLES) -> np.ndarray:
    """Pair index of every key (choose_hash of the `subtables` variant)."""
    L = lib()
    return np.array([L.orc_pair_d(int(k), subtables) for k in _u32(keys).tolist()], dtype=np.uint32)


def pack(keys, vals) -> np.ndarray:
    return (_u32(keys).astype(np.uint64) << np.uint64(32)) | _u32(vals).astype(np.uint64)


# ---------------------------------------------------------------- table with host policy

class OracleTable:
    """Sequential replay of a DyCuckoo table plus the reference host policy.

    Images are numpy uint64 arrays, one per subtable, ``nb[t] * 16`` slots each: the same
    layout ``dyc_dump`` writes.
    """

    def __init__(self, capacity_hint: int = 0, *, nb=None, subtables: int = 0, kick_bound: int = KICK_BOUND,
                 rescue_kicks: int = RESCUE_KICKS,
                 lower: float = 0.5, upper: float = 0.9, auto_resize: bool = False,
                 min_capacity: float = SUBTABLES * 1024 * 1024 * 1.2, max_fail_rounds: int = 8):
        if nb is None:
            self.d = subtables or SUBTABLES
            b = initial_buckets(capacity_hint, self.d)
            nb = [b] * self.d
        else:
            self.d = len(nb)
            if subtables and subtables != self.d:
                raise ValueError("nb lists %d subtables, subtables=%d" % (self.d, subtables))
        if self.d not in SUBTABLE_COUNTS:
            raise ValueError("subtables must be in 2..6")
        self.nb = [int(x) for x in nb]
        self.sub = [np.zeros(x * SLOTS, dtype=np.uint64) for x in self.nb]
        self.kick_bound = kick_bound
        self.rescue_kicks = rescue_kicks
        self.lower, self.upper = lower, upper
        self.auto_resize = auto_resize
        self.min_capacity = min_capacity
        self.max_fail_rounds = max_fail_rounds
        self.count = 0
        self.pending = np.zeros(0, dtype=np.uint64)  # entries owned but not placed
        self.resizes = []  # ("up"|"down", subtable, new_nb)
        self.kicks_total = 0
        self.fail_total = 0

    # -- image plumbing
    def _struct(self):
        s = _OrcTable()
        for t in range(self.d):
            s.sub[t] = self.sub[t].ctypes.data
            s.nb[t] = self.nb[t]
        s.d = self.d
        return s

    @property
    def capacity(self) -> int:
        return sum(self.nb) * SLOTS

    def load_images(self, images):
        self.sub = [np.ascontiguousarray(x, dtype=np.uint64).copy() for x in images]
        self.nb = [x.size // SLOTS for x in self.sub]
        self.d = len(self.sub)
        self.count = int(sum(int(np.count_nonzero(x)) for x in self.sub))

    def images(self):
        return [x.copy() for x in self.sub]

    def entries(self) -> np.ndarray:
        """All non-empty entries, sorted (the key->value multiset)."""
        allv = np.concatenate(self.sub + [self.pending])
        return np.sort(allv[allv != 0])

    def check_invariant(self) -> int:
        s = self._struct()
        return int(lib().orc_check_invariant(ctypes.byref(s)))

    # -- operations
    def _insert_entries(self, entries: np.ndarray, pre: int):
        """Continue eviction chains for entries already in hand (residual / pending path)."""
        L = lib()
        s = self._struct()
        fails = []
        kicks = ctypes.c_uint32(0)
        carry = ctypes.c_uint64(0)
        for e in entries.tolist():
            e = int(e)
            if pre == NO_TABLE:
                k = e >> 32
                p = NO_TABLE if (k & 1) == 0 else int(L.orc_get_pos_d(L.orc_pair_d(k, self.d), NO_TABLE, self.d))
            else:
                p = pre
            ok = L.orc_insert_entry(ctypes.byref(s), e, p, self.kick_bound, self.rescue_kicks,
                                    ctypes.byref(kicks),
                                    ctypes.byref(carry))
            self.kicks_total += kicks.value
            if not ok:
                fails.append(carry.value)
        return np.array(fails, dtype=np.uint64)

    def _place_pending(self, before: int = 0):
        """After an operation that found `before` entries pending and left more: upsize and
        re-insert every pending entry, at most max_fail_rounds rounds, while more than `before`
        remain. Entries that never fit stay as the stash without growing the table again."""
        rounds = 0
        while self.pending.size > before and rounds < self.max_fail_rounds:
            self.resize_up()
            pend, self.pending = self.pending, np.zeros(0, dtype=np.uint64)
            self.pending = self._insert_entries(pend, NO_TABLE)
            rounds += 1

    def insert(self, keys, vals):
        """Batch insert in input order (cuckoo_insert); returns (failures_before_retry, rejected)."""
        keys, vals = _u32(keys), _u32(vals)
        n = keys.size
        before = self.pending.size
        if self.auto_resize:
            while self.count + n > self.upper * self.capacity:
                self.resize_up()
        L = lib()
        s = self._struct()
        kicks = np.zeros(n, dtype=np.uint32)
        fail = np.zeros(max(n, 1), dtype=np.uint64)
        rej = ctypes.c_uint64(0)
        nfail = int(L.orc_insert(ctypes.byref(s), _p(keys), _p(vals), n, self.kick_bound,
                                 self.rescue_kicks, _p(kicks), _p(fail), fail.size, ctypes.byref(rej)))
        self.kicks_total += int(kicks.sum(dtype=np.uint64))
        self.fail_total += nfail
        self.count += n - int(rej.value)
        self.pending = np.concatenate([self.pending, fail[:nfail]])
        self._place_pending(before)
        self.last_kicks = kicks
        return nfail, int(rej.value)

    def insert_batch(self, keys, vals, batch_size: int):
        """hashAPI::hash_insert_batch (api.cpp:353-369): while count + batch_size exceeds
        upper * capacity, upsize (api_resize, api.cpp:175-181, then the call re-enters itself),
        then insert the slice. The bound always uses the full batch size, whatever the slice
        holds, and applies whatever auto_resize says."""
        while self.count + batch_size > self.upper * self.capacity:
            self.resize_up()
        return self.insert(keys, vals)

    def erase_batch(self, keys):
        """hashAPI::hash_delete_batch (api.cpp:466-506): erase the slice, then downsize once when
        count < lower * capacity and capacity > min_capacity (api.cpp:494-505). With auto_resize
        set, erase() has applied that same rule already, so it is not applied twice."""
        deleted = self.erase(keys)
        if (not self.auto_resize and self.count < self.lower * self.capacity
                and self.capacity > self.min_capacity):
            self.resize_down()
        return deleted

    def insert_mt(self, keys, vals, threads: int = 0):
        """Multithreaded insert (CPU timing baseline): same final key->value set as insert(),
        placement depends on thread timing. Returns (failures_before_retry, rejected)."""
        keys, vals = _u32(keys), _u32(vals)
        n = keys.size
        before = self.pending.size
        L = lib()
        s = self._struct()
        fail = np.zeros(max(n, 1), dtype=np.uint64)
        kicks, rej = ctypes.c_uint64(0), ctypes.c_uint64(0)
        nfail = int(L.orc_insert_mt(ctypes.byref(s), _p(keys), _p(vals), n, self.kick_bound, self.rescue_kicks,
                                    _p(fail), fail.size, ctypes.byref(kicks), ctypes.byref(rej),
                                    threads or int(L.orc_max_threads())))
        self.kicks_total += int(kicks.value)
        self.fail_total += nfail
        self.count += n - int(rej.value)
        self.pending = np.concatenate([self.pending, fail[:min(nfail, fail.size)]])
        self._place_pending(before)
        return nfail, int(rej.value)

    def find(self, keys):
        keys = _u32(keys)
        vals = np.zeros(keys.size, dtype=np.uint32)
        hit = np.zeros(keys.size, dtype=np.uint8)
        probes = ctypes.c_uint64(0)
        s = self._struct()
        lib().orc_find(ctypes.byref(s), _p(keys), keys.size, _p(vals), _p(hit), ctypes.byref(probes))
        self.last_probes = int(probes.value)
        # the pending entries are the table's stash, searched after pos2: the first entry (in
        # list order) holding the key answers
        miss, pk, first = self._stash_lookup(keys, hit == 0)
        if miss.size:
            vals[miss] = (self.pending[first] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            hit[miss] = 1
        return vals, hit

    def _stash_lookup(self, keys, candidates):
        """(indexes i among `candidates` whose key is in the stash, stash keys, index of the first
        stash entry of each such key)."""
        if not self.pending.size:
            return np.zeros(0, dtype=np.int64), None, None
        pkeys = (self.pending >> np.uint64(32)).astype(np.uint32)
        pk, first = np.unique(pkeys, return_index=True)
        idx = np.nonzero(candidates & (keys != 0) & (keys != 0xFFFFFFFF))[0]
        pos = np.searchsorted(pk, keys[idx])
        ok = pos < pk.size
        ok[ok] = pk[pos[ok]] == keys[idx[ok]]
        return idx[ok], pk, first[pos[ok]]

    def erase(self, keys):
        keys = _u32(keys)
        deleted = np.zeros(keys.size, dtype=np.uint8)
        s = self._struct()
        ndel = int(lib().orc_erase(ctypes.byref(s), _p(keys), keys.size, _p(deleted)))
        # a key found in neither bucket has every copy in the stash cleared, by its first
        # occurrence in the batch
        miss, _, _ = self._stash_lookup(keys, deleted == 0)
        if miss.size:
            _, firsts = np.unique(keys[miss], return_index=True)
            hits = miss[firsts]
            deleted[hits] = 1
            ndel += hits.size
            gone = np.isin((self.pending >> np.uint64(32)).astype(np.uint32), keys[hits])
            self.pending = self.pending[~gone]
        self.count -= ndel
        if (self.auto_resize and self.count < self.lower * self.capacity
                and self.capacity > self.min_capacity):
            self.resize_down()
        return deleted

    def resize_up(self) -> int:
        nb = (ctypes.c_uint32 * self.d)(*self.nb)
        t = int(lib().orc_choose_up_d(ctypes.byref(nb), self.d))
        new = np.zeros(self.nb[t] * 2 * SLOTS, dtype=np.uint64)
        lib().orc_upsize(_p(self.sub[t]), self.nb[t], _p(new), t)
        self.sub[t] = new
        self.nb[t] *= 2
        self.resizes.append(("up", t, self.nb[t]))
        return t

    def resize_down(self, defer_residuals: bool = False):
        """Halve the subtable chosen by api.cpp:147-153. Returns (t, residuals) or None when
        that subtable has an odd bucket count (pairwise compaction is undefined there)."""
        nb = (ctypes.c_uint32 * self.d)(*self.nb)
        t = int(lib().orc_choose_down_d(ctypes.byref(nb), self.d))
        if self.nb[t] % 2:
            return None
        new = np.zeros(self.nb[t] // 2 * SLOTS, dtype=np.uint64)
        res = np.zeros(self.nb[t] * SLOTS // 2 + 1, dtype=np.uint64)
        nres = int(lib().orc_downsize_compact(_p(self.sub[t]), self.nb[t], _p(new), _p(res),
                                               res.size))
        self.sub[t] = new
        self.nb[t] //= 2
        self.resizes.append(("down", t, self.nb[t]))
        res = res[:nres].copy()
        if not defer_residuals:
            before = self.pending.size
            fails = self._insert_entries(res, t)
            if fails.size:
                self.pending = np.concatenate([self.pending, fails])
                self._place_pending(before)
        return t, res


def count_probes(images, keys) -> int:
    """Buckets the pos1-then-pos2 find order reads for `keys` over a dumped image."""
    t = OracleTable(nb=[x.size // SLOTS for x in images])
    t.sub = [np.ascontiguousarray(x, dtype=np.uint64) for x in images]
    t.find(keys)
    return t.last_probes


# ---------------------------------------------------------------- wide keys

def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def fold64(keys) -> np.ndarray:
    """lo ^ h5(hi) of every 64-bit key (the 32-bit word the wide table hashes)."""
    L = lib()
    return np.array([L.orc_fold64(int(k)) for k in _u64(keys).tolist()], dtype=np.uint32)


class OracleWideTable:
    """Sequential replay of a wide-key DyCuckoo table (uint64 keys and values) plus the host
    policy of OracleTable: sizing, auto-resize, pending entries re-placed by upsizing and kept
    as a stash that find and erase search. Images are (keys, values) uint64 array pairs."""

    def __init__(self, capacity_hint: int = 0, *, nb=None, subtables: int = 0, kick_bound: int = KICK_BOUND,
                 rescue_kicks: int = RESCUE_KICKS, max_fail_rounds: int = 8, auto_resize: bool = False,
                 lower: float = 0.5, upper: float = 0.9, min_capacity: float = SUBTABLES * 1024 * 1024 * 1.2):
        if nb is None:
            self.d = subtables or SUBTABLES
            nb = [initial_buckets(capacity_hint, self.d)] * self.d
        else:
            self.d = len(nb)
        self.nb = [int(x) for x in nb]
        self.keys = [np.zeros(x * SLOTS, dtype=np.uint64) for x in self.nb]
        self.vals = [np.zeros(x * SLOTS, dtype=np.uint64) for x in self.nb]
        self.kick_bound, self.rescue_kicks, self.max_fail_rounds = kick_bound, rescue_kicks, max_fail_rounds
        self.auto_resize, self.lower, self.upper, self.min_capacity = auto_resize, lower, upper, min_capacity
        self.count = 0
        self.pend_keys = np.zeros(0, dtype=np.uint64)
        self.pend_vals = np.zeros(0, dtype=np.uint64)
        self.resizes = []

    def _struct(self):
        s = _OrcWide()
        for t in range(self.d):
            s.keys[t] = self.keys[t].ctypes.data
            s.vals[t] = self.vals[t].ctypes.data
            s.nb[t] = self.nb[t]
        s.d = self.d
        return s

    @property
    def capacity(self) -> int:
        return sum(self.nb) * SLOTS

    def load_images(self, keys, vals):
        self.keys = [_u64(x).copy() for x in keys]
        self.vals = [_u64(x).copy() for x in vals]
        self.nb = [x.size // SLOTS for x in self.keys]
        self.d = len(self.keys)
        self.count = int(sum(int(np.count_nonzero(x)) for x in self.keys))

    def entries(self):
        """All stored (key, value) pairs, table and stash, sorted by key then value."""
        k = np.concatenate(self.keys + [self.pend_keys])
        v = np.concatenate(self.vals + [self.pend_vals])
        live = k != 0
        k, v = k[live], v[live]
        order = np.lexsort((v, k))
        return k[order], v[order]

    def check_invariant(self) -> int:
        s = self._struct()
        return int(lib().orc_wide_check_invariant(ctypes.byref(s)))

    def _reinsert(self, keys, vals, pre):
        L = lib()
        s = self._struct()
        fk, fv = [], []
        kicks = ctypes.c_uint32(0)
        ck, cv = ctypes.c_uint64(0), ctypes.c_uint64(0)
        for k, v in zip(keys.tolist(), vals.tolist()):
            p = pre
            if pre == NO_TABLE and (k & 1):
                p = int(L.orc_get_pos_d(L.orc_pair_d(L.orc_fold64(k), self.d), NO_TABLE, self.d))
            if not L.orc_wide_insert_entry(ctypes.byref(s), k, v, p, self.kick_bound, self.rescue_kicks,
                                           ctypes.byref(kicks), ctypes.byref(ck), ctypes.byref(cv)):
                fk.append(ck.value)
                fv.append(cv.value)
        return np.array(fk, dtype=np.uint64), np.array(fv, dtype=np.uint64)

    def _place_pending(self, before: int = 0):
        rounds = 0
        while self.pend_keys.size > before and rounds < self.max_fail_rounds:
            self.resize_up()
            k, v = self.pend_keys, self.pend_vals
            self.pend_keys, self.pend_vals = self._reinsert(k, v, NO_TABLE)
            rounds += 1

    def insert(self, keys, vals):
        keys, vals = _u64(keys), _u64(vals)
        n = keys.size
        before = self.pend_keys.size
        if self.auto_resize:
            while self.count + n > self.upper * self.capacity:
                self.resize_up()
        s = self._struct()
        kicks = np.zeros(n, dtype=np.uint32)
        fk = np.zeros(max(n, 1), dtype=np.uint64)
        fv = np.zeros(max(n, 1), dtype=np.uint64)
        rej = ctypes.c_uint64(0)
        nfail = int(lib().orc_wide_insert(ctypes.byref(s), _p(keys), _p(vals), n, self.kick_bound, self.rescue_kicks,
                                          _p(kicks), _p(fk), _p(fv), fk.size, ctypes.byref(rej)))
        self.count += n - int(rej.value)
        self.pend_keys = np.concatenate([self.pend_keys, fk[:nfail]])
        self.pend_vals = np.concatenate([self.pend_vals, fv[:nfail]])
        self._place_pending(before)
        self.last_kicks = kicks
        return nfail, int(rej.value)

    def find(self, keys):
        keys = _u64(keys)
        vals = np.zeros(keys.size, dtype=np.uint64)
        hit = np.zeros(keys.size, dtype=np.uint8)
        s = self._struct()
        lib().orc_wide_find(ctypes.byref(s), _p(keys), keys.size, _p(vals), _p(hit))
        for i in np.nonzero(hit == 0)[0]:  # the stash, after pos2: its first entry of the key
            j = np.nonzero(self.pend_keys == keys[i])[0]
            if j.size and keys[i] not in (0, 2**64 - 1):
                vals[i] = self.pend_vals[j[0]]
                hit[i] = 1
        return vals, hit

    def erase(self, keys):
        keys = _u64(keys)
        deleted = np.zeros(keys.size, dtype=np.uint8)
        s = self._struct()
        ndel = int(lib().orc_wide_erase(ctypes.byref(s), _p(keys), keys.size, _p(deleted)))
        for i in np.nonzero(deleted == 0)[0]:  # the stash: every copy, by the key's first occurrence
            gone = self.pend_keys == keys[i]
            if gone.any() and keys[i] not in (0, 2**64 - 1):
                self.pend_keys, self.pend_vals = self.pend_keys[~gone], self.pend_vals[~gone]
                deleted[i] = 1
                ndel += 1
        self.count -= ndel
        if self.auto_resize and self.count < self.lower * self.capacity and self.capacity > self.min_capacity:
            self.resize_down()
        return deleted

    def resize_up(self) -> int:
        nb = (ctypes.c_uint32 * self.d)(*self.nb)
        t = int(lib().orc_choose_up_d(ctypes.byref(nb), self.d))
        nk = np.zeros(self.nb[t] * 2 * SLOTS, dtype=np.uint64)
        nv = np.zeros_like(nk)
        lib().orc_wide_upsize(_p(self.keys[t]), _p(self.vals[t]), self.nb[t], _p(nk), _p(nv), t)
        self.keys[t], self.vals[t] = nk, nv
        self.nb[t] *= 2
        self.resizes.append(("up", t, self.nb[t]))
        return t

    def resize_down(self, defer_residuals: bool = False):
        nb = (ctypes.c_uint32 * self.d)(*self.nb)
        t = int(lib().orc_choose_down_d(ctypes.byref(nb), self.d))
        if self.nb[t] % 2:
            return None
        half = self.nb[t] // 2
        nk = np.zeros(half * SLOTS, dtype=np.uint64)
        nv = np.zeros_like(nk)
        rk = np.zeros(half * SLOTS + 1, dtype=np.uint64)
        rv = np.zeros_like(rk)
        nres = int(lib().orc_wide_downsize_compact(_p(self.keys[t]), _p(self.vals[t]), self.nb[t], _p(nk), _p(nv),
                                                    _p(rk), _p(rv), rk.size))
        self.keys[t], self.vals[t] = nk, nv
        self.nb[t] = half
        self.resizes.append(("down", t, half))
        rk, rv = rk[:nres].copy(), rv[:nres].copy()
        if not defer_residuals and nres:
            before = self.pend_keys.size
            fk, fv = self._reinsert(rk, rv, t)
            if fk.size:
                self.pend_keys = np.concatenate([self.pend_keys, fk])
                self.pend_vals = np.concatenate([self.pend_vals, fv])
                self._place_pending(before)
        return t, rk, rv

