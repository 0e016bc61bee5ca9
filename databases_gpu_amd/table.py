"""Python host interface of the MI355X DyCuckoo table.

``DyCuckoo`` is a thin owner of a ``dyc_table*`` handle (include/dyc.h): device tensors in,
device tensors out, on the current torch stream.

``HashAPI`` mirrors the reference's host class ``hashAPI`` (GPU/dyhash/api.h:9-55) method for
method -- host arrays in, host arrays out, synchronous -- so call sequences of the reference's
drivers (GPU/dyhash/Main_backup_0910.cpp, GPU/warp/Main.cpp) carry over unchanged.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import SLOTS, DycError, check

_U32 = torch.int32  # torch has no uint32 arithmetic on GPU; bits are reinterpreted as uint32


def _as_u32_tensor(x, device) -> torch.Tensor:
    """Return a contiguous 4-byte tensor on `device` whose bits are the uint32 values of x."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype not in (torch.int32, torch.uint32):
            if t.dtype == torch.int64:
                t = (t & 0xFFFFFFFF).to(torch.int64)
                t = torch.where(t >= 2**31, t - 2**32, t).to(torch.int32)
            else:
                raise TypeError(f"keys/values must be 32-bit, got {t.dtype}")
        t = t.view(torch.int32) if t.dtype == torch.uint32 else t
        return t.to(device).contiguous()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint32)).view(np.int32)
    return torch.from_numpy(a).to(device)


def _host_u32(x) -> np.ndarray:
    """A C-contiguous numpy uint32 view of host data (no copy when it already is one)."""
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            raise ValueError("host-array calls take host data; use the device-tensor calls for GPU tensors")
        x = x.numpy()
    a = np.asarray(x)
    if a.dtype in (np.int32, np.uint32):
        a = a.view(np.uint32)
    else:
        a = a.astype(np.uint32)
    return np.ascontiguousarray(a)


def _stream_ptr(stream) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class DyCuckoo:
    """A dynamic cuckoo hash table of uint32 -> uint32 on one GPU: 4 subtables by default,
    2..6 with ``subtables=d`` (a configuration field)."""

    def __init__(self, capacity_hint: int, *, device=None, **config):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = _lib.default_config(**config)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_create(ctypes.byref(self.cfg), int(capacity_hint), ctypes.byref(h)), "dyc_create")
        self.h = h

    # -- lifecycle
    def close(self):
        if getattr(self, "h", None):
            self.lib.dyc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- operations on device tensors
    def insert(self, keys: torch.Tensor, vals: torch.Tensor, stream=None) -> None:
        """Insert keys[i] -> vals[i] (cuckoo_insert, cuckoo.cu:324-355)."""
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        with torch.cuda.device(self.device):
            check(self.lib.dyc_insert(self.h, k.data_ptr(), v.data_ptr(), k.numel(), _stream_ptr(stream)),
                  "dyc_insert")

    def insert_counted(self, keys: torch.Tensor, vals: torch.Tensor, stream=None) -> torch.Tensor:
        """Insert (eviction insert, input order) and return each key's eviction-chain length as
        int32: the reference's iterator_count (cuckoo.cu:349-351; api.cpp:219-243 reports its
        maximum and mean). Stats.last_max_chain holds the maximum."""
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        counts = torch.zeros(k.numel(), dtype=torch.int32, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_insert_counted(self.h, k.data_ptr(), v.data_ptr(), k.numel(), counts.data_ptr(),
                                              _stream_ptr(stream)), "dyc_insert_counted")
        return counts

    def insert_async(self, keys: torch.Tensor, vals: torch.Tensor, stream=None) -> None:
        """Enqueue an insert on `stream` without waiting (dyc_insert_async): inserts on several
        streams run at once; sync() (or any other call) completes them. The tensors must stay
        alive until then."""
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        self._async_hold = getattr(self, "_async_hold", []) + [k, v]
        with torch.cuda.device(self.device):
            check(self.lib.dyc_insert_async(self.h, k.data_ptr(), v.data_ptr(), k.numel(), _stream_ptr(stream)),
                  "dyc_insert_async")

    def sync(self) -> None:
        """Complete every insert_async in flight (dyc_sync)."""
        with torch.cuda.device(self.device):
            check(self.lib.dyc_sync(self.h), "dyc_sync")
        self._async_hold = []

    def upsert(self, keys: torch.Tensor, vals: torch.Tensor, stream=None) -> torch.Tensor:
        """Insert, or replace the value of keys already stored (keys distinct within the call).
        Returns per-key flags: 1 where an existing entry was updated."""
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        updated = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_upsert(self.h, k.data_ptr(), v.data_ptr(), k.numel(), updated.data_ptr(),
                                      _stream_ptr(stream)), "dyc_upsert")
        return updated

    def find(self, keys: torch.Tensor, vals: torch.Tensor | None = None, hit: torch.Tensor | None = None,
             stream=None):
        """Look keys up (cuckoo_search, cuckoo.cu:460-520). Returns (vals int32-as-uint32, hit uint8)."""
        k = _as_u32_tensor(keys, self.device)
        n = k.numel()
        if vals is None:
            vals = torch.empty(n, dtype=torch.int32, device=self.device)
        if hit is None:
            hit = torch.empty(n, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_find(self.h, k.data_ptr(), vals.data_ptr(), hit.data_ptr(), n, _stream_ptr(stream)),
                  "dyc_find")
        return vals, hit

    def erase(self, keys: torch.Tensor, deleted: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Delete keys (cuckoo_delete, cuckoo.cu:524-587). Returns per-key deleted flags (uint8)."""
        k = _as_u32_tensor(keys, self.device)
        if deleted is None:
            deleted = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_erase(self.h, k.data_ptr(), deleted.data_ptr(), k.numel(), _stream_ptr(stream)),
                  "dyc_erase")
        return deleted

    # -- operations on host arrays (pipelined host <-> device transfers)
    def insert_host(self, keys, vals, stream=None) -> None:
        """Insert from host arrays (numpy uint32, or anything convertible; memory maps work)."""
        k = _host_u32(keys)
        v = _host_u32(vals)
        if k.size != v.size:
            raise ValueError("keys and values differ in length")
        with torch.cuda.device(self.device):
            check(self.lib.dyc_insert_host(self.h, k.ctypes.data, v.ctypes.data, k.size, _stream_ptr(stream)),
                  "dyc_insert_host")

    def find_host(self, keys, stream=None):
        """Look host keys up; returns (values uint32, hit uint8) as numpy arrays."""
        k = _host_u32(keys)
        vals = np.zeros(k.size, dtype=np.uint32)
        hit = np.zeros(k.size, dtype=np.uint8)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_find_host(self.h, k.ctypes.data, vals.ctypes.data, hit.ctypes.data, k.size,
                                         _stream_ptr(stream)), "dyc_find_host")
        return vals, hit

    def erase_host(self, keys, stream=None) -> np.ndarray:
        k = _host_u32(keys)
        deleted = np.zeros(k.size, dtype=np.uint8)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_erase_host(self.h, k.ctypes.data, deleted.ctypes.data, k.size, _stream_ptr(stream)),
                  "dyc_erase_host")
        return deleted

    # -- checkpoint
    def checkpoint(self, path: str, stream=None) -> None:
        """Write the table (images, pending entries, counts, checksums) to `path`."""
        with torch.cuda.device(self.device):
            check(self.lib.dyc_checkpoint_write(self.h, str(path).encode(), _stream_ptr(stream)),
                  "dyc_checkpoint_write")

    @classmethod
    def restore(cls, path: str, *, device=None, **config) -> "DyCuckoo":
        """A table read back from a checkpoint file (its geometry; `config` for everything else)."""
        self = cls.__new__(cls)
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = _lib.default_config(**config)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_checkpoint_read(str(path).encode(), ctypes.byref(self.cfg), ctypes.byref(h)),
                  "dyc_checkpoint_read")
        self.h = h
        return self

    def resize_up(self, stream=None) -> None:
        with torch.cuda.device(self.device):
            check(self.lib.dyc_resize_up(self.h, _stream_ptr(stream)), "dyc_resize_up")

    def resize_down(self, stream=None) -> bool:
        """Halve one subtable; False when the chosen subtable has an odd bucket count."""
        with torch.cuda.device(self.device):
            rc = self.lib.dyc_resize_down(self.h, _stream_ptr(stream))
        if rc == _lib.DYC_ERR_ODD_BUCKETS:
            return False
        check(rc, "dyc_resize_down")
        return True

    # -- configuration
    def config(self) -> dict:
        c = _lib.Config()
        check(self.lib.dyc_get_config(self.h, ctypes.byref(c)), "dyc_get_config")
        return {name: getattr(c, name) for name, _ in c._fields_}

    def configure(self, **fields) -> None:
        """Change run-time fields (lanes, grids, eviction bounds, fill bounds); sizes stay."""
        c = _lib.Config()
        check(self.lib.dyc_get_config(self.h, ctypes.byref(c)), "dyc_get_config")
        for k, v in fields.items():
            if not hasattr(c, k):
                raise KeyError(f"unknown config field {k}")
            setattr(c, k, v)
        check(self.lib.dyc_set_config(self.h, ctypes.byref(c)), "dyc_set_config")

    # -- inspection
    def stats(self) -> dict:
        st = _lib.Stats()
        check(self.lib.dyc_get_stats(self.h, ctypes.byref(st)), "dyc_get_stats")
        return st.as_dict()

    @property
    def buckets(self):
        return self.stats()["buckets"]

    @property
    def subtables(self) -> int:
        return self.stats()["subtables"]

    @property
    def capacity(self) -> int:
        return self.stats()["capacity"]

    def __len__(self) -> int:
        return self.stats()["count"]

    def count_occupied(self, stream=None) -> int:
        out = ctypes.c_uint64()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_count_occupied(self.h, ctypes.byref(out), _stream_ptr(stream)), "dyc_count_occupied")
        return out.value

    def clear(self, stream=None) -> None:
        """Empty the table, keeping its subtable sizes."""
        with torch.cuda.device(self.device):
            check(self.lib.dyc_clear(self.h, _stream_ptr(stream)), "dyc_clear")

    def probe_count(self, keys, stream=None):
        """(buckets the pos1-then-pos2 find order reads for keys, keys resolved by pos1)."""
        k = _as_u32_tensor(keys, self.device)
        probes, hits = ctypes.c_uint64(), ctypes.c_uint64()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_probe_count(self.h, k.data_ptr(), k.numel(), ctypes.byref(probes), ctypes.byref(hits),
                                           _stream_ptr(stream)), "dyc_probe_count")
        return probes.value, hits.value

    def dump(self, stream=None):
        """Host copies of the subtables (uint64 slots, (key << 32) | value)."""
        out = []
        for t, nb in enumerate(self.buckets):
            a = np.empty(nb * SLOTS, dtype=np.uint64)
            with torch.cuda.device(self.device):
                check(self.lib.dyc_dump(self.h, t, a.ctypes.data, a.size, _stream_ptr(stream)), "dyc_dump")
            out.append(a)
        return out

    def load(self, images, stream=None) -> None:
        imgs = [np.ascontiguousarray(x, dtype=np.uint64) for x in images]
        d = self.subtables
        if len(imgs) != d or any(x.size % SLOTS or x.size == 0 for x in imgs):
            raise ValueError(f"need {d} non-empty images of whole buckets")
        nb = (ctypes.c_uint32 * d)(*[x.size // SLOTS for x in imgs])
        ptrs = (ctypes.c_void_p * d)(*[x.ctypes.data for x in imgs])
        with torch.cuda.device(self.device):
            check(self.lib.dyc_load(self.h, nb, ptrs, _stream_ptr(stream)), "dyc_load")

    def pending(self, stream=None) -> np.ndarray:
        n = ctypes.c_uint64()
        check(self.lib.dyc_pending(self.h, None, 0, ctypes.byref(n), _stream_ptr(stream)), "dyc_pending")
        a = np.empty(n.value, dtype=np.uint64)
        if n.value:
            with torch.cuda.device(self.device):
                check(self.lib.dyc_pending(self.h, a.ctypes.data, a.size, ctypes.byref(n), _stream_ptr(stream)),
                      "dyc_pending")
        return a


def _as_u64_tensor(x, device) -> torch.Tensor:
    """A contiguous 8-byte tensor on `device` whose bits are the uint64 values of x."""
    if isinstance(x, torch.Tensor):
        if x.dtype not in (torch.int64, torch.uint64):
            raise TypeError(f"wide keys/values must be 64-bit, got {x.dtype}")
        t = x.view(torch.int64) if x.dtype == torch.uint64 else x
        return t.to(device).contiguous()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64)).view(np.int64)
    return torch.from_numpy(a).to(device)


class DyCuckooWide:
    """A dynamic cuckoo hash table of uint64 -> uint64 on one GPU (the paper's wide-key layout:
    per bucket a line of 16 keys and a parallel line of 16 values). Tensors are int64 whose bits
    are the uint64 keys and values; keys 0 and 2^64 - 1 are reserved."""

    def __init__(self, capacity_hint: int, *, device=None, **config):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = _lib.default_config(**config)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_create(ctypes.byref(self.cfg), int(capacity_hint), ctypes.byref(h)),
                  "dyc_wide_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.dyc_wide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert(self, keys, vals, stream=None) -> None:
        k = _as_u64_tensor(keys, self.device)
        v = _as_u64_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_insert(self.h, k.data_ptr(), v.data_ptr(), k.numel(), _stream_ptr(stream)),
                  "dyc_wide_insert")

    def upsert(self, keys, vals, stream=None) -> torch.Tensor:
        """Insert, or replace the value of keys already stored (keys distinct within the call).
        Returns per-key flags: 1 where an existing entry was updated."""
        k = _as_u64_tensor(keys, self.device)
        v = _as_u64_tensor(vals, self.device)
        if k.numel() != v.numel():
            raise ValueError("keys and values differ in length")
        updated = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_upsert(self.h, k.data_ptr(), v.data_ptr(), k.numel(), updated.data_ptr(),
                                           _stream_ptr(stream)), "dyc_wide_upsert")
        return updated

    def find(self, keys, vals=None, hit=None, stream=None):
        """Returns (values int64-as-uint64, hit uint8)."""
        k = _as_u64_tensor(keys, self.device)
        n = k.numel()
        if vals is None:
            vals = torch.empty(n, dtype=torch.int64, device=self.device)
        if hit is None:
            hit = torch.empty(n, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_find(self.h, k.data_ptr(), vals.data_ptr(), hit.data_ptr(), n, _stream_ptr(stream)),
                  "dyc_wide_find")
        return vals, hit

    def erase(self, keys, stream=None) -> torch.Tensor:
        k = _as_u64_tensor(keys, self.device)
        deleted = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_erase(self.h, k.data_ptr(), deleted.data_ptr(), k.numel(), _stream_ptr(stream)),
                  "dyc_wide_erase")
        return deleted

    def resize_up(self, stream=None) -> None:
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_resize_up(self.h, _stream_ptr(stream)), "dyc_wide_resize_up")

    def resize_down(self, stream=None) -> bool:
        """False when the chosen subtable has an odd bucket count (refused)."""
        with torch.cuda.device(self.device):
            rc = self.lib.dyc_wide_resize_down(self.h, _stream_ptr(stream))
        if rc == _lib.DYC_ERR_ODD_BUCKETS:
            return False
        check(rc, "dyc_wide_resize_down")
        return True

    def clear(self, stream=None) -> None:
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_clear(self.h, _stream_ptr(stream)), "dyc_wide_clear")

    def stats(self) -> dict:
        st = _lib.Stats()
        check(self.lib.dyc_wide_get_stats(self.h, ctypes.byref(st)), "dyc_wide_get_stats")
        return st.as_dict()

    @property
    def buckets(self):
        return self.stats()["buckets"]

    def dump(self, stream=None):
        """([key arrays], [value arrays]) of the subtables, host uint64."""
        ks, vs = [], []
        for t, nb in enumerate(self.buckets):
            k = np.empty(nb * SLOTS, dtype=np.uint64)
            v = np.empty(nb * SLOTS, dtype=np.uint64)
            with torch.cuda.device(self.device):
                check(self.lib.dyc_wide_dump(self.h, t, k.ctypes.data, v.ctypes.data, k.size, _stream_ptr(stream)),
                      "dyc_wide_dump")
            ks.append(k)
            vs.append(v)
        return ks, vs

    def load(self, keys, vals, stream=None) -> None:
        ks = [np.ascontiguousarray(x, dtype=np.uint64) for x in keys]
        vs = [np.ascontiguousarray(x, dtype=np.uint64) for x in vals]
        d = self.stats()["subtables"]
        if len(ks) != d or len(vs) != d or any(a.size != b.size or a.size % SLOTS or not a.size for a, b in zip(ks, vs)):
            raise ValueError(f"need {d} pairs of equal, non-empty images of whole buckets")
        nb = (ctypes.c_uint32 * d)(*[x.size // SLOTS for x in ks])
        kp = (ctypes.c_void_p * d)(*[x.ctypes.data for x in ks])
        vp = (ctypes.c_void_p * d)(*[x.ctypes.data for x in vs])
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_load(self.h, nb, kp, vp, _stream_ptr(stream)), "dyc_wide_load")

    def pending(self, stream=None):
        n = ctypes.c_uint64()
        check(self.lib.dyc_wide_pending(self.h, None, None, 0, ctypes.byref(n), _stream_ptr(stream)), "dyc_wide_pending")
        k = np.empty(n.value, dtype=np.uint64)
        v = np.empty(n.value, dtype=np.uint64)
        if n.value:
            check(self.lib.dyc_wide_pending(self.h, k.ctypes.data, v.ctypes.data, n.value, ctypes.byref(n),
                                            _stream_ptr(stream)), "dyc_wide_pending")
        return k, v


class HashAPI:
    """Host-array facade with the reference's method names (GPU/dyhash/api.h:9-55).

    ``hashAPI(size)`` sizes from a slot-count hint (api.cpp:47-56); ``hash_insert`` /
    ``hash_search`` / ``hash_delete`` take host arrays and a size and run synchronously
    (api.cpp:184-303); the ``*_batch`` calls walk a pool staged by ``set_data_to_gpu`` in
    slices of ``batch_size`` (api.cpp:309-506), with the reference's resize policy
    (fill bounds ``lower``/``upper``) applied around each batch.
    """

    BATCH_SIZE = 100000  # test_uilt.h:5
    BATCH_NUM = 1000     # test_uilt.h:4

    def __init__(self, size: int, *, batch_size: int | None = None, **config):
        self.table = DyCuckoo(int(size), **config)
        self.batch_size = batch_size or self.BATCH_SIZE
        self._pool_k = self._pool_v = None
        self.batch_reset()

    # -- synchronous host-array operations (api.cpp:184-303)
    # (host arrays move through the library's pipelined host-pointer calls)
    def hash_insert(self, key, value, size: int) -> None:
        self.table.insert_host(np.asarray(key)[:size], np.asarray(value)[:size])

    def hash_search(self, key, ans, size: int):
        """Fills ans[:size] with the values (0 on a miss); returns the hit flags."""
        vals, hit = self.table.find_host(np.asarray(key)[:size])
        ans[:size] = vals
        return hit

    def hash_delete(self, key, ans, size: int):
        """Deletes key[:size]; ans[:size] receives the deleted flags (the reference leaves it unused)."""
        d = self.table.erase_host(np.asarray(key)[:size])
        if ans is not None:
            ans[:size] = d
        return d

    def resize_up(self) -> None:
        self.table.resize_up()

    def resize_low(self) -> bool:
        return self.table.resize_down()

    # -- batch API over a device pool (api.cpp:309-506)
    def set_data_to_gpu(self, key, value, size: int) -> bool:
        self._pool_k = _as_u32_tensor(np.asarray(key, dtype=np.uint32)[:size], self.table.device)
        self._pool_v = _as_u32_tensor(np.asarray(value, dtype=np.uint32)[:size], self.table.device)
        self._check = torch.empty(self.batch_size, dtype=torch.int32, device=self.table.device)
        self._hit = torch.empty(self.batch_size, dtype=torch.uint8, device=self.table.device)
        return True

    def batch_reset(self) -> None:
        self.batch_num_to_op = 0
        self.batch_search_num_to_op = 0
        self.batch_del_num_to_op = 0

    def _slice(self, i: int):
        lo = i * self.batch_size
        hi = min(lo + self.batch_size, self._pool_k.numel())
        if lo >= hi or i >= self.BATCH_NUM:
            raise IndexError(f"batch {i} is beyond the staged pool")
        return lo, hi

    def hash_insert_batch(self) -> None:
        """api.cpp:353-369: upsize while count + batch_size exceeds upper * capacity (the bound
        counts the full batch size whatever the slice holds, and applies whatever auto_resize
        says), then insert the slice."""
        lo, hi = self._slice(self.batch_num_to_op)
        upper = self.table.config()["upper_bound"]
        while True:
            st = self.table.stats()
            if st["count"] + self.batch_size <= upper * st["capacity"]:
                break
            self.table.resize_up()
        self.table.insert(self._pool_k[lo:hi], self._pool_v[lo:hi])
        self.batch_num_to_op += 1

    def hash_search_batch(self) -> None:
        lo, hi = self._slice(self.batch_search_num_to_op)
        self.table.find(self._pool_k[lo:hi], self._check[: hi - lo], self._hit[: hi - lo])
        self.batch_search_num_to_op += 1

    def hash_delete_batch(self) -> None:
        """api.cpp:466-506: erase the slice, then downsize once when count < lower * capacity
        and capacity > min_capacity (with auto_resize the erase has applied that rule itself)."""
        lo, hi = self._slice(self.batch_del_num_to_op)
        self.table.erase(self._pool_k[lo:hi])
        self.batch_del_num_to_op += 1
        cfg, st = self.table.config(), self.table.stats()
        if (not cfg["auto_resize"] and st["count"] < cfg["lower_bound"] * st["capacity"]
                and st["capacity"] > cfg["min_capacity"]):
            self.table.resize_down()

    def close(self):
        self.table.close()


__all__ = ["DyCuckoo", "DyCuckooGroup", "HashAPI", "DycError"]


class DyCuckooGroup:
    """Many independent tables served by one persistent launch per batch (dyc_group_*).

    A batch is a flat device tensor of keys plus host offsets (n_tables + 1 entries): table i's
    keys are keys[offsets[i]:offsets[i+1]].
    """

    def __init__(self, capacity_hints, *, device=None, **config):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = _lib.default_config(**config)
        hints = np.ascontiguousarray(np.asarray(capacity_hints, dtype=np.uint64))
        self.n = int(hints.size)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_group_create(ctypes.byref(self.cfg), self.n, hints.ctypes.data, ctypes.byref(h)),
                  "dyc_group_create")
        self.h = h

    def close(self):
        for p in list(getattr(self, "_plans", ())):  # a plan must go before its group
            p.close()
        if getattr(self, "h", None):
            self.lib.dyc_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _offsets(self, offsets) -> np.ndarray:
        o = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        if o.size != self.n + 1:
            raise ValueError(f"need {self.n + 1} offsets, got {o.size}")
        return o

    @property
    def launch(self) -> dict:
        out = _lib.GroupLaunch()
        check(self.lib.dyc_group_get_launch(self.h, ctypes.byref(out)), "dyc_group_get_launch")
        return out.as_dict()

    def set_launch(self, **fields) -> None:
        cur = _lib.GroupLaunch()
        check(self.lib.dyc_group_get_launch(self.h, ctypes.byref(cur)), "dyc_group_get_launch")
        for k, v in fields.items():
            setattr(cur, k, int(v))
        check(self.lib.dyc_group_set_launch(self.h, ctypes.byref(cur)), "dyc_group_set_launch")

    def insert(self, keys, vals, offsets, per_table: bool = False, stream=None) -> None:
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        o = self._offsets(offsets)
        fn = self.lib.dyc_group_insert_per_table if per_table else self.lib.dyc_group_insert
        with torch.cuda.device(self.device):
            check(fn(self.h, k.data_ptr(), v.data_ptr(), o.ctypes.data, _stream_ptr(stream)), "dyc_group_insert")

    def find(self, keys, offsets, vals=None, hit=None, per_table: bool = False, stream=None):
        k = _as_u32_tensor(keys, self.device)
        o = self._offsets(offsets)
        n = k.numel()
        if vals is None:
            vals = torch.empty(n, dtype=torch.int32, device=self.device)
        if hit is None:
            hit = torch.empty(n, dtype=torch.uint8, device=self.device)
        fn = self.lib.dyc_group_find_per_table if per_table else self.lib.dyc_group_find
        with torch.cuda.device(self.device):
            check(fn(self.h, k.data_ptr(), o.ctypes.data, vals.data_ptr(), hit.data_ptr(), _stream_ptr(stream)),
                  "dyc_group_find")
        return vals, hit

    def plan(self, keys, vals, offsets, lookups, lookup_offsets, out_vals=None, out_hit=None) -> "GroupPlan":
        """Capture the step clear -> insert(keys, vals) -> find(lookups) into a HIP graph
        (dyc_group_plan_create); GroupPlan.run() replays it with one launch. The tensors are
        fixed at capture (the plan keeps them alive); their contents may change between runs."""
        k = _as_u32_tensor(keys, self.device)
        v = _as_u32_tensor(vals, self.device)
        q = _as_u32_tensor(lookups, self.device)
        o = self._offsets(offsets)
        fo = self._offsets(lookup_offsets)
        if out_vals is None:
            out_vals = torch.empty(q.numel(), dtype=torch.int32, device=self.device)
        if out_hit is None:
            out_hit = torch.empty(q.numel(), dtype=torch.uint8, device=self.device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            torch.cuda.synchronize(self.device)
            check(self.lib.dyc_group_plan_create(self.h, k.data_ptr(), v.data_ptr(), o.ctypes.data, q.data_ptr(),
                                                 fo.ctypes.data, out_vals.data_ptr(), out_hit.data_ptr(),
                                                 ctypes.byref(h)), "dyc_group_plan_create")
        plan = GroupPlan(self, h, (k, v, q, out_vals, out_hit))
        if not hasattr(self, "_plans"):
            self._plans = weakref.WeakSet()
        self._plans.add(plan)
        return plan

    def erase(self, keys, offsets, deleted=None, stream=None):
        k = _as_u32_tensor(keys, self.device)
        o = self._offsets(offsets)
        if deleted is None:
            deleted = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_group_erase(self.h, k.data_ptr(), o.ctypes.data, deleted.data_ptr(), _stream_ptr(stream)),
                  "dyc_group_erase")
        return deleted

    def clear(self, stream=None) -> None:
        with torch.cuda.device(self.device):
            check(self.lib.dyc_group_clear(self.h, _stream_ptr(stream)), "dyc_group_clear")

    def table_stats(self, table: int) -> dict:
        st = _lib.Stats()
        check(self.lib.dyc_group_table_stats(self.h, int(table), ctypes.byref(st)), "dyc_group_table_stats")
        return st.as_dict()

    def dump(self, table: int, stream=None):
        out = []
        for t, nb in enumerate(self.table_stats(table)["buckets"]):
            a = np.empty(nb * SLOTS, dtype=np.uint64)
            with torch.cuda.device(self.device):
                check(self.lib.dyc_group_dump(self.h, int(table), t, a.ctypes.data, a.size, _stream_ptr(stream)),
                      "dyc_group_dump")
            out.append(a)
        return out

    def pending(self, stream=None):
        n = ctypes.c_uint64()
        check(self.lib.dyc_group_pending(self.h, None, None, 0, ctypes.byref(n), _stream_ptr(stream)),
              "dyc_group_pending")
        e = np.empty(n.value, dtype=np.uint64)
        tb = np.empty(n.value, dtype=np.uint32)
        if n.value:
            check(self.lib.dyc_group_pending(self.h, e.ctypes.data, tb.ctypes.data, n.value, ctypes.byref(n),
                                             _stream_ptr(stream)), "dyc_group_pending")
        return e, tb


class GroupPlan:
    """A captured group step (see DyCuckooGroup.plan)."""

    def __init__(self, group: DyCuckooGroup, h, tensors):
        self.group, self.h, self.tensors = group, h, tensors
        self.out_vals, self.out_hit = tensors[3], tensors[4]

    def run(self, stream=None) -> None:
        """One graph launch; waits for it and takes in the counters (as the eager insert does)."""
        with torch.cuda.device(self.group.device):
            check(self.group.lib.dyc_group_plan_run(self.h, _stream_ptr(stream)), "dyc_group_plan_run")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.group.lib.dyc_group_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DyCuckooWideGroup:
    """Many wide-key tables (uint64 -> uint64) served by one persistent launch per batch
    (dyc_wide_group_*). A batch is a flat device tensor of int64 keys plus host offsets
    (n_tables + 1 entries): table i's keys are keys[offsets[i]:offsets[i+1]]."""

    def __init__(self, capacity_hints, *, device=None, **config):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = _lib.default_config(**config)
        hints = np.ascontiguousarray(np.asarray(capacity_hints, dtype=np.uint64))
        self.n = int(hints.size)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_group_create(ctypes.byref(self.cfg), self.n, hints.ctypes.data, ctypes.byref(h)),
                  "dyc_wide_group_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.dyc_wide_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _offsets(self, offsets) -> np.ndarray:
        o = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        if o.size != self.n + 1:
            raise ValueError(f"need {self.n + 1} offsets, got {o.size}")
        return o

    def insert(self, keys, vals, offsets, per_table: bool = False, stream=None) -> None:
        k = _as_u64_tensor(keys, self.device)
        v = _as_u64_tensor(vals, self.device)
        o = self._offsets(offsets)
        fn = self.lib.dyc_wide_group_insert_per_table if per_table else self.lib.dyc_wide_group_insert
        with torch.cuda.device(self.device):
            check(fn(self.h, k.data_ptr(), v.data_ptr(), o.ctypes.data, _stream_ptr(stream)), "dyc_wide_group_insert")

    def find(self, keys, offsets, vals=None, hit=None, per_table: bool = False, stream=None):
        k = _as_u64_tensor(keys, self.device)
        o = self._offsets(offsets)
        if vals is None:
            vals = torch.empty(k.numel(), dtype=torch.int64, device=self.device)
        if hit is None:
            hit = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        fn = self.lib.dyc_wide_group_find_per_table if per_table else self.lib.dyc_wide_group_find
        with torch.cuda.device(self.device):
            check(fn(self.h, k.data_ptr(), o.ctypes.data, vals.data_ptr(), hit.data_ptr(), _stream_ptr(stream)),
                  "dyc_wide_group_find")
        return vals, hit

    def erase(self, keys, offsets, deleted=None, stream=None):
        k = _as_u64_tensor(keys, self.device)
        o = self._offsets(offsets)
        if deleted is None:
            deleted = torch.empty(k.numel(), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_group_erase(self.h, k.data_ptr(), o.ctypes.data, deleted.data_ptr(),
                                                _stream_ptr(stream)), "dyc_wide_group_erase")
        return deleted

    def clear(self, stream=None) -> None:
        with torch.cuda.device(self.device):
            check(self.lib.dyc_wide_group_clear(self.h, _stream_ptr(stream)), "dyc_wide_group_clear")

    def table_stats(self, table: int) -> dict:
        st = _lib.Stats()
        check(self.lib.dyc_wide_group_table_stats(self.h, int(table), ctypes.byref(st)), "dyc_wide_group_table_stats")
        return st.as_dict()

    def dump(self, table: int):
        """(key images, value images) of the table's subtables, host uint64 arrays."""
        st = self.table_stats(table)
        ks, vs = [], []
        for sub, nb in enumerate(st["buckets"]):
            k = np.zeros(nb * SLOTS, dtype=np.uint64)
            v = np.zeros(nb * SLOTS, dtype=np.uint64)
            check(self.lib.dyc_wide_group_dump(self.h, int(table), sub, k.ctypes.data, v.ctypes.data, k.size, None),
                  "dyc_wide_group_dump")
            ks.append(k)
            vs.append(v)
        return ks, vs

    def pending(self):
        """(keys, values, tables) of the entries waiting outside their tables' buckets."""
        n = ctypes.c_uint64()
        check(self.lib.dyc_wide_group_pending(self.h, None, None, None, 0, ctypes.byref(n), None),
              "dyc_wide_group_pending")
        k = np.zeros(n.value, dtype=np.uint64)
        v = np.zeros(n.value, dtype=np.uint64)
        t = np.zeros(n.value, dtype=np.uint32)
        if n.value:
            check(self.lib.dyc_wide_group_pending(self.h, k.ctypes.data, v.ctypes.data, t.ctypes.data, n.value,
                                                  ctypes.byref(n), None), "dyc_wide_group_pending")
        return k, v, t
