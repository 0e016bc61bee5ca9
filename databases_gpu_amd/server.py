"""Python face of the serving loop (dyc_server_* in include/dyc.h).

A server holds many resizable DyCuckoo tables in one HBM arena and serves queued requests --
the reference's batch calls (hash_insert / hash_search / hash_delete, GPU/dyhash/api.cpp:185-300)
of one table each -- by continuous batching: every tick fuses the admitted requests into one
insert, one find and one erase launch, and the next tick is enqueued while the current one runs.
With resident=1 one launch stays on the GPU instead and takes requests from a ring in host
memory as they are submitted (the device-resident loop).
Inputs and outputs are device tensors (int32 bits of uint32 keys/values, uint8 flags) that must
stay alive until their request completes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import OP_ERASE, OP_FIND, OP_INSERT, check  # noqa: F401


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("server requests take contiguous device tensors")
    return t.data_ptr()


class DyCuckooServer:
    def __init__(self, capacity_hints, *, hbm_budget: int = 0, tick_max_keys: int = 1 << 23,
                 tick_max_requests: int = 4096, pipeline_depth: int = 2, item_keys: int = 0, pool_cap: int = 65536,
                 one_launch: int = 1, spin_cap: int = 1 << 24, poll_backoff: int = 16, resident: int = 0,
                 resident_blocks: int = 0, resident_idle_ms: int = 100, **table_config):
        self.lib = _lib.load()
        hints = np.ascontiguousarray(np.asarray(capacity_hints, dtype=np.uint64))
        cfg = _lib.default_config(**table_config)
        sc = _lib.ServerConfig()
        check(self.lib.dyc_server_config_init(ctypes.byref(sc)), "dyc_server_config_init")
        sc.hbm_budget, sc.tick_max_keys, sc.tick_max_requests = hbm_budget, tick_max_keys, tick_max_requests
        sc.pipeline_depth, sc.item_keys, sc.pool_cap = pipeline_depth, item_keys, pool_cap
        sc.one_launch, sc.spin_cap, sc.poll_backoff = one_launch, spin_cap, poll_backoff
        sc.resident, sc.resident_blocks, sc.resident_idle_ms = resident, resident_blocks, resident_idle_ms
        h = ctypes.c_void_p()
        check(self.lib.dyc_server_create(ctypes.byref(cfg), ctypes.byref(sc), hints.size, hints.ctypes.data,
                                         ctypes.byref(h)), "dyc_server_create")
        self.h = h
        self.n_tables = int(hints.size)
        self._keep = {}  # ticket -> tensors the request uses, released when it completes

    def close(self):
        if self.h:
            self.lib.dyc_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- asynchronous requests
    def submit(self, table: int, op: int, keys: torch.Tensor, vals: torch.Tensor | None = None,
               out_vals: torch.Tensor | None = None, out_flags: torch.Tensor | None = None) -> int:
        # the server's stream does not wait for torch's: inputs must be complete before submit
        r = _lib.Request(table, op, _ptr(keys), _ptr(vals), _ptr(out_vals), _ptr(out_flags), keys.numel())
        tk = ctypes.c_uint64()
        check(self.lib.dyc_server_submit(self.h, ctypes.byref(r), ctypes.byref(tk)), "dyc_server_submit")
        self._keep[tk.value] = (keys, vals, out_vals, out_flags)
        return tk.value

    def insert(self, table, keys, vals) -> int:
        return self.submit(table, OP_INSERT, keys, vals)

    def find(self, table, keys, out_vals=None, out_flags=None) -> int:
        return self.submit(table, OP_FIND, keys, None, out_vals, out_flags)

    def erase(self, table, keys, out_flags=None) -> int:
        return self.submit(table, OP_ERASE, keys, None, None, out_flags)

    def poll(self, cap: int = 4096):
        """Advance the loop without blocking; returns [(ticket, status), ...] finished since."""
        tk = np.zeros(cap, dtype=np.uint64)
        stt = np.zeros(cap, dtype=np.int32)
        nd = ctypes.c_uint32()
        check(self.lib.dyc_server_poll(self.h, tk.ctypes.data, stt.ctypes.data, cap, ctypes.byref(nd)),
              "dyc_server_poll")
        out = [(int(tk[i]), int(stt[i])) for i in range(nd.value)]
        for t, _ in out:
            self._keep.pop(t, None)
        return out

    def drain(self):
        check(self.lib.dyc_server_drain(self.h), "dyc_server_drain")
        self._keep.clear()

    def replay(self, requests: np.ndarray, arrivals, mode: int = 0):
        """Replay a request trace in real time (requests: structured array of REQUEST_DTYPE whose
        pointers the caller keeps alive). mode 0 continuous batching with host-built ticks, 1
        synchronous per-request launches, 2 continuous batching through the device-resident loop.
        Returns (done_s, status) arrays."""
        requests = np.ascontiguousarray(requests)
        assert requests.dtype == np.dtype(_lib.REQUEST_DTYPE)
        arr = np.ascontiguousarray(np.asarray(arrivals, dtype=np.float64))
        n = requests.size
        torch.cuda.synchronize()  # the trace's inputs were written on torch's stream
        done = np.zeros(n, dtype=np.float64)
        status = np.zeros(n, dtype=np.int32)
        check(self.lib.dyc_server_replay(self.h, requests.ctypes.data, n, arr.ctypes.data, mode, done.ctypes.data,
                                         status.ctypes.data), "dyc_server_replay")
        return done, status

    # -- inspection
    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(self.lib.dyc_server_get_stats(self.h, ctypes.byref(st)), "dyc_server_get_stats")
        return st.as_dict()

    def table_stats(self, table: int) -> dict:
        st = _lib.Stats()
        check(self.lib.dyc_server_table_stats(self.h, table, ctypes.byref(st)), "dyc_server_table_stats")
        return st.as_dict()

    def dump(self, table: int):
        """The table's subtable images (host uint64 arrays); drains the server first."""
        st = self.table_stats(table)
        out = []
        for sub in range(st["subtables"]):
            img = np.zeros(st["buckets"][sub] * _lib.SLOTS, dtype=np.uint64)
            check(self.lib.dyc_server_dump(self.h, table, sub, img.ctypes.data, img.size), "dyc_server_dump")
            out.append(img)
        return out
