"""Launch-shape autotuner for DyCuckoo groups (and single tables).

The kernels take their shape at run time: lanes per key ("bucket width": how many lanes share
one 128-byte bucket read -- the bucket itself stays 16 slots), grid size (resident waves per
CU), keys per work item, the eviction bound (iterations under the reference victim rule
before the pseudo-random rescue takes over; results are the same, placement and speed differ),
and, for groups, occupancy: the persistent kernels exist in builds for at least 6 or 8 waves
per SIMD (fewer registers each) besides the compiler's own choice.
`tune_group` times every candidate on the caller's own batch (or a sample of its tables) with
HIP events and installs the fastest insert shape and the fastest find shape (the occupancy
builds are tried on the best shape of the grid). Results are cached in a JSON file keyed by
device name and a workload signature.
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np
import torch

CACHE_PATH = os.environ.get("DYC_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "dyc_autotune.json"))

INSERT_SPACE = {
    "insert_lanes": (1, 2, 4),
    "insert_blocks": (1024, 2048, 4096, 8192),
    "item_keys": (2048, 8192),
    "kick_bound": (15, 0),
}
FIND_SPACE = {
    "find_lanes": (4, 8, 16),
    "find_blocks": (1024, 2048, 4096, 8192),
    "item_keys": (2048, 8192),
}
WAVES_PER_EU = (6, 8)  # occupancy builds tried on the best grid shape (0 = the compiler's own)


def _time(fn, reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def _signature(offsets: np.ndarray) -> str:
    sizes = np.diff(offsets.astype(np.int64))
    q = np.percentile(sizes, [10, 50, 90]) if sizes.size else [0, 0, 0]
    return f"n{sizes.size}-tot{int(sizes.sum())}-p{int(q[0])}-{int(q[1])}-{int(q[2])}"


def _load_cache() -> dict:
    try:
        with open(CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: dict) -> None:
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        with open(CACHE_PATH, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass


def tune_group(group, keys: torch.Tensor, vals: torch.Tensor, offsets, lookups: torch.Tensor, lookup_offsets,
               reps: int = 2, use_cache: bool = True, log=None) -> dict:
    """Pick the fastest insert and find shapes for this batch on this device and install them.

    The group is cleared before every insert trial and left holding the batch afterwards.
    Returns {"insert": {...shape, "ms"}, "find": {...shape, "ms"}, "trials": n, "seconds": s}.
    """
    offsets = np.asarray(offsets, dtype=np.uint64)
    lookup_offsets = np.asarray(lookup_offsets, dtype=np.uint64)
    # the insert order changes what the insert launch does (the whole batch, or the few keys the
    # bulk fill leaves), so shapes tuned under one order are not reused under another
    order = getattr(getattr(group, "cfg", None), "insert_order", 0)
    key = f"{torch.cuda.get_device_name(group.device)}|group|{_signature(offsets)}|io{order}"
    cache = _load_cache() if use_cache else {}
    t0 = time.perf_counter()
    if key in cache:
        best = cache[key]
    else:
        best = {}
        trials = 0
        ins_best = (float("inf"), None)
        for combo in itertools.product(*INSERT_SPACE.values()):
            shape = dict(zip(INSERT_SPACE.keys(), combo), insert_waves_per_eu=0)
            group.set_launch(**shape)

            def run():
                group.clear()
                group.insert(keys, vals, offsets)

            run()
            ms = _time(run, reps)
            trials += 1
            if log:
                log(f"autotune insert {shape}: {ms:.3f} ms")
            if ms < ins_best[0]:
                ins_best = (ms, shape)
        for w in WAVES_PER_EU:
            shape = dict(ins_best[1], insert_waves_per_eu=w)
            group.set_launch(**shape)

            def run():
                group.clear()
                group.insert(keys, vals, offsets)

            run()
            ms = _time(run, reps)
            trials += 1
            if log:
                log(f"autotune insert {shape}: {ms:.3f} ms")
            if ms < ins_best[0]:
                ins_best = (ms, shape)
        group.set_launch(**ins_best[1])
        group.clear()
        group.insert(keys, vals, offsets)
        out_v = torch.empty(lookups.numel(), dtype=torch.int32, device=lookups.device)
        out_h = torch.empty(lookups.numel(), dtype=torch.uint8, device=lookups.device)
        find_best = (float("inf"), None)
        for combo in itertools.product(*FIND_SPACE.values()):
            shape = dict(zip(FIND_SPACE.keys(), combo), find_waves_per_eu=0)
            group.set_launch(**shape)
            ms = _time(lambda: group.find(lookups, lookup_offsets, out_v, out_h), reps)
            trials += 1
            if log:
                log(f"autotune find {shape}: {ms:.3f} ms")
            if ms < find_best[0]:
                find_best = (ms, shape)
        for w in WAVES_PER_EU:
            shape = dict(find_best[1], find_waves_per_eu=w)
            group.set_launch(**shape)
            ms = _time(lambda: group.find(lookups, lookup_offsets, out_v, out_h), reps)
            trials += 1
            if log:
                log(f"autotune find {shape}: {ms:.3f} ms")
            if ms < find_best[0]:
                find_best = (ms, shape)
        best = {"insert": dict(ins_best[1], ms=round(ins_best[0], 4)),
                "find": dict(find_best[1], ms=round(find_best[0], 4)), "trials": trials}
        if use_cache:
            cache[key] = best
            _store_cache(cache)
    # install: the find item size only matters for find, the insert one for insert; keep the
    # insert's item size (both launches read the same field) unless find measured better with
    # the same value
    shape = {k: v for k, v in best["insert"].items() if k != "ms"}
    shape.update({k: v for k, v in best["find"].items() if k not in ("ms", "item_keys")})
    group.set_launch(**shape)
    best = dict(best, seconds=round(time.perf_counter() - t0, 3), installed=shape)
    return best


TABLE_INSERT_SPACE = {
    "insert_lanes": (1, 2, 4),
    "insert_blocks": (16384, 32768, 65536),
    "kick_bound": (15, 0),
}
TABLE_FIND_SPACE = {
    "find_lanes": (4, 8, 16),
}


def tune_table(table, keys: torch.Tensor, vals: torch.Tensor, lookups, reps: int = 2,
               use_cache: bool = True, log=None, kick_bounds=(15,)) -> dict:
    """Pick the fastest insert and find shapes of one table for this batch and install them.

    `lookups` is one tensor of lookups or a sequence of them (say uniform and skewed lookups):
    the find shape chosen is the one with the least total time over all of them, since one
    shape serves every find of the table. `kick_bounds` lists the eviction bounds to try; the
    default keeps the reference's bound (15 iterations of its victim rule), so only the launch
    shape changes. The table is cleared before every insert trial and left holding the batch
    afterwards."""
    sets = [lookups] if isinstance(lookups, torch.Tensor) else list(lookups)
    space = dict(TABLE_INSERT_SPACE, kick_bound=tuple(kick_bounds))
    key = (f"{torch.cuda.get_device_name(table.device)}|table|n{keys.numel()}|cap{table.capacity}"
           f"|kb{'-'.join(map(str, space['kick_bound']))}|f{'-'.join(str(q.numel()) for q in sets)}")
    cache = _load_cache() if use_cache else {}
    t0 = time.perf_counter()
    if key in cache:
        best = cache[key]
    else:
        trials = 0
        ins_best = (float("inf"), None)
        for combo in itertools.product(*space.values()):
            shape = dict(zip(space.keys(), combo))
            table.configure(**shape)

            def run():
                table.clear()
                table.insert(keys, vals)

            run()
            ms = _time(run, reps)
            trials += 1
            if log:
                log(f"autotune insert {shape}: {ms:.3f} ms")
            if ms < ins_best[0]:
                ins_best = (ms, shape)
        table.configure(**ins_best[1])
        table.clear()
        table.insert(keys, vals)
        m = max(q.numel() for q in sets)
        out_v = torch.empty(m, dtype=torch.int32, device=sets[0].device)
        out_h = torch.empty(m, dtype=torch.uint8, device=sets[0].device)
        find_best = (float("inf"), None)
        for combo in itertools.product(*TABLE_FIND_SPACE.values()):
            shape = dict(zip(TABLE_FIND_SPACE.keys(), combo))
            table.configure(**shape)
            ms = sum(_time(lambda q=q: table.find(q, out_v[:q.numel()], out_h[:q.numel()]), reps) for q in sets)
            trials += 1
            if log:
                log(f"autotune find {shape}: {ms:.3f} ms")
            if ms < find_best[0]:
                find_best = (ms, shape)
        best = {"insert": dict(ins_best[1], ms=round(ins_best[0], 4)),
                "find": dict(find_best[1], ms=round(find_best[0], 4)), "trials": trials}
        if use_cache:
            cache[key] = best
            _store_cache(cache)
    shape = {k: v for k, v in best["insert"].items() if k != "ms"}
    shape.update({k: v for k, v in best["find"].items() if k != "ms"})
    table.configure(**shape)
    return dict(best, seconds=round(time.perf_counter() - t0, 3), installed=shape)
