#!/usr/bin/env python3
"""Benchmark of the DyCuckoo hot path on one MI355X.

Workload (the large bulk configuration): 4 subtables sized for 2^28 unique uint32 keys at load
factor 0.90 (about 2.4 GB of buckets), values splitmix32(key). One step =
  clear the table; bulk insert 2^28 (key, value) pairs; 2^28 finds at 50% hit with uniformly
  drawn lookups; 2^28 finds drawn Zipf(0.99) over the inserted keys.
All inputs are generated on the device before timing (resident in HBM). Before the steps the
launch shapes and the eviction bound are tuned on the batch (untimed; --no-tune skips it,
--reference-kick keeps the reference bound). Every phase is timed with HIP events on the stream
the library launches on. `value` is the step's operations (3 * 2^28) per second of step time, in
Mops/s.

Roofline: the algorithmic bytes per operation follow the reference algorithm and layout
(128-byte buckets, pos1 then pos2): find = 4 (key) + 4 (value) + 1 (hit flag) + 128 * p, with
p the exact mean number of buckets the find order reads (dyc_probe_count over the same lookups,
outside the timed region); insert = 8 (key, value) + i * (128 + 8), with i the mean number of
bucket exchanges per key from the insert kernel's counter. Achieved GB/s = bytes / kernel time;
peak = 8 TB/s.

cpu_baseline: the same DyCuckoo on the host (oracle/dyc_oracle.c: OpenMP over keys, 64-bit CAS
exchanges, every thread OpenMP is given) on a bounded sample of the same workload (2^26 keys by
default), with the sequential replay on 2^22 keys alongside, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Mops/s insert & find at 2^28 keys, % of HBM roofline; p99 batch latency"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=28, help="keys = 2^log2n (default 28)")
    ap.add_argument("--theta", type=float, default=0.90, help="load factor the table is sized for")
    ap.add_argument("--hit", type=float, default=0.5, help="hit rate of the uniform lookups")
    ap.add_argument("--zipf", type=float, default=0.99)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--find-lanes", type=int, default=8)
    ap.add_argument("--insert-lanes", type=int, default=1)
    ap.add_argument("--insert-order", type=int, default=3,
                    help="bulk: dyc_config.insert_order (3 = bulk fill, the default; 0 = eviction insert in input order)")
    ap.add_argument("--insert-blocks", type=int, default=0, help="insert grid (0 = library default)")
    ap.add_argument("--max-blocks", type=int, default=0, help="find grid cap (0 = library default)")
    ap.add_argument("--cpu-log2n", type=int, default=28,
                    help="CPU baseline (multithreaded): 2^k keys (default: the GPU workload's full size)")
    ap.add_argument("--cpu-seq-log2n", type=int, default=22, help="sequential CPU sample: 2^k keys")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--memsys-regions-mb", default="",
                    help="memsys: also measure over these buffer sizes in MB, comma-separated (e.g. 32,192,2386)")
    ap.add_argument("--subtables", type=int, default=4, help="subtable count d of the bulk table (2..6)")
    ap.add_argument("--st-list", default="2,3,4,5,6", help="subtables: the subtable counts swept")
    ap.add_argument("--st-kick", type=int, default=15, help="subtables: kick bound of every table")
    ap.add_argument("--workload", choices=("bulk", "dynamic", "batched", "serving", "resize", "memsys", "host",
                                           "subtables", "wide"),
                    default="bulk",
                    help="bulk: the large bulk configuration (default, the headline line); dynamic: the "
                         "resize stream (growth then shrink batches under fill bounds); batched: many uneven "
                         "tables in one persistent launch vs one launch per table; serving: the request loop "
                         "(continuous batching vs synchronous per-request launches) over a load sweep; resize: "
                         "single-subtable upsize and downsize of a 2^28-slot-class table; memsys: the device's "
                         "random bucket-read, CAS and store rates at the table's size (the kernels' ceilings); "
                         "host: the bulk workload from host arrays (PCIe included), pipelined vs copy-then-run; "
                         "subtables: insert and find throughput for each subtable count at a fixed table size; "
                         "wide: the bulk insert and find with 64-bit keys and values")
    ap.add_argument("--sv-requests", type=int, default=4096, help="serving: requests per run")
    ap.add_argument("--sv-loads", default="0.25,0.5,1,2,4,8,16,0",
                    help="serving: offered loads in G keys/s (0 = every request arrives at once)")
    ap.add_argument("--sv-modes", default="0,2,1",
                    help="serving: replay modes per load (0 host-built ticks, 2 device-resident loop, 1 synchronous)")
    ap.add_argument("--sv-resident-blocks", type=int, default=0,
                    help="serving: blocks of the resident loop launch (0 = half its resident capacity)")
    ap.add_argument("--sv-item-keys", type=int, default=0,
                    help="serving: keys per work item (0 = chosen per tick, or per request in the resident loop)")
    ap.add_argument("--sv-depth", type=int, default=2, help="serving: ticks in flight")
    ap.add_argument("--sv-tick-keys", type=int, default=1 << 23, help="serving: keys per tick")
    ap.add_argument("--sv-kick", type=int, default=0,
                    help="serving: eviction bound under the reference victim rule before the rescue walk (both "
                         "modes; results do not depend on it)")
    ap.add_argument("--sv-one-launch", type=int, default=1, help="serving: 1 one launch per tick, 0 per round")
    ap.add_argument("--sv-backoff", type=int, default=16, help="serving: poll backoff cap (x 256 cycles)")
    ap.add_argument("--bt-tables", type=int, default=8192, help="batched: number of tables")
    ap.add_argument("--bt-log2min", type=int, default=8, help="batched: log2 of the smallest table")
    ap.add_argument("--bt-log2max", type=int, default=20, help="batched: log2 of the largest table")
    ap.add_argument("--bt-theta", type=float, default=0.85, help="batched: per-table load factor")
    ap.add_argument("--bt-hit", type=float, default=0.6, help="batched: find hit rate")
    ap.add_argument("--bt-wide", action="store_true",
                    help="batched: uint64 keys and values (wide-key groups; no autotuning or graph form)")
    ap.add_argument("--bt-insert-order", type=int, default=-1, choices=(-1, 0, 3),
                    help="batched: the group insert (3: bulk fill per table chunk, then the eviction insert for "
                         "what overflows; 0: the eviction insert for every key; -1: 3 from 2^22 keys in the batch, "
                         "below that 0, whose single launch costs less than the fill's sequence). With 3 the "
                         "eviction insert is measured beside it (its launch shapes tuned for it); the graph form captures the same insert")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--kick-bound", type=int, default=None,
                    help="bulk: eviction bound of the table (with --no-tune it stays fixed; default: the library's 15, "
                         "or what the autotuner installs)")
    ap.add_argument("--reference-kick", action="store_true",
                    help="bulk: keep the reference eviction bound (15 iterations of its victim rule); by default "
                         "the autotuner also tries bound 0 (the rescue walk from the first kick), which changes "
                         "placement only -- stored sets and find results are the same")
    ap.add_argument("--dyn-batches", type=int, default=200, help="dynamic: batches (half grow, half shrink)")
    ap.add_argument("--dyn-log2ops", type=int, default=22, help="dynamic: 2^k operations per batch")
    ap.add_argument("--dyn-log2init", type=int, default=22, help="dynamic: 2^k keys live before the stream")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one plain copy kernel of 2^log2n int32 (4 B read + 4 B written per element) "
                         "before the steps, as the reference point for hardware byte counters")
    ap.add_argument("--traffic-csv", nargs="+", default=None, metavar="CSV",
                    help="no GPU: read rocprofv3 --pmc counter_collection CSVs of one bulk run made with "
                         "--calibrate-copy --steps 1 --warmup 0 (FETCH_SIZE, WRITE_SIZE, and optionally "
                         "TCC_HIT_sum/TCC_MISS_sum) and write HBM bytes per launch to --json-out")
    return ap.parse_args()


# ---------------------------------------------------------------- hardware byte counters
def traffic_from_csv(paths, log2n: int, cfg: dict) -> dict:
    """HBM bytes of the bulk step's launches from rocprofv3 --pmc counter CSVs (one counter set per
    file) of `bench.py --calibrate-copy --steps 1 --warmup 0`: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by the factor the same counter shows for the calibration copy kernel
    (2^log2n int32 read and written), summed per stage of the insert; L2 hit rate of each find from
    TCC_HIT_sum / TCC_MISS_sum when present."""
    import csv
    import collections

    disp = collections.OrderedDict()  # dispatch id -> {"name", counters...}
    for path in paths:
        for r in csv.DictReader(open(path)):
            d = disp.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"]})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    seq = [disp[k] for k in sorted(disp)]
    n = 1 << log2n
    i0 = next(i for i, d in enumerate(seq) if "pack_kernel<false" in d["name"])
    copy = [d for d in seq[:i0] if "neg" in d["name"].lower()][-1]  # the calibration copy, torch.neg
    fetch_f = 4 * n / (copy["FETCH_SIZE"] * 1024)
    write_f = 4 * n / (copy["WRITE_SIZE"] * 1024)

    def short(name):
        name = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("dyc::", "")
        return name.split("(")[0]

    def nbytes(d):
        return d.get("FETCH_SIZE", 0.0) * 1024 * fetch_f + d.get("WRITE_SIZE", 0.0) * 1024 * write_f

    # the step: clear, insert (first pack of a batch .. the counter read-back), uniform finds, Zipf finds
    i1 = next(i for i in range(i0, len(seq)) if "fold_post_kernel" in seq[i]["name"])
    stages = collections.OrderedDict()
    for d in seq[i0:i1 + 1]:
        st = stages.setdefault(short(d["name"]), {"launches": 0, "fetch_bytes": 0.0, "write_bytes": 0.0})
        st["launches"] += 1
        st["fetch_bytes"] += d.get("FETCH_SIZE", 0.0) * 1024 * fetch_f
        st["write_bytes"] += d.get("WRITE_SIZE", 0.0) * 1024 * write_f
    for st in stages.values():
        st["bytes_per_key"] = round((st["fetch_bytes"] + st["write_bytes"]) / n, 2)
        st["fetch_bytes"], st["write_bytes"] = int(st["fetch_bytes"]), int(st["write_bytes"])
    finds = [d for d in seq[i1 + 1:] if "probe_kernel" in d["name"]][:2]
    insert_bytes = sum(nbytes(d) for d in seq[i0:i1 + 1])
    out = {
        "_note": "HBM bytes per launch of the large bulk step (2^%d keys, load factor 0.90) from rocprofv3 --pmc "
                 "FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by the factor the same counter shows "
                 "for a plain copy kernel of 2^%d int32 (bench.py --calibrate-copy). FETCH_SIZE counts what L2 "
                 "fetches from the fabric, so a line the MALL serves counts as fetched. Written by bench.py "
                 "--traffic-csv." % (log2n, log2n),
        "calibration": {"fetch_factor": round(fetch_f, 5), "write_factor": round(write_f, 5)},
        "config": cfg,
        "insert": {"log2n": log2n, "bytes_per_launch": int(insert_bytes),
                   "bytes_per_op": round(insert_bytes / n, 2), "stages": stages},
    }
    for name, d in zip(("find_uniform", "find_zipf"), finds):
        e = {"log2n": log2n, "bytes_per_launch": int(nbytes(d)), "bytes_per_op": round(nbytes(d) / n, 2)}
        if "TCC_HIT_sum" in d and "TCC_MISS_sum" in d:
            e["l2_hit_rate"] = round(d["TCC_HIT_sum"] / max(1.0, d["TCC_HIT_sum"] + d["TCC_MISS_sum"]), 4)
        out[name] = e
    return out


def pct(xs, q):
    xs = np.asarray(xs, dtype=np.float64)
    return float(np.percentile(xs, q)) if xs.size else None


def cpu_baseline(log2n: int, theta: float, hit: float, zipf_s: float, seed: int, seq_log2n: int) -> dict:
    """The same workload on the host: multithreaded DyCuckoo (POSIX threads over key chunks,
    64-bit CAS exchanges) on 2^log2n keys over every host thread the box gives (OMP_NUM_THREADS),
    plus the sequential replay on 2^seq_log2n keys for information."""
    import oracle as O
    from databases_gpu_amd import workload as W

    def inputs(k):
        n = 1 << k
        keys_t = W.keys_range(0, n, seed, "cpu")
        keys = keys_t.numpy().view(np.uint32).copy()
        look = W.mixed_lookups(keys_t, n, hit, seed, n, seed + 7).numpy().view(np.uint32).copy()
        zl = W.zipf_lookups(keys_t, n, zipf_s, seed + 11).numpy().view(np.uint32).copy()
        return n, keys, W.splitmix32_np(keys), look, zl

    threads = int(O.lib().orc_max_threads())
    n, keys, vals, look, zl = inputs(log2n)
    tb = O.OracleTable(int(math.ceil(n / theta)), rescue_kicks=128)
    out_v = np.empty(n, dtype=np.uint32)
    out_h = np.empty(n, dtype=np.uint8)
    t0 = time.perf_counter()
    tb.insert_mt(keys, vals, threads)
    t1 = time.perf_counter()
    s = tb._struct()
    O.lib().orc_find_mt(O.ctypes.byref(s), O._p(look), n, O._p(out_v), O._p(out_h), threads)
    t2 = time.perf_counter()
    hits_ok = int(out_h.sum()) == int(round(hit * n))
    O.lib().orc_find_mt(O.ctypes.byref(s), O._p(zl), n, O._p(out_v), O._p(out_h), threads)
    t3 = time.perf_counter()
    hits_ok &= bool(out_h.all()) and tb.pending.size == 0 and tb.count == n
    del tb, keys, vals, look, zl, out_v, out_h

    ns, keys, vals, look, zl = inputs(seq_log2n)
    sq = O.OracleTable(int(math.ceil(ns / theta)), rescue_kicks=128)
    u0 = time.perf_counter()
    sq.insert(keys, vals)
    u1 = time.perf_counter()
    sq.find(look)
    u2 = time.perf_counter()
    sq.find(zl)
    u3 = time.perf_counter()
    return {
        "value": round(3 * n / (t3 - t0) / 1e6, 3),
        "unit": "Mops/s",
        "cores": threads,
        "kind": "port",
        "sample": f"2^{log2n} keys at load factor {theta}: multithreaded insert + 2^{log2n} uniform finds "
                  f"({hit:.0%} hit) + 2^{log2n} Zipf({zipf_s}) finds, {threads} POSIX threads",
        "insert_mops": round(n / (t1 - t0) / 1e6, 3),
        "find_uniform_mops": round(n / (t2 - t1) / 1e6, 3),
        "find_zipf_mops": round(n / (t3 - t2) / 1e6, 3),
        "seconds": round(t3 - t0, 3),
        "hits_ok": hits_ok,
        "sequential": {
            "keys": ns, "cores": 1,
            "value": round(3 * ns / (u3 - u0) / 1e6, 3),
            "insert_mops": round(ns / (u1 - u0) / 1e6, 3),
            "find_uniform_mops": round(ns / (u2 - u1) / 1e6, 3),
            "find_zipf_mops": round(ns / (u3 - u2) / 1e6, 3),
            "seconds": round(u3 - u0, 3),
        },
    }


def run_dynamic(args, dev) -> int:
    """The dynamic resize stream: batches of 2^dyn_log2ops operations, insert -> find -> delete
    in each batch; growth batches are 40/40/20 insert/find/delete, shrink batches 20/40/40; finds
    hit 70%; fill bounds [0.50, 0.90] trigger single-subtable upsize/downsize in line. The mix is
    honoured exactly, so the table peaks at initial + 0.2 * ops * batches/2 keys (2^26.4 with the
    defaults) and returns to the initial count. Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    stream = torch.cuda.current_stream()
    ops = 1 << args.dyn_log2ops
    init = 1 << args.dyn_log2init
    half = args.dyn_batches // 2
    log(f"dynamic stream: {args.dyn_batches} batches of 2^{args.dyn_log2ops} ops from 2^{args.dyn_log2init} keys")
    ik, iv, batches = W.dynamic_stream(init, ops, half, args.dyn_batches - half, args.seed, dev)
    torch.cuda.synchronize()
    # the random walk from the first collision (kick bound 0) is the measured faster policy; the
    # reference's bound of 15 evictions is kept under --reference-kick
    kick = args.kick_bound if args.kick_bound is not None else (15 if args.reference_kick else 0)
    t = D.DyCuckoo(int(math.ceil(init / args.theta)), auto_resize=1, lower_bound=0.5, upper_bound=args.theta,
                   min_capacity=0, insert_lanes=args.insert_lanes, find_lanes=args.find_lanes,
                   insert_blocks=args.insert_blocks, kick_bound=kick)
    t.insert(ik, iv)
    out_v = torch.empty(ops, dtype=torch.int32, device=dev)
    out_h = torch.empty(ops, dtype=torch.uint8, device=dev)
    out_d = torch.empty(ops, dtype=torch.uint8, device=dev)
    lat, peak, hits_ok = [], 0, True
    st0 = t.stats()
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i, b in enumerate(batches):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        t.insert(b.ins_keys, b.ins_vals)
        nf = b.find_keys.numel()
        t.find(b.find_keys, out_v[:nf], out_h[:nf])
        t.erase(b.del_keys, out_d[: b.del_keys.numel()])
        e1.record(stream)
        e1.synchronize()
        lat.append(e0.elapsed_time(e1))
        st = t.stats()
        peak = max(peak, st["count"])
        hits_ok &= int(out_d[: b.del_keys.numel()].sum()) == b.del_keys.numel()
        hits_ok &= st["count"] == b.live_after and st["pending"] == 0 and st["lost"] == 0
        if (i + 1) % 20 == 0:
            log(f"batch {i + 1}/{len(batches)}: {lat[-1]:.2f} ms, {st['count']} keys, buckets {st['buckets']}")
    wall = time.perf_counter() - wall0
    st = t.stats()
    total_ops = sum(b.ops for b in batches)
    total_ms = float(np.sum(lat))
    result = {
        "metric": "Mops/s dynamic resize stream, p50/p99 batch latency",
        "value": round(total_ops / (total_ms * 1e-3) / 1e6, 1),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": len(batches),
        "warmup": 0,
        "ms_per_step": round(total_ms / len(batches), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"dynamic resize stream: {len(batches)} batches of 2^{args.dyn_log2ops} ops "
                               f"(grow 40/40/20, shrink 20/40/40 insert/find/delete, finds 70% hit), "
                               f"fill bounds [0.50, {args.theta}]",
                   "initial_keys": init, "peak_keys": peak, "final_keys": st["count"], "kick_bound": kick,
                   "insert_lanes": args.insert_lanes, "insert_blocks": args.insert_blocks},
        "p50_ms": round(pct(lat, 50), 3),
        "p99_ms": round(pct(lat, 99), 3),
        "resizes_up": st["resizes_up"] - st0["resizes_up"],
        "resizes_down": st["resizes_down"] - st0["resizes_down"],
        "final_buckets": st["buckets"],
        "wall_s": round(wall, 3),
        "parity": "ok" if hits_ok else "mismatch",
    }
    t.close()
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if hits_ok else 1


def run_host(args, dev) -> int:
    """The bulk workload with the keys in host memory (the hashAPI calling pattern): 2^log2n keys
    inserted, then 2^log2n uniform finds at `hit`, results back in host arrays. Two ways:
    pipelined (dyc_insert_host / dyc_find_host: chunked pinned staging, copies overlapping
    kernels) and copy-then-run (whole arrays to the device, the device-pointer call, results
    back). Wall time per phase, PCIe included; every find result checked. Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    n = 1 << args.log2n
    keys = W.keys_range(0, n, args.seed, dev)
    look = W.mixed_lookups(keys, n, args.hit, args.seed, n, args.seed + 7)
    hk = keys.cpu().numpy().view(np.uint32)
    hv = W.splitmix32_np(hk)
    hq = look.cpu().numpy().view(np.uint32)
    present = np.isin(hq, hk)
    del keys, look
    torch.cuda.synchronize()
    t = D.DyCuckoo(int(math.ceil(n / args.theta)))
    res = {"pipelined": {"insert": [], "find": []}, "copy_then_run": {"insert": [], "find": []}}
    ok = True
    for s_ in range(args.warmup + args.steps):
        for mode in ("pipelined", "copy_then_run"):
            t.clear()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == "pipelined":
                t.insert_host(hk, hv)
            else:
                t.insert(torch.from_numpy(hk.view(np.int32)).to(dev), torch.from_numpy(hv.view(np.int32)).to(dev))
                torch.cuda.synchronize()
            t1 = time.perf_counter()
            if mode == "pipelined":
                v, h = t.find_host(hq)
            else:
                dv, dh = t.find(torch.from_numpy(hq.view(np.int32)).to(dev))
                v, h = dv.cpu().numpy().view(np.uint32), dh.cpu().numpy()
            t2 = time.perf_counter()
            ok &= bool(np.array_equal(h.astype(bool), present)) and bool(np.array_equal(v[present], W.splitmix32_np(hq[present])))
            if s_ >= args.warmup:
                res[mode]["insert"].append(t1 - t0)
                res[mode]["find"].append(t2 - t1)
        log(f"step {s_ + 1}: pipelined insert {res['pipelined']['insert'][-1] if res['pipelined']['insert'] else 0:.4f} s")
    out = {}
    for mode, r in res.items():
        ti, tf = float(np.median(r["insert"])), float(np.median(r["find"]))
        out[mode] = {"insert_s": round(ti, 4), "find_s": round(tf, 4), "insert_mops": round(n / ti / 1e6, 1),
                     "find_mops": round(n / tf / 1e6, 1), "mops": round(2 * n / (ti + tf) / 1e6, 1)}
    t.close()
    result = {
        "metric": "Mops/s insert & find from host arrays (PCIe included), pipelined vs copy-then-run",
        "value": out["pipelined"]["mops"], "unit": "Mops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"2^{args.log2n} unique keys at load factor {args.theta} from host memory, then "
                               f"2^{args.log2n} finds at {args.hit:.0%} hit, results to host memory",
                   "stage_keys": 1 << 21},
        "pipelined": out["pipelined"], "copy_then_run": out["copy_then_run"],
        "speedup": round(out["pipelined"]["mops"] / out["copy_then_run"]["mops"], 3),
        "parity": "ok" if ok else "mismatch",
    }
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if ok else 1


def run_memsys(args, dev) -> int:
    """The device's random-access ceilings at the size of the 2^28-key table (2.39 GB): random
    128-byte bucket reads/s (one lane per bucket), random 64-bit CAS/s and random 8-byte
    stores/s, 2^28 operations each, and the rate of per-wave atomic adds onto one address and onto
    64 (bench_memsys.so, built from bench_memsys.hip). Prints one JSON line with the ceilings the bulk kernels are compared
    against."""
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "bench_memsys.so"))
    lib.dyc_memsys_probe.restype = ctypes.c_int
    lib.dyc_memsys_probe.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    nb = 4 * 4660350  # buckets of the 2^28-key table at load factor 0.90
    out = (ctypes.c_double * 7)()
    rc = lib.dyc_memsys_probe(nb * 128, 1 << args.log2n, 3, out)
    if rc:
        print(f"dyc_memsys_probe failed: hipError {rc}", file=sys.stderr)
        return 1
    reads, cas, stores, half, read_cas, one_addr, spread64 = out[0], out[1], out[2], out[3], out[4], out[5], out[6]
    result = {
        "metric": "device random-access ceilings at 2.39 GB: bucket reads/s, 64-bit CAS/s, 8-byte stores/s",
        "value": round(reads / 1e9, 2), "unit": "G bucket reads/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"random slots of a {nb * 128} B buffer, 2^{args.log2n} operations per pattern"},
        "bucket_reads_per_s": reads, "bucket_read_gbs": round(reads * 128 / 1e9, 1),
        "cas_per_s": cas, "stores_per_s": stores, "half_bucket_reads_per_s": half, "read_then_cas_per_s": read_cas,
        "wave_atomics_one_address_per_s": one_addr, "wave_atomics_64_addresses_per_s": spread64,
    }
    regions = []
    for mb in [int(x) for x in args.memsys_regions_mb.split(",") if x.strip()]:
        # the same patterns confined to a smaller buffer: below the 256 MB last-level cache (MALL)
        # the lines stay on chip, which shows what locality alone would buy each pattern
        print(f"[memsys] region {mb} MB", file=sys.stderr, flush=True)
        rc = lib.dyc_memsys_probe(mb << 20, 1 << args.log2n, 3, out)
        if rc:
            print(f"dyc_memsys_probe failed: hipError {rc}", file=sys.stderr)
            return 1
        regions.append({"MB": mb, "bucket_reads_per_s": out[0], "cas_per_s": out[1], "stores_per_s": out[2],
                        "half_bucket_reads_per_s": out[3], "read_then_cas_per_s": out[4]})
    if regions:
        result["regions"] = regions
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0


def run_resize(args, dev) -> int:
    """Single-subtable resize at scale: a table sized for 2^log2n keys at load factor theta,
    filled to 45 %, then alternately upsized (subtable chosen by the reference rule doubles) and
    downsized back, steps times each. The upsize kernel reads the old subtable and writes both
    halves of the new one (24 B per old slot); the downsize kernel reads both halves of the old
    subtable and writes the new one (24 B per new slot), then residuals are re-inserted. Times
    are the API calls (allocation, kernel, counter read) with HIP events; the kernels alone are
    in the rocprofv3 trace of this command. Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    stream = torch.cuda.current_stream()
    n = 1 << args.log2n
    hint = int(math.ceil(n / args.theta))
    t = D.DyCuckoo(hint)
    fill = int(0.45 * t.capacity)
    keys = W.keys_range(0, fill, args.seed, dev)
    t.insert(keys, W.splitmix32(keys))
    torch.cuda.synchronize()
    log(f"resize: buckets {t.buckets}, {fill} keys")
    ups, downs, up_bytes, down_bytes = [], [], [], []
    for s_ in range(args.warmup + args.steps):
        nb0 = t.buckets
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        t.resize_up()
        e[1].record(stream)
        nb1 = t.buckets
        t.resize_down()
        e[2].record(stream)
        e[2].synchronize()
        nb2 = t.buckets
        grown = [i for i in range(len(nb0)) if nb1[i] != nb0[i]][0]
        shrunk = [i for i in range(len(nb1)) if nb2[i] != nb1[i]][0]
        if s_ >= args.warmup:
            ups.append(e[0].elapsed_time(e[1]))
            downs.append(e[1].elapsed_time(e[2]))
            up_bytes.append(24 * 16 * nb0[grown])
            down_bytes.append(24 * 16 * nb2[shrunk])
    st = t.stats()
    v, h = t.find(keys)
    ok = bool(h.bool().all()) and bool((v == W.splitmix32(keys)).all()) and st["count"] == fill and st["pending"] == 0
    up_ms, down_ms = float(np.median(ups)), float(np.median(downs))
    result = {
        "metric": "single-subtable upsize / downsize, GB/s of algorithmic bytes per call",
        "value": round(float(np.median(up_bytes)) / (up_ms * 1e-3) / 1e9, 1),
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"table sized for 2^{args.log2n} keys at load factor {args.theta}, filled to 45 %",
                   "buckets": st["buckets"], "keys": fill},
        "upsize": {"ms": round(up_ms, 3), "bytes": int(np.median(up_bytes)),
                   "gbs": round(float(np.median(up_bytes)) / (up_ms * 1e-3) / 1e9, 1)},
        "downsize": {"ms": round(down_ms, 3), "bytes": int(np.median(down_bytes)),
                     "gbs": round(float(np.median(down_bytes)) / (down_ms * 1e-3) / 1e9, 1)},
        "resizes_up": st["resizes_up"], "resizes_down": st["resizes_down"],
        "parity": "ok" if ok else "mismatch",
    }
    t.close()
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if ok else 1


def run_subtables(args, dev) -> int:
    """Subtable count sweep (the paper's "varying the number of tables", experiment.tex:213-220):
    for each d, a table of the same total size (hint = 2^log2n / theta slots, so each subtable
    holds 1/d of it) is cleared, bulk-inserted with 2^log2n keys and probed with 2^log2n finds at
    the --hit rate, steps times; the medians of each phase are reported per d. An insert that
    overflows upsizes the table (counted in resizes_up). Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    stream = torch.cuda.current_stream()
    n = 1 << args.log2n
    hint = int(math.ceil(n / args.theta))
    keys = W.keys_range(0, n, args.seed, dev)
    vals = W.splitmix32(keys)
    look = W.mixed_lookups(keys, n, args.hit, args.seed, n, args.seed + 7)
    present = torch.isin(look, keys)
    out_v = torch.empty(n, dtype=torch.int32, device=dev)
    out_h = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    per_d, ok = {}, True
    for d in [int(x) for x in args.st_list.split(",")]:
        t = D.DyCuckoo(hint, subtables=d, kick_bound=args.st_kick, find_lanes=args.find_lanes)
        ins, fnd, kicks, races = [], [], [], []
        for s_ in range(args.warmup + args.steps):
            t.clear()
            races0 = t.stats()["cas_retries"]
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record(stream)
            t.insert(keys, vals)
            e[1].record(stream)
            t.find(look, out_v, out_h)
            e[2].record(stream)
            e[2].synchronize()
            st0 = t.stats()
            if s_ >= args.warmup:
                ins.append(e[0].elapsed_time(e[1]))
                fnd.append(e[1].elapsed_time(e[2]))
                kicks.append(st0["last_kicks"])
                races.append(st0["cas_retries"] - races0)
        st = t.stats()
        good = st["count"] == n and st["pending"] == 0 and st["lost"] == 0 and bool((out_h.bool() == present).all()) \
            and bool((out_v[present] == W.splitmix32(look[present])).all())
        probes, _ = t.probe_count(look)
        ok &= good
        i_ms, f_ms = float(np.median(ins)), float(np.median(fnd))
        per_d[str(d)] = {"insert_ms": round(i_ms, 3), "insert_mops": round(n / i_ms / 1e3, 1),
                         "find_ms": round(f_ms, 3), "find_mops": round(n / f_ms / 1e3, 1),
                         "mean_exchanges": round(float(np.mean(kicks)) / n, 4), "mean_probes": round(probes / n, 4),
                         "cas_retries_per_op": round(float(np.mean(races)) / n, 5),
                         "kick_failures": st["kick_failures"],
                         "buckets": st["buckets"], "fill": round(n / st["capacity"], 4),
                         "resizes_up": st["resizes_up"], "parity": "ok" if good else "mismatch"}
        log(f"subtables {d}: insert {i_ms:.2f} ms ({n / i_ms / 1e3:.0f} Mops/s), find {f_ms:.2f} ms "
            f"({n / f_ms / 1e3:.0f} Mops/s), buckets {st['buckets']}, {'ok' if good else 'MISMATCH'}")
        t.close()
        torch.cuda.empty_cache()
    best = max(per_d.values(), key=lambda r: r["insert_mops"] + r["find_mops"])
    result = {
        "metric": "insert and find throughput per subtable count at a fixed table size, Mops/s",
        "value": round(2 * n / (best["insert_ms"] + best["find_ms"]) / 1e3, 1), "unit": "Mops/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"2^{args.log2n} unique u32 keys into a table of 2^{args.log2n}/{args.theta} slots "
                               f"split over d subtables, then 2^{args.log2n} finds at {args.hit:.0%} hit",
                   "kick_bound": args.st_kick, "find_lanes": args.find_lanes},
        "per_subtables": per_d,
        "parity": "ok" if ok else "mismatch",
    }
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if ok else 1


def run_wide(args, dev) -> int:
    """Wide keys: 2^log2n unique u64 keys (low word a seeded bijection of the index, high word a
    mix of it) and u64 values into a table sized for load factor theta, then 2^log2n finds at the
    --hit rate; clear + insert + find per step, medians reported. Algorithmic bytes: insert
    16 + x (128 key line + 3 x 8 key CAS / value exchange / key publish) per key (x = exchanges
    per key), find 17 + 128 p + 8 h per lookup (p = key lines read, h = hit rate). Prints one
    JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    stream = torch.cuda.current_stream()
    n = 1 << args.log2n
    idx = torch.arange(0, n, dtype=torch.int64, device=dev)

    def wide_keys(lo_idx):
        lo = W.u32_to_i64(W.keys_range(int(lo_idx[0]), int(lo_idx[-1]) + 1, args.seed, dev))
        hi = W.u32_to_i64(W.splitmix32(W.keys_range(int(lo_idx[0]), int(lo_idx[-1]) + 1, args.seed + 101, dev)))
        return (hi << 32) | lo  # int64 bits of the u64 key; low word unique and nonzero

    keys = wide_keys(idx)
    golden = -7046029254386353131  # 0x9E3779B97F4A7C15 as int64 bits
    vals = keys * golden + 12345
    m_hit = int(n * args.hit)
    perm = torch.randperm(n, device=dev, generator=torch.Generator(device=dev).manual_seed(args.seed))
    absent = wide_keys(torch.arange(n, 2 * n - m_hit, dtype=torch.int64, device=dev)) if m_hit < n else keys[:0]
    look = torch.cat([keys[perm[:m_hit]], absent])
    look = look[torch.randperm(n, device=dev, generator=torch.Generator(device=dev).manual_seed(args.seed + 1))]
    out_v = torch.empty(n, dtype=torch.int64, device=dev)
    out_h = torch.empty(n, dtype=torch.uint8, device=dev)
    del perm, idx
    torch.cuda.synchronize()
    t = D.DyCuckooWide(int(math.ceil(n / args.theta)), kick_bound=0 if not args.reference_kick else 15,
                       insert_order=args.insert_order)
    ins, fnd, kicks = [], [], []
    for s_ in range(args.warmup + args.steps):
        t.clear()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        t.insert(keys, vals)
        e[1].record(stream)
        t.find(look, out_v, out_h)
        e[2].record(stream)
        e[2].synchronize()
        st = t.stats()
        if st["count"] != n or st["pending"] or st["lost"]:
            raise RuntimeError(f"wide insert left the table inconsistent: {st}")
        if s_ >= args.warmup:
            ins.append(e[0].elapsed_time(e[1]))
            fnd.append(e[1].elapsed_time(e[2]))
            kicks.append(st["last_kicks"])
        log(f"wide step {s_ + 1}: insert {ins[-1] if s_ >= args.warmup else 0:.2f} ms, "
            f"find {fnd[-1] if s_ >= args.warmup else 0:.2f} ms")
    hits = int(out_h.sum())
    ok = hits == m_hit and bool((out_v[out_h.bool()] == look[out_h.bool()] * golden + 12345).all()) \
        and bool((out_v[~out_h.bool()] == 0).all())
    i_ms, f_ms = float(np.median(ins)), float(np.median(fnd))
    x = float(np.mean(kicks)) / n
    # key lines per lookup: 2 per miss, 1.5 per hit (pos1 holds half the keys: the start rule sends
    # odd keys to pos2 first), as dyc_probe_count measures on the 32-bit table (1.75 at 50 % hit)
    p_ = 2.0 * (1.0 - args.hit) + 1.5 * args.hit
    b_ins = (16 + x * (128 + 24)) * n
    b_find = (17 + 128 * p_ + 8 * args.hit) * n

    def roof(b, ms):
        gbs = b / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None}

    result = {
        "metric": "Mops/s insert & find with 64-bit keys and values",
        "value": round(2 * n / (i_ms + f_ms) / 1e3, 1), "unit": "Mops/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"2^{args.log2n} unique u64 keys and u64 values at load factor {args.theta}, then "
                               f"2^{args.log2n} finds at {args.hit:.0%} hit", "buckets": t.buckets,
                   "kick_bound": 15 if args.reference_kick else 0, "insert_order": args.insert_order},
        "insert": dict(roof(b_ins, i_ms), ms=round(i_ms, 3), mops=round(n / i_ms / 1e3, 1),
                       mean_exchanges=round(x, 4), bytes_per_op=round(b_ins / n, 1)),
        "find": dict(roof(b_find, f_ms), ms=round(f_ms, 3), mops=round(n / f_ms / 1e3, 1),
                     bytes_per_op=round(b_find / n, 1)),
        "parity": "ok" if ok else "mismatch",
    }
    t.close()
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if ok else 1


def run_serving(args, dev) -> int:
    """The serving loop: one resident table of 2^26 keys plus 256 tenant tables of 2^12-2^20 keys
    (log-uniform), requests of log-uniform size in [2^6, 2^18] keys arriving Poisson at a fixed
    offered load, 30% insert / 60% find (50% hit) / 10% delete batches, table chosen in
    proportion to its size, fill bounds [0.50, 0.90] with in-line resizes. For every offered load
    the same trace is replayed on fresh tables once per mode -- continuous batching with
    host-built ticks (mode 0), continuous batching through the device-resident loop (mode 2) and
    synchronous per-request launches (mode 1) -- and every output of each is compared with the
    trace's exact expected results. Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    loads = [float(x) for x in args.sv_loads.split(",") if x.strip()]
    modes = [int(x) for x in args.sv_modes.split(",") if x.strip()]
    mode_name = {0: "continuous", 1: "synchronous", 2: "resident"}
    log(f"serving: generating a {args.sv_requests}-request trace")
    tr = W.ServingTrace(args.sv_requests, seed=args.seed, device=dev, offered=None)
    ops = tr.ops
    mean_n = float(tr.req["n"].mean())
    log(f"serving: {ops} keys ({ops / args.sv_requests:.0f} per request), {tr.sizes.size} tables, "
        f"{int(tr.sizes.sum())} initial keys; ops mix {np.bincount(tr.req['op'], minlength=3).tolist()}")
    init = [tr.initial_keys(t) for t in range(tr.sizes.size)]
    init_v = [W.splitmix32(k) for k in init]
    ov = torch.zeros(ops, dtype=torch.int32, device=dev)
    of = torch.zeros(ops, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    budget = 16 << 30

    def fresh():
        srv = D.DyCuckooServer(tr.hints, hbm_budget=budget, pipeline_depth=args.sv_depth,
                               tick_max_keys=args.sv_tick_keys, insert_lanes=args.insert_lanes,
                               find_lanes=args.find_lanes, kick_bound=args.sv_kick, poll_backoff=args.sv_backoff,
                               one_launch=args.sv_one_launch, resident_blocks=args.sv_resident_blocks,
                               item_keys=args.sv_item_keys)
        for t in range(tr.sizes.size):
            srv.insert(t, init[t], init_v[t])
        srv.drain()
        return srv

    # warm-up: first launches of every kernel, both modes, on a throwaway server
    warm = W.ServingTrace(64, seed=args.seed + 99, device=dev, offered=None, resident=1 << 12, tenants=4,
                          tenant_lo=256, tenant_hi=1024, batch_lo=16, batch_hi=1024)
    wv = torch.zeros(warm.ops, dtype=torch.int32, device=dev)
    wf = torch.zeros(warm.ops, dtype=torch.uint8, device=dev)
    for mode in modes:
        ws = D.DyCuckooServer(warm.hints, hbm_budget=64 << 20, resident_blocks=args.sv_resident_blocks)
        ws.replay(warm.requests(wv, wf), warm.arrivals, mode=mode)
        ws.close()

    runs = []
    all_ok = True
    for load in loads:
        rng = np.random.default_rng(args.seed + 1000 + int(load * 1000))
        if load > 0:
            gaps = rng.exponential(mean_n / (load * 1e9), args.sv_requests)
            arrivals = np.cumsum(gaps) - gaps[0]
        else:
            arrivals = np.zeros(args.sv_requests)
        for mode in modes:
            srv = fresh()
            ov.zero_()
            of.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            done, status = srv.replay(tr.requests(ov, of), arrivals, mode=mode)
            wall = time.perf_counter() - t0
            torch.cuda.synchronize()
            ok = bool((status == 0).all()) and bool(torch.equal(of, tr.exp_flags)) and bool(torch.equal(ov, tr.exp_vals))
            all_ok &= ok
            lat_ms = (done - arrivals) * 1e3
            span = float(done.max() - arrivals[0])
            st = srv.stats()
            run = {"mode": mode_name[mode],
                   "offered_gkeys_s": load if load > 0 else None,
                   "mops": round(ops / span / 1e6, 1),
                   "p50_ms": round(pct(lat_ms, 50), 3), "p99_ms": round(pct(lat_ms, 99), 3),
                   "max_ms": round(float(lat_ms.max()), 3), "span_s": round(span, 4), "wall_s": round(wall, 4),
                   "ticks": st["ticks"] if mode == 0 else None, "rounds": st["rounds"] if mode == 0 else None,
                   "loop_launches": st["loop_launches"] if mode == 2 else None,
                   "loop_pauses": st["loop_pauses"] if mode == 2 else None,
                   "resizes_up": st["resizes_up"],
                   "resizes_down": st["resizes_down"], "deferred_memory": st["deferred_memory"],
                   "arena_peak_gb": round(st["arena_peak"] / 2**30, 3), "parity": "ok" if ok else "mismatch"}
            runs.append(run)
            log(f"load {load} G keys/s {run['mode']}: {run['mops']} Mops/s, p50 {run['p50_ms']} ms, "
                f"p99 {run['p99_ms']} ms, ticks {run['ticks']}, rounds {run['rounds']}, ups {run['resizes_up']}, "
                f"parity {run['parity']}")
            srv.close()
    sat = {r["mode"]: r for r in runs if r["offered_gkeys_s"] is None}
    head = sat.get("continuous") or runs[0]
    result = {
        "metric": "serving: Mops/s and p50/p99 batch latency, continuous batching vs synchronous per-batch launches",
        "value": head["mops"],
        "unit": "Mops/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "serving: resident table 2^26 keys + 256 tenants of 2^12-2^20 keys (log-uniform), "
                               "Poisson arrivals, batches log-uniform in [2^6, 2^18] keys, 30% insert / 60% find "
                               "(50% hit) / 10% delete, table chosen in proportion to size, fill bounds "
                               "[0.50, 0.90], in-line resizes",
                   "requests": args.sv_requests, "keys": ops, "tables": int(tr.sizes.size),
                   "initial_keys": int(tr.sizes.sum()), "pipeline_depth": args.sv_depth,
                   "tick_max_keys": args.sv_tick_keys, "hbm_budget_gb": budget / 2**30, "seed": args.seed,
                   "kick_bound": args.sv_kick, "poll_backoff": args.sv_backoff, "one_launch": args.sv_one_launch},
        "runs": runs,
        "saturation_speedup": (round(sat["continuous"]["mops"] / sat["synchronous"]["mops"], 2)
                               if "continuous" in sat and "synchronous" in sat else None),
        "resident_saturation_speedup": (round(sat["resident"]["mops"] / sat["synchronous"]["mops"], 2)
                                        if "resident" in sat and "synchronous" in sat else None),
        "parity": "ok" if all_ok else "mismatch",
    }
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if all_ok else 1


def run_batched(args, dev) -> int:
    """Batched uneven tables: bt_tables tables with key counts log-uniform in [2^bt_log2min,
    2^bt_log2max], each sized for load factor bt_theta; one step = clear, insert every table's
    keys, then find bt_tables segments of lookups (bt_hit of each table's own keys, the rest
    absent). Measured twice with the same kernels and launch shape: one persistent launch per
    batch (the group path) and one launch per table. Prints one JSON line."""
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import autotune
    from databases_gpu_amd import workload as W

    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(args.seed)
    sizes = np.exp(rng.uniform(np.log(2.0**args.bt_log2min), np.log(2.0**args.bt_log2max), size=args.bt_tables))
    sizes = sizes.astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(offs[-1])
    log(f"batched: {args.bt_tables} tables, {total} keys (2^{math.log2(total):.2f}), load factor {args.bt_theta}")
    keys = W.keys_range(0, total, args.seed, dev)
    vals = W.splitmix32(keys)
    if args.bt_wide:  # uint64 keys: the distinct low words above, high words of a second sequence
        hi = W.keys_range(total, 2 * total, args.seed + 1, dev)
        keys = (hi.to(torch.int64) << 32) | (keys.to(torch.int64) & 0xFFFFFFFF)
        vals = keys * 0x9E3779B1 + 12345
    hit_n = np.round(args.bt_hit * sizes).astype(np.int64)
    mark = torch.zeros(total + 1, dtype=torch.int32, device=dev)
    ones = torch.ones(args.bt_tables, dtype=torch.int32, device=dev)
    mark.index_add_(0, torch.from_numpy((offs[:-1] + hit_n.astype(np.uint64)).astype(np.int64)).to(dev), ones)
    mark.index_add_(0, torch.from_numpy(offs[1:].astype(np.int64)).to(dev), -ones)
    absent = torch.cumsum(mark, 0, dtype=torch.int32)[:total].bool()
    del mark
    miss = W.keys_range(total, 2 * total, args.seed, dev)
    if args.bt_wide:
        miss = (miss.to(torch.int64) << 32) | 0x5A5A5A5A  # high words no stored key has with this low word
    lookups = torch.where(absent, miss, keys)
    out_v = torch.empty(total, dtype=keys.dtype, device=dev)
    out_h = torch.empty(total, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    hints = (np.ceil(sizes / args.bt_theta) + 64).astype(np.uint64)
    order = args.bt_insert_order if args.bt_insert_order >= 0 else (3 if total >= (1 << 22) else 0)
    bulk = order == 3
    g = D.DyCuckooWideGroup(hints, insert_order=order) if args.bt_wide else D.DyCuckooGroup(hints, insert_order=order)
    tuned = None
    if not args.no_tune and not args.bt_wide:
        log("autotuning launch shapes on the batch")
        tuned = autotune.tune_group(g, keys, vals, offs, lookups, offs, log=None)
        log(f"autotune: {tuned['installed']} ({tuned.get('trials', 0)} trials, {tuned['seconds']} s)")

    def measure(per_table: bool, g=g):
        ins, fnd, tot, wall = [], [], [], []
        for s in range(args.warmup + args.steps):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            t0 = time.perf_counter()
            e[0].record(stream)
            g.clear()
            e[1].record(stream)
            g.insert(keys, vals, offs, per_table=per_table)
            e[2].record(stream)
            g.find(lookups, offs, out_v, out_h, per_table=per_table)
            e[3].record(stream)
            e[3].synchronize()
            if s >= args.warmup:
                ins.append(e[1].elapsed_time(e[2]))
                fnd.append(e[2].elapsed_time(e[3]))
                tot.append(e[1].elapsed_time(e[3]))
                wall.append((time.perf_counter() - t0) * 1e3)
        return ins, fnd, tot, wall

    ins_p, fnd_p, tot_p, wall_p = measure(False)
    filled = g.table_stats(0)["last_filled"]
    ok = bool((out_h.bool() == ~absent).all()) and bool((out_v[~absent] == vals[~absent]).all())
    evict = None
    if bulk:  # the eviction insert for every key, with launch shapes tuned for it
        if args.bt_wide:
            g0 = D.DyCuckooWideGroup(hints, insert_order=0)
        else:
            g0 = D.DyCuckooGroup(hints, insert_order=0)
            if tuned is not None:
                autotune.tune_group(g0, keys, vals, offs, lookups, offs, log=None)
        ins_e, fnd_e, tot_e, _ = measure(False, g0)
        ok &= bool((out_h.bool() == ~absent).all()) and bool((out_v[~absent] == vals[~absent]).all())
        g0.close()
        evict = {"insert_ms": round(float(np.median(ins_e)), 3), "find_ms": round(float(np.median(fnd_e)), 3),
                 "p50_ms": round(pct(tot_e, 50), 3), "p99_ms": round(pct(tot_e, 99), 3),
                 "mops": round(2 * total / (np.median(tot_e) * 1e-3) / 1e6, 1)}
        log(f"persistent, eviction insert: insert {evict['insert_ms']:.2f} ms, find {evict['find_ms']:.2f} ms")
    ok_graph, wall_g = True, [float("nan")]
    if not args.bt_wide:
        # the same step (clear, insert, find) captured once into a HIP graph and replayed
        plan = g.plan(keys, vals, offs, lookups, offs, out_v, out_h)
        wall_g = []
        for s in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            plan.run()
            if s >= args.warmup:
                wall_g.append((time.perf_counter() - t0) * 1e3)
        ok_graph = bool((out_h.bool() == ~absent).all()) and bool((out_v[~absent] == vals[~absent]).all())
        plan.close()
        log(f"step wall time: eager calls {np.median(wall_p):.3f} ms, graph replay {np.median(wall_g):.3f} ms")
    ok &= bool((out_h.bool() == ~absent).all()) and bool((out_v[~absent] == vals[~absent]).all())
    log(f"persistent: insert {np.median(ins_p):.2f} ms, find {np.median(fnd_p):.2f} ms")
    ins_t, fnd_t, tot_t, _ = measure(True)
    ok &= bool((out_h.bool() == ~absent).all()) and bool((out_v[~absent] == vals[~absent]).all())
    log(f"per table:  insert {np.median(ins_t):.2f} ms, find {np.median(fnd_t):.2f} ms")
    counts_ok = all(g.table_stats(i)["count"] == sizes[i] for i in range(0, args.bt_tables, max(1, args.bt_tables // 64)))
    ok &= counts_ok and ok_graph
    ops = 2 * total
    result = {
        "metric": "Mops/s batched uneven tables (insert all, then find all), persistent launch vs one launch per table",
        "value": round(ops / (np.median(tot_p) * 1e-3) / 1e6, 1),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.median(tot_p)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if args.bt_wide else "u32",
        "data": "synthetic",
        "config": {"workload": f"{args.bt_tables} tables, key counts log-uniform in [2^{args.bt_log2min}, "
                               f"2^{args.bt_log2max}], load factor {args.bt_theta}, finds {args.bt_hit:.0%} hit"
                               + (", uint64 keys and values" if args.bt_wide else ""),
                   "total_keys": total, "tables": args.bt_tables,
                   "launch": None if args.bt_wide else g.launch},
        "persistent": {"insert_ms": round(float(np.median(ins_p)), 3), "find_ms": round(float(np.median(fnd_p)), 3),
                       "p50_ms": round(pct(tot_p, 50), 3), "p99_ms": round(pct(tot_p, 99), 3),
                       "mops": round(ops / (np.median(tot_p) * 1e-3) / 1e6, 1),
                       "insert": "bulk fill per table chunk (insert_order 3), eviction insert of the overflow"
                                 if bulk else "eviction insert", "filled_by_bulk": filled},
        "persistent_eviction_insert": evict,
        "step_wall_ms": {"eager": round(float(np.median(wall_p)), 3),
                         "graph": None if args.bt_wide else round(float(np.median(wall_g)), 3),
                         "note": "clear + insert + find incl. host work and the final wait; graph = one replay of "
                                 "the captured step (dyc_group_plan_run)"},
        "per_table": {"insert_ms": round(float(np.median(ins_t)), 3), "find_ms": round(float(np.median(fnd_t)), 3),
                      "p50_ms": round(pct(tot_t, 50), 3), "p99_ms": round(pct(tot_t, 99), 3),
                      "mops": round(ops / (np.median(tot_t) * 1e-3) / 1e6, 1)},
        "speedup": round(float(np.median(tot_t) / np.median(tot_p)), 3),
        "speedup_same_insert": round(float(np.median(tot_t) / evict["p50_ms"]), 3) if evict else None,
        "autotune": tuned,
        "parity": "ok" if ok else "mismatch",
    }
    g.close()
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if ok else 1


def main() -> int:
    args = parse()
    if args.traffic_csv:
        cfg = {"insert_order": args.insert_order, "insert_lanes": args.insert_lanes,
               "insert_blocks": args.insert_blocks, "kick_bound": args.kick_bound, "find_lanes": args.find_lanes}
        out = traffic_from_csv(args.traffic_csv, args.log2n, cfg)
        text = json.dumps(out, indent=1)
        if args.json_out:
            with open(args.json_out, "w") as f:
                f.write(text + "\n")
        print(text)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the table does not shard "
              "across GPUs in this build)", file=sys.stderr)
        return 2
    import torch

    import databases_gpu_amd as D
    from databases_gpu_amd import workload as W

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    D.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    if args.workload == "dynamic":
        return run_dynamic(args, dev)
    if args.workload == "batched":
        return run_batched(args, dev)
    if args.workload == "serving":
        return run_serving(args, dev)
    if args.workload == "resize":
        return run_resize(args, dev)
    if args.workload == "memsys":
        return run_memsys(args, dev)
    if args.workload == "host":
        return run_host(args, dev)
    if args.workload == "subtables":
        return run_subtables(args, dev)
    if args.workload == "wide":
        return run_wide(args, dev)
    n = 1 << args.log2n
    hint = int(math.ceil(n / args.theta))
    log(f"generating inputs: n=2^{args.log2n}, theta={args.theta}")
    keys = W.keys_range(0, n, args.seed, dev)
    vals = W.splitmix32(keys)
    look_u = W.mixed_lookups(keys, n, args.hit, args.seed, n, args.seed + 7)
    look_z = W.zipf_lookups(keys, n, args.zipf, args.seed + 11)
    out_v = torch.empty(n, dtype=torch.int32, device=dev)
    out_h = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        torch.neg(keys, out=out_v)  # one elementwise kernel: reads 4n bytes, writes 4n bytes
        torch.cuda.synchronize()
        log(f"calibration copy kernel: {4 * n} bytes read, {4 * n} bytes written")

    extra = {} if args.kick_bound is None else {"kick_bound": args.kick_bound}
    t = D.DyCuckoo(hint, find_lanes=args.find_lanes, insert_lanes=args.insert_lanes,
                   insert_blocks=args.insert_blocks, max_blocks=args.max_blocks, subtables=args.subtables,
                   insert_order=args.insert_order, **extra)
    log(f"table: buckets {t.buckets}, capacity {t.capacity}, fill {n / t.capacity:.4f}")
    tuned = None
    if not args.no_tune:
        from databases_gpu_amd.autotune import tune_table

        tuned = tune_table(t, keys, vals, (look_u, look_z), log=log,
                           kick_bounds=(15,) if args.reference_kick else (15, 0))
        log(f"autotune installed {tuned['installed']} ({tuned['seconds']} s)")
    cfg = t.config()

    phases = ("clear", "insert", "find_uniform", "find_zipf")
    times = {p: [] for p in phases}
    step_ms = []
    kicks = []

    def step(record: bool):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(phases) + 1)]
        ev[0].record(stream)
        t.clear()
        ev[1].record(stream)
        t.insert(keys, vals)
        ev[2].record(stream)
        t.find(look_u, out_v, out_h)
        ev[3].record(stream)
        t.find(look_z, out_v, out_h)
        ev[4].record(stream)
        ev[4].synchronize()
        st = t.stats()
        if st["pending"] or st["lost"] or st["count"] != n:
            raise RuntimeError(f"insert left the table inconsistent: {st}")
        if record:
            for i, p in enumerate(phases):
                times[p].append(ev[i].elapsed_time(ev[i + 1]))
            step_ms.append(ev[0].elapsed_time(ev[4]))
            kicks.append(st["last_kicks"])

    for w in range(args.warmup):
        step(False)
        log(f"warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for s in range(args.steps):
        step(True)
        log(f"step {s + 1}/{args.steps}: {step_ms[-1]:.2f} ms "
            f"(insert {times['insert'][-1]:.2f}, find {times['find_uniform'][-1]:.2f}, "
            f"zipf {times['find_zipf'][-1]:.2f})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0

    # ---- parity of the last step's outputs against the generator (not timed)
    v, h = t.find(look_u, out_v, out_h)
    present = torch.isin(look_u, keys)
    parity_ok = bool((h.bool() == present).all()) and bool((v[present] == W.splitmix32(look_u[present])).all()) \
        and bool((v[~present] == 0).all())
    v, h = t.find(look_z, out_v, out_h)
    parity_ok &= bool(h.bool().all()) and bool((v == W.splitmix32(look_z)).all())
    log(f"parity vs generator: {'ok' if parity_ok else 'MISMATCH'}")

    # ---- algorithmic bytes
    probes_u, pos1_u = t.probe_count(look_u)
    probes_z, pos1_z = t.probe_count(look_z)
    p_u, p_z = probes_u / n, probes_z / n
    i_bar = float(np.mean(kicks)) / n
    bytes_find_u = (4 + 4 + 1 + 128 * p_u) * n
    bytes_find_z = (4 + 4 + 1 + 128 * p_z) * n
    bytes_insert = (8 + i_bar * 136) * n

    def med(p):
        return float(np.median(times[p]))

    def roof(bytes_, ms, traffic=None):
        gbs = bytes_ / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic}

    kernels = {
        "insert": dict(roof(bytes_insert, med("insert")), ms=round(med("insert"), 3),
                       mops=round(n / med("insert") / 1e3, 1), bytes_per_op=round(bytes_insert / n, 2),
                       mean_exchanges=round(i_bar, 4)),
        "find_uniform": dict(roof(bytes_find_u, med("find_uniform")), ms=round(med("find_uniform"), 3),
                             mops=round(n / med("find_uniform") / 1e3, 1), bytes_per_op=round(bytes_find_u / n, 2),
                             mean_probes=round(p_u, 5)),
        "find_zipf": dict(roof(bytes_find_z, med("find_zipf")), ms=round(med("find_zipf"), 3),
                          mops=round(n / med("find_zipf") / 1e3, 1), bytes_per_op=round(bytes_find_z / n, 2),
                          mean_probes=round(p_z, 5)),
    }
    # Fields below that this run cannot measure come from committed files, and say so: HBM bytes
    # per launch need a rocprofv3 --pmc run of their own (profiles/traffic.json, bench.py
    # --traffic-csv), and the device's random-access ceilings a memsys run (profiles/memsys.json).
    traffic_path = os.path.join(ROOT, "profiles", "traffic.json")
    tr = {}
    if os.path.exists(traffic_path):
        try:
            tr = json.load(open(traffic_path))
            tcfg = tr.get("config", {})
            # a measurement applies to the configuration it was taken in (the grid does not change the
            # bytes per key; the eviction bound and the insert order do, and so do the eviction
            # insert's lanes when that insert places every key; under the bulk fill it places only
            # the few keys both fill rounds leave, so its shape is named, not required to match)
            same = {"insert": ("kick_bound", "insert_order") + (() if cfg["insert_order"] == 3 else ("insert_lanes",)),
                    "find_uniform": ("find_lanes",), "find_zipf": ("find_lanes",)}
            for k in kernels:
                if k in tr and tr[k].get("log2n") == args.log2n and all(tcfg.get(f) == cfg[f] for f in same[k]):
                    kernels[k]["traffic"] = tr[k]["bytes_per_launch"]
                    src = "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; same configuration, " \
                          "stored, not measured in this run)"
                    if k == "insert" and tcfg.get("insert_lanes") != cfg["insert_lanes"]:
                        src = src[:-1] + "; its eviction stage ran %s lanes per key, this run's %s)" % (
                            tcfg.get("insert_lanes"), cfg["insert_lanes"])
                    kernels[k]["traffic_source"] = src
                    if "l2_hit_rate" in tr[k]:
                        kernels[k]["l2_hit_rate"] = tr[k]["l2_hit_rate"]
        except (OSError, ValueError, KeyError):
            tr = {}
    if cfg["insert_order"] == 3:
        # the bulk fill is a sequence of launches (pack, partition passes, fills, eviction insert of
        # the rest): `achieved`/`frac` keep the reference algorithm's bytes per key (8 + i * 136),
        # and the measured HBM rate of all its launches sits beside them
        ins = kernels["insert"]
        ins["launches"] = "bulk fill: pack, partition by chunk (2 digit passes), chunk fill, second round, " \
                          "eviction insert of the rest"
        ins["table_state"] = "cleared before each insert: the library knows the table holds zeros, so the " \
                             "dense fill does not read it (a table holding entries is read: about 0.25 ms more)"
        if ins["traffic"]:
            hbm = ins["traffic"] / (med("insert") * 1e-3) / 1e9
            ins["hbm_achieved"] = round(hbm, 1)
            ins["hbm_frac"] = round(hbm / HBM_PEAK_GBS, 4)
            ins["stages_bytes_per_key"] = {k: v["bytes_per_key"] for k, v in tr["insert"].get("stages", {}).items()}
    # Zipf lookups hit the same hot buckets over and over; L2 and the MALL serve those, so the
    # algorithmic bytes exceed what HBM moves and would read above the HBM peak. The kernel is
    # reported against the bytes HBM measured (bound "cache"), the algorithmic rate beside.
    zf = kernels["find_zipf"]
    zf["algorithmic_gbs"] = zf["achieved"]
    zf["bound"] = "cache"
    if zf["traffic"]:
        hbm = zf["traffic"] / (med("find_zipf") * 1e-3) / 1e9
        zf["achieved"], zf["frac"] = round(hbm, 1), round(hbm / HBM_PEAK_GBS, 4)
    else:
        zf["achieved"], zf["frac"] = None, None
    # the device's random-access ceilings at this table size (bench.py --workload memsys)
    memsys_path = os.path.join(ROOT, "profiles", "memsys.json")
    if os.path.exists(memsys_path):
        try:
            ms_ = json.load(open(memsys_path))
            ex = i_bar * n / (med("insert") * 1e-3)
            rc_ = ms_.get("read_then_cas_per_s", ms_["cas_per_s"])
            src = "profiles/memsys.json (bench.py --workload memsys; stored, not measured in this run)"
            if cfg["insert_order"] != 3:  # the eviction insert: one bucket read and one CAS per exchange
                kernels["insert"]["ceiling"] = {"kind": "random bucket read then 64-bit CAS per s", "rate": round(rc_ / 1e9, 2),
                                                "achieved": round(ex / 1e9, 2), "frac": round(ex / rc_, 3),
                                                "unit": "G/s", "source": src}
            rd = p_u * n / (med("find_uniform") * 1e-3)
            kernels["find_uniform"]["ceiling"] = {"kind": "random 128-B bucket reads/s",
                                                  "rate": round(ms_["bucket_reads_per_s"] / 1e9, 2),
                                                  "achieved": round(rd / 1e9, 2),
                                                  "frac": round(rd / ms_["bucket_reads_per_s"], 3), "unit": "G/s",
                                                  "source": src}
        except (OSError, ValueError, KeyError):
            pass
    # the roofline names one kernel, whose average launch time a kernel trace can confirm; the
    # bulk-fill insert (insert_order 3) is a sequence of launches and the Zipf finds are bound by
    # the caches, so with the bulk fill the roofline is the uniform finds'
    single = ("find_uniform",) if cfg["insert_order"] == 3 else ("insert", "find_uniform")
    dominant = max(single, key=med)
    roofline = {k: kernels[dominant][k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")}
    roofline["kernel"] = dominant
    if kernels[dominant].get("traffic_source"):
        roofline["traffic_source"] = kernels[dominant]["traffic_source"]

    ms_step = float(np.median(step_ms))
    value = 3 * n / (ms_step * 1e-3) / 1e6
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"large bulk: {args.subtables} subtables, 2^{args.log2n} unique u32 keys, load factor {args.theta}, "
                        f"bulk insert + 2^{args.log2n} finds at {args.hit:.0%} hit (uniform) + 2^{args.log2n} "
                        f"finds Zipf({args.zipf})",
            "keys": n, "theta": args.theta, "table_fill": round(n / t.capacity, 5), "buckets": t.buckets,
            "table_bytes": t.capacity * 8, "find_lanes": cfg["find_lanes"], "insert_lanes": cfg["insert_lanes"],
            "insert_blocks": cfg["insert_blocks"], "max_blocks": cfg["max_blocks"],
            "kick_bound": cfg["kick_bound"], "insert_order": cfg["insert_order"], "autotuned": tuned is not None,
            "seed": args.seed,
        },
        "roofline": roofline,
        "kernels": kernels,
        "p50_ms": round(pct(step_ms, 50), 3),
        "p99_ms": round(pct(step_ms, 99), 3),
        "phase_p99_ms": {p: round(pct(times[p], 99), 3) for p in phases},
        "wall_s": round(wall, 3),
        "parity": "ok" if parity_ok else "mismatch",
    }
    t.close()
    del keys, vals, look_u, look_z, out_v, out_h
    torch.cuda.empty_cache()
    if not args.no_cpu:
        log(f"cpu baseline: multithreaded on 2^{args.cpu_log2n} keys, sequential on 2^{args.cpu_seq_log2n}")
        result["cpu_baseline"] = cpu_baseline(args.cpu_log2n, args.theta, args.hit, args.zipf, args.seed,
                                              args.cpu_seq_log2n)
    else:
        result["cpu_baseline"] = None
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0 if parity_ok else 1


if __name__ == "__main__":
    sys.exit(main())
