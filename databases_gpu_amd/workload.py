"""Seeded synthetic inputs for the DyCuckoo workloads (keys, values, lookups), built on the device.

* Keys: ``key(i) = mix(i)`` where ``mix`` is a seeded bijection of the 32-bit words (an
  xor-multiply-xorshift chain, every step invertible), so ``key(0..n-1)`` are distinct and
  ``key(n..2n-1)`` are distinct from them -- absent lookups are disjoint by construction. The two
  reserved words (0 and 0xFFFFFFFF) are remapped to spare indices past the ones handed out.
* Values: ``splitmix32(key)`` (a 32-bit splitmix finaliser), so about half the values are
  >= 2^31 and 0 is possible.
* Zipf(s) lookups: ranks drawn by inverting the CDF, exact over the first 2^20 ranks and from the
  Euler-Maclaurin integral beyond, then mapped to inserted keys.
Everything is plain torch int64 arithmetic on the device; results are int32 tensors whose bits
are the uint32 values.
"""
from __future__ import annotations

import math

import numpy as np
import torch

M32 = 0xFFFFFFFF


def _mulmod32(x: torch.Tensor, c: int) -> torch.Tensor:
    """(x * c) mod 2^32 for x < 2^32 without leaving int64: split c's top bit off."""
    lo = c & 0x7FFFFFFF
    y = x * lo
    if c & 0x80000000:
        y = y + ((x << 31) & M32)
    return y & M32


def _mix32(x: torch.Tensor, seed: int) -> torch.Tensor:
    """Seeded bijection of uint32 words held in int64 tensors."""
    s1 = (seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & M32
    s2 = ((seed ^ 0x5851F42D) * 0x2545F491 + 0x1B873593) & M32
    x = (x ^ s1) & M32
    x = _mulmod32(x, 0x7FEB352D)
    x = x ^ (x >> 15)
    x = (x + s2) & M32
    x = _mulmod32(x, 0x846CA68B)
    x = x ^ (x >> 16)
    x = _mulmod32(x, 0x94D049BB)
    x = x ^ (x >> 13)
    return x


def _to_i32(x: torch.Tensor) -> torch.Tensor:
    x = x & M32
    return torch.where(x >= 2**31, x - 2**32, x).to(torch.int32)


def u32_to_i64(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.int64) & M32


SPARE_INDEX = 0xFFFFFFF0  # indexes at or above this are never handed out by keys_range


def _spare_keys(seed: int) -> tuple[int, int]:
    """Replacements for the two indexes whose mix is a reserved word (mix is a bijection, so
    exactly one index maps to 0 and one to 0xFFFFFFFF): mix of the top spare indexes."""
    out = []
    j = M32
    while len(out) < 2:
        v = int(_mix32(torch.tensor([j], dtype=torch.int64), seed)[0])
        if v not in (0, M32):
            out.append(v)
        j -= 1
        assert j >= SPARE_INDEX
    return out[0], out[1]


def keys_range(lo: int, hi: int, seed: int, device) -> torch.Tensor:
    """key(i) for i in [lo, hi) (hi <= 0xFFFFFFF0): distinct nonreserved uint32 words (as int32
    bits). Different ranges of one seed never share a key."""
    if hi > SPARE_INDEX:
        raise ValueError("key indexes must stay below 0xFFFFFFF0")
    i = torch.arange(lo, hi, dtype=torch.int64, device=device)
    k = _mix32(i, seed)
    bad = (k == 0) | (k == M32)
    if bool(bad.any()):
        r0, r1 = _spare_keys(seed)
        k = torch.where(k == 0, torch.full_like(k, r0), k)
        k = torch.where(k == M32, torch.full_like(k, r1), k)
    return _to_i32(k)


def splitmix32(keys: torch.Tensor) -> torch.Tensor:
    """values of the workloads: splitmix32(key)"""
    z = u32_to_i64(keys)
    z = (z + 0x9E3779B9) & M32
    z = _mulmod32(z ^ (z >> 16), 0x85EBCA6B)
    z = _mulmod32(z ^ (z >> 13), 0xC2B2AE35)
    z = z ^ (z >> 16)
    return _to_i32(z)


def splitmix32_np(keys: np.ndarray) -> np.ndarray:
    z = (keys.astype(np.uint64) + np.uint64(0x9E3779B9)) & np.uint64(M32)
    z = ((z ^ (z >> np.uint64(16))) * np.uint64(0x85EBCA6B)) & np.uint64(M32)
    z = ((z ^ (z >> np.uint64(13))) * np.uint64(0xC2B2AE35)) & np.uint64(M32)
    return (z ^ (z >> np.uint64(16))).astype(np.uint32)


def shuffle(x: torch.Tensor, seed: int) -> torch.Tensor:
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    return x[torch.randperm(x.numel(), generator=g, device=x.device)]


def mixed_lookups(present: torch.Tensor, n: int, hit_rate: float, key_seed: int, absent_lo: int,
                  seed: int) -> torch.Tensor:
    """n lookups: round(hit_rate*n) drawn without replacement from `present` (keys_range(0, m,
    key_seed)), the rest the absent keys keys_range(absent_lo, ..., key_seed) -- the same
    bijection at indexes past the inserted ones, hence disjoint -- shuffled together with `seed`."""
    dev = present.device
    nh = int(round(hit_rate * n))
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if nh <= present.numel():
        idx = torch.randperm(present.numel(), generator=g, device=dev)[:nh]
    else:
        idx = torch.randint(0, present.numel(), (nh,), generator=g, device=dev)
    hits = present[idx]
    miss = keys_range(absent_lo, absent_lo + (n - nh), key_seed, dev)
    return shuffle(torch.cat([hits, miss]), seed + 1)


def zipf_ranks(n_items: int, n_samples: int, s: float, seed: int, device) -> torch.Tensor:
    """Zipf(s) ranks in [0, n_items) by inverse CDF (rank 0 most popular)."""
    head = min(n_items, 1 << 20)
    r = torch.arange(1, head + 1, dtype=torch.float64, device=device)
    w = r.pow(-s)
    cdf_head = torch.cumsum(w, 0)
    h_head = float(cdf_head[-1])

    def tail_mass(k: float) -> float:  # sum_{j=head+1}^{k} j^-s by Euler-Maclaurin
        a = head + 0.5
        b = k + 0.5
        if abs(1 - s) < 1e-12:
            return math.log(b / a)
        return (b ** (1 - s) - a ** (1 - s)) / (1 - s)

    total = h_head + (tail_mass(n_items) if n_items > head else 0.0)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    u = torch.rand(n_samples, generator=g, dtype=torch.float64, device=device) * total
    in_head = u < h_head
    ranks = torch.searchsorted(cdf_head, u.clamp(max=h_head * (1 - 1e-15)))
    if n_items > head:
        # invert the tail integral: b^(1-s) = a^(1-s) + (u - h_head)(1-s)
        a = head + 0.5
        t = (u - h_head).clamp(min=0)
        if abs(1 - s) < 1e-12:
            b = a * torch.exp(t)
        else:
            b = (a ** (1 - s) + t * (1 - s)).clamp(min=1e-300).pow(1.0 / (1 - s))
        tail_rank = torch.floor(b - 0.5).clamp(head, n_items - 1).to(torch.int64)
        ranks = torch.where(in_head, ranks, tail_rank)
    return ranks.clamp(0, n_items - 1)


def zipf_lookups(present: torch.Tensor, n: int, s: float, seed: int) -> torch.Tensor:
    ranks = zipf_ranks(present.numel(), n, s, seed, present.device)
    return present[ranks]


class DynamicBatch:
    """One batch of the dynamic stream: inserts, then finds, then deletes (in that order)."""

    __slots__ = ("ins_keys", "ins_vals", "find_keys", "del_keys", "live_after")

    def __init__(self, ins_keys, ins_vals, find_keys, del_keys, live_after):
        self.ins_keys, self.ins_vals = ins_keys, ins_vals
        self.find_keys, self.del_keys = find_keys, del_keys
        self.live_after = live_after

    @property
    def ops(self) -> int:
        return self.ins_keys.numel() + self.find_keys.numel() + self.del_keys.numel()


def dynamic_stream(initial: int, batch_ops: int, grow_batches: int, shrink_batches: int, seed: int, device,
                   grow_mix=(0.4, 0.4, 0.2), shrink_mix=(0.2, 0.4, 0.4), find_hit=0.7):
    """The dynamic resize workload, pre-generated on the device.

    `initial` keys are live before the first batch (returned separately). Each growth batch of
    `batch_ops` operations is grow_mix = (insert, find, delete) shares: fresh keys inserted, finds
    at `find_hit` over the keys live after the batch's inserts (the rest absent keys), deletes of
    distinct keys live after the inserts. Shrink batches swap the insert and delete shares. Keys
    are keys_range(i, ..., seed) with i handed out in order, so every insert is fresh; absent
    lookups come from indexes far beyond any insert. Returns (initial_keys, initial_vals, batches).
    """
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    next_key = 0
    absent_base = 1 << 31
    absent_next = 0
    live = keys_range(0, initial, seed, device)
    next_key = initial
    init_keys, init_vals = live.clone(), splitmix32(live)
    batches = []
    for b in range(grow_batches + shrink_batches):
        mix = grow_mix if b < grow_batches else shrink_mix
        n_ins = int(round(mix[0] * batch_ops))
        n_find = int(round(mix[1] * batch_ops))
        n_del = batch_ops - n_ins - n_find
        ins = keys_range(next_key, next_key + n_ins, seed, device)
        next_key += n_ins
        live = torch.cat([live, ins])
        n_hit = int(round(find_hit * n_find))
        hit_idx = torch.randint(0, live.numel(), (n_hit,), generator=g, device=device)
        miss = keys_range(absent_base + absent_next, absent_base + absent_next + (n_find - n_hit), seed, device)
        absent_next += n_find - n_hit
        finds = torch.cat([live[hit_idx], miss])
        finds = finds[torch.randperm(finds.numel(), generator=g, device=device)]
        n_del = min(n_del, live.numel())
        perm = torch.randperm(live.numel(), generator=g, device=device)
        dels = live[perm[:n_del]]
        live = live[perm[n_del:]]
        batches.append(DynamicBatch(ins, splitmix32(ins), finds, dels, live.numel()))
    return init_keys, init_vals, batches


def keys_at(idx: torch.Tensor, seed: int) -> torch.Tensor:
    """key(i) for an int64 tensor of indexes (each < 0xFFFFFFF0): keys_range without the range."""
    k = _mix32(idx, seed)
    bad = (k == 0) | (k == M32)
    if bool(bad.any()):
        r0, r1 = _spare_keys(seed)
        k = torch.where(k == 0, torch.full_like(k, r0), k)
        k = torch.where(k == M32, torch.full_like(k, r1), k)
    return _to_i32(k)


class ServingTrace:
    """A request trace for the serving loop, with the exact result of every request.

    Tables: table 0 is the resident table, tables 1.. the tenants. Table t owns key indexes
    [base[t], base[t] + span[t]); its live keys are always the index window [lo, hi): inserts
    append fresh indexes at hi, erases remove the oldest live indexes from lo, finds draw
    round(find_hit * n) indexes uniformly from the window (with replacement) and the rest from
    an absent range no table owns. Request i targets a table chosen with probability
    proportional to the table's initial size, is an insert / find / erase with the `mix`
    probabilities (one reference batch call), and has a log-uniform size in [batch_lo, batch_hi].
    Arrivals are Poisson at `offered` keys per second (all at 0 when offered is None).

    Device tensors (int32 bits of uint32 words, uint8 flags), one payload position per key:
      keys, vals (splitmix32(key), what inserts store), exp_vals / exp_flags (a find's value
      and hit; an erase's deleted flag, always 1 here; 0 elsewhere).
    req: structured numpy array (table, op, n, off) in arrival order; arrivals: seconds.
    init: per table the initial key indexes are [base, base + size) -- see initial_keys().
    """

    def __init__(self, n_requests: int, seed: int, device, offered: float | None = None, resident: int = 1 << 26,
                 tenants: int = 256, tenant_lo: int = 1 << 12, tenant_hi: int = 1 << 20, batch_lo: int = 1 << 6,
                 batch_hi: int = 1 << 18, mix=(0.3, 0.6, 0.1), find_hit: float = 0.5, theta: float = 0.85,
                 resident_span: int = 1 << 29, tenant_span: int = 1 << 22):
        rng = np.random.default_rng(seed)
        self.seed = seed
        self.device = device
        sizes = [resident] + list(np.exp(rng.uniform(np.log(tenant_lo), np.log(tenant_hi), tenants)).astype(np.int64))
        self.sizes = np.asarray(sizes, dtype=np.int64)
        nt = self.sizes.size
        self.span = np.asarray([resident_span] + [tenant_span] * tenants, dtype=np.int64)
        self.base = np.concatenate([[0], np.cumsum(self.span)[:-1]]).astype(np.int64)
        absent_base = int(self.base[-1] + self.span[-1])
        if absent_base >= SPARE_INDEX:
            raise ValueError("table index spans exceed the 32-bit key space")
        self.hints = (np.ceil(self.sizes / theta) + 64).astype(np.uint64)
        p = self.sizes / self.sizes.sum()
        tables = rng.choice(nt, size=n_requests, p=p)
        ops = rng.choice(3, size=n_requests, p=np.asarray(mix, dtype=np.float64) / sum(mix))
        ns = np.exp(rng.uniform(np.log(batch_lo), np.log(batch_hi + 1), n_requests)).astype(np.int64)
        ns = np.clip(ns, batch_lo, batch_hi)
        lo = np.zeros(nt, dtype=np.int64)
        hi = self.sizes.copy()
        absent_next = 0
        # segments of the payload: (length, kind, a, b); kind 0: index a + j, kind 1: a + floor(u * b)
        seg_len, seg_kind, seg_a, seg_b, seg_hit = [], [], [], [], []
        req = np.zeros(n_requests, dtype=[("table", "<u4"), ("op", "<u4"), ("n", "<u8"), ("off", "<u8")])
        off = 0
        for i in range(n_requests):
            t, op, n = int(tables[i]), int(ops[i]), int(ns[i])
            live = int(hi[t] - lo[t])
            if op == 2 and live < n:  # erase at most what is live
                n = live
            if op == 0:
                if hi[t] + n > self.span[t]:
                    raise ValueError(f"table {t} outgrew its index span")
                seg_len.append(n), seg_kind.append(0), seg_a.append(self.base[t] + hi[t]), seg_b.append(0)
                seg_hit.append(0)
                hi[t] += n
            elif op == 2:
                seg_len.append(n), seg_kind.append(0), seg_a.append(self.base[t] + lo[t]), seg_b.append(0)
                seg_hit.append(1)
                lo[t] += n
            else:
                nh = int(round(find_hit * n)) if live > 0 else 0
                if nh:
                    seg_len.append(nh), seg_kind.append(1), seg_a.append(self.base[t] + lo[t]), seg_b.append(live)
                    seg_hit.append(1)
                if n - nh:
                    seg_len.append(n - nh), seg_kind.append(0), seg_a.append(absent_base + absent_next)
                    seg_b.append(0), seg_hit.append(0)
                    absent_next += n - nh
            req[i] = (t, op, n, off)
            off += n
        if absent_base + absent_next >= SPARE_INDEX:
            raise ValueError("absent key range exhausted")
        self.req = req
        self.total = off
        self.final_window = (lo, hi)
        if offered:
            gaps = rng.exponential(float(ns.mean()) / offered, n_requests)
            self.arrivals = np.cumsum(gaps) - gaps[0]
        else:
            self.arrivals = np.zeros(n_requests)
        self.offered = offered

        # payload on the device
        L = torch.as_tensor(np.asarray(seg_len, dtype=np.int64), device=device)
        total = int(L.sum()) if len(seg_len) else 0
        assert total == off
        if total:
            starts = torch.cumsum(L, 0) - L
            sid = torch.repeat_interleave(torch.arange(L.numel(), device=device), L)
            j = torch.arange(total, dtype=torch.int64, device=device) - starts[sid]
            kind = torch.as_tensor(np.asarray(seg_kind, dtype=np.int64), device=device)[sid]
            a = torch.as_tensor(np.asarray(seg_a, dtype=np.int64), device=device)[sid]
            b = torch.as_tensor(np.asarray(seg_b, dtype=np.int64), device=device)[sid]
            g = torch.Generator(device=device)
            g.manual_seed(seed + 1)
            u = torch.rand(total, generator=g, dtype=torch.float64, device=device)
            rnd = a + torch.minimum(torch.floor(u * b.to(torch.float64)).to(torch.int64), (b - 1).clamp(min=0))
            idx = torch.where(kind == 0, a + j, rnd)
            del u, rnd, j, a, b, kind
            self.keys = keys_at(idx, seed)
            del idx
            hit = torch.as_tensor(np.asarray(seg_hit, dtype=np.uint8), device=device)[sid]
            del sid
        else:
            self.keys = torch.zeros(0, dtype=torch.int32, device=device)
            hit = torch.zeros(0, dtype=torch.uint8, device=device)
        self.vals = splitmix32(self.keys)
        # expected outputs: finds -> (value, hit); erases -> deleted = 1; inserts -> 0
        op_of = torch.as_tensor(np.repeat(req["op"].astype(np.int64), req["n"].astype(np.int64)), device=device)
        self.exp_flags = torch.where(op_of == 0, torch.zeros_like(hit), hit)
        self.exp_vals = torch.where((op_of == 1) & (hit == 1), self.vals, torch.zeros_like(self.vals))

    @property
    def ops(self) -> int:
        return int(self.total)

    def initial_keys(self, t: int) -> torch.Tensor:
        i = torch.arange(int(self.base[t]), int(self.base[t] + self.sizes[t]), dtype=torch.int64, device=self.device)
        return keys_at(i, self.seed)

    def requests(self, out_vals: torch.Tensor, out_flags: torch.Tensor) -> np.ndarray:
        """The trace as dyc_request records pointing into this trace's payload and the given
        output tensors (int32 / uint8, one position per payload key)."""
        from ._lib import REQUEST_DTYPE

        r = np.zeros(self.req.size, dtype=REQUEST_DTYPE)
        off = self.req["off"].astype(np.uint64)
        op = self.req["op"]
        r["table"] = self.req["table"]
        r["op"] = op
        r["n"] = self.req["n"]
        r["keys"] = np.uint64(self.keys.data_ptr()) + off * np.uint64(4)
        r["vals"] = np.where(op == 0, np.uint64(self.vals.data_ptr()) + off * np.uint64(4), np.uint64(0))
        r["out_vals"] = np.where(op == 1, np.uint64(out_vals.data_ptr()) + off * np.uint64(4), np.uint64(0))
        r["out_flags"] = np.where(op != 0, np.uint64(out_flags.data_ptr()) + off, np.uint64(0))
        return r
