"""Inserts on several streams at once into one table (dyc_insert_async + dyc_sync): the
reference's multi-stream insert (GPU/old_version/cuckoo_gtest/api.h:27-31; HashStreamTEST.Four,
Main.cpp:72-113). Slot layout after concurrent inserts is not fixed, so each case checks the
key->value multiset of the table plus its pending list against the union of the batches, the
placement invariant of every stored entry on the oracle, counts, and find results."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def D():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import databases_gpu_amd as D

    D.load()
    return D


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def stored(t):
    img = t.dump()
    g = O.OracleTable(nb=[x.size // 16 for x in img])
    g.sub = [np.ascontiguousarray(x).copy() for x in img]
    assert g.check_invariant() == 0
    allv = np.concatenate([g.entries(), t.pending()])
    return np.sort(allv[allv != 0])


def run_streams(D, t, keys, vals, nstreams):
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    parts = np.array_split(np.arange(keys.numel()), nstreams)
    torch.cuda.synchronize()
    for s, idx in zip(streams, parts):
        lo, hi = int(idx[0]), int(idx[-1]) + 1
        t.insert_async(keys[lo:hi], vals[lo:hi], stream=s)
    t.sync()


def test_four_streams_of_2p20_keys_equal_the_union(D):
    from databases_gpu_amd import workload as W

    banner("4 streams x 2^20 keys in flight at once")
    dev = torch.device("cuda")
    n = 4 << 20
    keys = W.keys_range(0, n, 41, dev)
    vals = W.splitmix32(keys)
    t = D.DyCuckoo(int(np.ceil(n / 0.85)), insert_lanes=2, kick_bound=0)
    try:
        run_streams(D, t, keys, vals, 4)
        st = t.stats()
        assert st["count"] == n and st["pending"] == 0 and st["lost"] == 0
        assert np.array_equal(stored(t), np.sort(O.pack(u32(keys), u32(vals))))
        v, h = t.find(keys)
        assert bool(h.bool().all()) and bool((v == vals).all())
        absent = W.keys_range(n, n + (1 << 20), 41, dev)
        _, h = t.find(absent)
        assert not bool(h.bool().any())
    finally:
        t.close()


def test_streams_into_an_overfull_table_replace_failures(D):
    """4 streams x 2^17 keys into a table of 2^18 slots without auto_resize: entries that exhaust
    the eviction bound wait on each call's own list, and dyc_sync re-places them (upsize and
    re-insert) as dyc_insert does."""
    from databases_gpu_amd import workload as W

    banner("4 streams x 2^17 keys into 2^18 slots")
    dev = torch.device("cuda")
    n = 4 << 17
    keys = W.keys_range(0, n, 42, dev)
    vals = W.splitmix32(keys)
    t = D.DyCuckoo(1 << 18, kick_bound=0)
    try:
        run_streams(D, t, keys, vals, 4)
        st = t.stats()
        assert st["count"] == n and st["pending"] == 0 and st["lost"] == 0 and st["resizes_up"] > 0, st
        assert np.array_equal(stored(t), np.sort(O.pack(u32(keys), u32(vals))))
        v, h = t.find(keys)
        assert bool(h.bool().all()) and bool((v == vals).all())
    finally:
        t.close()


def test_other_calls_complete_inserts_in_flight(D):
    """A find issued while inserts are in flight on other streams sees them all (every call
    completes the inserts in flight first); with auto_resize, a call that would pass the upper
    bound completes the ones in flight and upsizes before it starts."""
    from databases_gpu_amd import workload as W

    banner("find after inserts in flight; auto_resize across streams")
    dev = torch.device("cuda")
    n = 3 << 18
    keys = W.keys_range(0, n, 43, dev)
    vals = W.splitmix32(keys)
    t = D.DyCuckoo(1 << 18, auto_resize=1, min_capacity=0)
    try:
        streams = [torch.cuda.Stream() for _ in range(3)]
        torch.cuda.synchronize()
        for i, s in enumerate(streams):
            t.insert_async(keys[i << 18:(i + 1) << 18], vals[i << 18:(i + 1) << 18], stream=s)
        v, h = t.find(keys)
        torch.cuda.synchronize()
        assert bool(h.bool().all()) and bool((v == vals).all())
        st = t.stats()
        assert st["count"] == n and st["resizes_up"] > 0 and st["pending"] == 0
        assert n <= 0.9 * st["capacity"]
        assert np.array_equal(stored(t), np.sort(O.pack(u32(keys), u32(vals))))
    finally:
        t.close()
