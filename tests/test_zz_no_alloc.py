"""dyc_insert_noalloc -- the insert that allocates nothing (its counter block is a library-static
device array) -- called through the C ABI and compared with dyc_insert, the entry point it is a
variant of, and with the oracle: stats, the key->value multiset of the dumped table, the
placement invariant, and find values and hit flags, bit-exact, over two calls in a row (the
second smaller), on two tables of one device (so the second table's calls find the static block
as the first table's calls left it). Also: a call beyond DYC_NOALLOC_MAX_KEYS is refused with the
table unchanged, and entries that exhaust the eviction bound in an overfull table stay in the
stash, where finds see them.
"""
import ctypes

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


def u32(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def dev_u32(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def unique_keys(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**32 - 1, size=int(n * 1.3) + 64, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(out)
    return out[:n]


def vals_of(k):
    from databases_gpu_amd.workload import splitmix32_np

    return splitmix32_np(k)


def noalloc(table, keys: torch.Tensor, vals: torch.Tensor) -> int:
    from databases_gpu_amd import _lib as L

    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return L.load().dyc_insert_noalloc(table.h, ctypes.c_void_p(keys.data_ptr()), ctypes.c_void_p(vals.data_ptr()),
                                       keys.numel(), stream)


def multiset(table) -> np.ndarray:
    img = table.dump()
    g = O.OracleTable(nb=[x.size // 16 for x in img])
    g.sub = [np.ascontiguousarray(x).copy() for x in img]
    assert g.check_invariant() == 0
    allv = np.concatenate([g.entries(), table.pending()])
    return np.sort(allv[allv != 0])


def test_noalloc_insert_equals_insert_on_two_calls(D):
    from databases_gpu_amd import _lib as L

    banner("dyc_insert_noalloc vs dyc_insert: 65536 keys, then 1000")
    pool = unique_keys(70000, 31)
    k1 = pool[:L.NOALLOC_MAX_KEYS].copy()
    k1[[5, 77]] = [0, 0xFFFFFFFF]  # reserved keys are rejected and counted
    k2 = pool[L.NOALLOC_MAX_KEYS:L.NOALLOC_MAX_KEYS + 1000]
    absent = unique_keys(5000, 32)
    absent = absent[~np.isin(absent, pool)][:2000]
    hint = int(np.ceil(70000 / 0.85))
    ref = O.OracleTable(hint, rescue_kicks=128)
    tables = [D.DyCuckoo(hint) for _ in range(3)]  # [dyc_insert, noalloc, noalloc on a second table]
    try:
        for step, keys in enumerate((k1, k2)):
            kd, vd = dev_u32(keys), dev_u32(vals_of(keys))
            tables[0].insert(kd, vd)
            for t in tables[1:]:
                rc = noalloc(t, kd, vd)
                assert rc == L.DYC_OK, L.load().dyc_last_error()
            ref.insert(keys, vals_of(keys))
            sts = [t.stats() for t in tables]
            for st in sts:
                assert st["count"] == ref.count and st["lost"] == 0, (step, st)
                assert st["buckets"] == sts[0]["buckets"]
            assert sts[1]["rejected"] == sts[0]["rejected"] == 2
            want = ref.entries()
            for t in tables:
                assert np.array_equal(multiset(t), want), f"call {step + 1}: multiset"
            q = np.concatenate([pool[:L.NOALLOC_MAX_KEYS + 1000], absent, np.array([0, 0xFFFFFFFF], np.uint32)])
            rv, rh = ref.find(q)
            for t in tables:
                v, h = t.find(dev_u32(q))
                torch.cuda.synchronize()
                assert np.array_equal(h.cpu().numpy(), rh) and np.array_equal(u32(v), rv), f"call {step + 1}: find"
    finally:
        for t in tables:
            t.close()


def test_noalloc_refuses_beyond_its_maximum(D):
    from databases_gpu_amd import _lib as L

    banner("dyc_insert_noalloc refuses 65537 keys")
    keys = unique_keys(L.NOALLOC_MAX_KEYS + 1, 33)
    t = D.DyCuckoo(1 << 18)
    try:
        rc = noalloc(t, dev_u32(keys), dev_u32(vals_of(keys)))
        assert rc == 1  # DYC_ERR_INVALID
        assert b"at most" in L.load().dyc_last_error()
        assert t.stats()["count"] == 0 and t.count_occupied() == 0
        rc = noalloc(t, dev_u32(keys[:100]), dev_u32(vals_of(keys[:100])))
        assert rc == L.DYC_OK and t.stats()["count"] == 100
    finally:
        t.close()


def test_noalloc_overfull_table_keeps_entries_in_the_stash(D):
    from databases_gpu_amd import _lib as L

    banner("dyc_insert_noalloc into an overfull table: 3000 keys, 1152 slots")
    keys = unique_keys(3000, 34)
    vals = vals_of(keys)
    t = D.DyCuckoo(1024)
    try:
        assert t.capacity < keys.size
        rc = noalloc(t, dev_u32(keys), dev_u32(vals))
        assert rc == L.DYC_OK, L.load().dyc_last_error()
        st = t.stats()
        assert st["count"] == keys.size and st["pending"] >= keys.size - t.capacity and st["lost"] == 0
        assert st["resizes_up"] == 0 and st["capacity"] == t.capacity
        assert np.array_equal(multiset(t), np.sort(O.pack(keys, vals)))
        v, h = t.find(dev_u32(keys))
        torch.cuda.synchronize()
        assert h.cpu().numpy().all() and np.array_equal(u32(v), vals)
    finally:
        t.close()
