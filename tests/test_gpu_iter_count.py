"""Per-key eviction-chain counts (dyc_insert_counted): the reference's iterator_count
(cuckoo.cu:349-351, written under ITERAOTOR_COUNT_FLAG; api.cpp:219-243 prints its maximum and mean).

A chain counts its exchanges, the final placement included, as the reference's warp_insert does
(count++ after every exchange, cuckoo.cu:427-429); a reserved key counts 0, and so does a chain
whose compare-and-swap retries ran out before its first exchange.

  single-key calls -- one key per call makes the GPU insert sequential, so every count and the
                      final image equal the oracle's sequential replay exactly
  one batch        -- counts are >= 1 except reserved keys, they sum to the call's exchanges
                      (stats last_kicks) and their maximum is stats last_max_chain; the stored set
                      equals the oracle's; the mean stays near the sequential replay's (parallel
                      chains interleave, so single counts differ)
"""
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


def unique_keys(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**32 - 1, size=int(n * 1.3) + 64, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(out)
    return out[:n]


def vals_of(k):
    from databases_gpu_amd.workload import splitmix32_np

    return splitmix32_np(k)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def test_single_key_calls_match_sequential_oracle(D):
    """1,800 one-key calls into 2,048 slots (fill 0.88): every count equals the oracle's."""
    banner("insert_counted: one key per call, fill 0.88")
    t = D.DyCuckoo(2048, auto_resize=0, rescue_kicks=128)
    nb = t.buckets
    tb = O.OracleTable(nb=nb, rescue_kicks=128)
    n = int(0.88 * sum(nb) * 16)
    keys = unique_keys(n, 71)
    vals = vals_of(keys)
    got = np.zeros(n, dtype=np.int64)
    kd, vd = dev(keys), dev(vals)
    for i in range(n):
        c = t.insert_counted(kd[i:i + 1], vd[i:i + 1])
        got[i] = int(c.item())
        st = t.stats()
        assert st["last_max_chain"] == got[i] and st["last_kicks"] == got[i]
    assert t.stats()["pending"] == 0
    tb.insert(keys, vals)
    assert tb.pending.size == 0
    want = tb.last_kicks.astype(np.int64)
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    assert want.max() > 1  # chains longer than one exchange were exercised
    img = t.dump()
    assert all(np.array_equal(a, b) for a, b in zip(img, tb.images()))
    t.close()


@pytest.mark.parametrize("lanes", [1, 8, 16])
def test_batch_counts_sum_max_and_reserved(D, lanes):
    n = 1 << 20
    banner(f"insert_counted: one batch of {n} keys at fill 0.85, {lanes} lanes per chain")
    keys = unique_keys(n, 72 + lanes)
    keys[::4099] = 0  # reserved keys
    keys[1::8191] = 0xFFFFFFFF
    vals = vals_of(keys)
    hint = int(np.ceil(n / 0.85))
    t = D.DyCuckoo(hint, auto_resize=0, insert_lanes=lanes, rescue_kicks=128)
    c = t.insert_counted(dev(keys), dev(vals)).cpu().numpy().astype(np.int64)
    st = t.stats()
    reserved = (keys == 0) | (keys == 0xFFFFFFFF)
    assert np.all(c[reserved] == 0)
    assert np.all(c[~reserved] >= 1)
    assert st["pending"] == 0
    assert int(c.sum()) == st["last_kicks"]
    assert int(c.max()) == st["last_max_chain"]
    tb = O.OracleTable(nb=t.buckets, rescue_kicks=128)
    tb.insert(keys, vals)
    assert tb.pending.size == 0
    img = t.dump()
    got = np.concatenate(img)
    assert np.array_equal(np.sort(got[got != 0]), tb.entries())
    seq_mean = tb.last_kicks[~reserved].mean()
    par_mean = c[~reserved].mean()
    assert 0.8 < par_mean / seq_mean < 1.25, (par_mean, seq_mean)
    # keys whose first bucket had room take exactly one exchange in both
    assert (c[~reserved] == 1).mean() > 0.5
    t.close()


def test_counts_of_chains_that_give_up(D):
    """A chain that exhausts its exchanges counts kick_bound + rescue_kicks (the reference counts
    MAX_CUCKOO_NUM, cuckoo.cu:394-433); its entry then waits on the pending list and is placed by
    an upsize, which is not counted again."""
    banner("insert_counted: capacity + 256 keys, kick_bound 2, rescue_kicks 3")
    t = D.DyCuckoo(64, auto_resize=0, kick_bound=2, rescue_kicks=3)
    cap = t.capacity
    keys = unique_keys(cap + 256, 75)
    vals = vals_of(keys)
    c = t.insert_counted(dev(keys), dev(vals)).cpu().numpy()
    st = t.stats()
    assert c.max() == 5 and st["last_max_chain"] == 5
    # 0: the chain parked before its first exchange (its CAS retries ran out under contention
    # for the few buckets; the reference's unconditional exchange has no such case)
    assert np.all(c >= 0) and np.all(c <= 5) and (c >= 1).mean() > 0.5
    assert st["kick_failures"] >= 256
    found = t.find(dev(keys))[1].cpu().numpy()
    assert found.all()
    t.close()
