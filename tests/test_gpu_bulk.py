"""GPU parity of the bulk-fill insert (insert_order 3): sort by bucket chunk, fill free slots
chunk by chunk, the overflow in its other bucket, the rest by the eviction insert. Slot layout
after a parallel insert is not fixed, so each case checks the key->value multiset of the
table (plus pending entries) against what was inserted, the placement invariant of every
stored entry on the oracle, counts, and find results."""
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


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def stored(t):
    """Sorted packed entries of the subtables and the pending list, and the oracle of the image."""
    images = t.dump()
    ref = O.OracleTable(nb=[x.size // 16 for x in images])
    ref.sub = [np.ascontiguousarray(x).copy() for x in images]
    ref.count = int(sum(np.count_nonzero(x) for x in images))
    ents = [x[x != 0] for x in images]
    ents.append(t.pending())  # packed entries
    return np.sort(np.concatenate(ents)), ref


def gfind(t, q):
    v, h = t.find(cu(q))
    torch.cuda.synchronize()
    return v.cpu().numpy().view(np.uint32), h.cpu().numpy()


@pytest.mark.parametrize("d", (2, 4, 6))
def test_bulk_fill_insert_multiset_invariant_and_find(D, d):
    n = 1 << 20
    banner(f"bulk fill insert 2^20 at 0.90, subtables={d}")
    keys = unique_keys(n, 1200 + d)
    vals = vals_of(keys)
    t = D.DyCuckoo(int(np.ceil(n / 0.90)), subtables=d, insert_order=3)
    t.insert(cu(keys), cu(vals))
    st = t.stats()
    assert st["count"] == n and st["lost"] == 0 and st["rejected"] == 0
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    assert np.array_equal(ents, np.sort(O.pack(keys, vals)))
    assert st["last_kicks"] >= n  # every placement counts at least once
    assert 0.95 * n < st["last_filled"] <= n  # the fill rounds placed nearly all of them
    gv, gh = gfind(t, keys)
    assert gh.all() and np.array_equal(gv, vals)
    # the first round's rule (fill_target) sends all keys to pos1 at d = 2, 2/3 at d = 4 and 0.7
    # at d = 6 (the insert's start rule: half); overflow and evictions move a few percent
    _, in_pos1 = t.probe_count(cu(keys))
    share = {2: 1.0, 4: 2 / 3, 6: 0.7}[d]
    assert in_pos1 > (share - 0.1) * n, (in_pos1, n)
    t.close()


def test_bulk_fill_into_a_partly_filled_resized_table(D):
    """Unequal subtables (one upsized), existing entries, reserved keys and duplicate copies in
    the batch: the fill takes only free slots, rejects the reserved keys, and keeps every copy."""
    banner("bulk fill into a partly filled table after an upsize")
    first = unique_keys(300_000, 1300)
    t = D.DyCuckoo(600_000, insert_order=0, auto_resize=False)
    t.insert(cu(first), cu(vals_of(first)))
    t.resize_up()
    nb = t.buckets
    assert len(set(nb)) == 2  # one subtable doubled
    t.configure(insert_order=3)
    fresh = unique_keys(700_000, 1301)
    fresh = fresh[~np.isin(fresh, first)][:250_000]
    dup = first[:5000]  # second copies of stored keys
    batch = np.concatenate([fresh, dup, np.array([0, 0xFFFFFFFF, 0], dtype=np.uint32)])
    rng = np.random.default_rng(1302)
    rng.shuffle(batch)
    assert batch.size >= 2 * sum(nb)  # large enough for the bulk fill
    bv = vals_of(batch)
    t.insert(cu(batch), cu(bv))
    st = t.stats()
    keep = (batch != 0) & (batch != 0xFFFFFFFF)
    assert st["count"] == first.size + int(keep.sum()) and st["lost"] == 0
    assert st["rejected"] == 3
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    want = np.sort(np.concatenate([O.pack(first, vals_of(first)), O.pack(batch[keep], bv[keep])]))
    assert np.array_equal(ents, want)
    gv, gh = gfind(t, fresh)
    assert gh.all() and np.array_equal(gv, vals_of(fresh))
    t.close()


def test_bulk_fill_into_fresh_filled_and_cleared_tables(D):
    """A fresh or cleared table is known to hold zeros only, and its first dense round does not
    read the table; any later batch reads it. Batch A into a fresh table, batch B on top (read
    path: A and B both stored), clear, batch B again (zeroed path: nothing of A survives)."""
    banner("bulk fill: fresh table, second batch on top, then clear and refill")
    n = 1 << 19
    keys = unique_keys(2 * n, 1500)
    a, b = keys[:n], keys[n:]
    t = D.DyCuckoo(int(np.ceil(2 * n / 0.9)), insert_order=3, auto_resize=False)
    assert n >= 2 * sum(t.buckets)  # both batches take the bulk fill
    t.insert(cu(a), cu(vals_of(a)))
    t.insert(cu(b), cu(vals_of(b)))
    st = t.stats()
    assert st["count"] == 2 * n and st["lost"] == 0
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    assert np.array_equal(ents, np.sort(O.pack(keys, vals_of(keys))))
    t.clear()
    t.insert(cu(b), cu(vals_of(b)))
    st = t.stats()
    assert st["count"] == n and st["lost"] == 0
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    assert np.array_equal(ents, np.sort(O.pack(b, vals_of(b))))
    gv, gh = gfind(t, keys)
    assert not gh[:n].any() and gh[n:].all() and np.array_equal(gv[n:], vals_of(b))
    t.close()


def test_bulk_fill_chunk_beyond_the_register_hold(D):
    """A dense-round chunk with more entries than its threads hold in registers (8192 for the
    1024-thread fill) counts them, then reads them again to place them. 9000 keys whose first
    bucket (the fill's round-1 target) lies in subtable 0 (one chunk: the subtables have fewer
    than 512 buckets) plus 2000 others: subtable 0 fills, the rest overflows into the other
    bucket of each pair."""
    import fillrule

    banner("bulk fill: one chunk of 9000 entries (beyond the register hold)")
    pool = unique_keys(60_000, 1550)
    start = fillrule.targets(pool, 4)
    keys = np.concatenate([pool[start == 0][:9000], pool[start != 0][:2000]])
    assert keys.size == 11_000
    rng = np.random.default_rng(1551)
    rng.shuffle(keys)
    vals = vals_of(keys)
    t = D.DyCuckoo(20_000, insert_order=3, auto_resize=False)
    nb = t.buckets
    assert max(nb) < 512 and keys.size >= 2 * sum(nb) and 9000 > 16 * nb[0]
    t.insert(cu(keys), cu(vals))
    st = t.stats()
    assert st["count"] == keys.size and st["lost"] == 0
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    assert np.array_equal(ents, np.sort(O.pack(keys, vals)))
    # subtable 0 is (nearly) full: a bucket may draw fewer than 16 of the 9000 start keys
    assert np.count_nonzero(t.dump()[0]) >= 0.99 * 16 * nb[0]
    gv, gh = gfind(t, keys)
    assert gh.all() and np.array_equal(gv, vals)
    t.close()


def test_bulk_fill_overflow_reaches_the_eviction_insert(D):
    """Many copies of a few keys overflow their first bucket and their other bucket, so all three
    stages run; every copy is stored or pending, and the placement invariant holds."""
    banner("bulk fill: overflow through both rounds and the eviction insert")
    base = unique_keys(3000, 1400)
    hot = base[:6]
    batch = np.concatenate([base, np.repeat(hot, 11)])  # 12 copies of each hot key
    rng = np.random.default_rng(1401)
    rng.shuffle(batch)
    t = D.DyCuckoo(8192, insert_order=3, auto_resize=False)
    assert batch.size >= 2 * sum(t.buckets)
    bv = vals_of(batch)
    t.insert(cu(batch), cu(bv))
    st = t.stats()
    assert st["count"] == batch.size and st["lost"] == 0
    ents, ref = stored(t)
    assert ref.check_invariant() == 0
    assert np.array_equal(ents, np.sort(O.pack(batch, bv)))
    gv, gh = gfind(t, base)
    assert gh.all() and np.array_equal(gv, vals_of(base))
    t.close()


def test_bulk_fill_matches_the_eviction_insert_on_the_same_batch(D):
    """The two insert paths store the same multiset and answer finds alike."""
    n = 1 << 19
    banner("bulk fill vs eviction insert, 2^19")
    keys = unique_keys(n, 1500)
    vals = vals_of(keys)
    out = []
    for order in (0, 3):
        t = D.DyCuckoo(int(np.ceil(n / 0.85)), insert_order=order)
        t.insert(cu(keys), cu(vals))
        ents, ref = stored(t)
        assert ref.check_invariant() == 0
        q = np.concatenate([keys[::3], unique_keys(50_000, 1501)])
        out.append((ents, *gfind(t, q), t.stats()["count"]))
        t.close()
    a, b = out
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] == n


def test_bulk_fill_partition_with_wide_chunk_ids(D):
    """A table of more than 2^16 chunks (4 x 8,388,618 buckets, 4.3 GB): the partition carries
    32-bit chunk ids over three digit passes. 80 M keys at 15 % fill: every key is found with its
    value, absent keys miss, and the device count of occupied slots equals the batch (n distinct
    keys found in a table holding n entries: the stored multiset is the batch)."""
    from databases_gpu_amd import workload as W

    banner("bulk fill, 65,540 chunks (32-bit chunk ids, 3 passes), 80 M keys")
    dev = torch.device("cuda")
    n = 80_000_000
    keys = W.keys_range(0, n, 21, dev)
    vals = W.splitmix32(keys)
    t = D.DyCuckoo(1 << 29, insert_order=3)
    try:
        assert sum(-(-b // 512) for b in t.buckets) > 65536
        t.insert(keys, vals)
        st = t.stats()
        assert st["count"] == n and st["pending"] == 0 and st["lost"] == 0, st
        assert t.count_occupied() == n
        v, h = t.find(keys)
        assert bool(h.bool().all()) and bool((v == vals).all())
        absent = W.keys_range(n, n + (1 << 22), 21, dev)
        v, h = t.find(absent)
        assert not bool(h.bool().any())
    finally:
        t.close()
