"""GPU parity of the wide-key table (dyc_wide_*: uint64 keys and values) against the oracle:
find bit-exact and erase byte-exact (key and value arrays) on oracle-built images, insert by
multiset + placement invariant + find, upsize byte-exact, downsize by multiset + invariant, the
stash, and the fill-bound policy over a grow / shrink stream."""
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


def keys64(n, seed, lo=1, hi=2**64 - 2):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(lo, hi, size=int(n * 1.3) + 64, dtype=np.uint64))
    rng.shuffle(out)
    return out[:n]


def absent(n, exclude, seed):
    a = keys64(2 * n + 64, seed)
    return a[~np.isin(a, exclude)][:n]


def vals_of(k):
    return (k * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0xD1B54A32D192ED03)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def gfind(t, q):
    v, h = t.find(cu(q))
    torch.cuda.synchronize()
    return v.cpu().numpy().view(np.uint64), h.cpu().numpy()


def oracle_of(t):
    ks, vs = t.dump()
    ref = O.OracleWideTable(nb=[x.size // 16 for x in ks], rescue_kicks=128)
    ref.load_images(ks, vs)
    return ref


def oracle_build(n, theta, d, seed):
    keys = keys64(n, seed)
    vals = vals_of(keys)
    tb = O.OracleWideTable(int(np.ceil(n / theta)), subtables=d, rescue_kicks=128)
    tb.insert(keys, vals)
    assert tb.pend_keys.size == 0
    return tb, keys, vals


@pytest.mark.parametrize("d", (2, 4, 6))
def test_wide_find_and_erase_on_oracle_image(D, d):
    banner(f"wide find/erase on an oracle image, subtables={d}")
    tb, keys, vals = oracle_build(100_000, 0.88, d, 700 + d)
    t = D.DyCuckooWide(1000, subtables=d)
    t.load(tb.keys, tb.vals)
    q = np.concatenate([keys[:50_000], absent(20_000, keys, 701), np.array([0, 2**64 - 1], np.uint64)])
    np.random.default_rng(d).shuffle(q)
    q = q[:-11]
    wv, wh = tb.find(q)
    gv, gh = gfind(t, q)
    assert np.array_equal(gh, wh) and np.array_equal(gv, wv)
    e = np.concatenate([keys[::3], absent(4000, keys, 702)])
    want = tb.erase(e)
    got = t.erase(cu(e)).cpu().numpy()
    assert np.array_equal(got, want)
    ks, vs = t.dump()
    assert all(np.array_equal(a, b) for a, b in zip(ks, tb.keys))
    assert all(np.array_equal(a, b) for a, b in zip(vs, tb.vals))
    assert t.stats()["count"] == tb.count
    t.close()


@pytest.mark.parametrize("d", (3, 4, 5))
def test_wide_insert_multiset_invariant_and_find(D, d):
    n = 1 << 20
    banner(f"wide insert 2^20 at 0.90, subtables={d}")
    keys = np.concatenate([keys64(n - 4096, 800 + d), keys64(4096, 810 + d, 1, 2**32 - 1)])  # some below 2^32
    keys = np.unique(keys)
    vals = vals_of(keys)
    t = D.DyCuckooWide(int(np.ceil(keys.size / 0.9)), subtables=d)
    t.insert(cu(keys), cu(vals))
    st = t.stats()
    assert st["count"] == keys.size and st["pending"] == 0 and st["lost"] == 0 and st["subtables"] == d
    ref = oracle_of(t)
    assert ref.check_invariant() == 0
    k, v = ref.entries()
    order = np.argsort(keys)
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    q = np.concatenate([keys[: keys.size // 2], absent(keys.size // 4, keys, 820 + d)])
    wv, wh = ref.find(q)
    gv, gh = gfind(t, q)
    assert np.array_equal(gh, wh) and np.array_equal(gv, wv)
    assert wh.sum() == keys.size // 2
    t.close()


@pytest.mark.parametrize("d", (2, 4, 6))
def test_wide_upsize_byte_exact_and_downsize_preserves(D, d):
    banner(f"wide upsize x{d + 1}, downsize x2, subtables={d}")
    tb, keys, vals = oracle_build(50_000, 0.85, d, 900 + d)
    t = D.DyCuckooWide(1000, subtables=d)
    t.load(tb.keys, tb.vals)
    for _ in range(d + 1):
        tb.resize_up()
        t.resize_up()
        assert t.buckets == tb.nb
        ks, vs = t.dump()
        assert all(np.array_equal(a, b) for a, b in zip(ks, tb.keys))
        assert all(np.array_equal(a, b) for a, b in zip(vs, tb.vals))
    order = np.argsort(keys)
    for _ in range(2):
        assert t.resize_down()
        tb.resize_down()
        assert t.buckets == tb.nb
        st = t.stats()
        assert st["pending"] == 0 and st["lost"] == 0
        g = oracle_of(t)
        assert g.check_invariant() == 0
        k, v = g.entries()
        assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    t.close()


def test_wide_stash_serves_keys_missing_from_both_buckets(D):
    banner("wide stash: 40 copies of one key")
    key = 0x0123456789ABCDEF
    t = D.DyCuckooWide(100, max_fail_rounds=0, rescue_kicks=16)
    vals = np.arange(1000, 1040, dtype=np.uint64)
    with pytest.raises(D.DycError) as ei:
        t.insert(cu(np.full(40, key, np.uint64)), cu(vals))
    assert ei.value.code == 4  # DYC_ERR_PENDING
    st = t.stats()
    assert st["pending"] == 8 and st["count"] == 40
    pk, pv = t.pending()
    assert np.all(pk == key) and len(set(pv.tolist())) == 8 and set(pv.tolist()) <= set(vals.tolist())
    one = cu(np.array([key], np.uint64))
    for _ in range(2):
        assert t.erase(one).cpu().tolist() == [1]
    ks, _ = t.dump()
    assert all(not x.any() for x in ks)
    gv, gh = gfind(t, np.array([key, key + 1], np.uint64))
    assert gh.tolist() == [1, 0] and int(gv[0]) in set(pv.tolist())
    assert t.erase(one).cpu().tolist() == [1]
    st = t.stats()
    assert st["pending"] == 0 and st["count"] == 37
    assert gfind(t, np.array([key], np.uint64))[1].tolist() == [0]
    t.close()


def test_wide_fill_bound_stream_matches_oracle_sizes(D):
    """Grow then shrink under auto_resize with bounds [0.5, 0.9]: the subtable sizes after every
    batch equal the oracle's (the policy depends only on counts), and every find is exact."""
    banner("wide auto-resize stream")
    kw = dict(auto_resize=True, lower=0.5, upper=0.9, min_capacity=0, rescue_kicks=128)
    tb = O.OracleWideTable(20_000, **kw)
    t = D.DyCuckooWide(20_000, auto_resize=1, lower_bound=0.5, upper_bound=0.9, min_capacity=0)
    keys = keys64(120_000, 950)
    vals = vals_of(keys)
    live = np.zeros(0, np.uint64)
    for b in range(6):
        k, v = keys[b * 20_000:(b + 1) * 20_000], vals[b * 20_000:(b + 1) * 20_000]
        tb.insert(k, v)
        t.insert(cu(k), cu(v))
        live = np.concatenate([live, k])
        assert t.buckets == tb.nb, b
    for b in range(5):
        e = live[b * 20_000:(b + 1) * 20_000]
        assert np.array_equal(t.erase(cu(e)).cpu().numpy(), tb.erase(e))
        assert t.buckets == tb.nb, ("shrink", b)
    rest = live[100_000:]
    gv, gh = gfind(t, np.concatenate([rest, live[:1000]]))
    assert gh[: rest.size].all() and not gh[rest.size:].any()
    assert np.array_equal(gv[: rest.size], vals_of(rest))
    assert t.stats()["count"] == tb.count == rest.size
    t.close()


@pytest.mark.parametrize("d", (2, 4))
def test_wide_upsert_updates_stored_keys_and_inserts_the_rest(D, d):
    """Upsert against a dict model: stored keys take the new value in place (no second copy),
    absent keys are inserted, the two reserved keys are rejected; counts, the placement
    invariant and finds agree."""
    n = 1 << 19
    banner(f"wide upsert 2^19 over 2^19, subtables={d}")
    keys = keys64(2 * n, 900 + d)
    old, fresh = keys[:n], keys[n:]
    t = D.DyCuckooWide(int(np.ceil(1.6 * n / 0.85)), subtables=d)
    t.insert(cu(old), cu(vals_of(old)))
    up_keys = np.concatenate([old[::2], fresh[: n // 2], np.array([0, 2**64 - 1], dtype=np.uint64)])
    rng = np.random.default_rng(910 + d)
    rng.shuffle(up_keys)
    up_vals = rng.integers(0, 2**64 - 1, size=up_keys.size, dtype=np.uint64)
    flags = t.upsert(cu(up_keys), cu(up_vals)).cpu().numpy()
    model = dict(zip(old.tolist(), vals_of(old).tolist()))
    for k, v in zip(up_keys.tolist(), up_vals.tolist()):
        if k not in (0, 2**64 - 1):
            model[k] = v
    assert np.array_equal(flags.astype(bool), np.isin(up_keys, old))
    st = t.stats()
    assert st["count"] == len(model) and st["pending"] == 0 and st["rejected"] == 2
    ref = oracle_of(t)
    assert ref.check_invariant() == 0
    mk = np.array(list(model.keys()), dtype=np.uint64)
    mv = np.array(list(model.values()), dtype=np.uint64)
    order = np.argsort(mk)
    k, v = ref.entries()  # one copy per key
    assert np.array_equal(k, mk[order]) and np.array_equal(v, mv[order])
    gv, gh = gfind(t, mk)
    assert gh.all() and np.array_equal(gv, mv)
    t.close()


def test_wide_upsert_reaches_keys_held_only_in_the_stash(D):
    """A key in neither of its buckets but in the stash gets the new value in every stash copy;
    a fresh key of the same call is inserted."""
    banner("wide upsert: stash-only key")
    key = 0x0123456789ABCDEF
    t = D.DyCuckooWide(100, max_fail_rounds=0, rescue_kicks=16)
    with pytest.raises(D.DycError) as ei:
        t.insert(cu(np.full(40, key, np.uint64)), cu(np.arange(1000, 1040, dtype=np.uint64)))
    assert ei.value.code == 4  # DYC_ERR_PENDING
    one = cu(np.array([key], np.uint64))
    for _ in range(2):  # clears the copies in pos1's bucket, then pos2's
        assert t.erase(one).cpu().tolist() == [1]
    assert t.stats()["pending"] == 8
    other = key + 2
    flags = t.upsert(cu(np.array([key, other], np.uint64)), cu(np.array([777, 778], np.uint64)))
    assert flags.cpu().tolist() == [1, 0]
    pk, pv = t.pending()
    assert np.all(pk == key) and np.all(pv == 777) and pk.size == 8
    gv, gh = gfind(t, np.array([key, other], np.uint64))
    assert gh.tolist() == [1, 1] and gv.tolist() == [777, 778]
    st = t.stats()
    assert st["count"] == 39 and st["pending"] == 8
    t.close()


def wide_stored(t):
    """(keys, values) of the subtables and the pending list, sorted by (key, value)."""
    ks, vs = t.dump()
    k = np.concatenate([x[x != 0] for x in ks])
    v = np.concatenate([y[x != 0] for x, y in zip(ks, vs)])
    pk, pv = t.pending()
    k = np.concatenate([k, np.asarray(pk, np.uint64)])
    v = np.concatenate([v, np.asarray(pv, np.uint64)])
    o = np.lexsort((v, k))
    return k[o], v[o]


@pytest.mark.parametrize("d", (2, 4))
def test_wide_bulk_fill_insert_multiset_invariant_and_find(D, d):
    """insert_order 3: chunk sort, fill of free slots, second round in the other bucket, eviction
    insert for the rest; reserved keys rejected."""
    n = 1 << 20
    banner(f"wide bulk fill insert 2^20 at 0.90, subtables={d}")
    keys = np.concatenate([keys64(n - 4096, 1600 + d), keys64(4096, 1610 + d, 1, 2**32 - 1)])
    keys = np.unique(keys)
    vals = vals_of(keys)
    batch = np.concatenate([keys, np.array([0, 2**64 - 1], dtype=np.uint64)])
    bv = np.concatenate([vals, np.array([5, 6], dtype=np.uint64)])
    t = D.DyCuckooWide(int(np.ceil(keys.size / 0.9)), subtables=d, insert_order=3)
    assert batch.size >= 2 * sum(t.buckets)
    t.insert(cu(batch), cu(bv))
    st = t.stats()
    assert st["count"] == keys.size and st["lost"] == 0 and st["rejected"] == 2
    assert 0.95 * keys.size < st["last_filled"] <= keys.size
    ref = oracle_of(t)
    assert ref.check_invariant() == 0
    k, v = wide_stored(t)
    order = np.lexsort((vals, keys))
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    gv, gh = gfind(t, keys)
    assert gh.all() and np.array_equal(gv, vals)
    t.close()


def test_wide_bulk_fill_into_fresh_filled_and_cleared_tables(D):
    """A fresh or cleared wide table holds zeros only, and its first dense round reads neither
    its key nor its value lines; any later batch reads them. Batch A into a fresh table, batch B
    on top (both stored, A's values intact), clear, batch B again (nothing of A survives)."""
    banner("wide bulk fill: fresh table, second batch on top, then clear and refill")
    n = 1 << 19
    keys = keys64(2 * n, 1650)
    a, b = keys[:n], keys[n:]
    vals = vals_of(keys)
    t = D.DyCuckooWide(int(np.ceil(2 * n / 0.9)), insert_order=3, auto_resize=False)
    assert n >= 2 * sum(t.buckets)  # both batches take the bulk fill
    t.insert(cu(a), cu(vals[:n]))
    t.insert(cu(b), cu(vals[n:]))
    st = t.stats()
    assert st["count"] == 2 * n and st["lost"] == 0
    assert oracle_of(t).check_invariant() == 0
    k, v = wide_stored(t)
    o = np.lexsort((vals, keys))
    assert np.array_equal(k, keys[o]) and np.array_equal(v, vals[o])
    t.clear()
    t.insert(cu(b), cu(vals[n:]))
    assert t.stats()["count"] == n
    assert oracle_of(t).check_invariant() == 0
    k, v = wide_stored(t)
    o = np.lexsort((vals[n:], b))
    assert np.array_equal(k, b[o]) and np.array_equal(v, vals[n:][o])
    gv, gh = gfind(t, keys)
    assert not gh[:n].any() and gh[n:].all() and np.array_equal(gv[n:], vals[n:])
    t.close()


def test_wide_bulk_fill_overflow_reaches_the_eviction_insert(D):
    banner("wide bulk fill: overflow through both rounds and the eviction insert")
    base = keys64(3000, 1700)
    batch = np.concatenate([base, np.repeat(base[:6], 11)])  # 12 copies of six keys
    rng = np.random.default_rng(1701)
    rng.shuffle(batch)
    bv = vals_of(batch) ^ np.arange(batch.size, dtype=np.uint64)  # copies carry distinct values
    t = D.DyCuckooWide(8192, insert_order=3, auto_resize=False)
    assert batch.size >= 2 * sum(t.buckets)
    t.insert(cu(batch), cu(bv))
    st = t.stats()
    assert st["count"] == batch.size and st["lost"] == 0
    assert oracle_of(t).check_invariant() == 0
    k, v = wide_stored(t)
    o = np.lexsort((bv, batch))
    assert np.array_equal(k, batch[o]) and np.array_equal(v, bv[o])
    single = base[6:]  # one copy each: find returns its value
    want = dict(zip(batch.tolist(), bv.tolist()))
    gv, gh = gfind(t, single)
    assert gh.all() and np.array_equal(gv, np.array([want[x] for x in single.tolist()], dtype=np.uint64))
    t.close()


def test_wide_bulk_fill_into_a_partly_filled_erased_resized_table(D):
    """The wide dense fill on a table that already holds entries, has erased slots and unequal
    subtables after an upsize: the first keys go in by the eviction insert (batches below the
    bulk threshold), a fifth of them are erased, one subtable is doubled, then one large batch of
    fresh keys, second copies of stored keys, re-inserts of erased keys and reserved keys takes
    the bulk fill. Multiset, invariant, counts and finds are checked against the expected set."""
    banner("wide bulk fill into a partly filled, erased, upsized table")
    first = keys64(300_000, 1800)
    t = D.DyCuckooWide(600_000, insert_order=3, auto_resize=False)
    nb0 = sum(t.buckets)
    for lo in range(0, first.size, 50_000):  # below 2 x buckets: the eviction insert
        part = first[lo:lo + 50_000]
        assert part.size < 2 * nb0
        t.insert(cu(part), cu(vals_of(part)))
    gone = first[::5]
    assert t.erase(cu(gone)).cpu().numpy().all()
    t.resize_up()
    nb = t.buckets
    assert len(set(nb)) == 2  # one subtable doubled
    kept = np.setdiff1d(first, gone)
    fresh = absent(250_000, first, 1801)
    dup = kept[:5000]  # second copies of stored keys
    back = gone[:3000]  # erased keys inserted again
    batch = np.concatenate([fresh, dup, back, np.array([0, 2**64 - 1, 0], dtype=np.uint64)])
    rng = np.random.default_rng(1802)
    rng.shuffle(batch)
    assert batch.size >= 2 * sum(nb)  # large enough for the bulk fill
    bv = vals_of(batch) ^ np.uint64(0x5555)
    t.insert(cu(batch), cu(bv))
    st = t.stats()
    keep = (batch != 0) & (batch != 2**64 - 1)
    assert st["count"] == kept.size + int(keep.sum()) and st["lost"] == 0
    assert oracle_of(t).check_invariant() == 0
    k, v = wide_stored(t)
    wk = np.concatenate([kept, batch[keep]])
    wv = np.concatenate([vals_of(kept), bv[keep]])
    o = np.lexsort((wv, wk))
    assert np.array_equal(k, wk[o]) and np.array_equal(v, wv[o])
    gv, gh = gfind(t, np.concatenate([fresh, back]))
    assert gh.all() and np.array_equal(gv, vals_of(np.concatenate([fresh, back])) ^ np.uint64(0x5555))
    gv, gh = gfind(t, gone[3000:3100])
    assert not gh.any()
    t.close()
