"""GPU parity of many wide-key tables served by one launch (dyc_wide_group_*), per table against
the wide-key oracle: each table's images (key and value arrays) satisfy the placement invariant
and hold exactly the inserted (key, value) multiset; finds equal the oracle's finds on those
images; erase flags and counts are exact; the tagged pending list serves its own table only."""
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


def keys64(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**64 - 2, size=int(n * 1.3) + 64, dtype=np.uint64))
    rng.shuffle(out)
    return out[:n]


def vals_of(k):
    return (k * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0xD1B54A32D192ED03)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def sizes_loguniform(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int64)


def oracle_of(g, i):
    ks, vs = g.dump(i)
    ref = O.OracleWideTable(nb=[x.size // 16 for x in ks])
    ref.load_images(ks, vs)
    return ref


@pytest.mark.parametrize("d,per_table", [(2, False), (4, False), (4, True)])
def test_wide_group_insert_find_match_oracle_per_table(D, d, per_table):
    banner(f"wide group insert/find, subtables={d}, per_table={per_table}")
    sizes = sizes_loguniform(29, 2**6, 2**13, 40 + d)
    sizes[3] = 0
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = keys64(int(offs[-1]), 41 + d)
    keys[: min(512, keys.size)] &= np.uint64(0xFFFFFFFF)  # some keys below 2^32 (fold(k) = k)
    keys[keys == 0] = 7
    keys = np.unique(keys)[: int(offs[-1])]
    np.random.default_rng(42).shuffle(keys)
    vals = vals_of(keys)
    hints = (np.ceil(np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooWideGroup(hints, subtables=d)
    g.insert(cu(keys), cu(vals), offs, per_table=per_table)
    # lookups: per table its own keys, then half as many of the next table's keys (absent here)
    q_parts, q_offs = [], [0]
    for i in range(len(sizes)):
        own = keys[offs[i]:offs[i + 1]]
        nxt = (i + 1) % len(sizes)
        other = keys[offs[nxt]:offs[nxt + 1]][: max(1, own.size // 2)]
        q = np.concatenate([own, other, np.array([0], dtype=np.uint64)])
        q_parts.append(q)
        q_offs.append(q_offs[-1] + q.size)
    q = np.concatenate(q_parts)
    v, h = g.find(cu(q), np.array(q_offs, dtype=np.uint64), per_table=per_table)
    v, h = v.cpu().numpy().view(np.uint64), h.cpu().numpy()
    for i in range(len(sizes)):
        st = g.table_stats(i)
        assert st["count"] == sizes[i] and st["subtables"] == d
        ref = oracle_of(g, i)
        assert ref.check_invariant() == 0
        own = keys[offs[i]:offs[i + 1]]
        ek, ev = ref.entries()
        o = np.lexsort((vals[offs[i]:offs[i + 1]], own))
        assert np.array_equal(ek, own[o]) and np.array_equal(ev, vals[offs[i]:offs[i + 1]][o])
        lo, hi = q_offs[i], q_offs[i + 1]
        rv, rh = ref.find(q[lo:hi])
        assert np.array_equal(h[lo:hi], rh) and np.array_equal(v[lo:hi], rv)
        assert rh[: own.size].all() and not rh[own.size:].any()
    g.close()


@pytest.mark.parametrize("per_table", [False, True])
def test_wide_group_erase_and_reserved_keys(D, per_table):
    banner(f"wide group erase + reserved keys, per_table={per_table}")
    sizes = sizes_loguniform(11, 2**8, 2**12, 50)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = keys64(int(offs[-1]), 51)
    keys[offs[2]] = np.uint64(2**64 - 1)  # reserved: rejected by insert
    vals = vals_of(keys)
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooWideGroup(hints)
    g.insert(cu(keys), cu(vals), offs, per_table=per_table)
    assert g.table_stats(2)["count"] == sizes[2] - 1 and g.table_stats(3)["count"] == sizes[3]
    e_parts, e_offs = [], [0]
    for i in range(len(sizes)):
        part = keys[int(offs[i]):int(offs[i]) + int(sizes[i]) // 2]
        e_parts.append(part)
        e_offs.append(e_offs[-1] + part.size)
    ek = np.concatenate(e_parts)
    d = g.erase(cu(ek), np.array(e_offs, dtype=np.uint64)).cpu().numpy()
    assert np.array_equal(d, (ek != np.uint64(2**64 - 1)).astype(np.uint8))
    for i in range(len(sizes)):
        assert g.table_stats(i)["count"] == sizes[i] - sizes[i] // 2
        assert oracle_of(g, i).check_invariant() == 0
    v, h = g.find(cu(keys), offs)
    h, v = h.cpu().numpy(), v.cpu().numpy().view(np.uint64)
    for i in range(len(sizes)):
        hh = h[offs[i]:offs[i + 1]]
        assert not hh[: sizes[i] // 2].any() and hh[sizes[i] // 2:].all()
        assert np.array_equal(v[offs[i]:offs[i + 1]][sizes[i] // 2:], vals[offs[i]:offs[i + 1]][sizes[i] // 2:])
    g.close()


@pytest.mark.parametrize("per_table", [False, True])
def test_wide_group_stash_serves_pending_entries_of_their_table(D, per_table):
    """40 copies of one key into table 1 (tables do not resize): 8 stay pending, tagged 1 (also
    when the single-table kernel parked them). After two erases clear the key's buckets, table 1's
    find still reaches a stash copy, table 0 does not see it, and one more erase clears the stash."""
    banner(f"wide group stash, per_table={per_table}")
    g = D.DyCuckooWideGroup([1000, 100, 1000])
    key = np.uint64(0x0123456789ABCDEF)
    offs = np.array([0, 0, 40, 40], dtype=np.uint64)
    vals = np.arange(500, 540, dtype=np.uint64)
    with pytest.raises(D.DycError) as ei:
        g.insert(cu(np.full(40, key, np.uint64)), cu(vals), offs, per_table=per_table)
    assert ei.value.code == 4  # DYC_ERR_PENDING
    assert g.table_stats(1)["count"] == 40
    pk, pv, pt = g.pending()
    assert pk.size == 8 and np.all(pt == 1) and np.all(pk == key)
    one = np.array([0, 0, 1, 1], dtype=np.uint64)
    for _ in range(2):
        assert g.erase(cu(np.array([key])), one).cpu().tolist() == [1]
    two = np.array([0, 1, 2, 2], dtype=np.uint64)
    v, h = g.find(cu(np.array([key, key])), two, per_table=per_table)
    assert h.cpu().tolist() == [0, 1] and int(v.cpu().numpy().view(np.uint64)[1]) in set(pv.tolist())
    assert g.erase(cu(np.array([key])), one).cpu().tolist() == [1]
    pk, _, _ = g.pending()
    assert pk.size == 0 and g.table_stats(1)["pending"] == 0
    assert g.find(cu(np.array([key])), one)[1].cpu().tolist() == [0]
    g.close()


@pytest.mark.parametrize("d", (2, 4))
def test_wide_group_bulk_fill_matches_oracle_per_table(D, d):
    """insert_order 3: each table's batch partitioned by its own chunks, chunk fill, a second
    round for the other bucket, the eviction insert of the rest (rounds chained on the device).
    A batch on top of it reads the tables; a refill after clear does not; every table equals the
    wide oracle each time, and last_filled says the fill placed nearly every key."""
    banner(f"wide group bulk fill, subtables={d}")
    sizes = sizes_loguniform(23, 2**8, 2**15, 140 + d)
    sizes[4] = 0
    total = int(sizes.sum())
    keys = keys64(2 * total, 141 + d)
    keys[: min(300, keys.size)] &= np.uint64(0xFFFFFFFF)  # some keys below 2^32 (fold(k) = k)
    keys[keys == 0] = 9
    keys = np.unique(keys)[: 2 * total]
    np.random.default_rng(142 + d).shuffle(keys)
    vals = vals_of(keys)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    hints = (np.ceil(2 * np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooWideGroup(hints, subtables=d, insert_order=3)

    def batch(lo):
        return keys[lo:lo + total], vals[lo:lo + total]

    def check(parts):
        for i in range(len(sizes)):
            st = g.table_stats(i)
            ks = np.concatenate([p[0][offs[i]:offs[i + 1]] for p in parts])
            vs = np.concatenate([p[1][offs[i]:offs[i + 1]] for p in parts])
            assert st["count"] == ks.size and st["subtables"] == d
            ref = oracle_of(g, i)
            assert ref.check_invariant() == 0
            ek, ev = ref.entries()
            o = np.lexsort((vs, ks))
            assert np.array_equal(ek, ks[o]) and np.array_equal(ev, vs[o])

    a, b = batch(0), batch(total)
    for part in (a, b):
        g.insert(cu(part[0]), cu(part[1]), offs)
        assert g.table_stats(0)["last_filled"] > 0.9 * total
    check([a, b])
    g.clear()
    g.insert(cu(b[0]), cu(b[1]), offs)
    assert g.table_stats(0)["last_filled"] > 0.9 * total
    check([b])
    v, h = g.find(cu(b[0]), offs)
    assert h.cpu().numpy().all() and np.array_equal(v.cpu().numpy().view(np.uint64), b[1])
    _, h = g.find(cu(a[0]), offs)
    assert not h.cpu().numpy().any()
    g.close()
