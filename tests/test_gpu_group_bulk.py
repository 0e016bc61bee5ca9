"""GPU parity of the group bulk fill (dyc_group_insert with insert_order 3), per table against
the oracle: each table's batch is partitioned by its own chunks in one pass, every chunk is
filled by one workgroup, and what overflows its first bucket goes through the group's eviction
insert. Slot layout after a parallel insert is not fixed, so each table is checked by the
key->value multiset of its image plus its pending entries, the placement invariant on the
oracle, counts, and finds; stats last_filled shows which path took the batch."""
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
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def sizes_loguniform(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int64)


def offsets_of(parts):
    return np.concatenate([[0], np.cumsum([p.size for p in parts])]).astype(np.uint64)


def check_tables(D, g, want_parts, d=4):
    """Table i holds exactly want_parts[i] (packed entries, duplicates included), placed by the
    invariant; its own keys are found with a value among their copies'."""
    pend, ptab = g.pending()
    for i, want in enumerate(want_parts):
        st = g.table_stats(i)
        assert st["count"] == want.size and st["subtables"] == d
        ref = O.OracleTable(nb=st["buckets"])
        ref.sub = g.dump(i)
        assert ref.check_invariant() == 0
        mine = pend[ptab == i]
        got = np.sort(np.concatenate([ref.entries(), mine]))
        assert np.array_equal(got, np.sort(want))


def find_own(g, parts):
    q = np.concatenate([p >> np.uint64(32) for p in parts]).astype(np.uint32)
    v, h = g.find(dev(q), offsets_of(parts))
    return q, v.cpu().numpy().view(np.uint32), h.cpu().numpy()


@pytest.mark.parametrize("d", (2, 4, 6))
def test_group_bulk_fill_matches_oracle_per_table(D, d):
    banner(f"group bulk fill: 41 tables of 2^6..2^16 keys, subtables={d}")
    sizes = sizes_loguniform(41, 2**6, 2**16, 60 + d)
    sizes[3] = 0  # an empty table
    keys = unique_keys(int(sizes.sum()), 61 + d)
    vals = vals_of(keys)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    hints = (np.ceil(np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints, subtables=d, insert_order=3)
    g.insert(dev(keys), dev(vals), offs)
    assert g.table_stats(0)["last_filled"] > 0.9 * keys.size  # the bulk fill took the batch
    assert g.table_stats(0)["lost"] == 0
    parts = [O.pack(keys[offs[i]:offs[i + 1]], vals[offs[i]:offs[i + 1]]) for i in range(len(sizes))]
    check_tables(D, g, parts, d)
    q, v, h = find_own(g, parts)
    assert h.all() and np.array_equal(v, vals_of(q))
    # absent keys: each table looks up the next table's keys
    nxt = [parts[(i + 1) % len(parts)] for i in range(len(parts))]
    qa = np.concatenate([p >> np.uint64(32) for p in nxt]).astype(np.uint32)
    _, ha = g.find(dev(qa), offsets_of(nxt))
    assert not ha.cpu().numpy().any()
    g.close()


def test_group_bulk_fill_reserved_keys_and_duplicates(D):
    """Reserved keys are rejected and counted; second copies of stored keys are kept, in their
    table only."""
    banner("group bulk fill: reserved keys and duplicate copies")
    sizes = sizes_loguniform(13, 2**8, 2**14, 70)
    keys = unique_keys(int(sizes.sum()), 71)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    raw = [keys[offs[i]:offs[i + 1]] for i in range(len(sizes))]
    raw[2] = np.concatenate([raw[2], np.array([0, 0xFFFFFFFF, 0], np.uint32)])
    raw[5] = np.concatenate([raw[5], raw[5][:100]])  # second copies
    rng = np.random.default_rng(72)
    for r in raw:
        rng.shuffle(r)
    bk = np.concatenate(raw)
    bv = vals_of(bk) ^ np.arange(bk.size, dtype=np.uint32)  # copies carry distinct values
    bo = offsets_of(raw)
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints, insert_order=3)
    g.insert(dev(bk), dev(bv), bo)
    st = g.table_stats(0)
    assert st["rejected"] == 3 and st["lost"] == 0
    assert st["last_filled"] > 0.9 * bk.size
    want = []
    for i in range(len(raw)):
        k, v = bk[bo[i]:bo[i + 1]], bv[bo[i]:bo[i + 1]]
        keep = (k != 0) & (k != 0xFFFFFFFF)
        want.append(O.pack(k[keep], v[keep]))
    check_tables(D, g, want)
    g.close()


def test_group_bulk_fill_on_filled_and_cleared_groups(D):
    """Batch A into fresh tables (known zero: not read), batch B on top (the tables are read),
    clear, batch B again: the images hold A and B, then B alone."""
    banner("group bulk fill: second batch on top, then clear and refill")
    sizes = sizes_loguniform(17, 2**9, 2**15, 80)
    keys = unique_keys(2 * int(sizes.sum()), 81)
    half = int(sizes.sum())
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    a = [keys[offs[i]:offs[i + 1]] for i in range(len(sizes))]
    b = [keys[half + offs[i]:half + offs[i + 1]] for i in range(len(sizes))]
    hints = (np.ceil(2 * sizes / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints, insert_order=3)
    for part in (a, b):
        k = np.concatenate(part)
        g.insert(dev(k), dev(vals_of(k)), offsets_of(part))
        assert g.table_stats(0)["last_filled"] > 0.9 * k.size
    both = [O.pack(np.concatenate([x, y]), vals_of(np.concatenate([x, y]))) for x, y in zip(a, b)]
    check_tables(D, g, both)
    g.clear()
    k = np.concatenate(b)
    g.insert(dev(k), dev(vals_of(k)), offsets_of(b))
    check_tables(D, g, [O.pack(y, vals_of(y)) for y in b])
    qa = np.concatenate(a)
    _, ha = g.find(dev(qa), offsets_of(a))
    assert not ha.cpu().numpy().any()
    g.close()


def test_group_bulk_fill_declines_tables_beyond_its_chunk_limit(D):
    """A table of more than 256 chunks (2^22 keys: 604 chunks) sends the whole batch to the
    eviction insert (last_filled 0) with the same results."""
    banner("group bulk fill declined: a table of 604 chunks")
    sizes = np.array([1 << 22, 3000, 5000])
    keys = unique_keys(int(sizes.sum()), 90)
    vals = vals_of(keys)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints, insert_order=3)
    g.insert(dev(keys), dev(vals), offs)
    assert g.table_stats(0)["last_filled"] == 0
    v, h = g.find(dev(keys), offs)
    assert h.cpu().numpy().all() and np.array_equal(v.cpu().numpy().view(np.uint32), vals)
    for i in range(3):
        assert g.table_stats(i)["count"] == sizes[i]
    g.close()


def test_group_bulk_fill_overfull_table_keeps_entries_in_the_stash(D):
    """A table given more keys than its slots: the fill places what fits, the eviction insert
    parks the rest on the group's pending list tagged with that table (DYC_ERR_PENDING: group
    tables do not resize); every key is still found in its own table."""
    banner("group bulk fill: one overfull table")
    sizes = np.array([4000, 2500, 6000])
    keys = unique_keys(int(sizes.sum()), 95)
    vals = vals_of(keys)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    hints = np.array([5000, 2000, 7000], dtype=np.uint64)  # table 1: fewer slots than keys
    g = D.DyCuckooGroup(hints, insert_order=3, kick_bound=4, rescue_kicks=8)
    assert 16 * sum(g.table_stats(1)["buckets"]) < sizes[1]
    with pytest.raises(D.DycError) as ei:
        g.insert(dev(keys), dev(vals), offs)
    assert ei.value.code == 4  # DYC_ERR_PENDING
    e, tb = g.pending()
    assert e.size > 0 and np.all(tb == 1)
    parts = [O.pack(keys[offs[i]:offs[i + 1]], vals[offs[i]:offs[i + 1]]) for i in range(3)]
    check_tables(D, g, parts)
    v, h = g.find(dev(keys), offs)
    assert h.cpu().numpy().all() and np.array_equal(v.cpu().numpy().view(np.uint32), vals)
    g.close()


def test_group_bulk_fill_graph_plan_replays_equal_eager_steps(D):
    """A plan of a group with insert_order 3 captures the whole bulk fill (both rounds and the
    eviction insert, chained on the device by the overflow counts). Four replays with the inputs
    rewritten and eager work in between each equal the eager bulk calls on a twin group (finds,
    counts, last_filled) and the oracle on sampled tables; an eager bulk insert that grows the
    shared workspace makes the plan refuse to run."""
    banner("group bulk fill: graph plan replays vs eager")
    sizes = sizes_loguniform(40, 2**8, 2**14, 110)
    sizes[9] = 0
    hints = (np.ceil(np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(sizes.sum())
    k0 = unique_keys(total, 111)
    kk, vv, qk = dev(k0), dev(vals_of(k0)), dev(k0)
    gp = D.DyCuckooGroup(hints, insert_order=3)
    ge = D.DyCuckooGroup(hints, insert_order=3)
    plan = gp.plan(kk, vv, offs, qk, offs)
    for r in range(4):
        keys = unique_keys(total, 120 + r)
        vals = vals_of(keys) ^ np.uint32(r)
        look = keys.copy()
        look[1::3] ^= np.uint32(0x5A5A5A5A)  # mostly absent keys
        kk.copy_(dev(keys))
        vv.copy_(dev(vals))
        qk.copy_(dev(look))
        torch.cuda.synchronize()
        _ = (kk.float() * 2).sum().item()  # eager work between replays
        plan.run()
        pv = plan.out_vals.cpu().numpy().view(np.uint32).copy()
        ph = plan.out_hit.cpu().numpy().copy()
        ge.clear()
        ge.insert(kk, vv, offs)
        ev, eh = ge.find(qk, offs)
        assert np.array_equal(ph, eh.cpu().numpy()) and np.array_equal(pv, ev.cpu().numpy().view(np.uint32))
        assert gp.table_stats(0)["last_filled"] > 0.9 * total and ge.table_stats(0)["last_filled"] > 0.9 * total
        for i in range(len(sizes)):
            assert gp.table_stats(i)["count"] == sizes[i] == ge.table_stats(i)["count"]
        for i in (0, 9, 21, 39):
            ref = O.OracleTable(nb=gp.table_stats(i)["buckets"])
            ref.sub = gp.dump(i)
            assert ref.check_invariant() == 0
            lo, hi = int(offs[i]), int(offs[i + 1])
            assert np.array_equal(ref.entries(), np.sort(O.pack(keys[lo:hi], vals[lo:hi])))
            rv, rh = ref.find(look[lo:hi])
            assert np.array_equal(ph[lo:hi], rh) and np.array_equal(pv[lo:hi][rh == 1], rv[rh == 1])
    # an eager bulk insert of a larger batch (5 % more keys per table, still fitting) grows the
    # shared workspace: the plan refuses to run
    big = np.ceil(sizes * 1.05).astype(np.int64)
    bk = unique_keys(int(big.sum()), 130)
    boffs = np.concatenate([[0], np.cumsum(big)]).astype(np.uint64)
    gp.clear()
    gp.insert(dev(bk), dev(vals_of(bk)), boffs)
    assert gp.table_stats(0)["last_filled"] > 0.9 * bk.size
    with pytest.raises(D.DycError) as ei:
        plan.run()
    assert ei.value.code == 1  # DYC_ERR_INVALID: make the plan again
    plan.close()
    gp.close()
    ge.close()
