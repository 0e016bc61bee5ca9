"""GPU parity of many tables served by one launch (dyc_group_*), per table against the oracle."""
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


def batch(sizes, seed):
    """Unique keys for every table (disjoint across tables), values splitmix32(key), offsets."""
    from databases_gpu_amd.workload import splitmix32_np

    total = int(sum(sizes))
    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(1, 2**32 - 1, size=int(total * 1.3) + 64, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(keys)
    keys = keys[:total]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return keys, splitmix32_np(keys), offs


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def sizes_loguniform(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).astype(np.int64)


@pytest.mark.parametrize("per_table", [False, True])
def test_group_insert_find_match_oracle_per_table(D, per_table):
    banner(f"group insert/find per_table={per_table}")
    sizes = sizes_loguniform(37, 2**6, 2**14, 1)
    sizes[5] = 0  # an empty table
    keys, vals, offs = batch(sizes, 2)
    # reference sizing floors hint/64 before rounding to a prime: add a bucket row of slack
    hints = (np.ceil(np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints)
    g.insert(dev(keys), dev(vals), offs, per_table=per_table)
    # lookups: per table its own keys, then as many keys of the next table (absent here)
    q_parts, q_offs, want = [], [0], []
    for i in range(len(sizes)):
        own = keys[offs[i]:offs[i + 1]]
        other = keys[offs[(i + 1) % len(sizes)]:offs[(i + 1) % len(sizes) + 1]][: max(1, own.size // 2)]
        q = np.concatenate([own, other, np.array([0], dtype=np.uint32)])
        q_parts.append(q)
        q_offs.append(q_offs[-1] + q.size)
    q = np.concatenate(q_parts)
    v, h = g.find(dev(q), np.array(q_offs, dtype=np.uint64), per_table=per_table)
    v, h = v.cpu().numpy().view(np.uint32), h.cpu().numpy()
    for i in range(len(sizes)):
        st = g.table_stats(i)
        assert st["count"] == sizes[i]
        assert st["buckets"] == [O.initial_buckets(int(hints[i]))] * 4
        ref = O.OracleTable(nb=st["buckets"])
        ref.sub = g.dump(i)
        assert ref.check_invariant() == 0
        own = keys[offs[i]:offs[i + 1]]
        assert np.array_equal(ref.entries(), np.sort(O.pack(own, vals[offs[i]:offs[i + 1]])))
        lo, hi = q_offs[i], q_offs[i + 1]
        rv, rh = ref.find(q[lo:hi])
        assert np.array_equal(v[lo:hi], rv) and np.array_equal(h[lo:hi], rh)
        assert rh[: own.size].all() and not rh[own.size:].any()
    g.close()


@pytest.mark.parametrize("per_table", [False, True])
def test_group_erase_and_reserved_keys(D, per_table):
    banner(f"group erase + reserved keys, per_table={per_table}")
    sizes = sizes_loguniform(11, 2**8, 2**12, 3)
    keys, vals, offs = batch(sizes, 4)
    keys[offs[2]] = 0xFFFFFFFF  # reserved: rejected by insert
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    g = D.DyCuckooGroup(hints, insert_lanes=4)
    g.insert(dev(keys), dev(vals), offs, per_table=per_table)
    assert g.table_stats(2)["count"] == sizes[2] - 1 and g.table_stats(3)["count"] == sizes[3]
    # erase the first half of every table
    e_parts, e_offs = [], [0]
    for i in range(len(sizes)):
        part = keys[int(offs[i]):int(offs[i]) + int(sizes[i]) // 2]
        e_parts.append(part)
        e_offs.append(e_offs[-1] + part.size)
    ek = np.concatenate(e_parts)
    d = g.erase(dev(ek), np.array(e_offs, dtype=np.uint64)).cpu().numpy()
    want = (ek != 0xFFFFFFFF).astype(np.uint8)
    assert np.array_equal(d, want)
    # table 2's reserved key sits in its erased half and was never stored, so every table keeps
    # sizes[i] - sizes[i] // 2 entries
    for i in range(len(sizes)):
        assert g.table_stats(i)["count"] == sizes[i] - sizes[i] // 2
    v, h = g.find(dev(keys), offs)
    h = h.cpu().numpy()
    for i in range(len(sizes)):
        hh = h[offs[i]:offs[i + 1]]
        assert not hh[: sizes[i] // 2].any() and hh[sizes[i] // 2:].all()
    g.close()


def test_group_launch_shapes_agree(D):
    """Every launch shape (lanes, grid, item size) gives the same results."""
    banner("group launch shapes")
    sizes = sizes_loguniform(64, 2**8, 2**13, 5)
    keys, vals, offs = batch(sizes, 6)
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    base = None
    for fl, il, fb, ib, items in [(8, 1, 0, 0, 4096), (4, 2, 64, 32, 512), (16, 4, 1, 1, 100000), (8, 8, 300, 7, 64)]:
        g = D.DyCuckooGroup(hints)
        g.set_launch(find_lanes=fl, insert_lanes=il, find_blocks=fb, insert_blocks=ib, item_keys=items)
        g.insert(dev(keys), dev(vals), offs)
        v, h = g.find(dev(keys), offs)
        got = (v.cpu().numpy().view(np.uint32), h.cpu().numpy())
        assert got[1].all() and np.array_equal(got[0], vals)
        if base is None:
            base = got
        g.close()


def test_group_stash_serves_pending_entries_of_their_table(D):
    """40 copies of one key into table 1 of a group (tables do not resize): 8 stay pending,
    tagged with table 1. Once two erases have cleared the key's two buckets, a find of table 1
    (both launch forms) still reaches a stash copy, table 0 does not see it, and one more erase
    clears the stash."""
    banner("group stash")
    g = D.DyCuckooGroup([1000, 100, 1000])
    key = 0x2468ACE1
    sizes = np.array([0, 40, 0])
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    vals = np.arange(500, 540, dtype=np.uint32)
    with pytest.raises(D.DycError) as ei:
        g.insert(dev(np.full(40, key, np.uint32)), dev(vals), offs)
    assert ei.value.code == 4  # DYC_ERR_PENDING
    e, tb = g.pending()
    assert e.size == 8 and np.all(tb == 1) and np.all((e >> np.uint64(32)) == key)
    stash_vals = set((e & np.uint64(0xFFFFFFFF)).tolist())
    one = np.array([0, 0, 1, 1], dtype=np.uint64)  # table 1 gets the single key
    for _ in range(2):
        assert g.erase(dev(np.array([key], np.uint32)), one).cpu().tolist() == [1]
    q = np.array([key, key], np.uint32)
    two = np.array([0, 1, 2, 2], dtype=np.uint64)  # one lookup in table 0, one in table 1
    for per_table in (False, True):
        v, h = g.find(dev(q), two, per_table=per_table)
        v, h = v.cpu().numpy().view(np.uint32), h.cpu().numpy()
        assert h.tolist() == [0, 1] and int(v[1]) in stash_vals
    assert g.erase(dev(np.array([key], np.uint32)), one).cpu().tolist() == [1]
    e, tb = g.pending()
    assert e.size == 0 and g.table_stats(1)["pending"] == 0
    v, h = g.find(dev(np.array([key], np.uint32)), one)
    assert h.cpu().tolist() == [0]
    g.close()


def test_group_occupancy_builds_agree(D):
    """The occupancy builds (at least 6 or 8 waves per SIMD) of the group kernels give the same
    results as the compiler's own, for every lane shape, including the ones that spill."""
    banner("group waves-per-EU builds")
    sizes = sizes_loguniform(40, 2**8, 2**12, 15)
    keys, vals, offs = batch(sizes, 16)
    hints = (np.ceil(sizes / 0.85) + 64).astype(np.uint64)
    for il, fl in [(1, 16), (2, 4), (4, 8)]:
        for w in (0, 6, 8):
            g = D.DyCuckooGroup(hints)
            g.set_launch(insert_lanes=il, find_lanes=fl, insert_waves_per_eu=w, find_waves_per_eu=w)
            assert g.launch["insert_waves_per_eu"] == w
            g.insert(dev(keys), dev(vals), offs)
            v, h = g.find(dev(keys), offs)
            assert h.cpu().numpy().all() and np.array_equal(v.cpu().numpy().view(np.uint32), vals)
            d = g.erase(dev(keys[::2]), np.array([0, keys[::2].size] + [keys[::2].size] * (len(sizes) - 1),
                                                  dtype=np.uint64))
            assert d.cpu().numpy()[: int(sizes[0] + 1) // 2].all()
            g.close()
    g = D.DyCuckooGroup(hints)
    with pytest.raises(D.DycError):
        g.set_launch(insert_waves_per_eu=5)
    g.close()


def test_group_graph_plan_replays_equal_eager_steps(D):
    """A captured step (clear -> insert -> find, one graph launch) replayed several times, with
    the inputs rewritten and eager work between the replays, equals the eager calls on a twin
    group every time (values, hits, counts, images), and the oracle on every table."""
    from databases_gpu_amd.workload import splitmix32_np

    banner("group graph plan: 4 replays vs eager")
    sizes = sizes_loguniform(50, 2**6, 2**12, 21)
    sizes[7] = 0
    hints = (np.ceil(np.maximum(sizes, 1) / 0.85) + 64).astype(np.uint64)
    keys0, vals0, offs = batch(sizes, 22)
    qk = dev(keys0)  # lookups: each table's own keys (rewritten below with absent ones mixed in)
    kk, vv = dev(keys0), dev(vals0)
    gp = D.DyCuckooGroup(hints, insert_lanes=2, find_lanes=4)
    ge = D.DyCuckooGroup(hints, insert_lanes=2, find_lanes=4)
    plan = gp.plan(kk, vv, offs, qk, offs)
    for r in range(4):
        keys, _, _ = batch(sizes, 100 + r)
        vals = splitmix32_np(keys) ^ np.uint32(r)
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
        for i in range(len(sizes)):
            assert gp.table_stats(i)["count"] == sizes[i] == ge.table_stats(i)["count"]
        for i in (0, 7, 13, 49):
            ref = O.OracleTable(nb=gp.table_stats(i)["buckets"])
            ref.sub = gp.dump(i)
            assert ref.check_invariant() == 0
            lo, hi = int(offs[i]), int(offs[i + 1])
            assert np.array_equal(ref.entries(), np.sort(O.pack(keys[lo:hi], vals[lo:hi])))
            rv, rh = ref.find(look[lo:hi])
            assert np.array_equal(ph[lo:hi], rh) and np.array_equal(pv[lo:hi][rh == 1], rv[rh == 1])
    plan.close()
    gp.close()
    ge.close()


def test_group_graph_plan_stash_and_moved_pending_list(D):
    """Entries that exhaust the bound in a replay are found by the same replay's find (the stash
    size is read on the device) and reported as DYC_ERR_PENDING; a plan whose group's pending
    list has since grown refuses to run instead of writing the old one."""
    banner("group graph plan: stash + moved pending list")
    g = D.DyCuckooGroup([1000, 100, 1000])
    key = 0x2468ACE1
    offs = np.array([0, 0, 40, 40], dtype=np.uint64)
    kk = dev(np.full(40, key, np.uint32))
    vv = dev(np.arange(500, 540, dtype=np.uint32))
    q = dev(np.array([key, key], np.uint32))
    qo = np.array([0, 1, 2, 2], dtype=np.uint64)
    plan = g.plan(kk, vv, offs, q, qo)
    for _ in range(2):
        with pytest.raises(D.DycError) as ei:
            plan.run()
        assert ei.value.code == 4
        assert plan.out_hit.cpu().tolist() == [0, 1]
        e, tb = g.pending()
        assert e.size == 8 and np.all(tb == 1)
    big = 70_000 * 20  # an eager insert this large grows the pending list
    sizes = np.array([big, 0, 0])
    g.clear()
    keys, vals, o2 = batch(sizes, 31)
    with pytest.raises(D.DycError):  # table 0 holds ~1000 slots: most of the batch cannot be placed
        g.insert(dev(keys), dev(vals), o2)
    with pytest.raises(D.DycError) as ei:
        plan.run()
    assert ei.value.code == 1 and "plan again" in str(ei.value)
    plan.close()
    g.close()
