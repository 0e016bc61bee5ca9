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
