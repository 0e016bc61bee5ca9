"""Pins of the oracle's table semantics: insert replay, find, erase, resize, policy.

Checks come from what the reference text fixes by hand (first placements, the victim rule's
first step), from invariants (every entry sits in bucket h_t(k) % P % nb[t] of a subtable t of
its pair), from a Python dict of the inserted set, and from the bucket arithmetic of resizing
(h % 2n is h % n or h % n + n; h % n == (h % 2n) % n).
"""
import numpy as np
import pytest

import oracle as O

NOT = O.NO_TABLE


def unique_keys(n, seed, lo=1, hi=2**32 - 2):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(lo, hi + 1, size=int(n * 1.3) + 16, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(out)
    assert out.size >= n
    return out[:n]


def splitmix32(k):
    """values of the workload: a 32-bit splitmix-style finaliser of the key"""
    z = (k.astype(np.uint64) + np.uint64(0x9E3779B9)) & np.uint64(0xFFFFFFFF)
    z = ((z ^ (z >> np.uint64(16))) * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    z = ((z ^ (z >> np.uint64(13))) * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    return (z ^ (z >> np.uint64(16))).astype(np.uint32)


def where(table, key):
    """(subtable, bucket, slot) of every copy of key in the image"""
    out = []
    for t in range(4):
        idx = np.nonzero((table.sub[t] >> np.uint64(32)) == np.uint64(key))[0]
        out += [(t, int(i) // 16, int(i) % 16) for i in idx]
    return out


# ------------------------------------------------------------------------ first placements
def test_first_insert_goes_to_pos1_for_even_and_pos2_for_odd(oracle_lib):
    L = O.lib()
    for key in (2, 10, 1234, 0x80000000, 3, 11, 12345, 0x7FFFFFFF):
        tb = O.OracleTable(nb=[6, 6, 6, 6])
        tb.insert([key], [99])
        p = L.orc_pair(key)
        t = L.orc_pos1(p) if key % 2 == 0 else L.orc_pos2(p)  # cuckoo.cu:336-341
        assert where(tb, key) == [(t, L.orc_bucket(key, t, 6), 0)]
        assert tb.last_kicks[0] == 1


def _keys_for_bucket(t, b, nb, even, count, start=2):
    L = O.lib()
    out, k = [], start
    while len(out) < count:
        p = L.orc_pair(k)
        first = L.orc_pos1(p) if k % 2 == 0 else L.orc_pos2(p)
        if (k % 2 == 0) == even and first == t and L.orc_bucket(k, t, nb) == b:
            out.append(k)
        k += 1
    return out


def test_bucket_fills_slot_by_slot_then_evicts_by_the_reference_rule(oracle_lib):
    """17 even keys whose first choice is the same bucket: slots 0..15 in order, then the 17th
    evicts the first resident whose pair index differs from the table number (cuckoo.cu:398-400,
    a pair index compared with a table number) and the victim moves to the other table of its
    pair."""
    L = O.lib()
    nb = 4
    ks = _keys_for_bucket(1, 2, nb, True, 17)
    tb = O.OracleTable(nb=[nb] * 4)
    tb.insert(ks[:16], list(range(100, 116)))
    for i, k in enumerate(ks[:16]):
        assert where(tb, k) == [(1, 2, i)]
    # the rule, worked by hand from the pair indexes of the residents
    pairs = [L.orc_pair(k) for k in ks[:16]]
    victim_slot = next(i for i, p in enumerate(pairs) if p != 1)
    victim = ks[victim_slot]
    tb.insert([ks[16]], [116])
    assert where(tb, ks[16]) == [(1, 2, victim_slot)]
    vt = L.orc_get_pos(L.orc_pair(victim), 1)
    assert vt != 1
    assert where(tb, victim) == [(vt, L.orc_bucket(victim, vt, nb), 0)]
    assert tb.last_kicks[0] == 2
    assert tb.check_invariant() == 0


def test_rescue_lane_takes_over_after_kick_bound(oracle_lib):
    """With kick_bound 0 every full bucket picks orc_rescue_lane(key, iteration)."""
    L = O.lib()
    nb = 4
    ks = _keys_for_bucket(1, 2, nb, True, 17)
    tb = O.OracleTable(nb=[nb] * 4, kick_bound=0, rescue_kicks=8)
    tb.insert(ks[:16], list(range(16)))
    lane = L.orc_rescue_lane(ks[16], 0)
    evicted = ks[lane]
    tb.insert([ks[16]], [16])
    assert where(tb, ks[16]) == [(1, 2, lane)]
    assert len(where(tb, evicted)) == 1 and where(tb, evicted)[0][0] != 1


# ------------------------------------------------------------------------ bulk semantics
@pytest.mark.parametrize("n,theta,rescue", [(20000, 0.85, 0), (50000, 0.90, 128), (30000, 0.95, 128)])
def test_bulk_insert_find_erase_against_dict(oracle_lib, n, theta, rescue):
    keys = unique_keys(n, seed=n)
    vals = splitmix32(keys)
    tb = O.OracleTable(int(np.ceil(n / theta)), rescue_kicks=rescue, max_fail_rounds=8)
    nfail, rej = tb.insert(keys, vals)
    assert rej == 0
    if rescue:
        assert nfail == 0, "the rescue phase places every entry at these fills"
    assert tb.pending.size == 0
    assert tb.check_invariant() == 0
    assert np.array_equal(tb.entries(), np.sort(O.pack(keys, vals)))
    # find: present keys and absent keys (disjoint by construction)
    absent = unique_keys(n, seed=n + 1)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys[: n // 2], absent[: n // 2]])
    v, h = tb.find(q)
    assert h[: n // 2].all() and not h[n // 2:].any()
    assert np.array_equal(v[: n // 2], vals[: n // 2]) and not v[n // 2:].any()
    # every present key costs 1 or 2 probes, every absent key 2
    assert n // 2 * 2 <= tb.last_probes <= q.size * 2
    # erase half, twice: the second time nothing is deleted
    d = tb.erase(keys[: n // 3])
    assert d.all()
    assert not tb.erase(keys[: n // 3]).any()
    v, h = tb.find(keys)
    assert not h[: n // 3].any() and h[n // 3:].all()
    assert np.array_equal(v[n // 3:], vals[n // 3:])
    assert tb.count == n - n // 3


@pytest.mark.parametrize("threads", [1, 4])
def test_multithreaded_insert_holds_the_sequential_set(oracle_lib, threads):
    """The CPU baseline's concurrent insert stores exactly the key->value set of the sequential
    replay (placement may differ), keeps the placement invariant and rejects reserved keys."""
    n = 1 << 17
    keys = unique_keys(n, seed=5150)
    keys[[7, 99]] = [0, 0xFFFFFFFF]
    vals = splitmix32(keys)
    seq = O.OracleTable(int(np.ceil(n / 0.9)), rescue_kicks=128)
    seq.insert(keys, vals)
    mt = O.OracleTable(int(np.ceil(n / 0.9)), rescue_kicks=128)
    nfail, rej = mt.insert_mt(keys, vals, threads=threads)
    assert rej == 2 and mt.pending.size == 0 and mt.count == n - 2
    assert mt.check_invariant() == 0
    assert np.array_equal(mt.entries(), seq.entries())
    v, h = mt.find(keys)
    ok = (keys != 0) & (keys != 0xFFFFFFFF)
    assert h[ok].all() and not h[~ok].any() and np.array_equal(v[ok], vals[ok])
    if threads == 1:  # one thread takes keys in input order: the sequential image exactly
        assert all(np.array_equal(a, b) for a, b in zip(mt.sub, seq.sub))


def test_reference_kick_bound_alone_fails_at_high_fill(oracle_lib):
    """With the reference rule only (rescue 0) a fill of 0.90 leaves entries in hand -- the
    reference drops them (cuckoo.cu:422-431); the oracle keeps them and upsizes."""
    n = 60000
    keys = unique_keys(n, seed=3)
    vals = splitmix32(keys)
    tb = O.OracleTable(int(np.ceil(n / 0.9)), rescue_kicks=0, max_fail_rounds=8)
    nfail, _ = tb.insert(keys, vals)
    assert nfail > 0
    assert tb.resizes and tb.resizes[0][0] == "up"
    assert tb.pending.size == 0
    assert np.array_equal(tb.entries(), np.sort(O.pack(keys, vals)))


def test_reserved_keys(oracle_lib):
    tb = O.OracleTable(nb=[4] * 4)
    nfail, rej = tb.insert([0, 5, 0xFFFFFFFF, 6], [1, 2, 3, 4])
    assert rej == 2 and tb.count == 2
    v, h = tb.find([0, 5, 0xFFFFFFFF, 6])
    assert list(h) == [0, 1, 0, 1] and list(v) == [0, 2, 0, 4]
    assert list(tb.erase([0, 0xFFFFFFFF, 5])) == [0, 0, 1]


def test_duplicate_inserts_are_stored_twice_and_erase_clears_the_first_bucket(oracle_lib):
    """No upsert (dy_hash.h:42): the same key twice gives two slots; erase clears every copy in
    the first bucket holding it (cuckoo.cu:550-563)."""
    tb = O.OracleTable(nb=[6] * 4)
    tb.insert([11] * 5, [9] * 5)
    copies = where(tb, 11)
    assert len(copies) == 5 and len({(t, b) for t, b, _ in copies}) == 1
    v, h = tb.find([11])
    assert h[0] == 1 and v[0] == 9
    assert tb.erase([11])[0] == 1
    assert where(tb, 11) == []


# ------------------------------------------------------------------------ resize
def _filled(n, hint, seed):
    keys = unique_keys(n, seed=seed)
    vals = splitmix32(keys)
    tb = O.OracleTable(hint, rescue_kicks=128)
    tb.insert(keys, vals)
    assert tb.pending.size == 0
    return tb, keys, vals


def test_upsize_is_lane_preserving_into_j_or_j_plus_n(oracle_lib):
    tb, keys, vals = _filled(20000, 30000, 5)
    before = [x.copy() for x in tb.sub]
    nb0 = list(tb.nb)
    t = tb.resize_up()
    assert t == 0 and tb.nb[0] == 2 * nb0[0]  # api.cpp:116-122 with equal sizes
    n = nb0[t]
    old, new = before[t].reshape(n, 16), tb.sub[t].reshape(2 * n, 16)
    for j in range(n):
        for i in range(16):
            e = int(old[j, i])
            if e:
                lo, hi = int(new[j, i]), int(new[j + n, i])
                assert (lo == e) != (hi == e)
    assert np.count_nonzero(new) == np.count_nonzero(old)
    assert tb.check_invariant() == 0
    assert np.array_equal(tb.entries(), np.sort(O.pack(keys, vals)))


def test_downsize_compaction_and_residuals(oracle_lib):
    tb, keys, vals = _filled(12000, 40000, 6)
    before = [x.copy() for x in tb.sub]
    t, res = tb.resize_down(defer_residuals=True)
    assert t == 3  # api.cpp:147-153 with equal sizes
    n = tb.nb[t]
    old, new = before[t].reshape(2 * n, 16), tb.sub[t].reshape(n, 16)
    want_res = []
    for g in range(n):
        for i in range(16):
            lo, cross = int(old[g, i]), int(old[g + n, 15 - i])
            assert int(new[g, i]) == (lo if lo else cross)
            if lo and cross:
                want_res.append(cross)
    assert sorted(res.tolist()) == sorted(want_res)
    # every compacted entry is in its bucket: h % n == (h % 2n) % n
    placed = np.concatenate([tb.sub[0], tb.sub[1], tb.sub[2], tb.sub[3], res])
    assert np.array_equal(np.sort(placed[placed != 0]), np.sort(O.pack(keys, vals)))
    assert tb.check_invariant() == 0


def test_downsize_then_residual_insert_preserves_the_set(oracle_lib):
    tb, keys, vals = _filled(15000, 40000, 8)
    tb.resize_down()
    tb.resize_down()
    assert tb.nb[3] * 2 == tb.nb[0] and tb.nb[2] * 2 == tb.nb[0]
    assert tb.pending.size == 0 and tb.check_invariant() == 0
    assert np.array_equal(tb.entries(), np.sort(O.pack(keys, vals)))
    v, h = tb.find(keys)
    assert h.all() and np.array_equal(v, vals)


def test_downsize_refuses_odd_bucket_counts(oracle_lib):
    """Pairwise compaction needs an even bucket count (cuckoo.cu:653,767-769): h % 2n maps onto
    h % n only then."""
    assert O.OracleTable(nb=[5, 5, 5, 5]).resize_down() is None  # chooses 3, odd
    tb = O.OracleTable(nb=[6, 6, 6, 3])  # chooses 2 (first nb[i] > nb[i+1]), even
    t, res = tb.resize_down()
    assert t == 2 and tb.nb == [6, 6, 3, 3] and res.size == 0


def test_resize_choice_sequences(oracle_lib):
    """Growth order 0,1,2,3,0,... and shrink order 3,2,1,0,3,... (api.cpp:116-122,147-153)."""
    tb = O.OracleTable(nb=[2, 2, 2, 2])
    ups = [tb.resize_up() for _ in range(6)]
    assert ups == [0, 1, 2, 3, 0, 1]
    assert tb.nb == [8, 8, 4, 4]
    tb = O.OracleTable(nb=[16, 16, 16, 16])
    downs = [tb.resize_down()[0] for _ in range(5)]
    assert downs == [3, 2, 1, 0, 3]
    assert tb.nb == [8, 8, 8, 4]
    # an up after a down refills the smaller subtable first
    assert tb.resize_up() == 3


def test_auto_resize_policy(oracle_lib):
    """upsize while count + n > upper * cap before an insert; one downsize after an erase that
    leaves count < lower * cap (api.cpp:175-181,353-369,490-505)."""
    tb = O.OracleTable(nb=[8, 8, 8, 8], auto_resize=True, upper=0.8, lower=0.5, min_capacity=0,
                       rescue_kicks=128)
    cap0 = tb.capacity  # 512 slots
    keys = unique_keys(2000, seed=11)
    vals = splitmix32(keys)
    tb.insert(keys[:400], vals[:400])  # 400 <= 409.6: no resize
    assert tb.capacity == cap0
    tb.insert(keys[400:500], vals[400:500])  # 500 > 409.6 -> up once (640 * .8 = 512)
    assert [r[0] for r in tb.resizes] == ["up"] and tb.capacity == 640
    tb.erase(keys[:250])  # 250 < 320 -> down once
    assert [r[0] for r in tb.resizes] == ["up", "down"]
    v, h = tb.find(keys[:500])
    assert not h[:250].any() and h[250:].all() and np.array_equal(v[250:], vals[250:500])


@pytest.mark.parametrize("auto", [False, True])
def test_batch_api_resize_policy(oracle_lib, auto):
    """hash_insert_batch bounds the fill with the full BATCH_SIZE whatever the slice holds
    (api.cpp:359-369); hash_delete_batch downsizes once under the lower bound (api.cpp:494-505),
    also when auto_resize applies the same rule inside the erase. Worked by hand: 4 x 8 buckets =
    512 slots, upper 0.8 -> 409.6; batch size 100."""
    tb = O.OracleTable(nb=[8, 8, 8, 8], auto_resize=auto, upper=0.8, lower=0.5, min_capacity=0,
                       rescue_kicks=128)
    keys = unique_keys(400, seed=12)
    vals = splitmix32(keys)
    sizes = []
    lo = 0
    for m in (100, 100, 100, 5, 10, 10):  # slices: the last three are short
        tb.insert_batch(keys[lo:lo + m], vals[lo:lo + m], 100)
        lo += m
        sizes.append(list(tb.nb))
    # 300 + 100 = 400 and 305 + 100 = 405 fit under 409.6; 315 + 100 = 415 does not (a slice-sized
    # bound, 315 + 10, would not have grown the table)
    assert sizes == [[8, 8, 8, 8]] * 5 + [[16, 8, 8, 8]]
    assert tb.count == 325
    tb.erase_batch(keys[:100])  # 225 < 0.5 * 640 -> down once (choose_down picks subtable 0)
    assert tb.nb == [8, 8, 8, 8] and tb.count == 225  # 225 < 256 still, but only one downsize
    tb.erase_batch(keys[100:200])  # 125 < 256 -> down: all equal, so the last subtable
    assert tb.nb == [8, 8, 8, 4]
    assert [r[0] for r in tb.resizes] == ["up", "down", "down"]
    v, h = tb.find(keys[:325])
    assert not h[:200].any() and h[200:].all() and np.array_equal(v[200:], vals[200:325])


def test_stash_answers_keys_missing_from_both_buckets(oracle_lib):
    """The pending list is the table's stash: 40 copies of one key fill its 32 bucket slots and
    leave 8 pending; after two erases clear its pos1 and pos2 buckets (the first bucket holding
    it each time, cuckoo.cu:541-563) the key is still found in the stash, and a third erase
    clears every stash copy."""
    tb = O.OracleTable(nb=[2, 2, 2, 2], max_fail_rounds=0, rescue_kicks=16)
    vals = np.arange(100, 140, dtype=np.uint32)
    tb.insert(np.full(40, 11, np.uint32), vals)
    assert tb.pending.size == 8 and tb.count == 40
    stash_vals = set((tb.pending & np.uint64(0xFFFFFFFF)).tolist())
    assert len(stash_vals) == 8 and stash_vals <= set(vals.tolist())
    for _ in range(2):
        assert tb.erase([11]).tolist() == [1]
    assert tb.count == 40 - 2 and sum(int(np.count_nonzero(x)) for x in tb.sub) == 0
    v, h = tb.find([11, 12])
    assert h.tolist() == [1, 0] and int(v[0]) == int(tb.pending[0] & np.uint64(0xFFFFFFFF))
    # a later insert that places all its own keys neither grows the table nor retries the stash
    tb.max_fail_rounds = 3
    tb.insert([12, 13], [5, 6])
    assert tb.nb == [2, 2, 2, 2] and tb.pending.size == 8 and tb.find([12, 13])[1].tolist() == [1, 1]
    assert tb.erase([12, 13]).tolist() == [1, 1]
    # two identical keys in one erase batch: the first clears the stash, the second finds nothing
    assert tb.erase([11, 11]).tolist() == [1, 0]
    assert tb.pending.size == 0 and tb.count == 40 - 3
    assert tb.find([11])[1].tolist() == [0]
