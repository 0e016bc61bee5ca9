"""Pins of the oracle's other subtable counts d = 2, 3, 5, 6 (diff_hash_num_pos_fun.cpp:14-188;
d = 4 is pinned in test_oracle_hashes.py / test_oracle_table.py).

Checks: each pair table lists every pair of distinct subtables exactly once (d = 2: both orders
of the one pair), read off the reference's comments by hand; the pair index is the pinned h5
reduced mod the pair count (k & 1 for d = 2); first placements and the victim vote worked by
hand on one-bucket tables; the placement invariant, a dict of the inserted set, resize choice
sequences and set preservation across resizes for every d.
"""
import itertools

import numpy as np
import pytest

import oracle as O

NOT = O.NO_TABLE
DS = (2, 3, 4, 5, 6)


def unique_keys(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**32 - 1, size=int(n * 1.3) + 16, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(out)
    return out[:n]


def where(table, key):
    out = []
    for t in range(table.d):
        idx = np.nonzero((table.sub[t] >> np.uint64(32)) == np.uint64(key))[0]
        out += [(t, int(i) // 16, int(i) % 16) for i in idx]
    return out


# ------------------------------------------------------------------------------- pair tables
# read off by hand: d = 2 diff_hash_num_pos_fun.cpp:21-29 (pos1 = p, pos2 = p ^ 1); d = 3 the
# shift form :43-55 (comments "xx1 12", "01 02"); d = 4 cuckoo.cu:196-213; d = 5 :126-146
# (comments "xx1 01 02 03 04", "23 24", "12 13 14", and pos 9 = (3,4)); d = 6 the switch of
# :158-188, pairs in lexicographic order
HAND = {
    2: [(0, 1), (1, 0)],
    3: [(0, 1), (1, 2), (0, 2)],
    4: [(1, 2), (0, 1), (2, 3), (0, 2), (1, 3), (0, 3)],
    5: [(1, 2), (0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (2, 4), (0, 4), (1, 4), (3, 4)],
    6: list(itertools.combinations(range(6), 2)),
}


@pytest.mark.parametrize("d", DS)
def test_pair_table_matches_the_reference_by_hand(oracle_lib, d):
    assert O.pair_table(d) == HAND[d]


@pytest.mark.parametrize("d", DS)
def test_pair_table_lists_every_pair_once(oracle_lib, d):
    """Each unordered pair of distinct subtables appears exactly once (twice, in both orders,
    for d = 2), and get_pos swaps between the two members."""
    L = O.lib()
    tab = O.pair_table(d)
    assert len(tab) == (2 if d == 2 else d * (d - 1) // 2)
    assert all(a != b and 0 <= a < d and 0 <= b < d for a, b in tab)
    unordered = sorted(tuple(sorted(x)) for x in tab)
    if d == 2:
        assert unordered == [(0, 1), (0, 1)]
    else:
        assert unordered == list(itertools.combinations(range(d), 2))
    for p, (a, b) in enumerate(tab):
        assert L.orc_get_pos_d(p, a, d) == b and L.orc_get_pos_d(p, b, d) == a
        assert L.orc_get_pos_d(p, NOT, d) == a  # cuckoo.cu:310: a fresh entry goes to pos1


def test_pair_table_check_detects_a_planted_mistake(oracle_lib):
    """The every-pair-once property rejects a table with one entry changed."""
    bad = list(HAND[5])
    bad[6] = (0, 4)  # duplicates pair 7
    assert sorted(tuple(sorted(x)) for x in bad) != list(itertools.combinations(range(5), 2))


@pytest.mark.parametrize("d", DS)
def test_pair_index_is_h5_mod_pairs(oracle_lib, d):
    """choose_hash: k & 1 for d = 2 (diff_hash_num_pos_fun.cpp:16-19), else h5(k) % C(d,2)
    (h5 itself is pinned by its inverse); every pair index is used."""
    k = unique_keys(40000, 5 + d)
    got = O.pair_of(k, d)
    if d == 2:
        want = k & np.uint32(1)
    else:
        want = O.hash_batch(k, 5) % np.uint32(d * (d - 1) // 2)
    assert np.array_equal(got, want)
    counts = np.bincount(got, minlength=len(HAND[d]))
    assert counts.min() > 0.85 * counts.mean()


def test_subtables_past_four_reuse_h1_and_h2(oracle_lib):
    """hash_function switches on num_table % 4 (cuckoo.cu:76-90)."""
    L = O.lib()
    for k in unique_keys(500, 9).tolist():
        assert L.orc_hash_t(k, 4) == L.orc_hash_t(k, 0) and L.orc_hash_t(k, 5) == L.orc_hash_t(k, 1)


# ------------------------------------------------------------------------- hand placements
def test_two_subtables_start_in_subtable_0_and_evict_slot_0(oracle_lib):
    """d = 2: even keys start in pos1(0) = 0 and odd keys in pos2(1) = 0, so every key is first
    tried in subtable 0; a full bucket gives up slot 0 (the __ballot(true) vote of
    diff_hash_num_pos_fun.cpp:3-6), and the victim moves to subtable 1."""
    tb = O.OracleTable(nb=[1, 1])
    ks = list(range(1, 18))  # odd and even
    tb.insert(ks[:16], [100 + k for k in ks[:16]])
    for i, k in enumerate(ks[:16]):
        assert where(tb, k) == [(0, 0, i)]
    tb.insert([ks[16]], [117])
    assert where(tb, ks[16]) == [(0, 0, 0)]
    assert where(tb, ks[0]) == [(1, 0, 0)]
    assert tb.last_kicks[0] == 2
    # a second overflow evicts the new slot-0 resident again
    tb.insert([18], [118])
    assert where(tb, 18) == [(0, 0, 0)] and where(tb, 17) == [(1, 0, 1)]
    assert tb.check_invariant() == 0


def _first_table(L, k, d):
    p = L.orc_pair_d(k, d)
    return L.orc_pos1_d(p, d) if k % 2 == 0 else L.orc_pos2_d(p, d)


@pytest.mark.parametrize("d", (3, 5, 6))
def test_fill_then_vote_by_hand(oracle_lib, d):
    """One-bucket subtables: 16 even keys that start in subtable t take slots 0..15; the 17th
    evicts the first resident whose pair index differs from t (cuckoo.cu:398-400), which moves
    to the other subtable of its pair."""
    L = O.lib()
    t = 1
    ks, k = [], 2
    while len(ks) < 17:
        if _first_table(L, k, d) == t:
            ks.append(k)
        k += 2
    tb = O.OracleTable(nb=[1] * d)
    tb.insert(ks[:16], list(range(16)))
    for i, key in enumerate(ks[:16]):
        assert where(tb, key) == [(t, 0, i)]
    pairs = [L.orc_pair_d(x, d) for x in ks[:16]]
    slot = next((i for i, p in enumerate(pairs) if p != t), 0)
    tb.insert([ks[16]], [16])
    assert where(tb, ks[16]) == [(t, 0, slot)]
    victim = ks[slot]
    vt = L.orc_get_pos_d(L.orc_pair_d(victim, d), t, d)
    assert vt != t and where(tb, victim) == [(vt, 0, 0)]


# ------------------------------------------------------------------------------ bulk, resize
@pytest.mark.parametrize("d", DS)
def test_bulk_insert_find_erase_against_dict(oracle_lib, d):
    n = 12000
    keys = unique_keys(n, 100 + d)
    vals = keys * np.uint32(3) + np.uint32(4)
    tb = O.OracleTable(int(np.ceil(n / 0.85)), subtables=d, rescue_kicks=128)
    assert tb.d == d and len(tb.nb) == d
    nfail, rej = tb.insert(keys, vals)
    assert rej == 0 and nfail == 0 and tb.pending.size == 0
    assert tb.check_invariant() == 0
    assert np.array_equal(tb.entries(), np.sort(O.pack(keys, vals)))
    v, h = tb.find(keys)
    assert h.all() and np.array_equal(v, vals)
    assert tb.erase(keys[: n // 2]).all()
    v, h = tb.find(keys)
    assert not h[: n // 2].any() and h[n // 2:].all()


@pytest.mark.parametrize("d", DS)
def test_resize_choice_sequences(oracle_lib, d):
    """Growth runs 0, 1, ..., d-1, 0, ... and shrink d-1, ..., 0, d-1 (api.cpp:116-122,147-153
    with TABLE_NUM = d)."""
    tb = O.OracleTable(nb=[4] * d)
    assert [tb.resize_up() for _ in range(d + 1)] == list(range(d)) + [0]
    tb = O.OracleTable(nb=[16] * d)
    assert [tb.resize_down()[0] for _ in range(d + 1)] == list(range(d - 1, -1, -1)) + [d - 1]


@pytest.mark.parametrize("d", DS)
def test_resizes_preserve_the_set(oracle_lib, d):
    n = 6000
    keys = unique_keys(n, 200 + d)
    vals = keys ^ np.uint32(0x5A5A5A5A)
    tb = O.OracleTable(int(n / 0.8), subtables=d, rescue_kicks=128)
    tb.insert(keys, vals)
    want = np.sort(O.pack(keys, vals))
    for _ in range(2):
        tb.resize_up()
        assert tb.check_invariant() == 0 and np.array_equal(tb.entries(), want)
    for _ in range(2):
        assert tb.resize_down() is not None
        assert tb.check_invariant() == 0 and np.array_equal(tb.entries(), want)


def test_initial_buckets_per_d(oracle_lib):
    """api.cpp:47-51 divides the hint by TABLE_NUM * 16: nextPrime(hint / (16 d)), made even."""
    for d, want in ((2, 32), (3, 24), (4, 18), (5, 14), (6, 12)):
        # 960 / (16 d) = 30, 20, 15, 12, 10 -> next primes 31, 23, 17, 13, 11 -> even
        assert O.initial_buckets(960, d) == want, d
