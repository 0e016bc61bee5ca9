"""Pins of the oracle's wide-key table (uint64 keys and values).

The wide table is the 32-bit table applied to fold(k) = lo ^ h5(hi). h5(0) = 0, so for keys
below 2^32 it must place every entry exactly where the (separately pinned) 32-bit oracle does:
that equivalence, checked slot by slot through evictions, the rescue walk and resizes for every
subtable count, pins the wide insert, find, erase and resize logic. Keys with high bits are
then checked by the placement invariant, a dict of the inserted set, and set preservation.
"""
import numpy as np
import pytest

import oracle as O

DS = (2, 3, 4, 5, 6)
M32 = np.uint64(0xFFFFFFFF)


def keys32(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**32 - 1, size=int(n * 1.3) + 16, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(out)
    return out[:n]


def keys64(n, seed):
    rng = np.random.default_rng(seed)
    out = np.unique(rng.integers(1, 2**64 - 2, size=int(n * 1.3) + 16, dtype=np.uint64))
    rng.shuffle(out)
    return out[:n]


def narrow_as_wide(tb):
    """(keys, values) images of a 32-bit oracle table"""
    return [x >> np.uint64(32) for x in tb.sub], [x & M32 for x in tb.sub]


def test_fold_is_lo_xor_h5_of_hi(oracle_lib):
    k = keys64(5000, 1)
    lo = (k & M32).astype(np.uint32)
    hi = (k >> np.uint64(32)).astype(np.uint32)
    assert np.array_equal(O.fold64(k), lo ^ O.hash_batch(hi, 5))
    small = keys32(500, 2).astype(np.uint64)
    assert np.array_equal(O.fold64(small), small.astype(np.uint32))  # h5(0) = 0
    assert int(O.hash_batch(np.array([0], np.uint32), 5)[0]) == 0


@pytest.mark.parametrize("d", DS)
def test_keys_below_2_32_place_exactly_as_the_32_bit_table(oracle_lib, d):
    """Same key sequence (values below 2^32) into both tables at fill 0.95 with the reference
    eviction bound and the rescue walk: identical images slot for slot, also after two upsizes
    and a downsize, and identical pending entries."""
    n = 6000
    keys = keys32(n, 10 + d)
    vals = keys ^ np.uint32(0x5A5A5A5A)
    hint = int(n / 0.95)
    nt = O.OracleTable(hint, subtables=d, rescue_kicks=64, max_fail_rounds=0)
    wt = O.OracleWideTable(hint, subtables=d, rescue_kicks=64, max_fail_rounds=0)
    nt.insert(keys, vals)
    wt.insert(keys.astype(np.uint64), vals.astype(np.uint64))
    assert wt.nb == nt.nb
    nk, nv = narrow_as_wide(nt)
    assert all(np.array_equal(a, b) for a, b in zip(wt.keys, nk))
    assert all(np.array_equal(a, b) for a, b in zip(wt.vals, nv))
    assert np.array_equal(wt.pend_keys, nt.pending >> np.uint64(32))
    for op in ("up", "up", "down"):
        getattr(nt, "resize_" + op)()
        getattr(wt, "resize_" + op)()
        nk, nv = narrow_as_wide(nt)
        assert wt.nb == nt.nb
        assert all(np.array_equal(a, b) for a, b in zip(wt.keys, nk)), op
        assert all(np.array_equal(a, b) for a, b in zip(wt.vals, nv)), op
    q = np.concatenate([keys[:1000], keys32(500, 99)])
    nv_, nh_ = nt.find(q)
    wv_, wh_ = wt.find(q.astype(np.uint64))
    assert np.array_equal(nh_, wh_) and np.array_equal(nv_.astype(np.uint64), wv_)
    e = keys[::7]
    assert np.array_equal(nt.erase(e), wt.erase(e.astype(np.uint64)))
    nk, nv = narrow_as_wide(nt)
    assert all(np.array_equal(a, b) for a, b in zip(wt.keys, nk))


@pytest.mark.parametrize("d", (2, 4, 6))
def test_wide_keys_against_a_dict(oracle_lib, d):
    n = 10000
    keys = keys64(n, 30 + d)
    vals = keys * np.uint64(3) + np.uint64(1)
    tb = O.OracleWideTable(int(n / 0.9), subtables=d, rescue_kicks=128)
    nfail, rej = tb.insert(keys, vals)
    assert rej == 0 and tb.pend_keys.size == 0 and tb.check_invariant() == 0
    k, v = tb.entries()
    order = np.argsort(keys)
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    absent = keys64(3000, 31 + d)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys[:4000], absent])
    fv, fh = tb.find(q)
    assert fh[:4000].all() and not fh[4000:].any() and np.array_equal(fv[:4000], vals[:4000])
    assert tb.erase(keys[:2500]).all() and not tb.erase(keys[:2500]).any()
    assert tb.count == n - 2500
    for _ in range(2):
        tb.resize_up()
    assert tb.resize_down() is not None
    assert tb.check_invariant() == 0
    k, v = tb.entries()
    rest = np.argsort(keys[2500:])
    assert np.array_equal(k, keys[2500:][rest]) and np.array_equal(v, vals[2500:][rest])


def test_reserved_wide_keys(oracle_lib):
    tb = O.OracleWideTable(1000)
    nfail, rej = tb.insert(np.array([0, 2**64 - 1, 5], np.uint64), np.array([1, 2, 3], np.uint64))
    assert rej == 2 and tb.count == 1
    v, h = tb.find(np.array([0, 2**64 - 1, 5], np.uint64))
    assert h.tolist() == [0, 0, 1] and int(v[2]) == 3


def test_wide_downsize_keeps_one_copy_of_an_entry_mirrored_in_both_halves(oracle_lib):
    """The reference compares the entry placed in a slot with its mirrored partner
    (cuckoo.cu:685-688): an identical (key, value) in both halves leaves one copy; a partner that
    differs in its value alone is a residual."""
    nb = 4
    keys = [np.zeros(nb * 16, np.uint64) for _ in range(4)]
    vals = [np.zeros(nb * 16, np.uint64) for _ in range(4)]
    t = 3  # the subtable a downsize of equal sizes chooses (api.cpp:147-153)
    keys[t][0 * 16 + 0] = 77        # bucket 0 slot 0
    vals[t][0 * 16 + 0] = 1
    keys[t][2 * 16 + 15] = 77       # bucket 2 slot 15: mirrors bucket 0 slot 0
    vals[t][2 * 16 + 15] = 1
    keys[t][1 * 16 + 3] = 88
    vals[t][1 * 16 + 3] = 1
    keys[t][3 * 16 + 12] = 88       # mirrors bucket 1 slot 3, different value
    vals[t][3 * 16 + 12] = 2
    tb = O.OracleWideTable(nb=[nb] * 4)
    tb.load_images(keys, vals)
    t2, rk, rv = tb.resize_down(defer_residuals=True)
    assert t2 == t and rk.tolist() == [88] and rv.tolist() == [2]
    assert tb.keys[t][0] == 77 and tb.keys[t][16 + 3] == 88
