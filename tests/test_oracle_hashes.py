"""Pins of the oracle's per-key arithmetic (hashes, bucket index, pair table, sizing).

None of these checks re-types the oracle's formula. The hashes are checked against inverse
functions derived independently from each step (every step of h1, h2, h4, h5 is a bijection on
32-bit words: odd multipliers, xor-shifts, and the add-xor step of h2 solved bit by bit), so a
dropped step, a wrong shift or a wrong constant in the oracle makes inverse(h(k)) != k. h3, the
modulus and the pair table are closed forms checked with Python integers and hand-derived tables.
"""
import numpy as np
import pytest

import oracle as O

M32 = (1 << 32) - 1
EDGE = np.array([1, 2, 3, 11, 0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFE, 0xFFFFFFFD, 59064253,
                 0x12345678, 0xDEADBEEF], dtype=np.uint32)


def _keys(n=100_000, seed=7):
    rng = np.random.default_rng(seed)
    return np.concatenate([EDGE, rng.integers(1, 2**32 - 1, size=n, dtype=np.uint64).astype(np.uint32)])


# ------------------------------------------------------------------ modular/xorshift inverses
def _inv(a):
    return pow(a % (1 << 32), -1, 1 << 32)


def _mul(x, c):
    return (x.astype(np.uint64) * np.uint64(c % (1 << 32))) & np.uint64(M32)


def _add(x, c):
    return (x.astype(np.uint64) + np.uint64(c % (1 << 32))) & np.uint64(M32)


def _unxsr(y, s):
    """inverse of x ^ (x >> s)"""
    y = y.astype(np.uint64)
    x = y.copy()
    for _ in range(32 // s + 1):
        x = y ^ (x >> np.uint64(s))
    return x & np.uint64(M32)


def _unxsl(y, s):
    """inverse of x ^ (x << s) (mod 2^32)"""
    y = y.astype(np.uint64)
    x = y.copy()
    for _ in range(32 // s + 1):
        x = (y ^ (x << np.uint64(s))) & np.uint64(M32)
    return x


def _unaddxor(y, c, s):
    """inverse of (x + c) ^ (x << s), solved bit by bit from bit 0 upwards"""
    y = y.astype(np.uint64)
    c = np.uint64(c)
    x = np.zeros_like(y)
    for i in range(32):
        low = np.uint64((1 << i) - 1)
        carry = (((x & low) + (c & low)) >> np.uint64(i)) & np.uint64(1)
        prev = (x >> np.uint64(i - s)) & np.uint64(1) if i >= s else np.zeros_like(y)
        bit = ((y >> np.uint64(i)) ^ prev ^ (c >> np.uint64(i)) ^ carry) & np.uint64(1)
        x |= bit << np.uint64(i)
    return x


def inv_h1(y):
    # steps: k*(2^15-1) - 1 ; xsr12 ; *5 ; xsr4 ; *2057 ; xsr16
    x = _unxsr(y, 16)
    x = _mul(x, _inv(2057))
    x = _unxsr(x, 4)
    x = _mul(x, _inv(5))
    x = _unxsr(x, 12)
    x = _mul(_add(x, 1), _inv((1 << 15) - 1))
    return x


def inv_h2(y):
    # steps: *4097 + C1 ; (^C2) xsr19 ; *33 + C3 ; (x + C4) ^ (x << 9) ; *9 + C5 ; (^C6) xsr16
    x = _unxsr(y.astype(np.uint64) ^ np.uint64(0xB55A4F09), 16)
    x = _mul(_add(x, -0xFD7046C5), _inv(9))
    x = _unaddxor(x, 0xD3A2646C, 9)
    x = _mul(_add(x, -0x165667B1), _inv(33))
    x = _unxsr(x ^ np.uint64(0xC761C23C), 19)
    x = _mul(_add(x, -0x7ED55D16), _inv(4097))
    return x


def inv_h4(y):
    # steps: (^61) xsr16 ; *9 ; xsr4 ; *0x27d4eb2d ; xsr15
    x = _unxsr(y, 15)
    x = _mul(x, _inv(0x27D4EB2D))
    x = _unxsr(x, 4)
    x = _mul(x, _inv(9))
    x = _unxsr(x, 16) ^ np.uint64(61)
    return x


def inv_h5(y):
    # steps: *(1-2^6) ; xsr17 ; *(1-2^9) ; xsl4 ; *(1-2^3) ; xsl10 ; xsr15
    x = _unxsr(y, 15)
    x = _unxsl(x, 10)
    x = _mul(x, _inv(1 - 8))
    x = _unxsl(x, 4)
    x = _mul(x, _inv(1 - 512))
    x = _unxsr(x, 17)
    x = _mul(x, _inv(1 - 64))
    return x


@pytest.mark.parametrize("which,inv", [(1, inv_h1), (2, inv_h2), (4, inv_h4), (5, inv_h5)])
def test_hash_is_the_bijection_its_steps_define(oracle_lib, which, inv):
    k = _keys()
    h = O.hash_batch(k, which)
    back = inv(h).astype(np.uint32)
    assert np.array_equal(back, k), f"h{which} does not invert"


def test_inverse_checks_detect_a_planted_mistake(oracle_lib):
    # a wrong shift in the forward hash must break the inverse: plant it on the CPU side
    k = _keys(2000)
    wrong = k.astype(np.uint64)
    wrong = (~wrong + (wrong << np.uint64(14))) & np.uint64(M32)  # 14 instead of 15
    wrong = wrong ^ (wrong >> np.uint64(12))
    wrong = (wrong * np.uint64(5)) & np.uint64(M32)
    wrong = wrong ^ (wrong >> np.uint64(4))
    wrong = (wrong * np.uint64(2057)) & np.uint64(M32)
    wrong = wrong ^ (wrong >> np.uint64(16))
    assert not np.array_equal(inv_h1(wrong).astype(np.uint32), k)


def test_h3_closed_form(oracle_lib):
    k = _keys(20000)
    got = O.hash_batch(k, 3)
    want = np.array([(((int(x) ^ 59064253) + 72355969) & M32) % 294967291 for x in k], dtype=np.uint32)
    assert np.array_equal(got, want)
    # hand values: the xor cancels; and a wrap past 2^32 before the modulus
    L = O.lib()
    assert L.orc_h3(59064253) == 72355969
    k_wrap = 0xFFFFFFFF ^ 59064253  # k ^ C = 0xFFFFFFFF, + D wraps to D - 1
    assert L.orc_h3(k_wrap) == (72355969 - 1) % 294967291


def test_subtable_hash_map_and_modulus(oracle_lib):
    """bucket(k, t, nb) = h_{t+1}(k) % 294967291 % nb (cuckoo.cu:80-89, dy_hash.h:24)."""
    L = O.lib()
    k = _keys(3000)
    hs = {t: O.hash_batch(k, t + 1) for t in range(4)}
    for nb in (2, 7, 1000, 36424, 4660218):
        for t in range(4):
            for x, h in zip(k[:600].tolist(), hs[t][:600].tolist()):
                assert L.orc_bucket(x, t, nb) == (h % 294967291) % nb
    assert O.PRIME == 294967291  # not 4294967291 (dy_hash.h:24)


def test_pair_table_exhaustive(oracle_lib):
    """pairs 0..5 = (1,2) (0,1) (2,3) (0,2) (1,3) (0,3): read off cuckoo.cu:196-213 by hand."""
    L = O.lib()
    want = {0: (1, 2), 1: (0, 1), 2: (2, 3), 3: (0, 2), 4: (1, 3), 5: (0, 3)}
    got = {p: (L.orc_pos1(p), L.orc_pos2(p)) for p in range(6)}
    assert got == want
    assert sorted(tuple(sorted(v)) for v in got.values()) == sorted(
        (a, b) for a in range(4) for b in range(a + 1, 4))
    for p, (a, b) in want.items():
        assert L.orc_get_pos(p, a) == b and L.orc_get_pos(p, b) == a
        assert L.orc_get_pos(p, O.NO_TABLE) == a  # fresh entries go to pos1 (cuckoo.cu:310)
    k = _keys(50000)
    pairs = O.hash_batch(k, 6)
    assert np.array_equal(pairs, O.hash_batch(k, 5) % 6)
    counts = np.bincount(pairs, minlength=6)
    assert counts.min() > 0.9 * counts.mean()  # every pair is used


def _sieve_next_prime(limit):
    n = limit + 200
    is_p = np.ones(n, dtype=bool)
    is_p[:2] = False
    for i in range(2, int(n**0.5) + 1):
        if is_p[i]:
            is_p[i * i::i] = False
    is_p[1] = True  # the reference's 6k+-1 test accepts 1 (api.cpp:4-11)
    nxt = np.empty(limit, dtype=np.int64)
    j = n - 1
    for i in range(n - 1, -1, -1):
        if is_p[i]:
            j = i
        if i < limit:
            nxt[i] = j
    return nxt


def test_next_prime_matches_sieve(oracle_lib):
    L = O.lib()
    nxt = _sieve_next_prime(20000)
    for i in range(20000):
        assert L.orc_next_prime(i) == nxt[i], i
    for big in (1048573, 4194301, 16777213, 4660217, 2**27 + 29):
        p = L.orc_next_prime(big)
        assert all(p % d for d in range(2, int(p**0.5) + 1))
        assert all(any(q % d == 0 for d in range(2, int(q**0.5) + 1)) for q in range(big, p))


def test_initial_buckets(oracle_lib):
    """api.cpp:47-56: s = nextPrime(hint/64), made even by +1 (prime from the sieve above)."""
    nxt = _sieve_next_prime(70000)
    for hint in (0, 1, 63, 64, 100, 640, 1 << 20, int(2**20 / 0.85), 4 * 10**6):
        s = int(nxt[hint // 64])
        want = s + 1 if s % 2 else s
        assert O.initial_buckets(hint) == want, hint
        assert want % 2 == 0
    assert O.initial_buckets(0) == 2 and O.initial_buckets(640) == 12  # 0 -> 1 -> 2; 10 -> 11 -> 12
