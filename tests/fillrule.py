"""The bulk fill's first-round target (Desc::fr and fill_target in dyc_device.hpp), restated for
the tests over the oracle's pair tables: per pair, all keys to pos1 (code 1), all to pos2 (code
2), or the start rule (code 0: odd keys pos2, cuckoo.cu:336-341); the first assignment a
depth-first search meets, trying pos1, the start rule, then pos2 for each pair in order, among
those that keep every subtable's share equal and send the most keys to pos1."""
import numpy as np

import oracle as O


def rule(d):
    pairs = O.pair_table(d)
    np_ = len(pairs)
    share = 2 * np_ // d
    load = [0] * d
    best = {"score": -1, "codes": None}
    codes = [0] * np_

    def go(i, score):
        if best["codes"] is not None and score + 2 * (np_ - i) <= best["score"]:
            return
        if i == np_:
            best["score"], best["codes"] = score, codes[:]
            return
        a, b = pairs[i]
        for x, c in ((2, 1), (1, 0), (0, 2)):
            load[a] += x
            load[b] += 2 - x
            if load[a] <= share and load[b] <= share:
                codes[i] = c
                go(i + 1, score + x)
            load[a] -= x
            load[b] -= 2 - x

    go(0, 0)
    return sum(c << (2 * p) for p, c in enumerate(best["codes"]))


def targets(keys, d):
    """The subtable round 1 of the bulk fill sends each key to."""
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    pairs = np.array(O.pair_table(d), dtype=np.int64)
    p = (k & 1).astype(np.int64) if d == 2 else O.hash_batch(k, 5).astype(np.int64) % len(pairs)
    r = (rule(d) >> (2 * p)) & 3
    second = (r == 2) | ((r == 0) & ((k & 1) == 1))
    return np.where(second, pairs[p, 1], pairs[p, 0])
