"""The large bulk workload on the GPU: 4 subtables, unique u32 keys, load factor 0.90, values
splitmix32(key), then finds at 50 % hit drawn uniformly and finds drawn Zipf(0.99) over the
inserted keys -- with the launch configuration the bulk benchmark runs (bulk-fill insert,
insert_order 3; find 8 lanes per key, insert 2 lanes, 65536 insert blocks, eviction bound 0).

* 2^28 keys (the full size): every uniform lookup hits iff its key was inserted, with the
  inserted value, and misses return 0; every Zipf lookup hits with its value; the device count of
  occupied slots is 2^28; erasing the uniform hits deletes exactly those keys (cuckoo.cu:460-520
  find, :524-587 delete). These are size-independent properties, checked on the device.
* 2^22 and 2^24 keys against the oracle: the dumped table's multiset equals the batch, every
  entry sits in a bucket of its pair (the placement invariant), find values and hit flags for
  both lookup distributions equal the oracle's find over the GPU's own image and the oracle's own
  replay of the insert, and the buckets the pos1-then-pos2 order reads (the count the
  algorithmic bytes of a find are computed from) equal the oracle's count over the same image.
"""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

BENCH_SHAPE = dict(find_lanes=8, insert_lanes=2, insert_blocks=65536, kick_bound=0, insert_order=3)


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def D():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import databases_gpu_amd as D

    D.load()
    return D


def u32(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def test_large_bulk_2p28_uniform_and_zipf(D):
    from databases_gpu_amd import workload as W

    n, seed = 1 << 28, 1
    banner("2^28 keys, theta 0.90, bulk fill; uniform 50% and Zipf(0.99) finds")
    dev = torch.device("cuda")
    keys = W.keys_range(0, n, seed, dev)
    vals = W.splitmix32(keys)
    t = D.DyCuckoo(int(np.ceil(n / 0.90)), **BENCH_SHAPE)
    t.insert(keys, vals)
    st = t.stats()
    assert st["count"] == n and st["pending"] == 0 and st["lost"] == 0 and st["rejected"] == 0, st
    assert t.count_occupied() == n
    out_v = torch.empty(n, dtype=torch.int32, device=dev)
    out_h = torch.empty(n, dtype=torch.uint8, device=dev)

    look = W.mixed_lookups(keys, n, 0.5, seed, n, seed + 7)
    v, h = t.find(look, out_v, out_h)
    present = torch.isin(look, keys)
    assert int(present.sum()) == n // 2
    assert bool((h.bool() == present).all()), "hit flags differ from membership"
    assert bool((v[present] == W.splitmix32(look[present])).all()), "values of hits"
    assert not bool(v[~present].bool().any()), "a miss returned a nonzero value"
    del v, h

    zl = W.zipf_lookups(keys, n, 0.99, seed + 11)
    v, h = t.find(zl, out_v, out_h)
    assert bool(h.bool().all()), "a Zipf lookup of an inserted key missed"
    assert bool((v == W.splitmix32(zl)).all()), "values of Zipf lookups"
    del v, h, zl

    d = t.erase(look)
    assert bool((d.bool() == present).all()), "deleted flags differ from membership"
    st = t.stats()
    assert st["count"] == n - n // 2
    _, h = t.find(look, out_v, out_h)
    assert not bool(h.bool().any())
    t.close()


@pytest.mark.parametrize("log2n", [22, 24])
def test_large_bulk_against_oracle(D, log2n):
    from databases_gpu_amd import workload as W

    n, seed = 1 << log2n, 3
    banner(f"2^{log2n} keys, theta 0.90, bulk fill, uniform + Zipf finds vs the oracle")
    dev = torch.device("cuda")
    keys = W.keys_range(0, n, seed, dev)
    vals = W.splitmix32(keys)
    hint = int(np.ceil(n / 0.90))
    t = D.DyCuckoo(hint, **BENCH_SHAPE)
    t.insert(keys, vals)
    st = t.stats()
    assert st["count"] == n and st["pending"] == 0 and st["lost"] == 0

    hk, hv = u32(keys), u32(vals)
    img = t.dump()
    g = O.OracleTable(nb=[x.size // 16 for x in img])
    g.sub = [np.ascontiguousarray(x).copy() for x in img]
    assert g.check_invariant() == 0
    assert np.array_equal(g.entries(), np.sort(O.pack(hk, hv)))

    ref = O.OracleTable(hint, rescue_kicks=128)
    ref.insert(hk, hv)
    assert ref.nb == list(t.buckets) and ref.pending.size == 0

    lookups = {
        "uniform": W.mixed_lookups(keys, n, 0.5, seed, n, seed + 7),
        "zipf": W.zipf_lookups(keys, n, 0.99, seed + 11),
    }
    for name, q in lookups.items():
        v, h = t.find(q)
        gv, gh = u32(v), h.cpu().numpy()
        hq = u32(q)
        ov, oh = g.find(hq)
        assert np.array_equal(gh, oh) and np.array_equal(gv, ov), f"{name}: find over the GPU image"
        rv, rh = ref.find(hq)
        assert np.array_equal(gh, rh) and np.array_equal(gv, rv), f"{name}: find vs the oracle's insert"
        probes, _ = t.probe_count(q)
        assert probes == O.count_probes(img, hq), f"{name}: probe count"
        if name == "zipf":
            assert gh.all()
        else:
            assert int(gh.sum()) == n // 2
    t.close()
