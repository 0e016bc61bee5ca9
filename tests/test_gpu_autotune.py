"""The launch-shape autotuner (databases_gpu_amd/autotune.py) on small batches: it installs the
shape it reports, leaves the table or group holding the batch with every find exact, and tunes a
table's find over several lookup sets at once (one shape serves all of them)."""
import numpy as np
import pytest
import torch

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


def test_tune_table_over_two_lookup_sets(D):
    from databases_gpu_amd import autotune
    from databases_gpu_amd import workload as W

    banner("autotune: table, uniform + Zipf lookup sets")
    n, dev = 1 << 20, torch.device("cuda")
    keys = W.keys_range(0, n, 5, dev)
    vals = W.splitmix32(keys)
    look_u = W.mixed_lookups(keys, n, 0.5, 5, n, 12)
    look_z = W.zipf_lookups(keys, n // 2, 0.99, 16)
    t = D.DyCuckoo(int(np.ceil(n / 0.9)))
    out = autotune.tune_table(t, keys, vals, (look_u, look_z), use_cache=False, kick_bounds=(15, 0))
    cfg = t.config()
    for k, v in out["installed"].items():
        assert cfg[k] == v, k
    assert out["find"]["find_lanes"] in (4, 8, 16) and out["trials"] == 18 + 3
    assert t.stats()["count"] == n  # left holding the batch
    v, h = t.find(look_z)
    assert bool(h.bool().all()) and bool((v == W.splitmix32(look_z)).all())
    v, h = t.find(look_u)
    present = torch.isin(look_u, keys)
    assert bool((h.bool() == present).all()) and bool((v[present] == W.splitmix32(look_u[present])).all())
    t.close()


def test_tune_group_installs_and_keeps_the_batch(D):
    from databases_gpu_amd import autotune

    banner("autotune: group of 64 tables")
    rng = np.random.default_rng(8)
    sizes = rng.integers(200, 20_000, size=64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    total = int(offs[-1])
    keys = np.unique(rng.integers(1, 2**32 - 1, size=int(total * 1.2), dtype=np.uint64).astype(np.uint32))
    keys = keys[rng.permutation(keys.size)][:total]
    vals = keys * np.uint32(7) + np.uint32(3)
    kt = torch.from_numpy(keys.view(np.int32)).cuda()
    vt = torch.from_numpy(vals.view(np.int32)).cuda()
    g = D.DyCuckooGroup((np.ceil(sizes / 0.85) + 64).astype(np.uint64))
    out = autotune.tune_group(g, kt, vt, offs, kt, offs, use_cache=False)
    lau = g.launch
    for k, v in out["installed"].items():
        assert lau[k] == v, k
    assert all(g.table_stats(i)["count"] == sizes[i] for i in range(64))
    fv, fh = g.find(kt, offs)
    assert bool(fh.bool().all()) and np.array_equal(fv.cpu().numpy().view(np.uint32), vals)
    g.close()
