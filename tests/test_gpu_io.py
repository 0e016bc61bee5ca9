"""GPU parity of the host-pointer entry points (pipelined transfers), checkpoints and dataset
ingest."""
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


def test_host_pointer_calls_match_oracle_across_chunks(D):
    from databases_gpu_amd.workload import splitmix32_np

    n = (1 << 22) + 123  # three staging chunks, the last one ragged
    banner(f"host-pointer insert/find/erase n={n}")
    keys = unique_keys(2 * n, 41)
    ins, absent = keys[:n], keys[n:]
    vals = splitmix32_np(ins)
    t = D.DyCuckoo(int(np.ceil(n / 0.85)))
    t.insert_host(ins, vals)
    st = t.stats()
    assert st["count"] == n and st["pending"] == 0
    q = np.concatenate([ins[::3], absent[: n // 3], np.array([0, 0xFFFFFFFF], dtype=np.uint32)])
    v, h = t.find_host(q)
    ref = O.OracleTable(nb=t.stats()["buckets"])
    ref.sub = t.dump()
    assert ref.check_invariant() == 0
    rv, rh = ref.find(q)
    assert np.array_equal(v, rv) and np.array_equal(h, rh)
    assert h[: ins[::3].size].all() and not h[ins[::3].size:].any()
    # the device-pointer path on the same table gives the same answers
    dv, dh = t.find(torch.from_numpy(q.view(np.int32)).cuda())
    assert np.array_equal(dv.cpu().numpy().view(np.uint32), v) and np.array_equal(dh.cpu().numpy(), h)
    d = t.erase_host(np.concatenate([ins[: n // 2], absent[:1000]]))
    assert d[: n // 2].all() and not d[n // 2:].any()
    assert t.stats()["count"] == n - n // 2
    t.close()


def test_host_insert_keeps_the_per_call_resize_policy(D):
    """With auto_resize the host path upsizes once per call for the whole batch, like the
    device path: the same calls give the same subtable sizes."""
    from databases_gpu_amd.workload import splitmix32_np

    banner("host-pointer policy vs device path")
    keys = unique_keys(3 << 20, 43)
    vals = splitmix32_np(keys)
    kw = dict(auto_resize=1, min_capacity=0)
    a, b = D.DyCuckoo(1 << 20, **kw), D.DyCuckoo(1 << 20, **kw)
    for lo in range(0, keys.size, 1 << 20):
        a.insert_host(keys[lo:lo + (1 << 20)], vals[lo:lo + (1 << 20)])
        b.insert(torch.from_numpy(keys[lo:lo + (1 << 20)].view(np.int32)).cuda(),
                 torch.from_numpy(vals[lo:lo + (1 << 20)].view(np.int32)).cuda())
        assert a.stats()["buckets"] == b.stats()["buckets"]
    sa = a.stats()
    assert sa["resizes_up"] > 0 and sa["count"] == keys.size
    a.close()
    b.close()


def test_checkpoint_round_trip_with_pending_entries(D, tmp_path):
    from databases_gpu_amd import io
    from databases_gpu_amd.workload import splitmix32_np

    banner("checkpoint write/read, pending entries, corruption")
    keys = unique_keys(200_000, 44)
    vals = splitmix32_np(keys)
    t = D.DyCuckoo(int(np.ceil(keys.size / 0.9)), max_fail_rounds=1)
    t.insert(torch.from_numpy(keys.view(np.int32)).cuda(), torch.from_numpy(vals.view(np.int32)).cuda())
    dup = np.full(40, 0x1234567, dtype=np.uint32)  # more copies than its two buckets hold
    with pytest.raises(D.DycError):
        t.insert(torch.from_numpy(dup.view(np.int32)).cuda(), torch.from_numpy(dup.view(np.int32)).cuda())
    st = t.stats()
    assert st["pending"] > 0
    pend = t.pending()
    path = tmp_path / "t.ckpt"
    t.checkpoint(path)
    hdr = io.checkpoint_header(path)
    imgs = t.dump()
    assert hdr["buckets"] == st["buckets"] and hdr["count"] == st["count"] and hdr["pending"] == pend.size
    # the checksums the library wrote equal an independent computation over the images
    for i in range(4):
        assert hdr["checksums"][i] == io.fnv1a64_words(imgs[i])
    assert hdr["pending_checksum"] == io.fnv1a64_words(pend)
    assert hdr["version"] == 2 and hdr["subtables"] == 4
    r = D.DyCuckoo.restore(path)
    assert r.stats()["buckets"] == st["buckets"] and r.stats()["count"] == st["count"]
    assert all(np.array_equal(x, y) for x, y in zip(r.dump(), imgs))
    assert np.array_equal(np.sort(r.pending()), np.sort(pend))
    q = torch.from_numpy(np.concatenate([keys[:5000], dup[:1]]).view(np.int32)).cuda()
    assert all(torch.equal(x, y) for x, y in zip(r.find(q), t.find(q)))
    r.close()
    # one flipped bit in a subtable image, then a truncated file: both refused
    raw = bytearray(path.read_bytes())
    raw[128 + 8 * 1000] ^= 0x10
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    with pytest.raises(D.DycError, match="CORRUPT"):
        D.DyCuckoo.restore(bad)
    bad.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(D.DycError, match="CORRUPT"):
        D.DyCuckoo.restore(bad)
    t.close()


def test_dataset_ingest_stores_index_plus_one(D, tmp_path):
    from databases_gpu_amd import io

    banner("dataset ingest")
    keys = unique_keys(3_000_001, 45)
    p = tmp_path / "data.dat"
    io.write_keys(p, keys)
    t = D.DyCuckoo(int(np.ceil(keys.size / 0.85)))
    assert io.ingest(t, p, chunk=1 << 20) == keys.size
    v, h = t.find_host(keys)
    assert h.all() and np.array_equal(v, np.arange(1, keys.size + 1, dtype=np.uint32))
    t.close()
