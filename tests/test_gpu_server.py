"""GPU parity of the serving loop (dyc_server_*).

Every request's results must be the same in the mix (continuous batching: many requests fused
per tick, ticks pipelined, resizes in-line) as alone (each request by itself with the plain
single-table kernels, in arrival order), and both must equal the trace's exact expected
results (the live-window model, itself pinned against the oracle in test_serving_trace.py).
"""
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


def populate(srv, tr):
    from databases_gpu_amd import workload as W

    keep = []
    for t in range(tr.sizes.size):
        k = tr.initial_keys(t)
        v = W.splitmix32(k)
        keep += [k, v]
        srv.insert(t, k, v)
    torch.cuda.synchronize()
    srv.drain()
    return keep


def run_mode(D, tr, mode, **kw):
    srv = D.DyCuckooServer(tr.hints, **kw)
    populate(srv, tr)
    ov = torch.zeros(tr.ops, dtype=torch.int32, device="cuda")
    of = torch.zeros(tr.ops, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    done, status = srv.replay(tr.requests(ov, of), tr.arrivals, mode=mode)
    torch.cuda.synchronize()
    return srv, ov, of, done, status


def check_contents(srv, tr):
    from databases_gpu_amd import workload as W

    lo, hi = tr.final_window
    for t in range(tr.sizes.size):
        imgs = srv.dump(t)
        ref = O.OracleTable(nb=[x.size // 16 for x in imgs])
        ref.sub = imgs
        assert ref.check_invariant() == 0, f"table {t}"
        live = W.keys_at(torch.arange(int(tr.base[t] + lo[t]), int(tr.base[t] + hi[t]), dtype=torch.int64), tr.seed)
        live = live.numpy().view(np.uint32)
        assert np.array_equal(ref.entries(), np.sort(O.pack(live, W.splitmix32_np(live)))), f"table {t}"
        assert srv.table_stats(t)["count"] == live.size


@pytest.mark.parametrize("mix,offered,one_launch", [((0.3, 0.6, 0.1), None, 1), ((0.2, 0.4, 0.4), None, 1),
                                                    ((0.3, 0.6, 0.1), 2e8, 1), ((0.3, 0.6, 0.1), None, 0)])
def test_mix_equals_alone_equals_expected(D, mix, offered, one_launch):
    from databases_gpu_amd import workload as W

    banner(f"server mix vs alone, mix={mix} offered={offered} one_launch={one_launch}")
    tr = W.ServingTrace(1200, seed=21, device="cuda", offered=offered, resident=1 << 14, tenants=16, tenant_lo=256,
                        tenant_hi=4096, batch_lo=1, batch_hi=1 << 11, mix=mix)
    kw = dict(min_capacity=0, tick_max_keys=1 << 14, hbm_budget=256 << 20, one_launch=one_launch)
    sa, ova, ofa, done_a, st_a = run_mode(D, tr, 0, **kw)
    assert (st_a == 0).all()
    assert torch.equal(ofa, tr.exp_flags) and torch.equal(ova, tr.exp_vals)
    check_contents(sa, tr)
    stats = sa.stats()
    assert stats["ticks"] < len(tr.req)  # requests were fused
    assert stats["resizes_up"] > 0
    if mix[2] > mix[0]:
        assert stats["resizes_down"] > 0
    sa.close()
    sb, ovb, ofb, done_b, st_b = run_mode(D, tr, 1, **kw)
    assert (st_b == 0).all()
    assert torch.equal(ofb, ofa) and torch.equal(ovb, ova)
    check_contents(sb, tr)
    sb.close()


def test_depth_one_and_fixed_items(D):
    from databases_gpu_amd import workload as W

    banner("server depth 1, item_keys 64, find lanes 4, insert lanes 2")
    tr = W.ServingTrace(600, seed=5, device="cuda", offered=None, resident=1 << 13, tenants=8, tenant_lo=256,
                        tenant_hi=2048, batch_lo=1, batch_hi=1 << 10)
    srv, ov, of, _, st = run_mode(D, tr, 0, min_capacity=0, pipeline_depth=1, item_keys=64, tick_max_requests=7,
                                  find_lanes=4, insert_lanes=2, hbm_budget=64 << 20, one_launch=0)
    assert (st == 0).all()
    assert torch.equal(of, tr.exp_flags) and torch.equal(ov, tr.exp_vals)
    assert srv.stats()["deferred_slots"] > 0
    check_contents(srv, tr)
    srv.close()


def test_submit_poll_and_overflow_pool(D):
    """Forty copies of one key overflow its two buckets: the extra copies wait in the overflow
    pool, where finds and erases see them, in the mix as alone."""
    banner("server submit/poll + overflow pool")
    K = 0x12345677
    outs = []
    for mode in ("mix", "alone"):
        srv = D.DyCuckooServer([4096, 4096, 4096], min_capacity=0, hbm_budget=64 << 20, max_fail_rounds=2)
        dup = torch.full((40,), K, dtype=torch.int64).to(torch.int32).cuda()
        vals = torch.full((40,), 77, dtype=torch.int32, device="cuda")
        q = torch.tensor([K, 5, K], dtype=torch.int64).to(torch.int32).cuda()
        other = torch.arange(1, 1001, dtype=torch.int32, device="cuda")
        f1v, f1h = torch.zeros(3, dtype=torch.int32, device="cuda"), torch.zeros(3, dtype=torch.uint8, device="cuda")
        e1 = torch.zeros(1, dtype=torch.uint8, device="cuda")
        f2v, f2h = torch.zeros(3, dtype=torch.int32, device="cuda"), torch.zeros(3, dtype=torch.uint8, device="cuda")
        f3h = torch.zeros(1000, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        reqs = [(1, 0, dup, vals, None, None), (2, 0, other, other, None, None), (1, 1, q, None, f1v, f1h),
                (1, 2, q[:1], None, None, e1), (1, 1, q, None, f2v, f2h), (2, 1, other, None, None, f3h)]
        if mode == "mix":
            tickets = [srv.submit(*r) for r in reqs]
            seen = {}
            while len(seen) < len(tickets):
                for tk, stt in srv.poll():
                    seen[tk] = stt
            assert sorted(seen) == tickets and all(v == 0 for v in seen.values())
        else:
            from databases_gpu_amd import _lib

            arr = np.zeros(len(reqs), dtype=_lib.REQUEST_DTYPE)
            for i, (t, op, k, v, ovp, ofp) in enumerate(reqs):
                arr[i] = (t, op, k.data_ptr(), v.data_ptr() if v is not None else 0,
                          ovp.data_ptr() if ovp is not None else 0, ofp.data_ptr() if ofp is not None else 0, k.numel())
            _, st = srv.replay(arr, np.zeros(len(reqs)), mode=1)
            assert (st == 0).all()
        torch.cuda.synchronize()
        stats = srv.stats()
        assert stats["pool_rounds"] > 0  # the overflow was re-placed (or kept) by the loop
        outs.append([x.cpu().numpy() for x in (f1v, f1h, e1, f2v, f2h, f3h)])
        srv.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    f1v, f1h, e1, f2v, f2h, f3h = outs[0]
    assert list(f1h) == [1, 0, 1] and list(f1v) == [77, 0, 77]
    assert e1[0] == 1
    assert f2h[0] == 1 and f2v[0] == 77  # copies remain in the other bucket and the pool
    assert f3h.all()


def test_admission_by_free_hbm(D):
    """With a budget that cannot hold the growth, inserts that need more space fail with
    DYC_ERR_OOM (in the mix after waiting, alone at once); everything else is exact."""
    banner("server admission by free HBM")
    hint = 8192
    nb = O.initial_buckets(hint)
    sub_bytes = -(-nb * 128 // 256) * 256
    budget = 2 * 4 * sub_bytes + 3 * nb * 128 + 4096  # two tables plus room for about one upsize
    results = []
    for mode in (0, 1):
        srv = D.DyCuckooServer([hint, hint], min_capacity=0, hbm_budget=budget)
        from databases_gpu_amd import _lib
        from databases_gpu_amd import workload as W

        keys = W.keys_range(0, 12 * 2048, 9, "cuda")
        vals = W.splitmix32(keys)
        fh = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        t1 = W.keys_range(1 << 20, (1 << 20) + 1000, 9, "cuda")
        fh1 = torch.zeros(1000, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        reqs = [(0, 0, keys[i * 2048:(i + 1) * 2048], vals[i * 2048:(i + 1) * 2048]) for i in range(12)]
        arr = np.zeros(15, dtype=_lib.REQUEST_DTYPE)
        for i, (t, op, k, v) in enumerate(reqs):
            arr[i] = (t, op, k.data_ptr(), v.data_ptr(), 0, 0, k.numel())
        t1v = W.splitmix32(t1)
        arr[12] = (1, 0, t1.data_ptr(), t1v.data_ptr(), 0, 0, 1000)
        arr[13] = (0, 1, keys.data_ptr(), 0, 0, fh.data_ptr(), keys.numel())
        arr[14] = (1, 1, t1.data_ptr(), 0, 0, fh1.data_ptr(), 1000)
        torch.cuda.synchronize()
        _, st = srv.replay(arr, np.zeros(15), mode=mode)
        torch.cuda.synchronize()
        ok = st[:12] == 0
        assert ok[0] and not ok.all(), st
        assert set(st[:12].tolist()) <= {0, 3} and (st[12:] == 0).all()
        want = np.repeat(ok, 2048).astype(np.uint8)
        assert np.array_equal(fh.cpu().numpy(), want)
        assert fh1.all()
        s = srv.stats()
        assert s["failed"] == int((~ok).sum()) and s["arena_peak"] <= budget
        if mode == 0:
            assert s["deferred_memory"] > 0
        results.append(st.copy())
        srv.close()
    assert np.array_equal(results[0], results[1])


def test_one_launch_with_unsupported_lanes_falls_back_to_rounds(D):
    from databases_gpu_amd import workload as W

    banner("server insert lanes 8 (rounds launched), find lanes 16")
    tr = W.ServingTrace(400, seed=8, device="cuda", offered=None, resident=1 << 12, tenants=4, tenant_lo=256,
                        tenant_hi=2048, batch_lo=1, batch_hi=1 << 10)
    srv, ov, of, _, st = run_mode(D, tr, 0, min_capacity=0, insert_lanes=8, find_lanes=16, hbm_budget=64 << 20)
    assert (st == 0).all()
    assert torch.equal(of, tr.exp_flags) and torch.equal(ov, tr.exp_vals)
    check_contents(srv, tr)
    srv.close()


def test_dependency_wait_bound_fails_the_tick_not_the_gpu(D):
    """A one-launch tick's dependency waits are bounded: with a bound of one poll, items that must
    wait give up, the launch still drains, and the server reports the tick's requests failed."""
    from databases_gpu_amd import workload as W

    banner("server spin_cap 1: bounded waits end the launch")
    tr = W.ServingTrace(800, seed=13, device="cuda", offered=None, resident=1 << 14, tenants=2, tenant_lo=256,
                        tenant_hi=1024, batch_lo=1 << 8, batch_hi=1 << 12)
    srv, ov, of, _, st = run_mode(D, tr, 0, min_capacity=0, spin_cap=1, hbm_budget=256 << 20)
    assert (st == 2).any()  # DYC_ERR_HIP: a wait hit its bound
    assert set(np.unique(st).tolist()) <= {0, 2}
    srv.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_reserved_keys_in_single_request_and_synchronous_inserts(D, mode):
    """A request alone on an idle server (mode 0) and the synchronous path (mode 1) both insert
    through the plain kernel; the table is credited with the entries that launch took in, so
    reserved keys (0 and 0xFFFFFFFF) in the batch are not counted."""
    from databases_gpu_amd import _lib

    banner(f"server: reserved keys in plain-kernel inserts, mode={mode}")
    rng = np.random.default_rng(77)
    keys = np.unique(rng.integers(1, 2**32 - 1, 50_000, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(keys)
    k0, k1 = keys[:20_000].copy(), keys[20_000:32_000].copy()
    k0[[5, 900]] = [0, 0xFFFFFFFF]
    k1[77] = 0
    t = [torch.from_numpy(a.view(np.int32)).cuda() for a in (k0, k0 ^ np.uint32(0x1234567), k1, k1 + np.uint32(9))]
    flags = torch.zeros(k0.size, dtype=torch.uint8, device="cuda")
    srv = D.DyCuckooServer([40_000, 40_000], min_capacity=0, hbm_budget=64 << 20)
    req = np.zeros(3, dtype=np.dtype(_lib.REQUEST_DTYPE))
    req[0] = (0, _lib.OP_INSERT, t[0].data_ptr(), t[1].data_ptr(), 0, 0, k0.size)
    req[1] = (1, _lib.OP_INSERT, t[2].data_ptr(), t[3].data_ptr(), 0, 0, k1.size)
    req[2] = (0, _lib.OP_ERASE, t[0].data_ptr(), 0, 0, flags.data_ptr(), k0.size)
    torch.cuda.synchronize()
    _, st = srv.replay(req, [0.0, 0.02, 0.04], mode=mode)  # spaced: each request alone
    assert (st == 0).all()
    assert srv.table_stats(1)["count"] == k1.size - 1
    f = flags.cpu().numpy()
    assert f.sum() == k0.size - 2 and f[5] == 0 and f[900] == 0
    assert srv.table_stats(0)["count"] == 0
    ref = O.OracleTable(nb=[x.size // 16 for x in srv.dump(1)])
    ref.sub = srv.dump(1)
    ok = k1 != 0
    assert np.array_equal(ref.entries(), np.sort(O.pack(k1[ok], (k1 + np.uint32(9))[ok])))
    srv.close()

