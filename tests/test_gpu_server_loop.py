"""GPU parity of the device-resident serving loop (replay mode 2, or resident=1 servers).

The loop must give every request the same results as host-built ticks (mode 0) and as each
request alone (mode 1), and all of them must equal the trace's exact expected results (the
live-window model, pinned against the oracle in test_serving_trace.py). Its waits are bounded:
the tests below force the bounds from the host (a one-poll dependency bound, a 1 ms idle bound)
and check that every launch ends on its own and the host reports what gave up.
"""
import time

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


@pytest.mark.parametrize("mix,offered,lanes", [((0.3, 0.6, 0.1), None, (1, 8)), ((0.2, 0.4, 0.4), None, (2, 4)),
                                               ((0.3, 0.6, 0.1), 2e8, (4, 16))])
def test_loop_equals_ticks_equals_expected(D, mix, offered, lanes):
    from databases_gpu_amd import workload as W

    banner(f"server loop vs ticks, mix={mix} offered={offered} insert/find lanes={lanes}")
    tr = W.ServingTrace(1200, seed=21, device="cuda", offered=offered, resident=1 << 14, tenants=16, tenant_lo=256,
                        tenant_hi=4096, batch_lo=1, batch_hi=1 << 11, mix=mix)
    kw = dict(min_capacity=0, tick_max_keys=1 << 14, hbm_budget=256 << 20, insert_lanes=lanes[0],
              find_lanes=lanes[1])
    sl, ovl, ofl, _, st_l = run_mode(D, tr, 2, **kw)
    assert (st_l == 0).all()
    assert torch.equal(ofl, tr.exp_flags) and torch.equal(ovl, tr.exp_vals)
    check_contents(sl, tr)
    stats = sl.stats()
    assert stats["loop_launches"] >= 1 and stats["loop_items"] >= len(tr.req)
    assert stats["resizes_up"] > 0
    if mix[2] > mix[0]:
        assert stats["resizes_down"] > 0
    sl.close()
    st_, ovt, oft, _, st_t = run_mode(D, tr, 0, **kw)
    assert (st_t == 0).all()
    assert torch.equal(oft, ofl) and torch.equal(ovt, ovl)
    st_.close()


def test_resident_server_submit_poll_and_overflow_pool(D):
    """A resident=1 server through submit/poll: forty copies of one key overflow its two buckets,
    the loop stops to re-place the pool, and finds and erases see the copies as alone."""
    banner("resident server submit/poll + overflow pool")
    K = 0x12345677
    outs = []
    for mode in ("loop", "alone"):
        srv = D.DyCuckooServer([4096, 4096, 4096], min_capacity=0, hbm_budget=64 << 20, max_fail_rounds=2,
                               resident=1 if mode == "loop" else 0)
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
        if mode == "loop":
            tickets = [srv.submit(*r) for r in reqs]
            seen = {}
            t0 = time.time()
            while len(seen) < len(tickets):
                for tk, stt in srv.poll():
                    seen[tk] = stt
                assert time.time() - t0 < 60
            assert sorted(seen) == tickets and all(v == 0 for v in seen.values())
            stats = srv.stats()
            assert stats["loop_launches"] >= 1 and stats["pool_rounds"] > 0
        else:
            from databases_gpu_amd import _lib

            arr = np.zeros(len(reqs), dtype=_lib.REQUEST_DTYPE)
            for i, (t, op, k, v, ovp, ofp) in enumerate(reqs):
                arr[i] = (t, op, k.data_ptr(), v.data_ptr() if v is not None else 0,
                          ovp.data_ptr() if ovp is not None else 0, ofp.data_ptr() if ofp is not None else 0, k.numel())
            _, st = srv.replay(arr, np.zeros(len(reqs)), mode=1)
            assert (st == 0).all()
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (f1v, f1h, e1, f2v, f2h, f3h)])
        srv.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    f1v, f1h, e1, f2v, f2h, f3h = outs[0]
    assert list(f1h) == [1, 0, 1] and list(f1v) == [77, 0, 77]
    assert e1[0] == 1
    assert f2h[0] == 1 and f2v[0] == 77
    assert f3h.all()


def test_loop_idle_bound_ends_the_launch_and_submission_restarts_it(D):
    """resident_idle_ms = 1: the launch ends by itself between requests spaced 30 ms apart, and
    each new submission starts it again; results stay exact."""
    from databases_gpu_amd import _lib
    from databases_gpu_amd import workload as W

    banner("resident loop: 1 ms idle bound, requests 30 ms apart")
    srv = D.DyCuckooServer([1 << 14], min_capacity=0, hbm_budget=64 << 20, resident=1, resident_idle_ms=1)
    keys = W.keys_range(0, 6000, 3, "cuda")
    vals = W.splitmix32(keys)
    fv = torch.zeros(6000, dtype=torch.int32, device="cuda")
    fh = torch.zeros(6000, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    req = np.zeros(6, dtype=_lib.REQUEST_DTYPE)
    for i in range(3):
        lo, hi = i * 2000, (i + 1) * 2000
        req[2 * i] = (0, _lib.OP_INSERT, keys[lo:hi].data_ptr(), vals[lo:hi].data_ptr(), 0, 0, 2000)
        req[2 * i + 1] = (0, _lib.OP_FIND, keys.data_ptr(), 0, fv.data_ptr(), fh.data_ptr(), 6000)
    _, st = srv.replay(req, [0.0, 0.03, 0.06, 0.09, 0.12, 0.15], mode=2)
    assert (st == 0).all()
    assert srv.stats()["loop_launches"] >= 4
    assert fh.cpu().numpy().all() and torch.equal(fv, vals)
    assert srv.table_stats(0)["count"] == 6000
    srv.close()


def test_loop_reads_inputs_rewritten_between_requests(D):
    """The loop outlives its callers' writes: a key buffer found once, then rewritten by the
    caller and submitted again, is read afresh (no stale lines from the first request)."""
    banner("resident loop: a rewritten input buffer is read afresh")
    srv = D.DyCuckooServer([1 << 16], min_capacity=0, hbm_budget=64 << 20, resident=1)
    from databases_gpu_amd import workload as W

    a = W.keys_range(0, 20_000, 4, "cuda")
    b = W.keys_range(1 << 20, (1 << 20) + 20_000, 4, "cuda")
    srv.insert(0, a, W.splitmix32(a))
    srv.insert(0, b, W.splitmix32(b) ^ 1)
    srv.drain()
    q = a.clone()
    outs = []
    for src in (a, b, a):
        q.copy_(src)
        fv = torch.zeros(20_000, dtype=torch.int32, device="cuda")
        fh = torch.zeros(20_000, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        srv.find(0, q, fv, fh)
        srv.drain()
        torch.cuda.synchronize()
        outs.append((fv.clone(), fh.clone()))
    assert all(bool(h.all()) for _, h in outs)
    assert torch.equal(outs[0][0], W.splitmix32(a)) and torch.equal(outs[2][0], W.splitmix32(a))
    assert torch.equal(outs[1][0], W.splitmix32(b) ^ 1)
    srv.close()


def test_loop_dependency_wait_bound_fails_requests_not_the_gpu(D):
    """The loop's dependency waits are bounded: with a bound of one poll, items that must wait for
    their table's earlier requests give up, the launch still ends, and the host reports those
    requests failed (DYC_ERR_HIP); nothing hangs."""
    from databases_gpu_amd import workload as W

    banner("resident loop spin_cap 1: bounded waits")
    tr = W.ServingTrace(800, seed=13, device="cuda", offered=None, resident=1 << 14, tenants=2, tenant_lo=256,
                        tenant_hi=1024, batch_lo=1 << 8, batch_hi=1 << 12)
    srv, ov, of, _, st = run_mode(D, tr, 2, min_capacity=0, spin_cap=1, hbm_budget=256 << 20)
    assert (st == 2).any()
    assert set(np.unique(st).tolist()) <= {0, 2}
    srv.close()


def test_loop_admission_by_free_hbm(D):
    """Inserts that need more space than the budget holds fail with DYC_ERR_OOM in the loop as in
    ticks and alone; everything else is exact."""
    from databases_gpu_amd import _lib
    from databases_gpu_amd import workload as W

    banner("resident loop admission by free HBM")
    hint = 8192
    nb = O.initial_buckets(hint)
    sub_bytes = -(-nb * 128 // 256) * 256
    budget = 2 * 4 * sub_bytes + 3 * nb * 128 + 4096
    results = []
    for mode in (2, 1):
        srv = D.DyCuckooServer([hint, hint], min_capacity=0, hbm_budget=budget)
        keys = W.keys_range(0, 12 * 2048, 9, "cuda")
        vals = W.splitmix32(keys)
        fh = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        t1 = W.keys_range(1 << 20, (1 << 20) + 1000, 9, "cuda")
        t1v = W.splitmix32(t1)
        fh1 = torch.zeros(1000, dtype=torch.uint8, device="cuda")
        arr = np.zeros(15, dtype=_lib.REQUEST_DTYPE)
        for i in range(12):
            k, v = keys[i * 2048:(i + 1) * 2048], vals[i * 2048:(i + 1) * 2048]
            arr[i] = (0, 0, k.data_ptr(), v.data_ptr(), 0, 0, 2048)
        arr[12] = (1, 0, t1.data_ptr(), t1v.data_ptr(), 0, 0, 1000)
        arr[13] = (0, 1, keys.data_ptr(), 0, 0, fh.data_ptr(), keys.numel())
        arr[14] = (1, 1, t1.data_ptr(), 0, 0, fh1.data_ptr(), 1000)
        torch.cuda.synchronize()
        _, st = srv.replay(arr, np.zeros(15), mode=mode)
        torch.cuda.synchronize()
        ok = st[:12] == 0
        assert ok[0] and not ok.all(), st
        assert set(st[:12].tolist()) <= {0, 3} and (st[12:] == 0).all()
        assert np.array_equal(fh.cpu().numpy(), np.repeat(ok, 2048).astype(np.uint8))
        assert fh1.all()
        s = srv.stats()
        assert s["failed"] == int((~ok).sum()) and s["arena_peak"] <= budget
        results.append(st.copy())
        srv.close()
    assert np.array_equal(results[0], results[1])
