"""The serving trace's expected results against the oracle: replaying every request in arrival
order on per-table oracle tables (with the fill-bound resize policy on) gives exactly the
trace's expected find values / hit flags and erase flags. CPU only."""
import numpy as np
import pytest
import torch

import oracle as O
from databases_gpu_amd import workload as W


def u32(t):
    return t.numpy().view(np.uint32)


@pytest.mark.parametrize("mix", [(0.3, 0.6, 0.1), (0.2, 0.4, 0.4)])
def test_trace_expected_results_match_oracle_replay(oracle_lib, mix):
    tr = W.ServingTrace(400, seed=11, device="cpu", offered=None, resident=1 << 13, tenants=6, tenant_lo=200,
                        tenant_hi=3000, batch_lo=1, batch_hi=1500, mix=mix)
    keys, vals = u32(tr.keys), u32(tr.vals)
    exp_v, exp_f = u32(tr.exp_vals), tr.exp_flags.numpy()
    tabs = []
    for t in range(tr.sizes.size):
        tb = O.OracleTable(int(tr.hints[t]), rescue_kicks=128, auto_resize=True, lower=0.5, upper=0.9,
                           min_capacity=0)
        k0 = u32(tr.initial_keys(t))
        tb.insert(k0, W.splitmix32_np(k0))
        tabs.append(tb)
    ops = np.bincount(tr.req["op"], minlength=3)
    assert ops.min() > 0
    for r in tr.req:
        t, op, n, off = int(r["table"]), int(r["op"]), int(r["n"]), int(r["off"])
        k = keys[off:off + n]
        if op == 0:
            assert not exp_f[off:off + n].any()
            tabs[t].insert(k, vals[off:off + n])
            assert tabs[t].pending.size == 0
        elif op == 1:
            v, h = tabs[t].find(k)
            assert np.array_equal(h, exp_f[off:off + n]) and np.array_equal(v, exp_v[off:off + n])
        else:
            d = tabs[t].erase(k)
            assert d.all() and np.array_equal(d, exp_f[off:off + n])
    # the final live windows are what the tables hold
    lo, hi = tr.final_window
    for t, tb in enumerate(tabs):
        live = u32(W.keys_at(torch.arange(int(tr.base[t] + lo[t]), int(tr.base[t] + hi[t]), dtype=torch.int64), tr.seed))
        assert np.array_equal(tb.entries(), np.sort(O.pack(live, W.splitmix32_np(live))))
    # the erase-heavy mix shrinks some table
    if mix[2] > mix[0]:
        assert any(r[0] == "down" for tb in tabs for r in tb.resizes)


def test_trace_poisson_arrivals_and_request_records():
    tr = W.ServingTrace(2000, seed=3, device="cpu", offered=1e8, resident=1 << 12, tenants=4, tenant_lo=100,
                        tenant_hi=1000, batch_lo=64, batch_hi=4096)
    gaps = np.diff(tr.arrivals)
    assert (gaps >= 0).all() and tr.arrivals[0] == 0
    mean_n = tr.req["n"].mean()
    # Poisson at `offered` keys/s: mean gap = mean batch / offered (loose: 2000 samples)
    assert abs(gaps.mean() / (mean_n / 1e8) - 1) < 0.1
    # record layout: pointers advance 4 bytes per key (1 for flags), inserts carry values only
    ov = torch.zeros(tr.ops, dtype=torch.int32)
    of = torch.zeros(tr.ops, dtype=torch.uint8)
    r = tr.requests(ov, of)
    assert r.dtype.itemsize == 48
    i = int(np.nonzero(tr.req["op"] == 1)[0][0])
    assert r["keys"][i] == tr.keys.data_ptr() + 4 * int(tr.req["off"][i])
    assert r["out_flags"][i] == of.data_ptr() + int(tr.req["off"][i]) and r["vals"][i] == 0
    j = int(np.nonzero(tr.req["op"] == 0)[0][0])
    assert r["vals"][j] == tr.vals.data_ptr() + 4 * int(tr.req["off"][j]) and r["out_vals"][j] == 0
