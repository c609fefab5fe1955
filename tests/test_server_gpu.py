"""Request server (mf_server_*) on the GPU: correlation rows against the
FP64 oracle, identical results alone and in a mix of concurrent requests,
admission under a small HBM budget, device-resident inputs, dead channels
and error reporting."""
import threading

import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _library(rng, n, lo=200, hi=2000):
    lens = rng.integers(lo, hi + 1, n)
    return [rng.standard_normal(int(L)) * np.hanning(int(L)) for L in lens]


def _channels(rng, n_ch, n):
    return [(rng.standard_normal(n) * rng.uniform(0.5, 3) + rng.uniform(-100, 100)).astype(np.float32)
            for _ in range(n_ch)]


def _check_against_oracle(srv, rid, tmpls, ids, chans, lags_per_row=400, rng=None, tol=1e-5):
    nb = max(len(tmpls[i]) for i in ids)
    worst = 0.0
    for ti, t in enumerate(ids):
        for c, x in enumerate(chans):
            r = srv.rho(rid, ti, c)
            assert r.size == x.size - nb + 1
            lags = rng.integers(0, r.size, lags_per_row) if rng is not None else None
            ref = O.pearson(tmpls[t], x, lags)
            got = r[lags] if lags is not None else r
            worst = max(worst, float(np.abs(got.astype(np.float64) - ref).max()))
    assert worst < tol, worst
    return worst


def test_single_request_matches_oracle():
    rng = np.random.default_rng(501)
    tmpls = _library(rng, 12)
    srv = capi.Server(np.float32, budget_bytes=1 << 30, max_batch=64, max_inflight=2)
    ids = srv.add_templates(tmpls)
    chans = _channels(rng, 5, 40000)
    pick = [ids[i] for i in (3, 0, 7, 11)]
    rid = srv.submit(chans, pick)
    srv.wait(rid)
    _check_against_oracle(srv, rid, tmpls, pick, chans, rng=rng)
    q, t = srv.latency(rid)
    assert 0 <= q <= t
    srv.release(rid)
    srv.close()


def test_results_identical_alone_and_in_the_mix():
    """Each request's rows are bit-identical whether it runs alone or is
    batched with concurrently arriving requests."""
    rng = np.random.default_rng(502)
    tmpls = _library(rng, 40)
    reqs = []
    for _ in range(24):
        n = int(rng.integers(6000, 60000))
        ch = _channels(rng, int(rng.integers(1, 6)), n)
        ids = sorted(rng.choice(40, int(rng.integers(1, 8)), replace=False).tolist())
        reqs.append((ch, ids))
    # alone: one request at a time
    alone = capi.Server(np.float32, budget_bytes=1 << 30, max_batch=1, max_inflight=1)
    alone.add_templates(tmpls)
    ref = []
    for ch, ids in reqs:
        rid = alone.submit(ch, ids)
        alone.wait(rid)
        ref.append([[alone.rho(rid, i, c) for c in range(len(ch))] for i in range(len(ids))])
        alone.release(rid)
    b_alone, r_alone, _ = alone.stats()
    assert b_alone == r_alone == len(reqs)
    alone.close()
    # mixed: submitted from two threads at once into a batching server
    mix = capi.Server(np.float32, budget_bytes=1 << 30, max_batch=64, max_inflight=2)
    mix.add_templates(tmpls)
    rids = [None] * len(reqs)

    def submit(idx):
        for k in idx:
            rids[k] = mix.submit(reqs[k][0], reqs[k][1])

    th = [threading.Thread(target=submit, args=(range(j, len(reqs), 2),)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, (ch, ids) in enumerate(reqs):
        mix.wait(rids[k])
        for i in range(len(ids)):
            for c in range(len(ch)):
                assert np.array_equal(mix.rho(rids[k], i, c), ref[k][i][c]), (k, i, c)
        mix.release(rids[k])
    batches, done, _ = mix.stats()
    assert done == len(reqs)
    assert batches < len(reqs), "requests were not batched"
    mix.close()
    # and the rows agree with the FP64 oracle
    srv = capi.Server(np.float32, budget_bytes=1 << 28, max_batch=8, max_inflight=2)
    srv.add_templates(tmpls)
    ch, ids = reqs[0]
    rid = srv.submit(ch, ids)
    srv.wait(rid)
    _check_against_oracle(srv, rid, tmpls, ids, ch, rng=rng)
    srv.close()


def test_window_stats_count_windows_and_constant_stretch():
    """mf_server_window_stats: the windows of a request are its rows'
    lengths; a stationary channel flags none, a channel with an exactly
    constant stretch flags at least the windows inside it (recomputed in
    full), and the rows stay within the FP32 contract of the oracle."""
    rng = np.random.default_rng(509)
    tmpls = _library(rng, 6, 300, 600)
    srv = capi.Server(np.float32, budget_bytes=1 << 30, max_batch=8, max_inflight=1)
    ids = srv.add_templates(tmpls)
    n = 30000
    chans = [(rng.standard_normal(n) * 2 + 50).astype(np.float32) for _ in range(2)]
    rid = srv.submit(chans, ids)
    srv.wait(rid)
    nb = max(len(t) for t in tmpls)
    ws = srv.window_stats()
    assert ws["windows"] == len(ids) * len(chans) * (n - nb + 1)
    assert ws["flagged"] == 0 and ws["full"] == 0, ws
    assert 0 <= ws["solo"] <= ws["windows"]
    srv.release(rid)
    g0, G = 10000, 3000
    chans[1][g0:g0 + G] = np.float32(57.0)
    rid = srv.submit(chans, ids)
    srv.wait(rid)
    ws2 = srv.window_stats()
    assert ws2["windows"] == 2 * ws["windows"]
    inside = sum(G - len(t) + 1 for t in tmpls)
    assert ws2["full"] >= inside and ws2["flagged"] >= ws2["full"], (ws2, inside)
    _check_against_oracle(srv, rid, tmpls, ids, chans, rng=rng)
    srv.release(rid)
    srv.close()


def test_admission_under_small_budget():
    rng = np.random.default_rng(503)
    tmpls = _library(rng, 6, 300, 600)
    # each request: 4 templates x 3 channels x ~100k outputs x 4 B ~ 4.8 MB
    srv = capi.Server(np.float32, budget_bytes=12 << 20, max_batch=16, max_inflight=2)
    ids = srv.add_templates(tmpls)
    chans = [_channels(rng, 3, 100000) for _ in range(6)]
    rids = [srv.submit(ch, ids[:4]) for ch in chans]
    for rid, ch in zip(rids, chans):
        srv.wait(rid)  # the next request is admitted once this one is released
        r = srv.rho(rid, 1, 2)
        assert r.size == 100000 - max(len(tmpls[i]) for i in ids[:4]) + 1
        srv.release(rid)
    assert srv.stats()[1] == 6
    with pytest.raises(RuntimeError):
        srv.submit(_channels(rng, 48, 100000), ids)  # larger than the whole budget
    srv.close()


def test_device_inputs_and_dead_channel():
    from hipmem import DeviceArray
    rng = np.random.default_rng(504)
    tmpls = _library(rng, 4, 250, 700)
    srv = capi.Server(np.float32, budget_bytes=1 << 28, max_batch=8, max_inflight=2)
    ids = srv.add_templates(tmpls)
    chans = _channels(rng, 3, 30000)
    chans[1][:] = 7.5  # dead trace: |min| == |max|
    dev = [DeviceArray(c) for c in chans]
    rid = srv.submit([d.ptr for d in dev], ids, on_device=True, lengths=[c.size for c in chans])
    srv.wait(rid)
    for i in range(len(ids)):
        assert not np.any(srv.rho(rid, i, 1))
    ok_ch = [chans[0], chans[2]]
    for i, t in enumerate(ids):
        for c, x in ((0, chans[0]), (2, chans[2])):
            ref = O.pearson(tmpls[t], x)[: x.size - max(len(tt) for tt in tmpls) + 1]
            assert np.abs(srv.rho(rid, i, c) - ref).max() < 1e-5
    del ok_ch
    srv.release(rid)
    srv.close()
    for d in dev:
        d.free()


def test_request_errors():
    rng = np.random.default_rng(505)
    srv = capi.Server(np.float32, budget_bytes=1 << 26, max_batch=4, max_inflight=1)
    ids = srv.add_templates(_library(rng, 2, 500, 500))
    with pytest.raises(ValueError):
        srv.submit(_channels(rng, 1, 1000), [ids[0], 99])   # unknown template
    with pytest.raises(ValueError):
        srv.submit(_channels(rng, 1, 300), ids)             # channel shorter than the template
    rid = srv.submit(_channels(rng, 1, 5000), ids)
    srv.wait(rid)
    with pytest.raises(ValueError):
        srv.rho(rid, 2, 0)
    srv.release(rid)
    with pytest.raises(ValueError):
        srv.wait(rid)                                       # released
    srv.close()


def test_cancel_pending_requests():
    """A request queued behind one that holds the whole budget is cancelled
    before launch: its waiters see MF_ERR_CANCELLED, the request behind it
    is admitted once memory frees up and matches the oracle.  A request
    that already ran cannot be cancelled."""
    rng = np.random.default_rng(506)
    tmpls = _library(rng, 3, 300, 600)
    n = 200_000
    # room for one request's rows (3 templates x 2 channels x n floats) at a time
    srv = capi.Server(np.float32, budget_bytes=3 * 2 * n * 4 * 1.5 + (2 << 20), max_batch=4, max_inflight=1)
    ids = srv.add_templates(tmpls)
    chans = [_channels(rng, 2, n) for _ in range(3)]
    a = srv.submit(chans[0], ids)
    srv.wait(a)                                   # a holds the arena until released
    b = srv.submit(chans[1], ids)
    c = srv.submit(chans[2], ids)
    assert not srv.done(b) and not srv.done(c)
    assert srv.cancel(b)
    assert srv.status(b) == capi.MF_ERR_CANCELLED
    with pytest.raises(RuntimeError):
        srv.rho(b, 0, 0)
    assert not srv.cancel(a)                      # already done
    srv.release(a)
    srv.wait(c)
    _check_against_oracle(srv, c, tmpls, ids, chans[2], lags_per_row=300, rng=rng)
    srv.release(b)
    srv.release(c)
    b2, r2, _ = srv.stats()
    assert r2 == 2
    srv.close()


def test_injected_allocation_failure_fails_one_batch_only():
    """The first batch's descriptor allocation fails (injected from the
    host): its request ends with MF_ERR_OUT_OF_MEMORY and the server keeps
    serving; the next request is correct."""
    rng = np.random.default_rng(507)
    tmpls = _library(rng, 4, 200, 900)
    srv = capi.Server(np.float32, budget_bytes=1 << 28, max_batch=8, max_inflight=1)
    ids = srv.add_templates(tmpls)
    chans = _channels(rng, 2, 30000)
    try:
        capi.debug_fail_allocations(1)
        a = srv.submit(chans, ids)
        assert srv.status(a) == capi.MF_ERR_OUT_OF_MEMORY
    finally:
        capi.debug_fail_allocations(0)
    srv.release(a)
    b = srv.submit(chans, ids)
    srv.wait(b)
    _check_against_oracle(srv, b, tmpls, ids, chans, lags_per_row=300, rng=rng)
    srv.release(b)
    srv.close()


def test_injected_allocation_failure_in_initialize():
    """A handle whose initialize fails on an allocation stays uninitialized
    (runtime errors on use, as the reference's before initialize); the next
    initialize succeeds and apply matches the oracle."""
    rng = np.random.default_rng(508)
    tpl = rng.standard_normal(300) * np.hanning(300)
    x = (rng.standard_normal(20000) + 30).astype(np.float32)
    h = capi.Handle(np.float32, "single")
    h.set_templates([tpl])
    h.set_signal_size(x.size)
    try:
        capi.debug_fail_allocations(2)
        with pytest.raises(RuntimeError, match="injected"):
            h.initialize()
    finally:
        left = capi.debug_fail_allocations(0)
    assert left == 1
    assert not h.is_initialized()
    with pytest.raises(RuntimeError):
        h.apply()
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    assert np.abs(h.rho(0) - O.pearson(tpl, x)).max() < 1e-5
    h.close()
