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
