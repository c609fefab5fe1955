"""Load generation for the request server (mf_server_*).

Three ways to drive the same seeded request sequence:

- run_poisson: open loop, Poisson arrivals at rate lam.  A submitter thread
  issues requests at their arrival times; a collector thread waits for them
  in order, records submit-to-completion latency, keeps an audit sample of
  correlation rows, and releases each request's memory.
- run_closed_loop: `depth` requests kept outstanding, for capacity.
- run_reference_form: the plain one-request-at-a-time form a user of the
  reference API gets, one matched-filter instance per (request, template):
  initialize (template spectrum), set the request's channels, apply, wait.

Channel windows are device-resident (slices of resident channel streams),
so no host-to-device copy is in the timed path.
"""
from __future__ import annotations

import queue
import threading
import time

import numpy as np

from . import capi


class ResidentStreams:
    """The workload's channel streams in HBM (one device buffer each)."""

    def __init__(self, streams, alloc):
        # alloc(np.ndarray) -> object with .ptr (device address) kept alive here
        self.bufs = [alloc(s) for s in streams]
        self.itemsize = streams[0].dtype.itemsize

    def window(self, chans, start):
        return [self.bufs[c].ptr + start * self.itemsize for c in chans]


def corr_samples(wl, req) -> int:
    chans, _, n, ids = req
    nb = max(len(wl.library[t]) for t in ids)
    return len(chans) * len(ids) * (n - nb + 1)


def _submit(srv, res, lib_ids, req):
    chans, start, n, ids = req
    return srv.submit(res.window(chans, start), [lib_ids[t] for t in ids], on_device=True,
                      lengths=[n] * len(chans))


def _pct(v, p):
    return float(np.percentile(np.asarray(v), p)) if len(v) else float("nan")


def run_poisson(srv, wl, res, lib_ids, lam: float, n_req: int, seed: int, audit_every: int = 100):
    """Open-loop run; returns latency percentiles, throughput and audit rows
    [(request, template index, channel index, rho row)]."""
    rng = np.random.default_rng(seed)
    reqs = [wl.sample_request(rng) for _ in range(n_req)]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
    q: queue.Queue = queue.Queue()
    lat, queued, audit = [], [], []
    done_samples = [0]
    errors = []
    t_first = [None]
    t_last = [None]

    def submitter():
        try:
            t0 = time.perf_counter()
            t_first[0] = t0
            for k, req in enumerate(reqs):
                target = t0 + arrivals[k]
                while True:
                    dt = target - time.perf_counter()
                    if dt <= 0:
                        break
                    if dt > 0.002:
                        time.sleep(dt - 0.001)
                q.put((k, _submit(srv, res, lib_ids, req)))
        except Exception as e:  # surfaced by the caller
            errors.append(e)
        finally:
            q.put(None)

    def collector():
        try:
            while True:
                item = q.get()
                if item is None:
                    break
                k, rid = item
                srv.wait(rid)
                t_last[0] = time.perf_counter()
                qm, tm = srv.latency(rid)
                lat.append(tm)
                queued.append(qm)
                done_samples[0] += corr_samples(wl, reqs[k])
                if k % audit_every == 0:
                    chans, start, n, ids = reqs[k]
                    ti, ci = k % len(ids), k % len(chans)
                    audit.append((reqs[k], ti, ci, srv.rho(rid, ti, ci)))
                srv.release(rid)
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=submitter), threading.Thread(target=collector)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errors:
        raise errors[0]
    wall = t_last[0] - t_first[0]
    return {
        "lambda_per_s": lam,
        "requests": n_req,
        "p50_ms": _pct(lat, 50),
        "p99_ms": _pct(lat, 99),
        "mean_ms": float(np.mean(lat)),
        "p50_queue_ms": _pct(queued, 50),
        "offered_corr_samples_per_s": float(np.mean([corr_samples(wl, r) for r in reqs]) * lam),
        "corr_samples_per_s": done_samples[0] / wall,
        "wall_s": wall,
    }, audit


def run_closed_loop(srv, wl, res, lib_ids, n_req: int, depth: int, seed: int):
    """Keeps `depth` requests outstanding; returns corr samples/s."""
    rng = np.random.default_rng(seed)
    reqs = [wl.sample_request(rng) for _ in range(n_req)]
    outstanding = []
    total = 0
    t0 = time.perf_counter()
    nxt = 0
    while nxt < n_req or outstanding:
        while nxt < n_req and len(outstanding) < depth:
            outstanding.append((nxt, _submit(srv, res, lib_ids, reqs[nxt])))
            nxt += 1
        k, rid = outstanding.pop(0)
        srv.wait(rid)
        total += corr_samples(wl, reqs[k])
        srv.release(rid)
    wall = time.perf_counter() - t0
    return {"requests": n_req, "depth": depth, "corr_samples_per_s": total / wall, "wall_s": wall}


def run_reference_form(wl, res, n_req: int, seed: int, dtype=np.float32):
    """One engine instance per (request, template), as the reference API is
    used: MultiChannel MatchedFilter with the template on every channel,
    initialize, set the channels, apply, wait.  Returns corr samples/s."""
    rng = np.random.default_rng(seed)
    reqs = [wl.sample_request(rng) for _ in range(n_req)]
    total = 0
    t0 = time.perf_counter()
    for req in reqs:
        chans, start, n, ids = req
        ptrs = res.window(chans, start)
        for t in ids:
            h = capi.Handle(dtype, "multi")
            h.set_templates([wl.library[t]] * len(chans))
            h.set_signal_size(n)
            h.initialize()
            for c, p in enumerate(ptrs):
                h.set_signal_device(c, p, n)
            h.apply()
            h.synchronize()
            h.close()
            total += len(chans) * (n - len(wl.library[t]) + 1)
    wall = time.perf_counter() - t0
    return {"requests": n_req, "corr_samples_per_s": total / wall, "wall_s": wall}
