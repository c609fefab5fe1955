#!/usr/bin/env python3
"""Throughput of the matched-filter detection path on one MI355X.

Workload (default): 16 templates x 24 channels, 8,640,000 FP32 samples per
channel (one day at 100 Hz), 512-sample templates, integer moveouts in
[0, 200], DC offsets U(-100, 100); one batched engine instance holds the
24 channels once and the 384 template rows (16 template events x 24
channels).  One step = correlation + normalization of all rows, the
moveout stack of each of the 16 events, and PeakFinder (threshold 0.5,
minimum distance 512) on each stack.  Inputs are resident in HBM before
the timed region.

Prints ONE JSON line (metric / value / roofline / cpu_baseline ...).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "corr samples/s (tmpl×chan×samp) at |Δρ|≤1e-5; p50/p99 request latency"
HBM_PEAK = 8.0e12


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload: str, nfft: int, kernel: str):
    """HBM bytes per launch of the correlation kernel for this workload and
    plan, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summaries under
    profiles/ (counter values scaled by a copy-kernel calibration).  None
    when no summary matches."""
    import glob
    import re
    pat = re.compile(r"traffic:\s+workload=(\S+)\s+dtype=(\S+)\s+nfft=(\d+)\s+kernel=(\S+)\s+bytes=(\d+)")
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*.md"))):
        with open(path, encoding="utf-8") as f:
            for m in pat.finditer(f.read()):
                if m.group(1) == workload and m.group(2) == "f32" and int(m.group(3)) == nfft \
                        and m.group(4) == kernel:
                    return int(m.group(5)), os.path.relpath(path, here)
    return None, None


def serve_section(args, torch):
    """Served stream of detection requests: resident library of 1000
    templates (200-2000 samples), 48 resident one-hour channel streams,
    requests of 60 s - 1 h windows on 1-48 channels x 1-100 templates.
    Measures the batching server at capacity and under Poisson arrivals
    (p50/p99 latency), against the one-request-at-a-time forms."""
    from mflib_amd import capi, loadgen, synth
    log("serving: generating library and channel streams")
    wl = synth.serving_stream()

    class _Dev:
        def __init__(self, a):
            self.t = torch.from_numpy(a).to("cuda:0")
            self.ptr = self.t.data_ptr()

    res = loadgen.ResidentStreams(wl.streams, _Dev)
    torch.cuda.synchronize()
    free, _ = torch.cuda.mem_get_info()
    budget = min(200e9, free - 24e9)
    srv = capi.Server(np.float32, budget_bytes=budget, max_batch=args.serve_batch, max_inflight=2)
    lib_ids = srv.add_templates(wl.library)
    log(f"serving: budget {budget / 1e9:.0f} GB, library {len(lib_ids)} templates")
    loadgen.run_closed_loop(srv, wl, res, lib_ids, 16, 8, seed=1)  # warm-up
    cap = loadgen.run_closed_loop(srv, wl, res, lib_ids, args.serve_requests, 64, seed=2)
    log(f"serving: batched capacity {cap['corr_samples_per_s'] / 1e9:.1f} G corr samples/s")
    runs, audit = [], []
    for lam in args.serve_lambda:
        n_req = max(50, int(lam * args.serve_seconds))  # at least 50 requests: 5 s at 10/s
        r, a = loadgen.run_poisson(srv, wl, res, lib_ids, lam, n_req, seed=3 + int(lam))
        runs.append(r)
        audit += a
        log(f"serving: lambda {lam:g}/s: p50 {r['p50_ms']:.2f} ms, p99 {r['p99_ms']:.2f} ms, "
            f"{r['corr_samples_per_s'] / 1e9:.1f} G corr samples/s")
    b, n, _ = srv.stats()
    ws = _window_fracs(srv.window_stats())
    log(f"serving: flagged windows {100.0 * ws['flagged_frac']:.4f}%, solo {100.0 * ws['solo_frac']:.1f}%")
    srv.close()
    one = capi.Server(np.float32, budget_bytes=min(budget, 64e9), max_batch=1, max_inflight=1)
    lib1 = one.add_templates(wl.library)
    loadgen.run_closed_loop(one, wl, res, lib1, 8, 1, seed=1)
    single = loadgen.run_closed_loop(one, wl, res, lib1, max(20, args.serve_requests // 4), 1, seed=2)
    one.close()
    log(f"serving: one request at a time {single['corr_samples_per_s'] / 1e9:.1f} G corr samples/s")
    ref = loadgen.run_reference_form(wl, res, args.serve_ref_requests, seed=2)
    log(f"serving: reference form {ref['corr_samples_per_s'] / 1e9:.2f} G corr samples/s")
    out = {
        "workload": "served stream: library 1000 templates (200-2000 samples, resident spectra), "
                    "48 resident 1 h channels at 100 Hz, requests = 60 s-1 h window (log-uniform) "
                    "x 1-48 channels x 1-100 templates, FP32, Poisson arrivals",
        "budget_bytes": budget,
        "batched_capacity_corr_samples_per_s": cap["corr_samples_per_s"],
        "one_at_a_time_server_corr_samples_per_s": single["corr_samples_per_s"],
        "reference_form_corr_samples_per_s": ref["corr_samples_per_s"],
        "batched_vs_one_at_a_time": cap["corr_samples_per_s"] / single["corr_samples_per_s"],
        "batched_vs_reference_form": cap["corr_samples_per_s"] / ref["corr_samples_per_s"],
        "poisson": runs,
        "batches": b,
        "requests_completed": n,
        "window_stats": ws,
    }
    return out, (wl, audit)


def single_channel_section(args, torch):
    """One hour at 100 Hz, FP64, one 400-sample template injected 20 times:
    SingleChannel MatchedFilter<double> apply + PeakFinder(0.7, 200), with
    the signal resident in HBM.  Returns the GPU figures and the state the
    CPU-baseline leg checks and times."""
    from mflib_amd import capi, synth
    x, t, pos = synth.single_channel_hour()
    n = x.size
    h = capi.Handle(np.float64, "single")
    h.set_templates([t])
    h.set_signal_size(n)
    h.initialize()
    xd = torch.from_numpy(x).to("cuda:0")
    torch.cuda.synchronize()
    h.set_signal_device(0, xd.data_ptr(), n)
    nout = h.filtered_length(0)
    for _ in range(3):  # warm-up, including the peak finder's one-time workspace
        h.apply()
        h.synchronize()
        capi.peaks_device(h.rho_device_ptr(0), nout, np.float64, 0.7, 200)
    reps = max(50, args.steps * 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        h.apply()
        h.synchronize()  # the peak finder reads the row from its own stream
        pk = capi.peaks_device(h.rho_device_ptr(0), nout, np.float64, 0.7, 200)
    dt = (time.perf_counter() - t0) / reps
    corr_ms = h.timings()[0]
    rho = h.rho(0)
    hits = sum(1 for p in pos if pk.size and np.min(np.abs(pk - p)) <= 2)
    out = {
        "workload": "single channel: 360,000 FP64 samples (1 h at 100 Hz), one 400-sample template "
                    "injected 20 times, apply + PeakFinder(0.7, 200)",
        "corr_samples_per_apply": nout,
        "plan_nfft_L": list(h.plan(0)),
        "ms_per_request": dt * 1e3,
        "corr_kernel_ms": corr_ms,
        "corr_samples_per_s": nout / dt,
        "peaks": int(pk.size),
        "injections_recovered": hits,
        "injections": len(pos),
    }
    h.close()
    del xd
    log(f"single channel hour: {dt * 1e3:.3f} ms per apply+peaks (corr kernel {corr_ms:.3f} ms), "
        f"{nout / dt / 1e9:.2f} G corr samples/s, {hits}/{len(pos)} injections")
    return out, (x, t, rho, pk)


def single_channel_cpu(state, threads):
    """The oracle's restated SingleChannel MatchedFilter<double> apply and
    PeakFinder on the host (median of 5 runs of 20 repeats, after one
    warm-up run), and |drho| of the GPU row against the FP64 two-pass
    Pearson oracle."""
    import oracle as O
    x, t, rho, pk = state
    reps = 20
    box = {}

    def run():
        for _ in range(reps):
            box["rho"], _, _ = O.mf_apply([t], [x], mode="single", dtype=np.float64, threads=threads)
            box["pk"] = O.peak_finder(box["rho"][0], 0.7, 200)

    dt, runs = _median_timed(run)
    dt /= reps
    ref_pk = box["pk"]
    lags = np.random.default_rng(0x51).integers(0, rho.size, 20000)
    exact = O.pearson(t, x, lags)
    return {
        "value": rho.size / dt, "unit": "corr samples/s", "cores": threads, "kind": "port",
        "sample": f"the whole hour ({rho.size} corr samples): median of 5 runs of {reps} repeats of apply + "
                  "PeakFinder, after 1 warm-up run",
        "ms_per_request": dt * 1e3,
        "host": host_info(threads),
    }, {
        "max_abs_drho_vs_fp64": float(np.abs(rho[lags] - exact).max()),
        "audited_samples": int(lags.size),
        "peaks_equal_to_oracle_pipeline": bool(np.array_equal(np.sort(pk), np.sort(ref_pk))),
    }


def _window_fracs(ws):
    """Shares of a handle's windows (mf_get_window_stats) that were
    recomputed exactly (flagged), flagged by the fast path's error bound,
    recomputed in full (cancelled variance), and computed by solo rows."""
    w = max(ws["windows"], 1)
    return {"windows": ws["windows"], "flagged_frac": ws["flagged"] / w,
            "bound_flag_frac": (ws["flagged"] - ws["full"]) / w, "full_recompute_frac": ws["full"] / w,
            "solo_frac": ws["solo"] / w}


def uneven_section(args, torch):
    """Uneven batched workload: 200 templates of 200-3000 samples, each on
    1-50 of 50 channels lasting 1-24 h; moveouts 0-500; 0-30 injections per
    template.  One batched engine instance (FFT plans from the measured
    autotuner, length buckets, apply replayed from a HIP graph) against the
    plain per-template loop (one instance per template, as the reference's
    one-signal-size-per-instance API needs)."""
    from collections import Counter
    from mflib_amd import capi, synth
    log("uneven batch: generating 200 templates x 1-50 channels x 1-24 h")
    wl = synth.uneven_batch()
    # the measured autotuner, or (for profiling, where the profiler skews
    # the calibration launches) the built-in MI355X cost table
    table = None if args.no_autotune else capi.autotune(np.float32)
    rows = wl.rows()
    ns = [s.size for s in wl.signals]
    dev = [torch.from_numpy(s).to("cuda:0") for s in wl.signals]
    torch.cuda.synchronize()
    h = capi.Handle(np.float32, "batched")
    h.set_templates([wl.templates[t][c] for t, c in rows], shifts=[wl.moveouts[t][c] for t, c in rows],
                    weights=[1.0] * len(rows))
    h.set_batch_layout(ns, [c for _, c in rows], [t for t, _ in rows], len(wl.templates))
    h.set_graph_launch(True)
    t0 = time.perf_counter()
    h.initialize()
    init_s = time.perf_counter() - t0
    for c, d in enumerate(dev):
        h.set_signal_device(c, d.data_ptr(), ns[c])
    n_corr = sum(h.filtered_length(r) for r in range(len(rows)))
    plans = Counter(h.plan(r)[0] for r in range(len(rows)))
    for _ in range(2):
        h.apply()
        h.detect(0.4, 500)
    steps = max(3, args.steps // 2)
    corr = []
    t0 = time.perf_counter()
    for _ in range(steps):
        h.apply()
        peaks = h.detect(0.4, 500)
        corr.append(h.timings()[0])
    dt = (time.perf_counter() - t0) / steps
    corr_ms = float(np.median(corr))
    hits = 0
    for t in range(len(wl.templates)):
        det = np.asarray(peaks[t])
        hits += sum(1 for o in wl.origins[t] if det.size and np.min(np.abs(det - o)) <= 2)
    n_inj = sum(len(o) for o in wl.origins)
    # bytes the correlation launch must move: channel reads (once per lane, at
    # least once per channel) and the rho writes
    alg_bytes = 4.0 * n_corr + 4.0 * sum(ns)
    sample_rows = list(range(0, len(rows), max(1, len(rows) // 12)))
    batched_rows = {r: h.rho(r) for r in sample_rows}
    # the stack of every group at each of its injection origins (for the
    # FP64 audit of the stacks and of what the detector can find)
    stack_at_origins = {}
    for t in range(len(wl.templates)):
        if len(wl.origins[t]):
            ptr, _ = h.stack_device(t)
            stack_at_origins[t] = _device_gather_f32(ptr, wl.origins[t])
    ws = h.window_stats()
    h.close()
    log(f"uneven batch: {len(rows)} rows, {n_corr / 1e9:.2f} G corr samples, {dt * 1e3:.1f} ms per step "
        f"(corr {corr_ms:.1f} ms), {n_corr / dt / 1e9:.1f} G corr samples/s, plans {dict(plans)}; "
        f"flagged windows {100.0 * ws['flagged'] / ws['windows']:.3f}%, solo {100.0 * ws['solo'] / ws['windows']:.1f}%")
    # plain per-template loop
    row_index = {rc: i for i, rc in enumerate(rows)}
    plain_rows = {}
    want = {rows[r]: r for r in sample_rows}
    t0 = time.perf_counter()
    total = 0
    for t in range(len(wl.templates)):
        chans = sorted(wl.templates[t])
        n = min(ns[c] for c in chans)
        p = capi.Handle(np.float32, "multi")
        p.set_templates([wl.templates[t][c] for c in chans], shifts=[wl.moveouts[t][c] for c in chans],
                        weights=[1.0] * len(chans))
        p.set_signal_size(n)
        p.initialize()
        for i, c in enumerate(chans):
            p.set_signal_device(i, dev[c].data_ptr(), n)
        p.apply()
        p.detect(0.4, 500)
        total += len(chans) * p.filtered_length(0)
        for i, c in enumerate(chans):
            if (t, c) in want:
                plain_rows[want[(t, c)]] = p.rho(i)
        p.close()
    plain_dt = time.perf_counter() - t0
    diff = max(float(np.abs(batched_rows[r].astype(np.float64) - plain_rows[r]).max()) for r in sample_rows)
    del row_index
    log(f"uneven batch: plain per-template loop {plain_dt:.2f} s, {total / plain_dt / 1e9:.2f} G corr samples/s")
    state = (wl, sample_rows, batched_rows, stack_at_origins, peaks)
    return {
        "workload": "uneven batch: 200 templates (200-3000 samples) x 1-50 of 50 channels (1-24 h at 100 Hz), "
                    "moveouts 0-500, 0-30 injections per template, FP32, stack + PeakFinder(0.4, 500)",
        "rows": len(rows),
        "corr_samples_per_step": n_corr,
        "ms_per_step": dt * 1e3,
        "corr_kernel_ms": corr_ms,
        "corr_samples_per_s": n_corr / dt,
        "corr_hbm_GBps_algorithmic": alg_bytes / (corr_ms * 1e-3) / 1e9,
        "initialize_s": init_s,
        "fft_sizes_rows": {str(k): v for k, v in sorted(plans.items())},
        "autotune_table": table,
        "plain_per_template_loop_corr_samples_per_s": total / plain_dt,
        "batched_vs_plain": (n_corr / dt) / (total / plain_dt),
        "max_abs_drho_batched_vs_plain": diff,
        "injections_recovered": hits,
        "injections": n_inj,
        "window_stats": _window_fracs(ws),
    }, state


def _device_gather_f32(ptr: int, idx):
    """Values of a device float32 array at the given indices (one small copy
    per index through the HIP runtime torch loaded)."""
    import ctypes
    path = "libamdhip64.so"
    with open("/proc/self/maps") as f:       # the runtime torch already loaded
        for line in f:
            if "libamdhip64.so" in line.split()[-1]:
                path = line.split()[-1]
                break
    hip = ctypes.CDLL(path)
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(len(idx), dtype=np.float32)
    for k, i in enumerate(np.asarray(idx).tolist()):
        if hip.hipMemcpy(out.ctypes.data + 4 * k, ptr + 4 * int(i), 4, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
    return out


def uneven_audit(state):
    """FP64 oracle audit of the uneven batch: rho of the sampled rows at
    seeded lags, and every group's stack at its injection origins against
    the oracle's stack value there (mean over the template's channels of
    |rho| at origin + moveout).  The oracle's stack values also say which
    injections the detector can find at all: the generator injects at a
    peak amplitude of 2-8 x the noise RMS, so many bursts stay below the
    threshold even in the exact stack."""
    import oracle as O
    wl, sample_rows, batched_rows, stack_at_origins, peaks = state
    rows = wl.rows()
    rng = np.random.default_rng(0xA0D4)
    worst, n_aud = 0.0, 0
    for r in sample_rows:
        t, c = rows[r]
        g = batched_rows[r]
        lags = rng.integers(0, g.size, 2000)
        ref = O.pearson(wl.templates[t][c], wl.signals[c], lags)
        worst = max(worst, float(np.abs(g[lags].astype(np.float64) - ref).max()))
        n_aud += lags.size
    st_err, above, above_found, margin, margin_found = 0.0, 0, 0, 0, 0
    for t, got in stack_at_origins.items():
        chans = sorted(wl.templates[t])
        ref = np.zeros(len(wl.origins[t]))
        for c in chans:
            lags = np.asarray(wl.origins[t], dtype=np.int64) + wl.moveouts[t][c]
            ref += np.abs(O.pearson(wl.templates[t][c], wl.signals[c], lags))
        ref /= len(chans)
        st_err = max(st_err, float(np.abs(got.astype(np.float64) - ref).max()))
        det = np.asarray(peaks[t])
        for o, v in zip(wl.origins[t], ref):
            hit = int(det.size > 0 and np.min(np.abs(det - o)) <= 2)
            if v > 0.4:
                above += 1
                above_found += hit
            if v > 0.45:
                margin += 1
                margin_found += hit
    return {"max_abs_drho_vs_fp64": worst, "audited_rows": len(sample_rows), "audited_samples": n_aud,
            "max_abs_dstack_at_origins_vs_fp64": st_err,
            "injections_with_fp64_stack_above_threshold_0.4": above,
            "of_those_recovered": above_found,
            "injections_with_fp64_stack_above_0.45": margin, "of_those_recovered_0.45": margin_found}


def streaming_section(args, torch, wl, day_row0):
    """Streaming: the multi-channel day pushed in 60 s chunks (6,000 samples
    per channel per push) into a streaming handle of the same layout; per
    push: 24 device chunks in (one batched push), all 384 rows, 16 stacks
    and PeakFinder.
    Audit: the row tails of template event 0 on channel 0 against the
    full-day row of the same engine (only the FFT blocking differs)."""
    from mflib_amd import capi
    E, C = wl.n_events, wl.n_channels
    chunk, pushes = 6000, 40

    def handle():
        h = capi.Handle(np.float32, "batched")
        h.set_templates([wl.templates[e][c] for e in range(E) for c in range(C)],
                        shifts=[int(wl.moveouts[e, c]) for e in range(E) for c in range(C)], weights=[1.0] * (E * C))
        h.set_batch_layout([chunk] * C, [c for e in range(E) for c in range(C)],
                           [e for e in range(E) for c in range(C)], E)
        h.set_streaming(True)
        h.initialize()
        return h

    h = handle()
    carry = h.stream_carry()
    nb = len(wl.templates[0][0])
    dev = [torch.from_numpy(wl.signals[c][:chunk * pushes]).to("cuda:0") for c in range(C)]
    torch.cuda.synchronize()
    lat, tails = [], []
    for p in range(pushes):
        t0 = time.perf_counter()
        h.push_chunks_device([dev[c][p * chunk:].data_ptr() for c in range(C)], chunk)
        h.apply()
        h.detect(0.5, 512)
        lat.append((time.perf_counter() - t0) * 1e3)
        tails.append(h.rho(0)[-chunk:])
    got = np.concatenate(tails)
    starts = np.arange(chunk * pushes) - nb + 1
    ok = starts >= 0
    drho = float(np.abs(got[ok] - day_row0[starts[ok]]).max())
    lat = np.array(lat[2:])                        # the first pushes include allocation
    n_corr = E * C * chunk
    h.close()
    # the same pushes with peaks continuous across pushes (mf_stream_peaks_all)
    h = handle()
    lat_sp, stacks, got = [], [[] for _ in range(E)], [[] for _ in range(E)]
    for p in range(pushes):
        t0 = time.perf_counter()
        h.push_chunks_device([dev[c][p * chunk:].data_ptr() for c in range(C)], chunk)
        h.apply()
        for g, pk in enumerate(h.stream_peaks_all(0.5, 512, last=(p == pushes - 1))):
            got[g].append(pk)
        lat_sp.append((time.perf_counter() - t0) * 1e3)
        for g in range(E):
            stacks[g].append(h.shift_and_stack(g))
    # every group's stream peaks against PeakFinder over the stream's whole stack
    same = True
    for g in range(E):
        lead = carry - (len(stacks[g][0]) - chunk)
        whole = np.concatenate([st[-chunk:] for st in stacks[g]])[lead:]
        same = same and list(np.concatenate(got[g])) == list(capi.peaks(whole, 0.5, 512))
    h.close()
    lat_sp = np.array(lat_sp[2:])
    res = {"workload": "streaming: 16 templates x 24 channels, 60 s chunks (6000 samples) of the multi-channel "
                       "day, 384 rows + 16 stacks + PeakFinder per push",
           "chunk_samples": chunk, "pushes": pushes, "carry_samples": carry,
           "ms_per_push_p50": float(np.percentile(lat, 50)), "ms_per_push_p99": float(np.percentile(lat, 99)),
           "realtime_factor": 60e3 / float(np.percentile(lat, 50)),
           "new_corr_samples_per_s": n_corr / (float(np.percentile(lat, 50)) * 1e-3),
           "max_abs_drho_vs_day_row": drho,
           "continuous_peaks_ms_per_push_p50": float(np.percentile(lat_sp, 50)),
           "continuous_peaks_ms_per_push_p99": float(np.percentile(lat_sp, 99)),
           "continuous_peaks_equal_whole_stack_peak_finder": bool(same)}
    log(f"streaming: {res['ms_per_push_p50']:.2f} ms per 60 s push (p99 {res['ms_per_push_p99']:.2f}), "
        f"{res['realtime_factor']:.0f}x real time; |drho| vs day row {drho:.2e}")
    log(f"streaming with peaks continuous across pushes (16 groups, one call): "
        f"{res['continuous_peaks_ms_per_push_p50']:.2f} ms per push (p99 "
        f"{res['continuous_peaks_ms_per_push_p99']:.2f}); equal to PeakFinder over the whole stacks: {same}")
    return res


def precision_section(args, torch):
    """Precision stress: 256 templates of 128-256 samples against 4 channels
    of 2,000,000 FP32 samples with a 1e4 DC offset and constant gaps, on
    three paths: FFT FP32 (wave-level kernel), and the time-domain MFMA
    path with FP16 and with BF16 hi/lo operands.  Keeps rows and peaks for
    the FP64-oracle audit in the CPU leg."""
    from mflib_amd import capi, synth
    log("precision stress: generating 4 channels x 2M samples, 256 templates")
    wl = synth.precision_stress()
    C, T = len(wl.signals), len(wl.templates)
    n = wl.signals[0].size
    dev = [torch.from_numpy(s).to("cuda:0") for s in wl.signals]
    torch.cuda.synchronize()
    rng = np.random.default_rng(0x93)
    audit_rows = [(int(rng.integers(0, C)), int(rng.integers(0, T))) for _ in range(6)]
    # lags: random, inside gaps and across gap edges
    lag_sets = {}
    for c, t in audit_rows:
        g = wl.gaps[c]
        lags = list(rng.integers(0, n - 256, 1500))
        for s0, L in g[:8]:
            lags += [s0 - 100, s0 - 3, s0, s0 + 5, s0 + L // 2, s0 + L - 130, s0 + L - 2, s0 + L + 10]
        lag_sets[(c, t)] = np.clip(np.array(lags, dtype=np.int64), 0, n - 257)
    paths = {}
    for name, variant in (("fft_fp32", 0), ("mfma_fp16_split2", 3), ("mfma_bf16_split3", 4)):
        h = capi.Handle(np.float32, "batched")
        h.set_templates([wl.templates[t] for c in range(C) for t in range(T)])
        h.set_batch_layout([n] * C, [c for c in range(C) for _ in range(T)], [c for c in range(C) for _ in range(T)], C)
        h.set_kernel_variant(variant)
        h.initialize()
        for c in range(C):
            h.set_signal_device(c, dev[c].data_ptr(), n)
        h.apply()
        h.synchronize()
        ms = []
        for _ in range(3):
            h.apply()
            h.synchronize()
            ms.append(h.timings()[0])
        n_corr = sum(h.filtered_length(r) for r in range(C * T))
        corr_ms = float(np.median(ms))
        ws = h.window_stats()
        rows = {}
        for c, t in audit_rows:
            r = h.rho(c * T + t)
            pk = capi.peaks_device(h.rho_device_ptr(c * T + t), r.size, np.float32, 0.6, 200)
            rows[(c, t)] = (r[lag_sets[(c, t)]].astype(np.float64), pk)
        if name == "fft_fp32":
            # Detector reduce step over each channel's 256 template rows
            glen = h.filtered_length(0)
            d_det = torch.empty(glen, dtype=torch.float32, device="cuda:0")
            d_id = torch.empty(glen, dtype=torch.int32, device="cuda:0")
            red_ms = []
            for rep in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for c in range(C):
                    h.reduce_device(c, True, d_det.data_ptr(), d_id.data_ptr(), glen)
                e1.record()
                torch.cuda.synchronize()
                red_ms.append(e0.elapsed_time(e1))
            red = float(np.median(red_ms))
            nred = C * T * glen
            # audit: the first 20,000 lags of channel 1 (rows T..2T-1), ids are template indices
            sl = 20_000
            det1, id1 = h.reduce(1, True)
            reduce_audit = (np.stack([h.rho(T + t)[:sl] for t in range(T)]), det1[:sl].copy(), id1[:sl] - T)
            reduce_info = {"ms_all_4_channels": red, "corr_samples_reduced_per_s": nred / (red * 1e-3),
                           "hbm_GBps": (nred * 4.0 + C * glen * 8.0) / (red * 1e-3) / 1e9,
                           "hbm_frac_of_8TBps": (nred * 4.0 + C * glen * 8.0) / (red * 1e-3) / 8e12}
            log(f"detector reduce (256 templates x 4 channels, |max|): {red:.2f} ms, "
                f"{reduce_info['hbm_GBps']:.0f} GB/s")
            # per-detection products (Detector::detect's refinement and scaling
            # factors) for 16,384 detections on random rows and lags, |rho|
            prng = np.random.default_rng(0x51)
            npk, alpha = 16384, 51
            p_tid = prng.integers(0, C * T, npk).astype(np.int32)
            p_lag = prng.integers(alpha + 1, glen - alpha - 2, npk).astype(np.int32)
            dts = np.full(C * T, 0.01)
            products_info = {"detections": npk}
            for kind, interp in (("quadratic", capi.INTERP_QUADRATIC), ("lanczos_51_100", capi.INTERP_LANCZOS)):
                got = h.detection_products(p_lag, p_tid, dts, True, interp, alpha, 100)
                walls = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    got = h.detection_products(p_lag, p_tid, dts, True, interp, alpha, 100)
                    walls.append((time.perf_counter() - t0) * 1e3)
                wall = float(np.median(walls))
                products_info[kind] = {"ms_per_call": wall, "detections_per_s": npk / (wall * 1e-3)}
                log(f"detection products {kind}: {wall:.2f} ms per call of {npk} detections")
            # the Detector's default (SINGLE) policy over all 1024 rows: one
            # batch against the per-template loop (timed on 32 rows, scaled)
            h.detect_templates(0.5, 200, True)
            t0 = time.perf_counter()
            lists = h.detect_templates(0.5, 200, True)
            single_batch = (time.perf_counter() - t0) * 1e3
            t0 = time.perf_counter()
            same = all(np.array_equal(h.detect_template(r, 0.5, 200, True), lists[r]) for r in range(32))
            single_loop = (time.perf_counter() - t0) * 1e3 / 32 * (C * T)
            products_info["single_policy_peak_lists"] = {
                "rows": C * T, "ms_batched": single_batch, "ms_per_template_loop_scaled": single_loop,
                "peaks": int(sum(len(x) for x in lists)), "first_32_rows_identical": bool(same)}
            log(f"detector SINGLE policy, {C * T} rows: {single_batch:.1f} ms batched, "
                f"{single_loop:.1f} ms as a per-template loop")
            # audit sample: 24 detections, their lags' neighbourhoods and windows
            sample = []
            for k in range(24):
                r, pk = int(p_tid[k]), int(p_lag[k])
                c, t = divmod(r, T)
                lt = len(wl.templates[t])
                sample.append((h.rho(r)[pk - alpha:pk + alpha + 2].astype(np.float64), c, t, pk,
                               wl.signals[c][pk:pk + lt].astype(np.float64),
                               {kk: float(v[k]) for kk, v in got.items()}))
            products_audit = sample
        plan = h.plan(0)
        h.close()
        paths[name] = {"corr_kernel_ms": corr_ms, "corr_samples_per_s": n_corr / (corr_ms * 1e-3),
                       "plan_row0_nfft": plan[0], "window_stats": _window_fracs(ws), "rows": rows}
        fr = paths[name]["window_stats"]
        log(f"precision stress {name}: corr {corr_ms:.2f} ms, {n_corr / (corr_ms * 1e-3) / 1e9:.1f} G corr samples/s, "
            f"flagged windows {100.0 * fr['flagged_frac']:.3f}% (error bound {100.0 * fr['bound_flag_frac']:.3f}%, "
            f"cancelled variance {100.0 * fr['full_recompute_frac']:.3f}%), solo {100.0 * fr['solo_frac']:.1f}%")
    # normalization scans: window scales of every audited (channel, template
    # length) from FP64 prefix scans and from compensated FP32 scans, kept at
    # the audit lags; throughput on a 64M-sample signal
    scans = {}
    for scan in ("f64", "f32c"):
        got = {}
        for c, t in audit_rows:
            lt = len(wl.templates[t])
            d_out = torch.empty(n - lt + 1, dtype=torch.float32, device="cuda:0")
            capi.window_scale_device(dev[c].data_ptr(), n, lt, scan, d_out.data_ptr())
            got[(c, t)] = d_out.cpu().numpy()[lag_sets[(c, t)]].astype(np.float64)
        big = torch.randn(1 << 26, dtype=torch.float32, device="cuda:0") + 1.0e4
        d_big = torch.empty(big.numel() - 255, dtype=torch.float32, device="cuda:0")
        capi.window_scale_device(big.data_ptr(), big.numel(), 256, scan, d_big.data_ptr())
        ms = float(np.median([capi.window_scale_device(big.data_ptr(), big.numel(), 256, scan, d_big.data_ptr())
                              for _ in range(5)]))
        nbytes = 4.0 * big.numel() + 4.0 * d_big.numel()
        scans[scan] = {"rows": got, "kernel_ms_64M_windows_Lt256": ms,
                       "windows_per_s": d_big.numel() / (ms * 1e-3),
                       "hbm_GBps": nbytes / (ms * 1e-3) / 1e9, "hbm_frac_of_8TBps": nbytes / (ms * 1e-3) / 8e12}
        del big, d_big
        log(f"normalization scans {scan}: {ms:.3f} ms for 64M windows, {nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
    out = {
        "workload": "precision stress: 256 templates (128-256 samples, 2-20 Hz) x 4 channels x 2,000,000 FP32 "
                    "samples, DC 1e4, 20 constant gaps of 500-5000 samples per channel",
        "corr_samples_per_apply": n_corr,
        "paths": {k: {kk: vv for kk, vv in v.items() if kk != "rows"} for k, v in paths.items()},
        "normalization_scans": {k: {kk: vv for kk, vv in v.items() if kk != "rows"} for k, v in scans.items()},
        "detector_reduce": reduce_info,
        "detection_products": products_info,
    }
    return out, (wl, audit_rows, lag_sets, paths, scans, reduce_audit, products_audit)


def precision_cpu(state):
    """FP64 oracle audit of the precision-stress rows of every path, and
    peak-index agreement against PeakFinder(0.6, 200) on the oracle rows."""
    import oracle as O
    wl, audit_rows, lag_sets, paths, scans, reduce_audit, products_audit = state
    res = {name: {"max_abs_drho_vs_fp64": 0.0, "peak_sets_equal": 0, "rows": len(audit_rows)} for name in paths}
    for scan in scans:
        res["scan_" + scan] = {"max_rel_scale_err": 0.0, "max_abs_drho_vs_fp64": 0.0, "rows": len(audit_rows),
                               "note": "rho = exact FP64 numerator x this scale (the scan's own error only)"}
    for k, (c, t) in enumerate(audit_rows):
        lags = lag_sets[(c, t)]
        ref = O.pearson(wl.templates[t], wl.signals[c], lags)
        sref = O.window_scale(wl.signals[c], len(wl.templates[t]), lags)
        for scan, p in scans.items():
            ok = sref > 0
            rel = np.abs(p["rows"][(c, t)][ok] / sref[ok] - 1.0)
            r = res["scan_" + scan]
            # windows that are exactly constant in FP32 (inside the gaps):
            # rho is 0 there; a scan that leaves a positive bracket gives a
            # spurious finite scale, which the fused kernel's flags catch
            r["constant_windows"] = r.get("constant_windows", 0) + int((~ok).sum())
            r["constant_windows_with_nonzero_scale"] = (r.get("constant_windows_with_nonzero_scale", 0)
                                                        + int((p["rows"][(c, t)][~ok] != 0).sum()))
            r["max_rel_scale_err"] = max(r["max_rel_scale_err"], float(rel.max()) if rel.size else 0.0)
            r["max_abs_drho_vs_fp64"] = max(r["max_abs_drho_vs_fp64"],
                                            float((np.abs(ref[ok]) * rel).max()) if rel.size else 0.0)
        full = None
        if k < 2:
            full = O.pearson(wl.templates[t], wl.signals[c])[: wl.signals[c].size - 256 + 1]
            ref_pk = O.peak_finder(full, 0.6, 200)
        for name, p in paths.items():
            vals, pk = p["rows"][(c, t)]
            res[name]["max_abs_drho_vs_fp64"] = max(res[name]["max_abs_drho_vs_fp64"],
                                                    float(np.abs(vals - ref).max()))
            if full is not None:
                res[name]["peak_sets_equal"] += int(np.array_equal(np.sort(pk), np.sort(ref_pk)))
    rows1, det1, id1 = reduce_audit
    rdet, rid = O.reduce_correlations(rows1, True)
    res["detector_reduce"] = {"lags_checked": int(rdet.size), "templates": int(rows1.shape[0]),
                              "det_bit_identical": bool(np.array_equal(rdet, det1)),
                              "id_identical": bool(np.array_equal(rid, id1))}
    # products (Lanczos 51/100, |rho|): the GPU's grid point reaches the
    # oracle interpolant's maximum; scaling factors vs the oracle
    worst_gap, worst_rel, ok_flags = 0.0, 0.0, True
    for seg, c, t, pk, win, got in products_audit:
        _, prof = O.lanczos_shift(seg, 51, 0.01, True, 51, 100)   # the segment holds every tap
        j = int(round((got["shift"] / 0.01 + 0.5) * 100)) - 1
        worst_gap = max(worst_gap, float(prof.max() - prof[j]))
        ref = O.amplitude_factors(wl.templates[t], win)
        ok_flags &= (ref is not None) == bool(int(got["flags"]) & 1)
        if ref is not None:
            worst_rel = max(worst_rel, abs(got["alpha_gr"] / ref[0] - 1), abs(got["alpha_sr"] / ref[1] - 1))
    res["detection_products"] = {"detections_checked": len(products_audit),
                                 "max_lanczos_gap_to_oracle_max": worst_gap,
                                 "max_rel_err_scaling_factors": worst_rel, "amplitude_flags_equal": bool(ok_flags)}
    for name in res:
        if name in ("detector_reduce", "detection_products"):
            continue
        if name in paths:
            res[name]["peak_rows_checked"] = 2
        res[name]["audited_samples"] = int(sum(lag_sets[r].size for r in audit_rows))
    return res


def serve_audit(wl, audit):
    """FP64 oracle check of the audited served rows (1 in 100 requests, one
    (template, channel) row each, 2000 seeded lags per row)."""
    import oracle as O
    rng = np.random.default_rng(0x5E7)
    worst, n = 0.0, 0
    for (chans, start, nwin, ids), ti, ci, row in audit:
        x = wl.streams[chans[ci]][start:start + nwin]
        lags = rng.integers(0, row.size, 2000)
        ref = O.pearson(wl.library[ids[ti]], x, lags)
        worst = max(worst, float(np.abs(row[lags].astype(np.float64) - ref).max()))
        n += lags.size
    return {"max_abs_drho_vs_fp64": worst, "audited_rows": len(audit), "audited_samples": n}


def host_info(threads: int) -> dict:
    """Host facts the CPU figures depend on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads_used": threads, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def _median_timed(fn, runs: int = 5):
    """One untimed warm-up call, then the median wall time of `runs` calls."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(wl, threads: int, accuracy_rows, gpu_rho_rows, sample_samples: int = 720_000):
    """Reference-faithful CPU restatement (oracle) on a bounded sample: one
    template event on all 24 channels over the first `sample_samples`
    samples (2 h at 100 Hz; the restated sizing picks the same nfft 8192 as
    for the whole day) in the reference's MatchedFilter<float> multi-channel
    form; median of 5 runs after one warm-up.  Also audits the GPU rows
    against the FP64 Pearson oracle at seeded lags."""
    import oracle as O
    O.lib()
    ev = 0
    tmpl = wl.templates[ev]
    sigs = [np.ascontiguousarray(s[:sample_samples]) for s in wl.signals]
    box = {}

    def ref_threading():
        box["rho"], box["nfft"], box["L"] = O.mf_apply(tmpl, sigs, mode="multi", dtype=np.float32, threads=threads)

    dt, runs = _median_timed(ref_threading)
    rho, nfft = box["rho"], box["nfft"]
    n_corr = rho.shape[0] * rho.shape[1]
    # all-cores form: the same event as 24 independent one-channel instances,
    # one per worker, each with one OpenMP thread (ctypes releases the GIL)
    from concurrent.futures import ThreadPoolExecutor

    def all_cores():
        with ThreadPoolExecutor(max_workers=threads) as ex:
            box["parts"] = list(ex.map(lambda c: O.mf_apply([tmpl[c]], [sigs[c]], mode="multi", dtype=np.float32,
                                                            threads=1)[0], range(len(tmpl))))

    dt_all, runs_all = _median_timed(all_cores)
    all_same = all(np.array_equal(box["parts"][c][0], rho[c]) for c in range(len(tmpl)))
    del box["parts"]
    # accuracy audit of the GPU result vs the FP64 oracle
    rng = np.random.default_rng(0xA0D17)
    worst = 0.0
    n_aud = 0
    for (e, c), g in zip(accuracy_rows, gpu_rho_rows):
        lags = rng.integers(0, len(g), 4000)
        ref = O.pearson(wl.templates[e][c], wl.signals[c], lags)
        worst = max(worst, float(np.abs(g[lags].astype(np.float64) - ref).max()))
        n_aud += len(lags)
    return {
        "value": n_corr / dt,
        "unit": "corr samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"1 of {wl.n_events} template events, {len(tmpl)} channels x the first {sample_samples} "
                   f"samples ({n_corr} corr samples), through the restated MatchedFilter<float> apply "
                   f"(own radix-2 FFT, nfft={nfft} as for the whole day, serial block loop as "
                   f"mkl_sequential, OpenMP normalization on {threads} threads); median of 5 runs after "
                   "1 warm-up. MKL is not available, so an MKL build of the reference would be faster"),
        "seconds_median": dt,
        "seconds_runs": runs,
        "host": host_info(threads),
        "all_cores": {"value": n_corr / dt_all, "cores": threads, "seconds_median": dt_all,
                      "seconds_runs": runs_all,
                      "sample": f"the same event as {len(tmpl)} independent one-channel instances on {threads} "
                                "worker threads, one OpenMP thread each (a stronger baseline than the "
                                "reference's own threading); median of 5 runs after 1 warm-up",
                      "rows_identical_to_reference_threading": bool(all_same)},
    }, {"max_abs_drho_vs_fp64": worst, "audited_samples": n_aud}


def hbm_probe() -> int:
    """Plain device streams of 1 and 4 GiB, one JSON line each: a copy (read +
    write), a sum (read only) and the maximum over 256 rows (read, 1/256
    written; the access pattern of the Detector reduce and of the stack).
    Each is warmed up once and timed over 5 calls with HIP events."""
    import torch

    def timed(f, reps=5):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    for gib in (1, 4):
        n = gib * (1 << 30) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        rows = a.view(256, -1)
        for name, f, nbytes in (("copy", lambda: b.copy_(a), 8 * n),
                                ("sum", lambda: a.sum(), 4 * n),
                                ("max over 256 rows", lambda: torch.amax(rows, dim=0), 4 * n + 4 * n // 256)):
            ms = timed(f)
            print(json.dumps({"hbm_probe": name, "GiB": gib, "ms": round(ms, 4),
                              "TB_per_s": round(nbytes / (ms * 1e-3) / 1e12, 3)}), flush=True)
        del a, b, rows
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--samples", type=int, default=8_640_000, help="samples per channel")
    ap.add_argument("--nfft", type=int, default=0, help="force the FFT length (0 = autotuner)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="uneven batch: plan from the built-in cost table (reproducible under a profiler)")
    ap.add_argument("--variant", type=int, default=0, help="correlation kernel variant (0 = best)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven batched workload")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming workload")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-stress workload")
    ap.add_argument("--precision-only", action="store_true", help="only the precision-stress workload")
    ap.add_argument("--uneven-only", action="store_true", help="only the uneven batched workload")
    ap.add_argument("--serve-only", action="store_true", help="only the served-stream measurement")
    ap.add_argument("--serve-lambda", type=float, nargs="+", default=[10.0, 100.0, 500.0, 1000.0, 2000.0])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-requests", type=int, default=400, help="requests of the capacity run")
    ap.add_argument("--serve-ref-requests", type=int, default=6)
    ap.add_argument("--serve-batch", type=int, default=256)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy BYTES device-to-device 3 times (calibrates rocprof HBM counters)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only time plain device streams of 1 and 4 GiB (copy, sum, max over 256 rows): "
                         "the practical HBM rates the stack and reduce kernels compare with")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); sharding the path across GPUs "
              "(by template rows or time blocks) is not implemented", file=sys.stderr)
        return 2
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3, "ms_per_copy": ms,
                          "copy_GBps_read_plus_write": 2 * a.numel() * 4 / (ms * 1e-3) / 1e9}))
        return 0

    if args.hbm_probe:
        return hbm_probe()

    import torch  # before the engine library: one HIP runtime per process
    from mflib_amd import capi, synth

    if args.precision_only:
        out, state = precision_section(args, torch)
        if not args.no_cpu_baseline:
            out["accuracy"] = precision_cpu(state)
        print(json.dumps({"precision_stress": out}), flush=True)
        return 0
    if args.uneven_only:
        uneven, ustate = uneven_section(args, torch)
        if not args.no_cpu_baseline:
            uneven["accuracy"] = uneven_audit(ustate)
        print(json.dumps({"uneven_batch": uneven}), flush=True)
        return 0
    if args.serve_only:
        serving, (swl, audit) = serve_section(args, torch)
        if not args.no_cpu_baseline:
            serving["accuracy"] = serve_audit(swl, audit)
        print(json.dumps({"serving": serving}), flush=True)
        return 0

    log("generating the synthetic day (16 templates x 24 channels)")
    t0 = time.perf_counter()
    wl = synth.multichannel_day(n_samples=args.samples)
    log(f"generated in {time.perf_counter() - t0:.1f} s")
    E, C = wl.n_events, wl.n_channels
    tlen = len(wl.templates[0][0])
    n = wl.signals[0].size

    h = capi.Handle(np.float32, "batched")
    tmpl = [wl.templates[e][c] for e in range(E) for c in range(C)]
    shifts = [int(wl.moveouts[e, c]) for e in range(E) for c in range(C)]
    h.set_templates(tmpl, shifts=shifts, weights=[1.0] * (E * C))
    h.set_batch_layout([n] * C, [c for e in range(E) for c in range(C)],
                       [e for e in range(E) for c in range(C)], E)
    if args.nfft:
        h.set_fft_length(args.nfft)
    if args.variant:
        h.set_kernel_variant(args.variant)
    t0 = time.perf_counter()
    h.initialize()
    log(f"initialized in {time.perf_counter() - t0:.2f} s; plan nfft/L = {h.plan(0)}")
    dev = torch.device("cuda:0")
    for c in range(C):
        x = torch.from_numpy(wl.signals[c]).to(dev)
        h.set_signal_device(c, x.data_ptr(), n)
        del x
    torch.cuda.synchronize()
    n_out = h.filtered_length(0)
    n_corr = E * C * n_out
    def step():
        h.apply()
        return h.detect(0.5, 512)

    log(f"warmup {args.warmup} step(s)")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} step(s)")
    corr_ms = []
    stack_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        peaks = step()
        c_ms, s_ms = h.timings()
        corr_ms.append(c_ms)
        stack_ms.append(s_ms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = n_corr / (dt / args.steps)
    corr_avg = float(np.mean(corr_ms))
    bytes_per_launch = n_corr * (4.0 / E + 4.0)
    achieved = bytes_per_launch / (corr_avg * 1e-3)
    nfft = h.plan(0)[0]
    # FP32 kernel choice: the wave-level kernel exists and is the default for
    # M = nfft/2 <= 2048; the block-level kernel otherwise
    wave = args.variant in (0, 2) and nfft <= 4096
    kernel = "corr2_kernel" if wave else "corr_kernel"
    traffic, traffic_src = measured_traffic(f"{E}x{C}x{n}", nfft, kernel)
    hits = 0
    for e in range(E):
        det = np.asarray(peaks[e])
        hits += sum(1 for o in wl.origins[e] if det.size and np.min(np.abs(det - o)) <= 2)
    ws = h.window_stats()
    log(f"{ms_per_step:.2f} ms/step, corr kernels {corr_avg:.2f} ms, "
        f"{value / 1e9:.1f} G corr samples/s; detections {hits}/{E * len(wl.origins[0])}; "
        f"flagged windows {100.0 * ws['flagged'] / ws['windows']:.4f}%, solo {100.0 * ws['solo'] / ws['windows']:.1f}%")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "corr samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded band-passed noise with injected templates)",
        "config": {
            "workload": "multi-channel detection: 16 templates x 24 channels x 8.64M samples "
                        "(1 day at 100 Hz), 512-sample templates, FP32, moveouts U[0,200], "
                        "stack + PeakFinder(0.5, 512)",
            "templates": E, "channels": C, "samples_per_channel": n, "template_length": tlen,
            "corr_samples_per_step": n_corr, "nfft": h.plan(0)[0], "block_length": h.plan(0)[1],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
            "traffic_source": traffic_src,
            "kernel": (f"{kernel}<float,{nfft // 2}{',true' if kernel == 'corr2_kernel' else ''}> "
                       "(FFT correlation + fused normalization)"),
            "algorithmic_bytes_per_corr_sample": 4.0 / E + 4.0,
            "avg_launch_ms": corr_avg,
        },
        "stack_ms_all_groups": float(np.mean(stack_ms)),
        "window_stats": _window_fracs(ws),
        "detections_recovered": hits,
        "detections_injected": E * len(wl.origins[0]),
    }
    audit_rows = [h.rho(e * C + c) for e, c in [(e, c) for e in (0, E - 1) for c in (0, C // 2, C - 1)]]
    h.close()
    if not args.no_stream:
        out["streaming"] = streaming_section(args, torch, wl, audit_rows[0])
    sc_out, sc_state = single_channel_section(args, torch)
    out["single_channel"] = sc_out
    serve_state = None
    if not args.no_serve:
        serving, serve_state = serve_section(args, torch)
        out["serving"] = serving
        lat = min(serving["poisson"], key=lambda r: abs(r["lambda_per_s"] - 1000.0))
        out["latency"] = {"p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"],
                          "lambda_per_s": lat["lambda_per_s"], "workload": "served stream (see serving)"}
    uneven_state = None
    if not args.no_uneven:
        out["uneven_batch"], uneven_state = uneven_section(args, torch)
    prec_state = None
    if not args.no_precision:
        out["precision_stress"], prec_state = precision_section(args, torch)
    if not args.no_cpu_baseline:
        log("CPU baseline (bounded sample) and accuracy audit")
        # the host threads of this GPU's share of the box: OMP_NUM_THREADS
        # when the environment sets it, else every CPU this process may use
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        rows = [(e, c) for e in (0, E - 1) for c in (0, C // 2, C - 1)]
        cb, acc = cpu_baseline(wl, threads, rows, audit_rows)
        out["cpu_baseline"] = cb
        out["accuracy"] = acc
        if serve_state is not None:
            out["serving"]["accuracy"] = serve_audit(*serve_state)
        if prec_state is not None:
            out["precision_stress"]["accuracy"] = precision_cpu(prec_state)
        if uneven_state is not None:
            out["uneven_batch"]["accuracy"] = uneven_audit(uneven_state)
        sc_cpu, sc_acc = single_channel_cpu(sc_state, threads)
        out["single_channel"]["cpu_baseline"] = sc_cpu
        out["single_channel"]["accuracy"] = sc_acc
    print(json.dumps(out), flush=True)
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
