"""Writes tests/golden/golden_vectors.npz: small seeded inputs and the CPU
oracle's outputs for them, frozen as fixtures.

The reference holds no fixture files (its tests draw unseeded random
numbers, testing/utilities.cpp:83-84), so these vectors come from the
oracle in oracle/, whose functions are pinned by the reference's
known-answer tests and by closed forms in tests/test_oracle.py and
tests/test_products_oracle.py.  The fixture freezes those outputs: the CPU
suite checks that the oracle still reproduces them, and the GPU suite
checks the engine against them.  The PeakFinder case carries the
reference's own expected index lists (testing/peakFinder.cpp:16-59).

Run from the repository root after building the oracle:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
import refcases as R  # noqa: E402

SEED = 0x4D464C00
N_MULTI = 6000                   # samples per channel, multi-channel case
LENS_MULTI = (150, 256, 400)     # one template per channel
SHIFTS = (0, 7, 19)              # moveouts in samples
WEIGHTS = (1.0, 0.5, 0.75)
EVENTS = (1000, 3500, 5000)      # injection lags of the stack (before moveout)
N_SINGLE = 5000
LENS_SINGLE = (120, 333, 257)    # uneven templates on one channel
GAP = (2000, 2600)               # constant stretch of channel 1 and of the single channel
SIZING = ((256, 20000), (400, 360399), (512, 8640511), (1000, 100000), (3000, 1000000), (128, 2000))


def _tapered(rng, n):
    return rng.standard_normal(n) * np.hanning(n)


def build():
    rng = np.random.default_rng(SEED)
    g = {}

    # multi-channel: one template per channel, DC offsets, a constant gap, injected events
    tpls = [_tapered(rng, L) for L in LENS_MULTI]
    sigs = []
    for c, dc in enumerate((0.0, 50.0, -1000.0)):
        x = rng.standard_normal(N_MULTI) * 0.5 + dc
        if c == 1:
            x[GAP[0]:GAP[1]] = dc + 0.25
        for e in EVENTS:
            p = e + SHIFTS[c]
            x[p:p + LENS_MULTI[c]] += 2.0 * tpls[c]
        sigs.append(x)
    nb = max(LENS_MULTI)
    nout = N_MULTI - nb + 1
    rho64 = np.stack([O.pearson(t, s)[:nout] for t, s in zip(tpls, sigs)])
    sigs32 = [s.astype(np.float32) for s in sigs]
    rho32 = np.stack([O.pearson(t, s)[:nout] for t, s in zip(tpls, sigs32)])
    for i, t in enumerate(tpls):
        g[f"multi_template_{i}"] = t
        g[f"multi_signal_{i}"] = sigs[i]
    g["multi_shifts"] = np.array(SHIFTS, np.int32)
    g["multi_weights"] = np.array(WEIGHTS)
    g["multi_rho_f64"] = rho64
    g["multi_rho_f32in"] = rho32
    g["multi_stack_abs"] = O.shift_and_stack(rho64, SHIFTS, WEIGHTS, absval=True)
    g["multi_stack_raw"] = O.shift_and_stack(rho64, SHIFTS, WEIGHTS, absval=False)
    g["multi_stack_peaks"] = O.peak_finder(g["multi_stack_abs"], 0.3, 200).astype(np.int32)

    # single channel: uneven templates cut from the signal, the reduce and the products
    x = rng.standard_normal(N_SINGLE) + 3.0
    x[GAP[0]:GAP[1]] = 3.0
    cuts = (40, 1700, 3900)
    stpls = [x[c:c + L].copy() for c, L in zip(cuts, LENS_SINGLE)]
    nbs = max(LENS_SINGLE)
    nouts = N_SINGLE - nbs + 1
    srho = np.stack([O.pearson(t, x)[:nouts] for t in stpls])
    g["single_signal"] = x
    for i, t in enumerate(stpls):
        g[f"single_template_{i}"] = t
    g["single_rho_f64"] = srho
    for absval in (0, 1):
        det, who = O.reduce_correlations(srho, absval=bool(absval))
        g[f"single_reduce_det_{absval}"] = det
        g[f"single_reduce_id_{absval}"] = who
    # per-detection products at each template's own cut (rho = 1) and at the
    # largest local maxima of its row elsewhere
    peaks, tids = [], []
    for i in range(len(stpls)):
        cand = O.peak_finder(srho[i], 0.0, 150)
        top = cand[np.argsort(-srho[i][cand], kind="stable")[:5]]
        for p in sorted(set(int(v) for v in top) | {cuts[i]}):
            peaks.append(p)
            tids.append(i)
    g["single_peaks"] = np.array(peaks, np.int32)
    g["single_peak_template"] = np.array(tids, np.int32)
    dt = 0.01
    quad, lanc, prof, gr, sr, ok = [], [], [], [], [], []
    for p, i in zip(peaks, tids):
        quad.append(O.quadratic_shift(srho[i], p, dt))
        s, pr = O.lanczos_shift(srho[i], p, dt)
        lanc.append(s)
        prof.append(pr)
        f = O.amplitude_factors(stpls[i], x[p:p + len(stpls[i])]) if p + len(stpls[i]) <= N_SINGLE else None
        ok.append(f is not None)
        gr.append(f[0] if f else 0.0)
        sr.append(f[1] if f else 0.0)
    g["single_dt"] = np.array(dt)
    g["single_quadratic_shift"] = np.array(quad)
    g["single_lanczos_shift"] = np.array(lanc)
    g["single_lanczos_profile"] = np.stack(prof)
    g["single_alpha_gr"] = np.array(gr)
    g["single_alpha_sr"] = np.array(sr)
    g["single_amplitude_ok"] = np.array(ok)

    # PeakFinder: the reference's Gaussian-sine train and its expected lists
    sig, thr, md, ref13, ref21 = R.peak_finder_signal()
    g["pf_signal"] = sig
    g["pf_threshold"] = np.array(thr)
    g["pf_min_dist"] = np.array(md, np.int32)
    g["pf_expected_min_dist"] = np.array(ref13, np.int32)
    g["pf_expected_no_min_dist"] = np.array(ref21, np.int32)

    # FFT and block sizing
    rows = [(nb_, nx, *O.window_length(nb_, nx)) for nb_, nx in SIZING]
    g["sizing"] = np.array(rows, np.int64)
    return g


def main():
    g = build()
    path = os.path.join(HERE, "golden_vectors.npz")
    np.savez_compressed(path, **g)
    print(f"wrote {path}: {len(g)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
