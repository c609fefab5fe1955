"""Inputs of the reference's own tests, regenerated exactly.

The reference ships no fixture files; its deterministic tests build their
inputs analytically.  These functions rebuild them so the same checks can
run against the oracle (CPU) and the GPU engine.
"""
from __future__ import annotations

import numpy as np


def basic_test_inputs():
    """testing/matchedFilter.cpp:62-95 (basicTestDouble / basicTestFloat)."""
    n = 1000
    tmax = 8.0
    dt = tmax / (n - 1)
    f0 = 4.0 / tmax
    amp = 2.0
    t = np.arange(n) * dt
    x = amp * np.exp(-0.1 * t) * np.sin(2 * np.pi * f0 * t)
    x2 = amp * np.exp(-0.15 * t) * np.cos(2 * np.pi * f0 * t)
    nb = int(2.0 / dt + 0.5) + 1
    tb = np.arange(nb) * dt
    b = np.sin(2 * np.pi * f0 * tb)
    b2 = np.cos(2 * np.pi * f0 * tb)
    return x, x2, b, b2


def shift_and_stack_inputs():
    """testing/matchedFilter.cpp:275-352 (shiftAndStack).  Returns
    (templates, signals, travel_times, onsets, fs, origin_time)."""
    tmax, df = 120.0, 100.0
    dt = 1 / df
    n = int(tmax / dt + 0.5) + 1
    tdur = 4.0
    tsize = int(tdur / dt + 0.5) + 1
    toff = [1.0, 2.0, 1.5]
    nextra = [int(o / dt + 0.5) for o in toff]
    nalloc = max(tsize + e for e in nextra)
    tpl = [np.zeros(nalloc) for _ in range(3)]
    sig = [np.zeros(n) for _ in range(3)]
    origin = 8.0
    tt = [20.0, 40.0, 80.0]
    ids = [int((origin + t) / dt + 0.5) for t in tt]
    for i in range(n):
        s1 = 0.001
        if i % 2 == 0:
            s1 = -s1
        for k in range(3):
            sig[k][i] = s1
    amps = [(2.0, 1.0), (1.5, 1.5), (1.2, 2.5)]
    for i in range(tsize):
        t = i * dt
        vals = [a * np.exp(-d * t) * np.sin(2 * np.pi * t / tdur) for a, d in amps]
        if i == 0 or i == tsize - 1:
            vals = [0.0, 0.0, 0.0]
        for k in range(3):
            tpl[k][i + nextra[k]] = vals[k]
            sig[k][ids[k] + i] = vals[k]
    templates = [tpl[k][: tsize + nextra[k]].copy() for k in range(3)]
    return templates, sig, tt, toff, df, origin


def moveout_shift(travel: float, onset: float, fs: float) -> int:
    """nshift = int((travel - onset) * fs + 0.5) (src/matchedFilter.cpp:580-584)."""
    return int((travel - onset) * fs + 0.5)


def gaussian_sine(tdur, dt, sigma, sine_freq):
    """testing/peakFinder.cpp:146-171."""
    t0, t1 = -tdur / 2, tdur / 2
    npts = int((t1 - t0) / dt + 0.5) + 1
    res = np.zeros(npts)
    for i in range(npts):
        t = t0 + i * dt
        e = (t / sigma) * (t / sigma)
        s = 2 * np.pi * (t - t0) / tdur * sine_freq
        res[i] = np.exp(-(e * e)) * np.cos(s)
    res = res / np.max(np.abs(res))
    res[0] = 0
    res[-1] = 0
    return res


def peak_finder_signal():
    """testing/peakFinder.cpp:16-59: the Gaussian-sine train and KATs."""
    n = 4000
    tdur, dt = 4.0, 0.01
    tpl = gaussian_sine(tdur, dt, tdur * 0.3, 12.0)
    sig = np.zeros(n)
    ihalf = len(tpl) // 2
    sig[: len(tpl) - ihalf] = tpl[ihalf:]
    for frac in (0.25, 0.5, 0.75):
        i1 = int(n * frac) - ihalf
        sig[i1:i1 + len(tpl)] = tpl
    i1 = n - ihalf
    sig[i1:i1 + ihalf] = tpl[1:ihalf + 1]
    min_dist = int(0.5 / dt + 1) + 1
    ref13 = [0, 67, 933, 1000, 1067, 1933, 2000, 2067, 2933, 3000, 3067, n - 1 - 67, n - 1]
    ref21 = [0, 33, 67]
    for c in (1000, 2000, 3000):
        ref21 += [c - 67, c - 33, c, c + 33, c + 67]
    ref21 += [n - 1 - 67, n - 1 - 33, n - 1]
    return sig, 0.8, min_dist, ref13, ref21


def single_channel_inputs(rng):
    """testing/singleChannel.cpp:104-160 with a seeded generator:
    U(2,3) signal of 4000 samples, templates cut at 25/1124/3601 of
    lengths 130, 140, 125."""
    n = 4000
    wave = rng.uniform(2, 3, n)
    tpls = [wave[25:25 + 130].copy(), wave[1124:1124 + 140].copy(), wave[3601:3601 + 125].copy()]
    return wave, tpls


def detector_inputs(seed: int = 0):
    """testing/detector.cpp:100-227 (singleChannelDetection.detector): five
    decaying-sine templates of 4 s at 100 Hz; templates 0, 1, 2 and 4 are
    written, phase-shifted and scaled by 1/4 (even i) or 1/2 (odd i), three
    times each over a 120 s trace of uniform noise in [-0.05, 0.05]
    (template 3 is left out on purpose).  The reference draws the noise
    unseeded; any seed serves.  Returns (signal, templates, dt, expected),
    expected = (onset, phase-shifted onset, scale, template) of the 12
    detections in time order."""
    df = 100.0
    dt = 1 / df
    template_dur, signal_dur = 4.0, 120.0
    n_signal = int(signal_dur / dt + 0.5) + 1
    n_template = int(template_dur / dt + 0.5) + 1
    amp = [1.0, 2.0, 0.5, 0.75, 3.0]
    freq = [10.0, 5.0, 7.5, 12.5, 2.5]
    decay = [2.8, 2.9, 2.6, 2.2, 2.3]
    phi = [0.001, 0.002, 0.004, -0.003, -0.002]
    inserts = {0: [4, 28, 90], 1: [8, 24, 100], 2: [33, 66, 86], 4: [16, 44, 106]}

    def make(npts, a, f, dec, shift):          # createTemplateSignal, 285-301
        t = np.arange(npts) * dt
        return a * np.exp(-t * dec) * np.sin(2 * np.pi * f * (t - shift))

    rng = np.random.default_rng(seed)
    signal = rng.uniform(-0.05, 0.05, n_signal)
    templates, expected = [], []
    for i in range(5):
        templates.append(make(n_template, amp[i], freq[i], decay[i], 0.0))
        scale = 1.0 / (4.0 if i % 2 == 0 else 2.0)
        w = make(n_template, scale * amp[i], freq[i], decay[i], phi[i])
        for toff in inserts.get(i, []):
            k = int(toff / dt + 0.5)
            signal[k:k + n_template] = w          # written, not added (214-217)
            expected.append((float(toff), toff + phi[i], scale, i))
    expected.sort(key=lambda e: e[0])
    return signal, templates, dt, expected
