"""The reference's API on the GPU: the C++ test program (restated gtest
cases) and pymflib's matched-filter cases (python/unit_tests.py:180-300),
checked against the FP64 Pearson definition and the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from refcases import peak_finder_signal, shift_and_stack_inputs, moveout_shift

pytestmark = pytest.mark.gpu
pymflib = pytest.importorskip("mflib_amd.pymflib", reason="pymflib not built (python -m mflib_amd.build)")

HERE = os.path.dirname(os.path.abspath(__file__))


def dumb_xc(template, signal):
    """python/unit_tests.py:184-196: the windowed Pearson definition."""
    template = template - np.average(template)
    tn = np.linalg.norm(template)
    nb, nx = len(template), len(signal)
    xc = np.zeros(nx - nb + 1)
    for i in range(len(xc)):
        y = signal[i:i + nb] - np.average(signal[i:i + nb])
        xc[i] = np.dot(template, y) / (tn * np.linalg.norm(y))
    return xc


def test_cpp_api_program():
    exe = os.path.join(HERE, "cpp", "bin", "test_facade")
    print("running", exe, flush=True)
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110)
    sys.stdout.write(res.stdout)
    assert res.returncode == 0, res.stdout


def test_multichannel_matched_filtering():
    """python/unit_tests.py:198-244 (tolerance 1e-14)."""
    signal_size, tmax = 3000, 8
    f0 = 4 / tmax
    dt = tmax / (signal_size - 1)
    times = np.linspace(0, tmax, signal_size, endpoint=False)
    s1 = np.exp(-0.1 * times) * np.sin(2 * np.pi * f0 * times)
    s2 = np.exp(-0.15 * times) * np.cos(2 * np.pi * f0 * times)
    nb = int(2 / dt + 0.5)
    tt = np.linspace(0, 2, nb, endpoint=False)
    t1, t2 = np.sin(2 * np.pi * f0 * tt), np.cos(2 * np.pi * f0 * tt)
    params = pymflib.MultiChannel.MatchedFilterParameters()
    for t in (t1, t2):
        w = pymflib.WaveformTemplate()
        w.set_signal(t)
        w.set_sampling_rate(1 / dt)
        params.add_template(w)
    params.set_signal_size(signal_size)
    mf = pymflib.MultiChannel.MatchedFilter()
    mf.initialize(params)
    mf.set_signal(0, s1)
    mf.set_signal(1, s2)
    assert not mf.have_matched_filtered_signals()
    mf.apply()
    assert mf.have_matched_filtered_signals()
    assert np.max(np.abs(mf.get_matched_filtered_signal(0) - dumb_xc(t1, s1))) < 1e-14
    assert np.max(np.abs(mf.get_matched_filtered_signal(1) - dumb_xc(t2, s2))) < 1e-14


def test_single_channel_matched_filtering():
    """python/unit_tests.py:248-290 with a seeded generator (tolerance 1e-14)."""
    rng = np.random.default_rng(0x51C)
    fs = 100.0
    n = int(fs * 60 * 3) + 1
    tn = int(fs * 10) + 1
    signal = rng.uniform(-3, 4, n)
    cuts = (int(n * 0.25), int(n * 0.05), int(n * 0.8))
    params = pymflib.SingleChannel.MatchedFilterParameters()
    params.set_signal_size(n)
    for i in cuts:
        w = pymflib.WaveformTemplate()
        w.set_sampling_rate(fs)
        w.set_signal(signal[i:i + tn])
        params.add_template(w)
    assert params.get_number_of_templates() == 3
    mf = pymflib.SingleChannel.MatchedFilter()
    mf.initialize(params)
    mf.set_signal(signal)
    mf.apply()
    for k, i in enumerate(cuts):
        ref = dumb_xc(signal[i:i + tn], signal)
        got = mf.get_matched_filtered_signal(k)
        assert np.max(np.abs(got - ref)) < 1e-14, k
        assert abs(got[i] - 1) < 1e-14


def test_shift_and_stack_extension_matches_oracle():
    templates, signals, tt, onsets, fs, origin = shift_and_stack_inputs()
    params = pymflib.MultiChannel.MatchedFilterParameters()
    for k in range(3):
        w = pymflib.WaveformTemplate()
        w.set_signal(templates[k])
        w.set_sampling_rate(fs)
        w.set_phase_onset_time_in_signal(onsets[k])
        w.set_phase_travel_time(tt[k])
        params.add_template(w)
    params.set_signal_size(len(signals[0]))
    mf = pymflib.MultiChannel.MatchedFilter()
    mf.initialize(params)
    for k in range(3):
        mf.set_signal(k, signals[k])
    mf.apply()
    rows = np.stack([mf.get_matched_filtered_signal(k) for k in range(3)])
    shifts = [moveout_shift(tt[k], onsets[k], fs) for k in range(3)]
    ref = O.shift_and_stack(rows, shifts, [1.0, 1.0, 1.0], absval=True)
    got = mf.shift_and_stack()
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)  # same loop order in FP64: bit-exact
    assert abs(got.max() - 1) < 1e-3 and abs(np.argmax(got) / fs - origin) <= 0.5 / fs


def test_peak_finder_kats():
    sig, thr, min_dist, ref13, ref21 = peak_finder_signal()
    pf = pymflib.PeakFinder()
    pf.set_threshold(thr)
    pf.set_minimum_peak_distance(min_dist)
    pf.set_signal(sig)
    pf.apply()
    assert pf.get_peak_indices().tolist() == ref13
    pf.set_minimum_peak_distance(0)
    pf.apply()
    assert pf.get_peak_indices().tolist() == ref21


def test_pymflib_detector_kat():
    """testing/detector.cpp:100-283 through pymflib: SingleChannel.Detector
    (REDUCED policy, Lanczos) on a SingleChannel.MatchedFilter; the
    detections' times, templates, interpolated times and scaling factors."""
    from refcases import detector_inputs
    signal, tpls, dt, expected = detector_inputs(1)
    sc = pymflib.SingleChannel
    parms = sc.MatchedFilterParameters()
    nsp = pymflib.NetworkStationPhase()
    nsp.set_station("CCUT")
    for i, t in enumerate(tpls):
        w = pymflib.WaveformTemplate()
        w.set_signal(t)
        w.set_sampling_rate(1 / dt)
        w.set_identifier([nsp, i])
        parms.add_template(w)
    parms.set_signal_size(len(signal))
    mf = sc.MatchedFilter()
    mf.initialize(parms)
    mf.set_signal(signal)
    mf.apply()
    ip = sc.DetectionTimeInterpolationParameters()
    ip.enable_lanczos()
    dp = sc.DetectorParameters()
    dp.set_detection_time_interpolation_parameters(ip)
    dp.set_minimum_detection_spacing(int(4.0 / 2 / dt))
    dp.set_matched_filtered_signal_detector_policy(pymflib.MatchedFilteredSignalDetectorPolicy.reduce)
    dp.enable_save_detected_waveform()
    det = sc.Detector()
    det.initialize(dp)
    det.detect(mf)
    assert det.get_number_of_detections() == len(expected) == 12
    R = pymflib.RelativeMagnitudeType
    for d, (onset, shifted, scale, it) in zip(det, expected):
        assert abs(d.get_detection_time() - onset) < 1e-10
        assert d.get_template_identifier()[1] == it
        assert abs(shifted - d.get_interpolated_detection_time()) < dt / 2
        assert abs(d.get_amplitude_scaling_factor(R.gibbons_ringdal_2006) - scale) < 1e-2
        assert abs(d.get_amplitude_scaling_factor(R.schaff_richards_2014) - scale) < 1e-2
        k = int(round(onset / dt))
        assert np.array_equal(d.get_detected_signal(), signal[k:k + len(tpls[it])])


def test_pymflib_dlpack_device_exchange():
    """set_signal from a GPU array (torch, through DLPack) gives the rows a
    host array gives, bit for bit; get_matched_filtered_signal_device is a
    zero-copy view that torch.from_dlpack reads, overwritten in place by the
    next apply()."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(17)
    n = 50_000
    tpls = [rng.standard_normal(300), rng.standard_normal(250)]
    sigs = [rng.standard_normal(n) + 3.0 for _ in tpls]

    def build():
        parms = pymflib.MultiChannel.MatchedFilterParameters()
        for t in tpls:
            w = pymflib.WaveformTemplate()
            w.set_signal(t)
            w.set_sampling_rate(100.0)
            parms.add_template(w)
        parms.set_signal_size(n)
        mf = pymflib.MultiChannel.MatchedFilter()
        mf.initialize(parms)
        return mf

    host, dev = build(), build()
    for i, x in enumerate(sigs):
        host.set_signal(i, x)
        dev.set_signal(i, torch.from_numpy(x).to("cuda"))
    host.apply()
    dev.apply()
    for i in range(len(tpls)):
        ref = host.get_matched_filtered_signal(i)
        assert np.array_equal(dev.get_matched_filtered_signal(i), ref)
        view = dev.get_matched_filtered_signal_device(i)
        t = torch.from_dlpack(view)
        assert t.is_cuda and t.dtype == torch.float64 and t.numel() == len(ref) == len(view)
        assert np.array_equal(t.cpu().numpy(), ref)
        # two views share the engine's buffer
        assert torch.from_dlpack(dev.get_matched_filtered_signal_device(i)).data_ptr() == t.data_ptr()
    # the view aliases the row: the next apply() rewrites it in place
    t0 = torch.from_dlpack(dev.get_matched_filtered_signal_device(0))
    dev.set_signal(0, torch.from_numpy(sigs[0][::-1].copy()).to("cuda"))
    dev.apply()
    assert np.array_equal(t0.cpu().numpy(), dev.get_matched_filtered_signal(0))
    with pytest.raises(ValueError):
        dev.set_signal(0, torch.zeros(n, dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        dev.set_signal(0, torch.zeros((2, n // 2), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        dev.set_signal(0, torch.zeros(n - 1, dtype=torch.float64, device="cuda"))
    # a host torch tensor takes the host path
    dev.set_signal(1, torch.from_numpy(sigs[1]))
    dev.apply()
    assert np.array_equal(dev.get_matched_filtered_signal(1), host.get_matched_filtered_signal(1))


def test_pymflib_single_channel_device_signal_feeds_detector_waveforms():
    """A single-channel signal given as a GPU array: the detector's saved
    waveforms (read through getSignalPointer's host copy, fetched from HBM
    on demand) equal the signal."""
    torch = pytest.importorskip("torch")
    from refcases import detector_inputs
    signal, tpls, dt, expected = detector_inputs(2)
    sc = pymflib.SingleChannel
    parms = sc.MatchedFilterParameters()
    for t in tpls:
        w = pymflib.WaveformTemplate()
        w.set_signal(t)
        w.set_sampling_rate(1 / dt)
        parms.add_template(w)
    parms.set_signal_size(len(signal))
    mf = sc.MatchedFilter()
    mf.initialize(parms)
    mf.set_signal(torch.from_numpy(signal).to("cuda"))
    mf.apply()
    dp = sc.DetectorParameters()
    dp.set_minimum_detection_spacing(200)
    dp.set_matched_filtered_signal_detector_policy(pymflib.MatchedFilteredSignalDetectorPolicy.reduce)
    dp.enable_save_detected_waveform()
    det = sc.Detector()
    det.initialize(dp)
    det.detect(mf)
    assert len(det) == len(expected)
    for d, (onset, _, _, it) in zip(det, expected):
        k = int(round(onset / dt))
        assert np.array_equal(d.get_detected_signal(), signal[k:k + len(tpls[it])])
    view = mf.get_matched_filtered_signal_device(0)
    assert np.array_equal(torch.from_dlpack(view).cpu().numpy(), mf.get_matched_filtered_signal(0))
