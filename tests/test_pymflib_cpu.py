"""pymflib host-side classes on the CPU: the reference's own Python cases
for NetworkStationPhase, WaveformTemplate and MatchedFilterParameters
(python/unit_tests.py:8-178), pickling and error conventions.  No GPU
call is made."""
import pickle

import numpy as np
import pytest

pymflib = pytest.importorskip("mflib_amd.pymflib", reason="pymflib not built (python -m mflib_amd.build)")


def test_network_station_phase():
    nsp = pymflib.NetworkStationPhase()
    nsp.set_network("UU")
    nsp.set_station("ALT")
    nsp.set_channel("EHN")
    nsp.set_location_code("01")
    nsp.set_phase("P")
    assert (nsp.get_network(), nsp.get_station(), nsp.get_channel(), nsp.get_location_code(),
            nsp.get_phase()) == ("UU", "ALT", "EHN", "01", "P")
    p = pickle.loads(pickle.dumps(nsp, 2))
    assert p == nsp and hash(p) == hash(nsp)
    p.set_phase("S")
    assert p != nsp


def test_waveform_template():
    template_signal = np.arange(1, 400, 1)
    nsp = pymflib.NetworkStationPhase()
    nsp.set_network("UU")
    nsp.set_station("FORU")
    nsp.set_channel("HHZ")
    nsp.set_location_code("01")
    nsp.set_phase("P")
    wt = pymflib.WaveformTemplate()
    wt.set_signal(template_signal)
    assert np.max(np.abs(wt.get_signal() - template_signal)) == 0
    wt.set_sampling_rate(100)
    assert wt.get_sampling_rate() == 100
    wt.set_shift_and_stack_weight(0.8)
    assert wt.get_shift_and_stack_weight() == 0.8
    wt.set_phase_travel_time(9)
    assert wt.get_phase_travel_time() == 9
    wt.set_phase_onset_time_in_signal(0.5)
    assert wt.get_phase_onset_time_in_signal() == 0.5
    wt.set_magnitude(6)
    assert wt.get_magnitude() == 6
    wt.set_polarity(pymflib.Polarity.compression)
    assert wt.get_polarity() == pymflib.Polarity.compression
    wt.set_identifier([nsp, 4])
    nid, num = wt.get_identifier()
    assert nid.get_station() == "FORU" and num == 4
    for w in (wt, pickle.loads(pickle.dumps(wt, 2))):
        assert np.max(np.abs(w.get_signal() - template_signal)) == 0
        assert w.get_sampling_rate() == 100 and w.get_shift_and_stack_weight() == 0.8
        assert w.get_phase_travel_time() == 9 and w.get_phase_onset_time_in_signal() == 0.5
        assert w.get_magnitude() == 6 and w.get_polarity() == pymflib.Polarity.compression
        assert w.get_identifier()[0] == nsp and w.get_identifier()[1] == 4
    wt.clear()
    assert wt.get_shift_and_stack_weight() == 1


def test_waveform_template_errors():
    wt = pymflib.WaveformTemplate()
    with pytest.raises(RuntimeError):
        wt.get_signal()
    with pytest.raises(ValueError):
        wt.set_sampling_rate(0)
    with pytest.raises(ValueError):
        wt.set_shift_and_stack_weight(1.2)
    with pytest.raises(ValueError):
        wt.set_phase_travel_time(-1)
    wt.set_signal(np.ones(10))
    with pytest.raises(RuntimeError):
        wt.set_phase_onset_time_in_signal(0.1)  # no sampling rate yet
    wt.set_sampling_rate(100)
    # src/waveformTemplate.cpp:287-303 bounds the onset by (npts-1)*fs
    wt.set_phase_onset_time_in_signal(9 * 100)
    with pytest.raises(ValueError):
        wt.set_phase_onset_time_in_signal(9 * 100 + 1)
    # setting the signal or rate forgets the onset time (157, 206)
    wt.set_sampling_rate(50)
    with pytest.raises(RuntimeError):
        wt.get_phase_onset_time_in_signal()


def test_matched_filter_parameters():
    wf1, wf2 = pymflib.WaveformTemplate(), pymflib.WaveformTemplate()
    params = pymflib.MultiChannel.MatchedFilterParameters()
    t1, t2 = np.arange(1, 100, 1), np.arange(1, 95, 1)
    for wf, t in ((wf1, t1), (wf2, t2)):
        wf.set_signal(t)
        wf.set_sampling_rate(100)
        wf.set_shift_and_stack_weight(0.8)
        wf.set_phase_onset_time_in_signal(0.05)
        wf.set_phase_travel_time(4)
        params.add_template(wf)
    params.set_signal_size(1024)
    assert params.get_number_of_templates() == 2
    assert params.get_fft_length() == 512
    assert params.get_block_length() == 512 - 99 + 1
    params.set_fft_length(550)
    assert params.get_fft_length() == 550
    for i, t in enumerate((t1, t2)):
        back = params.get_template(i)
        assert back.get_sampling_rate() == 100 and back.get_shift_and_stack_weight() == 0.8
        assert back.get_phase_onset_time_in_signal() == 0.05 and back.get_phase_travel_time() == 4
        assert np.max(np.abs(back.get_signal() - t)) == 0
    with pytest.raises(ValueError):
        params.get_template(2)
    with pytest.raises(ValueError):
        params.set_fft_length(50)  # shorter than the longest template
    params.clear_templates()
    assert params.get_number_of_templates() == 0
    with pytest.raises(RuntimeError):
        params.get_fft_length()


def test_parameters_sampling_rate_mismatch():
    params = pymflib.SingleChannel.MatchedFilterParameters()
    a, b = pymflib.WaveformTemplate(), pymflib.WaveformTemplate()
    a.set_signal(np.ones(20))
    a.set_sampling_rate(100)
    b.set_signal(np.ones(20))
    b.set_sampling_rate(100.01)
    params.add_template(a)
    with pytest.raises(ValueError):
        params.add_template(b)
    no_rate = pymflib.WaveformTemplate()
    no_rate.set_signal(np.ones(5))
    with pytest.raises(RuntimeError):
        params.add_template(no_rate)


def test_matched_filter_requires_initialize():
    mf = pymflib.MultiChannel.MatchedFilter()
    with pytest.raises(RuntimeError):
        mf.set_signal(0, np.zeros(10))
    with pytest.raises(RuntimeError):
        mf.apply()
    with pytest.raises(ValueError):
        mf.initialize(pymflib.MultiChannel.MatchedFilterParameters())


def test_peak_finder_setters():
    pf = pymflib.PeakFinder()
    assert abs(pf.get_threshold() - 0.7) < 1e-15  # src/peakFinder.cpp:24
    with pytest.raises(ValueError):
        pf.set_threshold(1.01)
    pf.set_minimum_peak_distance(7)
    assert pf.get_minimum_peak_distance() == 7
    with pytest.raises(ValueError):
        pf.set_signal(np.zeros(0))


def test_cpp_test_program_built():
    import os
    exe = os.path.join(os.path.dirname(__file__), "cpp", "bin", "test_facade")
    assert os.access(exe, os.X_OK), "tests/cpp/bin/test_facade missing (python -m mflib_amd.build)"


def test_detector_parameters_defaults_and_checks():
    """src/singleChannel/detectorParameters.cpp:9-12,81-90 and
    detectionTimeInterpolationParameters.cpp:9-13,76-101."""
    sc = pymflib.SingleChannel
    p = sc.DetectorParameters()
    assert p.get_detection_threshold() == 0.7
    assert p.get_maxima_policy() == pymflib.MaxmimumMatchedFilterPolicy.maximum
    assert p.get_matched_filtered_signal_detector_policy() == pymflib.MatchedFilteredSignalDetectorPolicy.single
    assert p.get_minimum_detection_spacing() == 0
    assert p.want_amplitude_scaling_factor() and not p.want_detected_waveform()
    for bad in (0.0, -0.1, 1.01):
        with pytest.raises(ValueError):
            p.set_detection_threshold(bad)
    p.set_detection_threshold(0.55)
    p.set_minimum_detection_spacing(120)
    p.enable_save_detected_waveform()
    p.disable_save_amplitude_scaling_factor()
    ip = sc.DetectionTimeInterpolationParameters()
    assert ip.get_type() == pymflib.DetectionTimeInterpolationType.quadratic
    assert ip.get_lanczos_alpha() == 51 and ip.get_lanczos_number_of_interpolation_points() == 100
    with pytest.raises(ValueError):
        ip.set_lanczos_alpha(0)
    with pytest.raises(ValueError):
        ip.set_lanczos_number_of_interpolation_points(0)
    ip.enable_lanczos()
    ip.set_lanczos_alpha(19)
    p.set_detection_time_interpolation_parameters(ip)
    got = p.get_detection_time_interpolation_parameters()
    assert got.get_type() == pymflib.DetectionTimeInterpolationType.lanczos and got.get_lanczos_alpha() == 19
    p.clear()   # keeps the interpolation settings, as the reference's clear() does
    assert p.get_detection_threshold() == 0.7 and p.get_minimum_detection_spacing() == 0
    assert p.get_detection_time_interpolation_parameters().get_lanczos_alpha() == 19


def test_detection_value_object():
    """src/singleChannel/detection.cpp:124-474: setter checks and unset getters."""
    d = pymflib.SingleChannel.Detection()
    R = pymflib.RelativeMagnitudeType
    assert not d.have_detection_time() and not d.have_amplitude_scaling_factor(R.gibbons_ringdal_2006)
    with pytest.raises(ValueError):
        d.get_detection_time()
    with pytest.raises(RuntimeError):
        d.get_amplitude_scaling_factor(R.schaff_richards_2014)
    with pytest.raises(ValueError):
        d.set_correlation_coefficient(1.5)
    with pytest.raises(ValueError):
        d.set_detection_time(-1.0)
    with pytest.raises(ValueError):
        d.set_amplitude_scaling_factor(0.0, R.gibbons_ringdal_2006)
    d.set_correlation_coefficient(-0.8)
    d.set_detection_time(12.5)
    d.set_interpolated_detection_time(12.503)
    d.set_amplitude_scaling_factor(0.25, R.schaff_richards_2014)
    assert d.get_amplitude_scaling_factor(R.schaff_richards_2014) == 0.25
    assert abs(d.get_magnitude_perturbation(R.schaff_richards_2014) - np.log10(0.25)) < 1e-15
    assert not d.have_amplitude_scaling_factor(R.gibbons_ringdal_2006)
    nsp = pymflib.NetworkStationPhase()
    nsp.set_network("UU")
    d.set_template_identifier([nsp, 7])
    assert d.get_template_identifier()[1] == 7 and d.get_template_identifier()[0] == nsp
    d.set_detected_signal(np.arange(5.0))
    assert np.array_equal(d.get_detected_signal(), np.arange(5.0))
    d.set_polarity(pymflib.Polarity.dilatation)
    assert d.get_polarity() == pymflib.Polarity.dilatation


def test_detector_requires_initialize():
    det = pymflib.SingleChannel.Detector()
    mf = pymflib.SingleChannel.MatchedFilter()
    with pytest.raises(ValueError):
        det.detect(mf)
    assert det.get_number_of_detections() == 0 and len(det) == 0
    with pytest.raises(IndexError):
        det[0]


def test_set_signal_accepts_host_arrays_of_any_kind():
    """Host arrays (numpy of any dtype, lists) take the host path; only a
    GPU array goes through DLPack.  Without an initialized filter both fail
    the reference's way."""
    mf = pymflib.MultiChannel.MatchedFilter()
    with pytest.raises(RuntimeError):
        mf.set_signal(0, np.zeros(10))
    with pytest.raises(RuntimeError):
        mf.set_signal(0, [0.0, 1.0])
    assert hasattr(pymflib, "DeviceRow")


def test_relative_magnitude_state_checks():
    """src/singleChannel/relativeMagnitude.cpp:273-359: use before
    initialize and an empty template fail the reference's way (no GPU call)."""
    mag = pymflib.SingleChannel.RelativeMagnitude()
    assert not mag.is_initialized() and not mag.have_detected_signal()
    with pytest.raises(RuntimeError):
        mag.get_detected_signal_length()
    with pytest.raises(RuntimeError):
        mag.compute_amplitude_scaling_factor(pymflib.RelativeMagnitudeType.gibbons_ringdal_2006)
    with pytest.raises(RuntimeError):
        mag.initialize(pymflib.WaveformTemplate())   # no signal on the template


def test_detection_pickles():
    sc = pymflib.SingleChannel
    R = pymflib.RelativeMagnitudeType
    d = sc.Detection()
    nsp = pymflib.NetworkStationPhase()
    nsp.set_station("CCUT")
    d.set_template_identifier([nsp, 3])
    d.set_correlation_coefficient(0.91)
    d.set_detection_time(4.0)
    d.set_interpolated_detection_time(4.001)
    d.set_travel_time(2.5)
    d.set_amplitude_scaling_factor(0.25, R.schaff_richards_2014)
    d.set_polarity(pymflib.Polarity.compression)
    d.set_detected_signal(np.linspace(0, 1, 7))
    e = pickle.loads(pickle.dumps(d))
    assert e.get_correlation_coefficient() == 0.91 and e.get_template_identifier()[1] == 3
    assert e.get_template_identifier()[0] == nsp
    assert e.get_detection_time() == 4.0 and e.get_interpolated_detection_time() == 4.001
    assert e.get_travel_time() == 2.5 and not e.have_phase_onset_time()
    assert e.get_amplitude_scaling_factor(R.schaff_richards_2014) == 0.25
    assert not e.have_amplitude_scaling_factor(R.gibbons_ringdal_2006)
    assert e.get_polarity() == pymflib.Polarity.compression
    assert np.array_equal(e.get_detected_signal(), np.linspace(0, 1, 7))
    empty = pickle.loads(pickle.dumps(sc.Detection()))
    assert not empty.have_detection_time() and not empty.have_detected_signal()
