"""GPU per-detection products (mf_detection_products) against the oracle's
restatement of Detector::detect's time refinement and RelativeMagnitude
(src/singleChannel/detector.cpp:38-152, 537-664;
src/singleChannel/relativeMagnitude.cpp:35-58, 273-388), evaluated on the
GPU's own correlation rows and signal samples:
  - the clamped rho and the quadratic shift are bit-identical (same FP64
    operations, no contraction on either side);
  - the Lanczos shift is a grid point where the oracle's interpolant reaches
    its maximum to 1e-12 (the two sides' sin() may differ in the last bit,
    which can only reorder near-equal grid points), and is the oracle's own
    point wherever that maximum is unique by more than 1e-9;
  - the amplitude flags are equal and the scaling factors agree to 1e-11
    relative (FP64 sums in another order).
Plus the reference's detector known-answer test (testing/detector.cpp:100-283)
end to end on the GPU."""
import numpy as np
import pytest

import oracle as O
import refcases
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _check_against_oracle(h, signals, tpls, row_channel, peaks, tids, dts, absval, interp, alpha=51, points=100):
    got = h.detection_products(peaks, tids, dts, absval=absval, interpolation=interp, lanczos_alpha=alpha,
                               lanczos_points=points)
    rows = {}
    for k, (p, it) in enumerate(zip(peaks, tids)):
        if it not in rows:
            rows[it] = h.rho(it).astype(np.float64)
        row = rows[it]
        dt = 1.0 if dts is None else dts[it]
        ctx = f"k={k} lag={p} template={it} interp={interp} absval={absval}"
        assert got["rho"][k] == max(-1.0, min(1.0, row[p])), ctx
        if interp == capi.INTERP_QUADRATIC:
            assert got["shift"][k] == O.quadratic_shift(row, p, dt, absval), ctx
        elif interp == capi.INTERP_LANCZOS:
            ref, prof = O.lanczos_shift(row, p, dt, absval, alpha, points)
            j = int(round((got["shift"][k] / dt + 0.5) * points)) - 1     # grid index of the GPU's pick
            assert 0 <= j < points - 1, ctx
            assert got["shift"][k] == dt * (-0.5 + (1.0 / points) * (j + 1)), ctx
            top = prof.max()
            assert prof[j] >= top - 1e-12 * max(1.0, abs(top)), ctx
            if np.sort(prof)[-2] < top - 1e-9 * max(1.0, abs(top)):
                assert got["shift"][k] == ref, ctx
        else:
            assert got["shift"][k] == 0.0, ctx
        t = tpls[it]
        sig = signals[row_channel[it]].astype(np.float64)
        if p + len(t) > len(sig):
            assert got["flags"][k] == capi.PRODUCT_TRUNCATED, ctx
            continue
        ref = O.amplitude_factors(t, sig[p:p + len(t)])
        if ref is None:
            assert got["flags"][k] == 0, ctx
        else:
            assert got["flags"][k] == capi.PRODUCT_AMPLITUDE, ctx
            assert got["alpha_gr"][k] == pytest.approx(ref[0], rel=1e-11, abs=1e-11 * ref[1]), ctx
            assert got["alpha_sr"][k] == pytest.approx(ref[1], rel=1e-11), ctx
    return got


def _single_channel(dtype, seed):
    rng = np.random.default_rng(seed)
    n = 30_000
    x = rng.standard_normal(n) * 0.4 + 3.0
    lens = [150, 211, 96, 211, 180]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    for k, t in enumerate(tpls):
        for pos in rng.integers(0, n - 300, 3):
            x[pos:pos + len(t)] += (1.5 + k) * t / np.linalg.norm(t)
    x[12_000:13_500] = 2.5                           # a constant gap: windows there have no amplitude
    x = x.astype(dtype)
    h = capi.Handle(dtype, "single")
    h.set_templates(tpls)
    h.set_signal_size(n)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    return h, [x], tpls, [0] * len(tpls)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("absval", [False, True])
@pytest.mark.parametrize("interp", [capi.INTERP_QUADRATIC, capi.INTERP_LANCZOS, capi.INTERP_NONE])
def test_products_single_channel(dtype, absval, interp):
    h, sigs, tpls, chan = _single_channel(dtype, 7 + absval)
    nout = h.filtered_length(0)
    rng = np.random.default_rng(interp)
    lags = np.concatenate([[0, 1, nout - 2, nout - 1, 12_100, 12_500], rng.integers(0, nout, 150)])
    peaks, tids = [], []
    # the detector's own peaks on the reduced trace, then arbitrary lags on every row
    pk, pt = h.detect_reduced(0.3, 50, absval=absval)
    peaks += list(pk)
    tids += list(pt)
    for k, lag in enumerate(lags):
        peaks.append(int(lag))
        tids.append(k % len(tpls))
    dts = np.array([0.01, 0.02, 0.005, 0.01, 0.025])
    got = _check_against_oracle(h, sigs, tpls, chan, np.array(peaks), np.array(tids), dts, absval, interp,
                                alpha=19 if interp == capi.INTERP_LANCZOS and absval else 51)
    assert (got["flags"] == 0).sum() >= 2              # the two lags inside the constant gap
    h.close()


def test_products_multi_channel_reads_each_rows_channel():
    rng = np.random.default_rng(21)
    n = 20_000
    tpls = [rng.standard_normal(L) for L in (200, 160, 200)]
    sigs = []
    for c, t in enumerate(tpls):
        x = rng.standard_normal(n) * (0.1 + 0.1 * c) + 10.0 * c
        x[5000 + 100 * c:5000 + 100 * c + len(t)] += 4 * t
        sigs.append(x.astype(np.float32))
    h = capi.Handle(np.float32, "multi")
    h.set_templates(tpls)
    h.set_signal_size(n)
    h.initialize()
    for c, x in enumerate(sigs):
        h.set_signal(c, x)
    h.apply()
    nout = h.filtered_length(0)
    peaks = np.array([5000, 5100, 5200, 0, nout - 1, 777, 777, 777])
    tids = np.array([0, 1, 2, 0, 2, 0, 1, 2])
    got = _check_against_oracle(h, sigs, tpls, [0, 1, 2], peaks, tids, None, False, capi.INTERP_QUADRATIC)
    assert got["alpha_gr"][:3] == pytest.approx([4.0, 4.0, 4.0], rel=0.05)
    h.close()


def test_detector_kat_on_gpu():
    """testing/detector.cpp:100-283 through the engine: reduce, PeakFinder
    and the products, Lanczos with the default alpha 51 and 100 points."""
    signal, tpls, dt, expected = refcases.detector_inputs(0)
    h = capi.Handle(np.float64, "single")
    h.set_templates(tpls)
    h.set_signal_size(len(signal))
    h.initialize()
    h.set_signal(0, signal)
    h.apply()
    peaks, tids = h.detect_reduced(0.7, int(4.0 / 2 / dt), absval=False)
    assert len(peaks) == len(expected) == 12
    dts = np.full(len(tpls), dt)
    got = _check_against_oracle(h, [signal], tpls, [0] * 5, peaks, tids, dts, False, capi.INTERP_LANCZOS)
    for k, (onset, shifted, scale, it) in enumerate(expected):
        assert abs(peaks[k] * dt - onset) < 1e-10
        assert tids[k] == it
        assert abs(shifted - (peaks[k] * dt + got["shift"][k])) < dt / 2
        assert abs(got["alpha_gr"][k] - scale) < 1e-2 and abs(got["alpha_sr"][k] - scale) < 1e-2
        assert got["rho"][k] > 0.7
    h.close()


def test_products_argument_errors():
    h, _, tpls, _ = _single_channel(np.float32, 3)
    nout = h.filtered_length(0)
    with pytest.raises(ValueError):
        h.detection_products([5], [len(tpls)])                # template out of range
    with pytest.raises(ValueError):
        h.detection_products([nout], [0])                     # lag out of range
    with pytest.raises(ValueError):
        h.detection_products([-1], [0])
    with pytest.raises(ValueError):
        h.detection_products([5], [0], interpolation=7)
    with pytest.raises(ValueError):
        h.detection_products([5], [0], interpolation=capi.INTERP_LANCZOS, lanczos_alpha=0)
    with pytest.raises(ValueError):
        h.detection_products([5], [0], interpolation=capi.INTERP_LANCZOS, lanczos_points=0)
    with pytest.raises(ValueError):
        h.detection_products([5], [0], sampling_periods=np.zeros(len(tpls)))
    with pytest.raises(RuntimeError):                          # beyond the bounded Lanczos work
        h.detection_products([5], [0], interpolation=capi.INTERP_LANCZOS, lanczos_alpha=5000, lanczos_points=1000)
    assert h.detection_products([], [])["rho"].size == 0
    h.close()
    h2 = capi.Handle(np.float32, "single")
    h2.set_templates(tpls)
    h2.set_signal_size(30_000)
    h2.initialize()
    with pytest.raises(RuntimeError):                          # before apply
        h2.detection_products([5], [0])
    h2.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_amplitude_factors_batch_matches_oracle(dtype):
    """mf_amplitude_factors (RelativeMagnitude on the GPU) against the
    oracle's restatement, window by window, including a constant window."""
    rng = np.random.default_rng(5 + (dtype == np.float64))
    lt, nw = 333, 40
    t = rng.standard_normal(lt)
    wins = rng.standard_normal((nw, lt)) * rng.uniform(0.1, 5, (nw, 1)) + rng.uniform(-50, 50, (nw, 1))
    wins[3] = 2.0 * t + 1.0
    wins[7] = 4.25                                   # constant: no factors
    wins = wins.astype(dtype)
    gr, sr, ok = capi.amplitude_factors(t, wins, dtype)
    for k in range(nw):
        ref = O.amplitude_factors(t, wins[k].astype(np.float64))
        if ref is None:
            assert not ok[k], k
            continue
        assert ok[k]
        assert gr[k] == pytest.approx(ref[0], rel=1e-11, abs=1e-11 * ref[1]), k
        assert sr[k] == pytest.approx(ref[1], rel=1e-11), k
    assert not ok[7] and ok.sum() == nw - 1
    if dtype == np.float64:
        assert gr[3] == pytest.approx(2.0, rel=1e-12) and sr[3] == pytest.approx(2.0, rel=1e-12)


def test_pymflib_relative_magnitude():
    """python/singleChannel/relativeMagnitude.cpp's names over the GPU path:
    testing/relativeMagnitude.cpp's formulas (get_dmag_1/2) at 1e-12."""
    pymflib = pytest.importorskip("mflib_amd.pymflib")
    rng = np.random.default_rng(9)
    t = rng.uniform(-2, 1, 400)
    w = pymflib.WaveformTemplate()
    w.set_signal(t)
    w.set_sampling_rate(100.0)
    mag = pymflib.SingleChannel.RelativeMagnitude()
    mag.initialize(w)
    assert mag.is_initialized() and mag.get_detected_signal_length() == 400
    y = rng.uniform(-3, 8, 400)
    mag.set_detected_signal(y)
    assert mag.have_detected_signal()
    xd, yd = t - t.mean(), y - y.mean()
    R = pymflib.RelativeMagnitudeType
    d1 = np.log10(abs(xd @ yd) / (xd @ xd))
    d2 = np.log10(np.sqrt((yd @ yd) / 400) / np.sqrt((xd @ xd) / 400))
    assert abs(mag.compute_magnitude_perturbation(R.gibbons_ringdal_2006) - d1) < 1e-12
    assert abs(mag.compute_magnitude_perturbation(R.schaff_richards_2014) - d2) < 1e-12
    with pytest.raises(ValueError):
        mag.set_detected_signal(np.full(400, 1.5))    # constant
    assert not mag.have_detected_signal()


@pytest.mark.parametrize("points,alpha", [(1, 51), (2, 51), (3, 1), (7, 2)])
def test_lanczos_degenerate_grids_match_oracle(points, alpha):
    """Grids of 0, 1, 2 and 6 shifts (points - 1), and windows of one or two
    lobes: the GPU's pick equals the oracle's (a single grid point, or none
    at all, leaves nothing to tie)."""
    h, sigs, tpls, chan = _single_channel(np.float64, 13)
    nout = h.filtered_length(0)
    lags = np.array([0, 1, 2, nout // 3, nout - 2, nout - 1], dtype=np.int32)
    tids = np.array([0, 1, 2, 3, 4, 0], dtype=np.int32)
    got = h.detection_products(lags, tids, None, absval=False, interpolation=capi.INTERP_LANCZOS,
                               lanczos_alpha=alpha, lanczos_points=points)
    for k, (p, it) in enumerate(zip(lags, tids)):
        ref, prof = O.lanczos_shift(h.rho(int(it)).astype(np.float64), int(p), 1.0, False, alpha, points)
        if points <= 2:
            assert got["shift"][k] == ref == 0.0
        elif np.sort(prof)[-2] < prof.max() - 1e-9 * max(1.0, abs(prof.max())):
            assert got["shift"][k] == ref, (k, p)
    h.close()
