"""The oracle's per-detection products (Detector::detect's time refinement
and RelativeMagnitude, src/singleChannel/detector.cpp:38-152, 537-664;
src/singleChannel/relativeMagnitude.cpp:35-58, 273-388), pinned by closed
forms and by the reference's own detector known-answer test
(testing/detector.cpp:100-283) run through the oracle pipeline."""
import math

import numpy as np
import pytest

import oracle as O
import refcases


# ---- quadraticRefinement -------------------------------------------------
@pytest.mark.parametrize("p", [7.3, 6.81, 7.0, 7.49])
def test_quadratic_vertex_of_exact_parabola(p):
    i = np.arange(20.0)
    xc = 0.9 - 0.01 * (i - p) ** 2
    k = int(round(p))
    assert O.quadratic_shift(xc, k) == pytest.approx(p - k, abs=1e-12)
    # dt scales the offset into seconds
    assert O.quadratic_shift(xc, k, dt=0.01) == pytest.approx(0.01 * (p - k), abs=1e-14)


def test_quadratic_minimum_is_negated_and_non_extremum_is_zero():
    i = np.arange(20.0)
    xc = -0.5 + 0.02 * (i - 9.2) ** 2          # a trough: negated, same vertex
    assert O.quadratic_shift(xc, 9) == pytest.approx(0.2, abs=1e-12)
    ramp = np.linspace(-1, 1, 20)              # lag 9 neither max nor min of its three
    assert O.quadratic_shift(ramp, 9) == 0.0
    par = 0.9 - 0.01 * (i - 0.3) ** 2
    assert O.quadratic_shift(par, 0) == 0.0    # end points
    assert O.quadratic_shift(par, 19) == 0.0


def test_quadratic_absolute_value_policy():
    i = np.arange(20.0)
    xc = -(0.9 - 0.01 * (i - 11.4) ** 2)       # a negative peak
    assert O.quadratic_shift(xc, 11, absval=True) == pytest.approx(0.4, abs=1e-12)
    # without |.| the same lags are a trough, negated: same vertex
    assert O.quadratic_shift(xc, 11, absval=False) == pytest.approx(0.4, abs=1e-12)


def test_quadratic_flat_top_is_guarded():
    xc = np.array([0.0, 0.5, 0.5, 0.5, 0.0])
    # y1 = y2 = y3: both neighbours pulled 100 eps below y2 -> vertex at 0
    assert O.quadratic_shift(xc, 2) == 0.0


# ---- lanczosRefinement ---------------------------------------------------
def _bandlimited(p, n=400, f=0.02, width=60.0):
    i = np.arange(float(n))
    return np.cos(2 * np.pi * f * (i - p)) * np.exp(-((i - p) / width) ** 2)


@pytest.mark.parametrize("p", [200.3, 199.75, 200.46, 200.12])
def test_lanczos_finds_fractional_peak(p):
    xc = _bandlimited(p)
    k = int(round(p))
    s, prof = O.lanczos_shift(xc, k, dt=1.0, alpha=51, points=100)
    assert len(prof) == 99
    assert abs(k + s - p) <= 0.01 + 1e-12        # one grid step (1/points)
    s2, _ = O.lanczos_shift(xc, k, dt=0.01, alpha=51, points=100)
    assert s2 == pytest.approx(0.01 * s, rel=1e-12)


def test_lanczos_interpolant_reproduces_samples_at_zero_shift():
    # grid point 50 of 100 is shift 0: x is the lag itself, xi are integers,
    # and the windowed sinc is 1 at 0 and sin(pi k)/(pi k) ~ 1e-17 elsewhere
    rng = np.random.default_rng(3)
    xc = rng.standard_normal(300)
    for k in (0, 1, 57, 150, 298, 299):
        _, prof = O.lanczos_shift(xc, k, alpha=19, points=100)
        assert prof[49] == pytest.approx(xc[k], abs=1e-13)
    # off-row taps count as zero: a row of one nonzero sample
    e = np.zeros(300)
    e[150] = 1.0
    _, prof = O.lanczos_shift(e, 150, alpha=19, points=100)
    xs = -0.5 + np.arange(1, 100) / 100.0
    sinc = lambda x: np.where(np.abs(x) > 1e-13, np.sin(np.pi * x) / (np.pi * np.where(x == 0, 1, x)), 1.0)
    assert np.allclose(prof, sinc(xs) * sinc(xs / 19), atol=1e-14)


def test_lanczos_absolute_value_and_degenerate_grids():
    p = 200.3
    xc = -_bandlimited(p)
    s_abs, _ = O.lanczos_shift(xc, 200, absval=True, alpha=51, points=100)
    assert abs(200 + s_abs - p) <= 0.01 + 1e-12
    assert O.lanczos_shift(xc, 200, alpha=51, points=1)[0] == 0.0   # no grid point
    assert O.lanczos_shift(xc, 200, alpha=51, points=2)[0] == 0.0   # the only point is shift 0


# ---- RelativeMagnitude ----------------------------------------------------
def test_amplitude_factors_closed_forms():
    rng = np.random.default_rng(5)
    t = rng.standard_normal(400)
    gr, sr = O.amplitude_factors(t, 2.5 * t + 7.0)     # scaled copy plus DC
    assert gr == pytest.approx(2.5, rel=1e-13) and sr == pytest.approx(2.5, rel=1e-13)
    gr, sr = O.amplitude_factors(t, -0.5 * t)          # |.|: no sign flip
    assert gr == pytest.approx(0.5, rel=1e-13) and sr == pytest.approx(0.5, rel=1e-13)
    # t plus a residual orthogonal to the demeaned template
    td = t - t.mean()
    e = rng.standard_normal(400)
    e -= e.mean()
    e -= (e @ td) / (td @ td) * td
    gr, sr = O.amplitude_factors(t, t + e)
    assert gr == pytest.approx(1.0, rel=1e-12)
    assert sr == pytest.approx(math.sqrt(1 + (e @ e) / (td @ td)), rel=1e-12)
    assert O.amplitude_factors(t, np.full(400, 3.25)) is None       # constant window


def test_amplitude_factors_match_reference_test_formulas():
    """testing/relativeMagnitude.cpp:20-124: log10 of the factors against
    get_dmag_1 / get_dmag_2 at 1e-12."""
    rng = np.random.default_rng(11)
    x = rng.uniform(-2, 1, 400)
    for lo, hi in ((-3, 8), (-7, 1)):
        y = rng.uniform(lo, hi, 400)
        xd, yd = x - x.mean(), y - y.mean()
        d1 = math.log10(abs(xd @ yd) / (xd @ xd))
        d2 = math.log10(math.sqrt((yd @ yd) / 400) / math.sqrt((xd @ xd) / 400))
        gr, sr = O.amplitude_factors(x, y)
        assert abs(math.log10(gr) - d1) < 1e-12 and abs(math.log10(sr) - d2) < 1e-12


# ---- the reference's detector known-answer test ----------------------------
def test_detector_kat_through_oracle():
    """testing/detector.cpp:100-283: single-channel MatchedFilter<double>,
    REDUCED policy, threshold 0.7 (the default), Lanczos with the default
    alpha 51 and 100 points, minimum spacing 200.  Every detection time is
    exact, the template matches, the interpolated time is within dt/2 of
    the phase-shifted onset and both scaling factors within 1e-2 of the
    written scale."""
    signal, tpls, dt, expected = refcases.detector_inputs(0)
    rho, _, _ = O.mf_apply(tpls, [signal], mode="single")
    det, ids = O.reduce_correlations(rho, absval=False)
    peaks = O.peak_finder(det, 0.7, int(4.0 / 2 / dt))
    assert len(peaks) == len(expected) == 12
    for p, (onset, shifted, scale, it) in zip(peaks, expected):
        assert abs(p * dt - onset) < 1e-10
        assert ids[p] == it
        shift, _ = O.lanczos_shift(rho[ids[p]], p, dt=dt, alpha=51, points=100)
        assert abs(shifted - (p * dt + shift)) < dt / 2
        gr, sr = O.amplitude_factors(tpls[it], signal[p:p + len(tpls[it])])
        assert abs(gr - scale) < 1e-2 and abs(sr - scale) < 1e-2
