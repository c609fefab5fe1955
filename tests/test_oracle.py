"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each oracle function is checked against the reference's own known answers
where they exist and otherwise against what the mathematics fixes
(numpy brute force, closed forms, invariants), chosen so that a dropped
term, a wrong index or a transposed operand fails.
"""
import numpy as np
import pytest

import oracle as O
import refcases as R


# ---------------------------------------------------------------- sizing
def test_window_length_reference_kats():
    """testing/matchedFilterParameters.cpp:19-31,82-89 and python/unit_tests.py:159."""
    assert O.window_length(4500, 75000) == (32768, 28269)
    assert O.window_length(4500, 86400 * 100) == (65536, 61037)
    assert O.window_length(101, 1024) == (512, 412)


def test_window_length_python_kat_uses_signal_size():
    """python/unit_tests.py:121-159: nb=99, signal 1024 -> get_fft_length 512
    (getFFTLength passes nx = signal size, src/matchedFilterParameters.cpp:268-284)."""
    assert O.window_length(99, 1024)[0] == 512


def _brute_window(nb, nx):
    cost = [18, 59, 138, 303, 660, 1441, 3150, 6875, 14952, 32373, 69762, 149647, 319644,
            680105, 1441974, 3047619, 6422736, 13500637, 28311786, 59244791, 59244791 * 2.09]
    best = None
    for i in range(21):
        n = 2 << i
        if n <= nb - 1:
            continue
        L = n - (nb - 1)
        obj = np.ceil(nx / L) * cost[i]
        if best is None or obj < best[0]:
            best = (obj, n, L)
    return best[1], best[2]


@pytest.mark.parametrize("nb,nx", [(2, 3), (400, 360399), (512, 8640511), (128, 2000127),
                                   (3000, 8642999), (200, 6199), (2000, 361999), (1, 10)])
def test_window_length_brute_force(nb, nx):
    assert O.window_length(nb, nx) == _brute_window(nb, nx)


def test_window_length_rejects_bad_input():
    with pytest.raises(ValueError):
        O.window_length(0, 100)
    with pytest.raises(ValueError):
        O.window_length(100, 100)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [2, 4, 8, 64, 1024, 8192])
def test_own_real_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N)
    X = O.rfft(x)
    assert np.allclose(X, np.fft.rfft(x), atol=1e-11 * np.sqrt(N))
    assert np.allclose(O.irfft(X, N) / N, x, atol=1e-13)


# ---------------------------------------------------------------- Pearson
def _corrcoef_brute(b, x):
    nb = len(b)
    return np.array([np.corrcoef(b, x[i:i + nb])[0, 1] for i in range(len(x) - nb + 1)])


def test_naive_pearson_matches_corrcoef_brute_force():
    rng = np.random.default_rng(3)
    b = rng.standard_normal(17)
    x = rng.standard_normal(300) + 5
    ref = _corrcoef_brute(b, x)
    assert np.abs(O.naive_pearson(b, x) - ref).max() < 1e-13
    assert np.abs(O.pearson(b, x) - ref).max() < 1e-13
    assert np.abs(O.pearson(b, x.astype(np.float32)) - _corrcoef_brute(b, x.astype(np.float32).astype(np.float64))).max() < 1e-12


def test_pearson_closed_forms():
    rng = np.random.default_rng(4)
    x = rng.standard_normal(1000)
    b = x[123:123 + 64].copy()
    r = O.pearson(b, x)
    assert abs(r[123] - 1.0) < 1e-14                 # a copied window correlates perfectly
    assert np.all(np.abs(r) <= 1 + 1e-14)
    r2 = O.pearson(b, 3.5 * x - 7.0)                  # affine invariance of the signal
    assert np.abs(r - r2).max() < 1e-13
    r3 = O.pearson(-b + 2.0, x)                       # template sign flip / offset
    assert np.abs(r + r3).max() < 1e-13
    lags = np.array([0, 5, 123, 936])
    assert np.array_equal(O.pearson(b, x, lags), r[lags])


def test_pearson_constant_window_is_zero():
    x = np.concatenate([np.random.default_rng(0).standard_normal(200), np.full(300, 2.5)])
    r = O.pearson(np.arange(50.0), x)
    assert np.all(r[200:450] == 0.0)


def test_window_scale_closed_forms():
    """1/||w - mean|| pinned by closed forms: a ramp window has
    sum (w - mean)^2 = nb (nb^2 - 1) / 12 times the slope squared, a
    constant window has scale 0, and the scale is invariant to offsets and
    scales as 1/|c| under x -> c x."""
    nb = 37
    x = (np.arange(500, dtype=np.float32) * 0.25).astype(np.float32) + 3.0
    s = O.window_scale(x, nb)
    want = 1.0 / np.sqrt(nb * (nb * nb - 1) / 12.0 * 0.25 ** 2)
    assert np.abs(s / want - 1).max() < 1e-13
    c = np.concatenate([np.zeros(100), np.full(200, 7.0), np.zeros(50)]).astype(np.float32)
    sc = O.window_scale(c, nb)
    assert np.all(sc[100:300 - nb + 1] == 0.0)
    rng = np.random.default_rng(9)
    y = rng.standard_normal(400).astype(np.float32)
    s1 = O.window_scale(y, nb)
    s2 = O.window_scale((y * np.float32(4.0)).astype(np.float32), nb)   # exact power-of-two scaling
    assert np.abs(s1 / (4.0 * s2) - 1).max() < 1e-14
    lags = np.array([0, 17, 363])
    assert np.array_equal(O.window_scale(y, nb, lags), s1[lags])


def test_window_scale_agrees_with_pearson_oracle():
    """rho = <t~, w - mean(w)> * scale with t~ the demeaned unit-norm
    template must equal the (separately pinned) Pearson oracle."""
    rng = np.random.default_rng(10)
    nb = 50
    x = (rng.standard_normal(3000) * 2.0 + 40.0).astype(np.float32)
    b = rng.standard_normal(nb)
    t = b - b.mean()
    t /= np.linalg.norm(t)
    w = np.lib.stride_tricks.sliding_window_view(x.astype(np.float64), nb)
    num = (w - w.mean(axis=1, keepdims=True)) @ t
    rho = num * O.window_scale(x, nb)
    assert np.abs(rho - O.pearson(b, x)).max() < 1e-12


def test_demean_normalize_reverse_closed_form():
    b = np.array([1.0, 2.0, 4.0, 8.0])
    d = b - b.mean()
    ref = (d / np.linalg.norm(d))[::-1]
    assert np.allclose(O.demean_normalize_reverse(b), ref, atol=1e-16)
    assert abs(O.demean_normalize_reverse(b).sum()) < 1e-15


# ---------------------------------------------------------------- apply restatement
def test_restated_apply_double_basic_test():
    """testing/matchedFilter.cpp:58-168 tolerance 1e-13."""
    x, x2, b, b2 = R.basic_test_inputs()
    rho, nfft, L = O.mf_apply([b, b2, b, b2], [x, x2, x, x2])
    assert (nfft, L) == O.window_length(len(b), 1000 + len(b) - 1)
    for i, (t, s) in enumerate([(b, x), (b2, x2), (b, x), (b2, x2)]):
        assert np.abs(rho[i] - O.naive_pearson(t, s)).max() < 1e-13


def test_restated_apply_float_basic_test():
    """testing/matchedFilter.cpp:172-271 tolerance 2e-6."""
    x, x2, b, b2 = R.basic_test_inputs()
    bf, b2f = b.astype(np.float32), b2.astype(np.float32)
    xf, x2f = x.astype(np.float32), x2.astype(np.float32)
    rho, _, _ = O.mf_apply([bf, b2f], [xf, x2f], dtype=np.float32)
    assert np.abs(rho[0] - O.pearson(bf, xf)).max() < 2e-6
    assert np.abs(rho[1] - O.pearson(b2f, x2f)).max() < 2e-6


def test_restated_apply_single_channel_uneven():
    """testing/singleChannel.cpp:104-160 tolerance 1e-12, output length n - nb_max + 1."""
    wave, tpls = R.single_channel_inputs(np.random.default_rng(9))
    rho, _, _ = O.mf_apply(tpls, [wave], mode="single")
    assert rho.shape == (3, 4000 - 140 + 1)
    for k in range(3):
        assert np.abs(rho[k] - O.naive_pearson(tpls[k], wave)[:rho.shape[1]]).max() < 1e-12


def test_restated_apply_shift_and_stack_case():
    """testing/matchedFilter.cpp:275-465: rho <= 1e-11, stack max 1 +/- 1e-3 at 8 s."""
    tp, sg, tt, toff, df, origin = R.shift_and_stack_inputs()
    rho, _, _ = O.mf_apply(tp, sg)
    for k in range(3):
        assert np.abs(rho[k] - O.naive_pearson(tp[k], sg[k])[:rho.shape[1]]).max() <= 1e-11
    shifts = [R.moveout_shift(tt[k], toff[k], df) for k in range(3)]
    assert shifts == [1900, 3800, 7850]
    st = O.shift_and_stack(rho, shifts, [1, 1, 1])
    i = int(np.argmax(st))
    assert abs(st[i] - 1) < 1e-3 and abs(i / df - origin) < 1 / df / 2


def test_restated_apply_float_quirk_differs_for_uneven_lengths():
    """The float initialize demeans over the padded max length
    (src/matchedFilter.cpp:725-726): for uneven lengths the quirk result is
    not the Pearson coefficient; without the quirk it is."""
    wave, tpls = R.single_channel_inputs(np.random.default_rng(2))
    wf = wave.astype(np.float32)
    ok, _, _ = O.mf_apply(tpls, [wf], mode="single", dtype=np.float32)
    qk, _, _ = O.mf_apply(tpls, [wf], mode="single", dtype=np.float32, float_quirk=True)
    ref = O.pearson(tpls[2], wf)[:ok.shape[1]]
    assert np.abs(ok[2] - ref).max() < 5e-4
    assert np.abs(qk[2] - ref).max() > 1e-2


def test_restated_apply_dead_trace_rows_skip_normalisation():
    """Dead traces keep their (near-zero) numerator: src/matchedFilter.cpp:1126."""
    rng = np.random.default_rng(1)
    rho, _, _ = O.mf_apply([rng.standard_normal(50)], [np.full(600, 7.0)])
    assert np.abs(rho).max() < 1e-12


# ---------------------------------------------------------------- stack
def test_shift_and_stack_closed_form():
    rng = np.random.default_rng(5)
    rho = rng.uniform(-1, 1, (4, 100))
    sh = [0, 3, 7, 1]
    w = [0.5, 1.0, 0.25, 0.75]
    st = O.shift_and_stack(rho, sh, w, absval=True)
    ref = sum(wi / sum(w) * np.abs(rho[i, s:s + 93]) for i, (s, wi) in enumerate(zip(sh, w)))
    assert len(st) == 93
    assert np.abs(st - ref).max() < 1e-15
    st2 = O.shift_and_stack(rho, sh, w, absval=False)
    ref2 = sum(wi / sum(w) * rho[i, s:s + 93] for i, (s, wi) in enumerate(zip(sh, w)))
    assert np.abs(st2 - ref2).max() < 1e-15


# ---------------------------------------------------------------- peaks
@pytest.mark.parametrize("absval", [False, True])
def test_reduce_correlations_brute_force(absval):
    """Detector reduce: det = clamp(max over templates), id = the first
    template reaching the max (numpy argmax also returns the first)."""
    rng = np.random.default_rng(11 + absval)
    rho = rng.uniform(-1.2, 1.2, (7, 300))
    rho[3, 10] = rho[1, 10] = 5.0            # tie: the first template wins
    rho[:, 20] = -0.25                        # all equal: template 0
    det, idx = O.reduce_correlations(rho, absval)
    v = np.abs(rho) if absval else rho
    assert np.array_equal(idx, np.argmax(v, axis=0))
    assert np.array_equal(det, np.clip(v.max(axis=0), -1.0, 1.0))
    assert idx[10] == 1 and idx[20] == 0
    d32, i32 = O.reduce_correlations(rho.astype(np.float32), absval, length=250)
    assert np.array_equal(i32, np.argmax(v.astype(np.float32)[:, :250], axis=0))


def test_reduce_correlations_single_row_is_clamped_copy():
    x = np.linspace(-2, 2, 41)
    det, idx = O.reduce_correlations(x[None, :])
    assert np.array_equal(det, np.clip(x, -1, 1)) and not idx.any()


def test_peak_finder_reference_kats():
    """testing/peakFinder.cpp:16-143."""
    sig, thr, md, ref13, ref21 = R.peak_finder_signal()
    assert list(O.peak_finder(sig, thr, md)) == ref13
    assert list(O.peak_finder(sig, thr, 0)) == ref21
    assert len(O.peak_finder(np.full(4000, 0.4), thr, md)) == 0


def _brute_peaks(x, thr, md):
    n = len(x)
    cand = []
    for i in range(n):
        if not x[i] > thr:
            continue
        if n == 1:
            ok = True
        elif i == 0:
            ok = x[1] < x[0]
        elif i == n - 1:
            ok = x[n - 1] > x[n - 2]
        else:
            ok = x[i] >= x[i - 1] and x[i + 1] < x[i]
        if ok:
            cand.append(i)
    if md <= 0:
        return cand
    order = sorted(cand, key=lambda i: (-x[i], i))
    keep = []
    for i in order:
        if not keep or min(abs(i - k) for k in keep) > md:
            keep.append(i)
    return sorted(keep)


@pytest.mark.parametrize("seed", range(6))
def test_peak_finder_brute_force(seed):
    rng = np.random.default_rng(seed)
    x = np.round(rng.uniform(0, 1, 400), 1)  # many plateaus and ties
    for md in (0, 1, 2, 5, 40):
        assert list(O.peak_finder(x, 0.5, md)) == _brute_peaks(x, 0.5, md)


def test_peak_finder_edges():
    assert list(O.peak_finder(np.array([0.9]), 0.5, 0)) == [0]
    assert list(O.peak_finder(np.array([0.9, 0.1]), 0.5, 0)) == [0]
    assert list(O.peak_finder(np.array([0.1, 0.9]), 0.5, 0)) == [1]
    assert list(O.peak_finder(np.array([0.9, 0.9, 0.1]), 0.5, 0)) == [1]   # plateau keeps last
    assert list(O.peak_finder(np.array([0.5, 0.5]), 0.5, 0)) == []         # strict threshold
