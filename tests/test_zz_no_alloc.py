"""mf_peaks_device_noalloc: PeakFinder on a device signal with its
candidates in the library's fixed device scratch array.  Checked against
mf_peaks_device (which allocates its workspace) and the oracle's
PeakFinder (src/peakFinder.cpp:468-586), on two calls in a row where the
second is smaller, so stale scratch contents from the first would show."""
import numpy as np
import pytest

import oracle as O
from hipmem import DeviceArray
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _signal(n, seed):
    rng = np.random.default_rng(seed)
    x = np.convolve(rng.standard_normal(n), np.hanning(17), mode="same")
    x /= np.abs(x).max()
    return np.round(x, 3)  # plateaus and ties


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_peaks_match_allocating_entry_point(dtype):
    big = _signal(400_000, 11).astype(dtype)
    small = _signal(30_001, 12).astype(dtype)
    for x, thr, md in ((big, 0.2, 25), (small, 0.4, 3), (small, 0.0, 0)):
        print(f"noalloc peaks: n={x.size} dtype={np.dtype(dtype).name} thr={thr} minDist={md}", flush=True)
        d = DeviceArray(x)
        try:
            got = capi.peaks_device_noalloc(d.ptr, x.size, dtype, thr, md)
            want = capi.peaks_device(d.ptr, x.size, dtype, thr, md)
        finally:
            d.free()
        assert np.array_equal(got, want)
        assert np.array_equal(got, O.peak_finder(x, thr, md))
        assert len(got) > 0


def test_noalloc_peaks_single_sample_and_limits():
    one = np.array([0.9])
    d = DeviceArray(one)
    try:
        assert list(capi.peaks_device_noalloc(d.ptr, 1, np.float64, 0.5, 10)) == [0]
        assert list(capi.peaks_device_noalloc(d.ptr, 1, np.float64, 0.95, 10)) == []
        with pytest.raises(ValueError, match="scratch limit"):
            capi.peaks_device_noalloc(d.ptr, capi.NOALLOC_MAX_SAMPLES + 1, np.float64, 0.5, 10)
        with pytest.raises(ValueError):
            capi.peaks_device_noalloc(d.ptr, 1, np.float64, 1.5, 10)
    finally:
        d.free()


def test_noalloc_peaks_at_the_size_limit():
    """The largest accepted input, white noise with a local maximum about
    every third sample, then a small one: the second call must not see the
    first call's candidates."""
    n = capi.NOALLOC_MAX_SAMPLES
    rng = np.random.default_rng(5)
    x = np.round(rng.uniform(0, 1, n), 4).astype(np.float32)
    print(f"noalloc peaks at the limit: n={n}", flush=True)
    d = DeviceArray(x)
    try:
        got = capi.peaks_device_noalloc(d.ptr, n, np.float32, 0.0, 0)
        assert np.array_equal(got, O.peak_finder(x, 0.0, 0))
        assert len(got) > n // 4
        got = capi.peaks_device_noalloc(d.ptr, n, np.float32, 0.0, 2)
        want = capi.peaks_device(d.ptr, n, np.float32, 0.0, 2)
        assert np.array_equal(got, want)
        got2 = capi.peaks_device_noalloc(d.ptr, 1000, np.float32, 0.0, 2)
        assert np.array_equal(got2, O.peak_finder(x[:1000], 0.0, 2))
    finally:
        d.free()
