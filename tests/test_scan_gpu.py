"""GPU window scales from prefix scans (mf_window_scale_device): FP64 scans
and compensated FP32 (float-float) scans against the FP64 two-pass oracle
scale 1/||w - mean(w)|| (the Pearson scale of normalizeSignal,
src/matchedFilter.cpp:94-166).

Prefix scans lose accuracy where a window's spread is small next to the
anchored samples of its block (the cancellation of Lt S2 - S1^2).  The
bound checked is relative error <= 2e-7 (FP32 output rounding and the
final square root) + 30 kappa u (1 + 2 sqrt(span / Lt)), with
kappa = (sum of d^2 over the block's scanned span) / ||w - mean(w)||^2,
d = x - anchor, span = 2048 + Lt - 1, and u = 2^-53 (FP64 scans) or 2^-44
(float-float scans).  The 30 covers the scan's depth (two 6-level wave
scans, the offset add and the window difference); the square-root term
bounds the S1^2 part of the bracket by Cauchy-Schwarz.
"""
import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu

CHUNK = 2048  # outputs per block of the scan kernel


def _signal(n, seed):
    rng = np.random.default_rng(seed)
    x = 1.0e4 + np.convolve(rng.standard_normal(n + 8), np.ones(8) / 8.0, mode="valid")[:n]
    g0 = n // 3
    # a gap at level 1 with 1e-6 noise (kept by FP32 at that level): blocks
    # anchored inside it see small d, blocks anchored outside see d ~ -1e4
    x[g0:g0 + 3000] = 1.0 + 1.0e-6 * rng.standard_normal(3000)
    return x.astype(np.float32)


def _kappa(x, lt, n_out, ref):
    x64 = x.astype(np.float64)
    k = np.empty(n_out)
    for c0 in range(0, n_out, CHUNK):
        c1 = min(c0 + CHUNK, n_out)
        span = x64[c0:c1 + lt - 1] - x64[c0]
        k[c0:c1] = np.sum(span * span)
    with np.errstate(divide="ignore"):
        return k * ref * ref   # ref = 1/||w - mean||


@pytest.mark.parametrize("lt", [2, 37, 256, 2048])
def test_window_scales_match_oracle_within_scan_bound(lt):
    from hipmem import DeviceArray
    n = 60_000
    x = _signal(n, 100 + lt)
    n_out = n - lt + 1
    ref = O.window_scale(x, lt)
    kappa = _kappa(x, lt, n_out, ref)
    dx = DeviceArray(x)
    ds = DeviceArray(np.zeros(n_out, np.float32))
    try:
        for scan, u in (("f64", 2.0 ** -53), ("f32c", 2.0 ** -44)):
            capi.window_scale_device(dx.ptr, n, lt, scan, ds.ptr)
            got = ds.get().astype(np.float64)
            ok = ref > 0
            rel = np.abs(got[ok] / ref[ok] - 1.0)
            span = CHUNK + lt - 1
            bound = 2e-7 + 30.0 * kappa[ok] * u * (1.0 + 2.0 * np.sqrt(span / lt))
            worst = np.argmax(rel / bound)
            assert np.all(rel <= bound), (scan, lt, rel[worst], bound[worst])
            # well-conditioned windows (outside the gap) are at output precision
            span_ratio = (CHUNK + lt - 1) / lt
            good = ok & (kappa < 4.0 * span_ratio)
            assert good.any()
            assert np.abs(got[good] / ref[good] - 1.0).max() < 3e-7, scan
    finally:
        dx.free()
        ds.free()


def test_window_scale_rejects_bad_arguments():
    from hipmem import DeviceArray
    dx = DeviceArray(np.ones(100, np.float32))
    ds = DeviceArray(np.zeros(100, np.float32))
    try:
        for lt, scan_code in ((1, 0), (2049, 0), (101, 0), (10, 2)):
            rc = capi.lib().mf_window_scale_device(capi._vp(dx.ptr), 100, lt, scan_code, capi._vp(ds.ptr), None)
            assert rc == capi.MF_ERR_INVALID_ARGUMENT, (lt, scan_code, rc)
    finally:
        dx.free()
        ds.free()
