"""GPU window scales from prefix scans (mf_window_scale_device): FP64 scans
and compensated FP32 (float-float) scans against the FP64 two-pass oracle
scale 1/||w - mean(w)|| (the Pearson scale of normalizeSignal,
src/matchedFilter.cpp:94-166).

Prefix scans lose accuracy where a window's spread is small next to the
anchored samples of its block (the cancellation of Lt S2 - S1^2).  The
bound checked is relative error <= 2e-7 (FP32 output rounding and the
final square root) + 30 kappa u (1 + 2 sqrt(span / Lt)), with
kappa = (sum of d^2 over the block's scanned span) / ||w - mean(w)||^2,
d = x - anchor, span = chunk + Lt - 1, and u = 2^-53 (FP64 scans) or 2^-44
(float-float scans).  The 30 covers the scan's depth (up to 16 sequential
adds per thread, a 6-level wave scan, the wave offsets and the window
difference); the square-root term bounds the S1^2 part of the bracket by
Cauchy-Schwarz.
"""
import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu

def _chunk(lt):
    """Outputs per block of the scan kernel: one span of 256 x K samples
    (K = 8 for Lt up to 256, 16 up to 2048) covers chunk + Lt - 1."""
    maxlt, k = (256, 8) if lt <= 256 else (2048, 16)
    return (256 * k - maxlt + 1) // 4 * 4


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
    ch = _chunk(lt)
    for c0 in range(0, n_out, ch):
        c1 = min(c0 + ch, n_out)
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
            span = _chunk(lt) + lt - 1
            bound = 2e-7 + 30.0 * kappa[ok] * u * (1.0 + 2.0 * np.sqrt(span / lt))
            worst = np.argmax(rel / bound)
            assert np.all(rel <= bound), (scan, lt, rel[worst], bound[worst])
            # well-conditioned windows (outside the gap) are at output precision
            span_ratio = (_chunk(lt) + lt - 1) / lt
            good = ok & (kappa < 4.0 * span_ratio)
            assert good.any()
            assert np.abs(got[good] / ref[good] - 1.0).max() < 3e-7, scan
    finally:
        dx.free()
        ds.free()


def test_window_scale_unaligned_signal_matches_aligned():
    """A signal pointer off 16-byte alignment takes the scalar-load path;
    its scales equal those of an aligned copy of the same samples."""
    from hipmem import DeviceArray
    x = _signal(30_001, 5)
    lt = 100
    dx = DeviceArray(x)                 # x[1:] starts 4 bytes past an aligned address
    dy = DeviceArray(x[1:].copy())      # the same samples, aligned
    n = x.size - 1
    a = DeviceArray(np.zeros(n - lt + 1, np.float32))
    b = DeviceArray(np.zeros(n - lt + 1, np.float32))
    try:
        for scan in ("f64", "f32c"):
            capi.window_scale_device(dx.ptr + 4, n, lt, scan, a.ptr)
            capi.window_scale_device(dy.ptr, n, lt, scan, b.ptr)
            assert np.array_equal(a.get(), b.get()), scan
    finally:
        for d in (dx, dy, a, b):
            d.free()


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
