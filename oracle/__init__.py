"""ctypes binding of the CPU oracle (oracle/liborc_mf.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.  See mf_oracle.h for what each
function restates (reference file:line) and mf_oracle.cpp's header for
what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc_mf.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_window_length.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_next_pow2.argtypes = [ctypes.c_int]
        L.orc_demean_normalize_reverse_f64.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_demean_normalize_reverse_f32.argtypes = [ctypes.c_int, _fp, _fp]
        L.orc_naive_pearson_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.orc_pearson_oracle_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp,
                                             ctypes.c_int, _llp, _dp]
        L.orc_pearson_oracle_f32x.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _fp,
                                              ctypes.c_int, _llp, _dp]
        L.orc_window_scale_f32x.argtypes = [ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, _llp, _dp]
        L.orc_mf_apply_f64.argtypes = [ctypes.c_int, ctypes.c_int, _ip,
                                       ctypes.POINTER(_dp), ctypes.c_int,
                                       ctypes.POINTER(_dp), ctypes.c_int, _dp, _ip, _ip]
        L.orc_mf_apply_f32.argtypes = [ctypes.c_int, ctypes.c_int, _ip,
                                       ctypes.POINTER(_dp), ctypes.c_int,
                                       ctypes.POINTER(_fp), ctypes.c_int, ctypes.c_int,
                                       _fp, _ip, _ip]
        L.orc_shift_and_stack_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_longlong,
                                              ctypes.c_int, _ip, _dp, ctypes.c_int,
                                              _dp, _ip]
        L.orc_shift_and_stack_f32.argtypes = [ctypes.c_int, _fp, ctypes.c_longlong,
                                              ctypes.c_int, _ip, _dp, ctypes.c_int,
                                              _fp, _ip]
        L.orc_reduce_correlations_f32.argtypes = [ctypes.c_int, _fp, ctypes.c_longlong, ctypes.c_int,
                                                  ctypes.c_int, _fp, _ip]
        L.orc_reduce_correlations_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_longlong, ctypes.c_int,
                                                  ctypes.c_int, _dp, _ip]
        L.orc_peak_finder_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_double,
                                          ctypes.c_int, _ip, _ip]
        L.orc_peak_finder_f32.argtypes = [ctypes.c_int, _fp, ctypes.c_float,
                                          ctypes.c_int, _ip, _ip]
        L.orc_rfft_f64.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_quadratic_shift.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_lanczos_shift.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_amplitude_factors.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_irfft_f64.argtypes = [ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected its input (rc={rc})")


def window_length(nb: int, nx: int):
    """(nfft, L) as computeOptimalFFTAndBlockLength returns them."""
    a, b = ctypes.c_int(), ctypes.c_int()
    _check(lib().orc_window_length(nb, nx, ctypes.byref(a), ctypes.byref(b)),
           "window_length")
    return a.value, b.value


def next_pow2(n: int) -> int:
    return lib().orc_next_pow2(n)


def demean_normalize_reverse(b, dtype=np.float64):
    b = np.ascontiguousarray(b, dtype=dtype)
    out = np.empty_like(b)
    if dtype == np.float64:
        _check(lib().orc_demean_normalize_reverse_f64(len(b), _ptr(b, _dp), _ptr(out, _dp)), "dnr")
    else:
        _check(lib().orc_demean_normalize_reverse_f32(len(b), _ptr(b, _fp), _ptr(out, _fp)), "dnr")
    return out


def naive_pearson(b, x):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x) - len(b) + 1)
    _check(lib().orc_naive_pearson_f64(len(b), _ptr(b, _dp), len(x), _ptr(x, _dp),
                                       _ptr(out, _dp)), "naive_pearson")
    return out


def pearson(b, x, lags=None):
    """FP64 two-pass Pearson oracle (compensated sums).  x may be float32
    (promoted exactly) or float64.  Optional subset of lags."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    if x.dtype == np.float32:
        x = np.ascontiguousarray(x)
        fn, xp = lib().orc_pearson_oracle_f32x, _ptr(x, _fp)
    else:
        x = np.ascontiguousarray(x, dtype=np.float64)
        fn, xp = lib().orc_pearson_oracle_f64, _ptr(x, _dp)
    if lags is None:
        out = np.empty(len(x) - len(b) + 1)
        _check(fn(len(b), _ptr(b, _dp), len(x), xp, 0, None, _ptr(out, _dp)), "pearson")
    else:
        lags = np.ascontiguousarray(lags, dtype=np.int64)
        out = np.empty(len(lags))
        _check(fn(len(b), _ptr(b, _dp), len(x), xp, len(lags), _ptr(lags, _llp),
                  _ptr(out, _dp)), "pearson")
    return out


def window_scale(x, nb: int, lags=None):
    """FP64 two-pass 1/||w - mean(w)|| of the windows x[i:i+nb] of a
    float32 signal (the Pearson scale of normalizeSignal); optional lags."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if lags is None:
        out = np.empty(len(x) - nb + 1)
        _check(lib().orc_window_scale_f32x(nb, len(x), _ptr(x, _fp), 0, None, _ptr(out, _dp)), "window_scale")
    else:
        lags = np.ascontiguousarray(lags, dtype=np.int64)
        out = np.empty(len(lags))
        _check(lib().orc_window_scale_f32x(nb, len(x), _ptr(x, _fp), len(lags), _ptr(lags, _llp),
                                           _ptr(out, _dp)), "window_scale")
    return out


def mf_apply(templates, signals, mode: str = "multi", dtype=np.float64,
             float_quirk: bool = False, threads: int = 0):
    """Reference-faithful MatchedFilter<T>::apply.  Returns (rho[nT, n-nb+1],
    nfft, L)."""
    templates = [np.ascontiguousarray(t, dtype=np.float64) for t in templates]
    nT = len(templates)
    lens = np.array([len(t) for t in templates], dtype=np.int32)
    tptrs = (_dp * nT)(*[_ptr(t, _dp) for t in templates])
    m = 1 if mode == "multi" else 0
    if m == 1:
        assert len(signals) == nT
    sigs = [np.ascontiguousarray(s, dtype=dtype) for s in (signals if m == 1 else signals[:1])]
    n = len(sigs[0])
    nb = int(lens.max())
    out = np.empty((nT, n - nb + 1), dtype=dtype)
    nfft, L = ctypes.c_int(), ctypes.c_int()
    if dtype == np.float64:
        sp = (_dp * len(sigs))(*[_ptr(s, _dp) for s in sigs])
        rc = lib().orc_mf_apply_f64(m, nT, _ptr(lens, _ip), tptrs, n, sp, threads,
                                    _ptr(out, _dp), ctypes.byref(nfft), ctypes.byref(L))
    else:
        sp = (_fp * len(sigs))(*[_ptr(s, _fp) for s in sigs])
        rc = lib().orc_mf_apply_f32(m, nT, _ptr(lens, _ip), tptrs, n, sp,
                                    1 if float_quirk else 0, threads,
                                    _ptr(out, _fp), ctypes.byref(nfft), ctypes.byref(L))
    _check(rc, "mf_apply")
    return out, nfft.value, L.value


def shift_and_stack(rho, shifts, weights, absval: bool = True):
    rho = np.ascontiguousarray(rho)
    nT, nOut = rho.shape
    shifts = np.ascontiguousarray(shifts, dtype=np.int32)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    out = np.empty(max(nOut - int(shifts.max()), 0), dtype=rho.dtype)
    ln = ctypes.c_int()
    if rho.dtype == np.float64:
        rc = lib().orc_shift_and_stack_f64(nT, _ptr(rho, _dp), nOut, nOut, _ptr(shifts, _ip),
                                           _ptr(weights, _dp), int(absval), _ptr(out, _dp),
                                           ctypes.byref(ln))
    else:
        rc = lib().orc_shift_and_stack_f32(nT, _ptr(rho, _fp), nOut, nOut, _ptr(shifts, _ip),
                                           _ptr(weights, _dp), int(absval), _ptr(out, _fp),
                                           ctypes.byref(ln))
    _check(rc, "shift_and_stack")
    return out[:ln.value]


def reduce_correlations(rho, absval: bool = False, length: int | None = None):
    """Detector reduce + clamp over the rows of rho (2-D, float32 or
    float64): returns (det, id) of the first `length` lags."""
    rho = np.ascontiguousarray(rho)
    if rho.dtype not in (np.float32, np.float64):
        rho = rho.astype(np.float64)
    nT, ld = rho.shape
    n = ld if length is None else int(length)
    det = np.empty(n, dtype=rho.dtype)
    idx = np.empty(n, dtype=np.int32)
    if rho.dtype == np.float32:
        rc = lib().orc_reduce_correlations_f32(nT, _ptr(rho, _fp), ld, n, int(absval), _ptr(det, _fp), _ptr(idx, _ip))
    else:
        rc = lib().orc_reduce_correlations_f64(nT, _ptr(rho, _dp), ld, n, int(absval), _ptr(det, _dp), _ptr(idx, _ip))
    _check(rc, "reduce_correlations")
    return det, idx


def quadratic_shift(xc, peak: int, dt: float = 1.0, absval: bool = False) -> float:
    """quadraticRefinement of Detector::detect on the FP64 row xc at lag peak."""
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    out = np.zeros(1)
    _check(lib().orc_quadratic_shift(len(xc), _ptr(xc, _dp), int(peak), float(dt), int(absval), _ptr(out, _dp)),
           "quadratic_shift")
    return float(out[0])


def lanczos_shift(xc, peak: int, dt: float = 1.0, absval: bool = False, alpha: int = 51, points: int = 100):
    """lanczosRefinement of Detector::detect: returns (shift, profile), the
    profile holding the interpolant at the points-1 grid shifts."""
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    out = np.zeros(1)
    prof = np.zeros(max(points - 1, 1))
    _check(lib().orc_lanczos_shift(len(xc), _ptr(xc, _dp), int(peak), float(dt), int(absval), int(alpha),
                                   int(points), _ptr(out, _dp), _ptr(prof, _dp)), "lanczos_shift")
    return float(out[0]), prof[:max(points - 1, 0)]


def amplitude_factors(tmpl, window):
    """RelativeMagnitude's (Gibbons-Ringdal, Schaff-Richards) scaling factors
    of window against tmpl; None when the window is constant."""
    t = np.ascontiguousarray(tmpl, dtype=np.float64)
    y = np.ascontiguousarray(window, dtype=np.float64)
    assert len(t) == len(y)
    gr, sr = np.zeros(1), np.zeros(1)
    ok = ctypes.c_int()
    _check(lib().orc_amplitude_factors(len(t), _ptr(t, _dp), _ptr(y, _dp), _ptr(gr, _dp), _ptr(sr, _dp),
                                       ctypes.byref(ok)), "amplitude_factors")
    return (float(gr[0]), float(sr[0])) if ok.value else None


def peak_finder(x, threshold: float, min_dist: int):
    x = np.ascontiguousarray(x)
    idx = np.empty(max(len(x), 1), dtype=np.int32)
    npk = ctypes.c_int()
    if x.dtype == np.float64:
        rc = lib().orc_peak_finder_f64(len(x), _ptr(x, _dp), threshold, min_dist,
                                       _ptr(idx, _ip), ctypes.byref(npk))
    else:
        x = x.astype(np.float32)
        rc = lib().orc_peak_finder_f32(len(x), _ptr(x, _fp), np.float32(threshold), min_dist,
                                       _ptr(idx, _ip), ctypes.byref(npk))
    _check(rc, "peak_finder")
    return idx[:npk.value].copy()


def rfft(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x) + 2)
    _check(lib().orc_rfft_f64(len(x), _ptr(x, _dp), _ptr(out, _dp)), "rfft")
    return out[0::2] + 1j * out[1::2]


def irfft(X, N):
    Xi = np.empty(2 * (N // 2 + 1))
    Xi[0::2] = np.real(X)
    Xi[1::2] = np.imag(X)
    out = np.empty(N)
    _check(lib().orc_irfft_f64(N, _ptr(Xi, _dp), _ptr(out, _dp)), "irfft")
    return out
