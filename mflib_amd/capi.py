"""ctypes binding of the C ABI in include/mflib_gpu.h (libmflib_gpu.so).

This is the layer the GPU parity tests and bench.py drive.  It raises
ValueError for MF_ERR_INVALID_ARGUMENT and RuntimeError for every other
failure, mirroring the std::invalid_argument / std::runtime_error split of
the reference.  There is no CPU fallback: loading fails loudly when the
library is missing, and every compute call fails when no GPU is present.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MFLIB_GPU_LIB") or os.path.join(_HERE, "lib", "libmflib_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mflib_gpu.h")

MF_OK = 0
MF_ERR_INVALID_ARGUMENT = -1
MF_ERR_BAD_STATE = -2
MF_ERR_OUT_OF_MEMORY = -3
MF_ERR_DEVICE = -4
MF_ERR_UNSUPPORTED = -5
MF_ERR_CANCELLED = -6
PRECISION = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}
MODES = {"single": 0, "multi": 1, "batched": 2}

_lib = None
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "mf_last_error": ([], ctypes.c_char_p),
    "mf_version": ([_ip, _ip, _ip], ctypes.c_int),
    "mf_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mf_destroy": ([_vp], ctypes.c_int),
    "mf_clear": ([_vp], ctypes.c_int),
    "mf_set_signal_size": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_set_templates": ([_vp, ctypes.c_int, _ip, ctypes.POINTER(_dp), _ip, _dp], ctypes.c_int),
    "mf_set_stack_absolute_values": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_set_batch_layout": ([_vp, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int], ctypes.c_int),
    "mf_set_fft_length": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_set_kernel_variant": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_set_graph_launch": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_initialize": ([_vp], ctypes.c_int),
    "mf_is_initialized": ([_vp, _ip], ctypes.c_int),
    "mf_get_number_of_templates": ([_vp, _ip], ctypes.c_int),
    "mf_get_signal_length": ([_vp, _ip], ctypes.c_int),
    "mf_get_filtered_length": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_get_plan": ([_vp, ctypes.c_int, _ip, _ip], ctypes.c_int),
    "mf_set_signal": ([_vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "mf_set_signal_device": ([_vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "mf_zero_signal": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_get_signal": ([_vp, ctypes.c_int, _vp, ctypes.c_int], ctypes.c_int),
    "mf_apply": ([_vp], ctypes.c_int),
    "mf_synchronize": ([_vp], ctypes.c_int),
    "mf_have_matched_filtered_signals": ([_vp, _ip], ctypes.c_int),
    "mf_get_rho": ([_vp, ctypes.c_int, _vp, ctypes.c_int], ctypes.c_int),
    "mf_get_rho_device": ([_vp, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "mf_shift_and_stack": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_shift_and_stack_device": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_detect": ([_vp, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int, _ip], ctypes.c_int),
    "mf_get_stack_device": ([_vp, ctypes.c_int, ctypes.POINTER(_vp), _ip], ctypes.c_int),
    "mf_get_number_of_groups": ([_vp, _ip], ctypes.c_int),
    "mf_template_spectrum": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_peaks": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip,
                  ctypes.c_int, _ip], ctypes.c_int),
    "mf_peaks_device": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip,
                         ctypes.c_int, _ip], ctypes.c_int),
    "mf_peaks_device_noalloc": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip,
                                 ctypes.c_int, _ip], ctypes.c_int),
    "mf_compute_optimal_fft_and_block_length": ([ctypes.c_int, ctypes.c_int, _ip, _ip], ctypes.c_int),
    "mf_window_scale_device": ([_vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _vp, _dp], ctypes.c_int),
    "mf_reduce_group": ([_vp, ctypes.c_int, ctypes.c_int, _vp, _ip, ctypes.c_int, _ip], ctypes.c_int),
    "mf_reduce_group_device": ([_vp, ctypes.c_int, ctypes.c_int, _vp, _ip, ctypes.c_int, _ip], ctypes.c_int),
    "mf_detect_reduced": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, _ip, ctypes.c_int,
                           _ip], ctypes.c_int),
    "mf_get_device": ([_vp, _ip], ctypes.c_int),
    "mf_set_streaming": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_get_stream_carry": ([_vp, _ip], ctypes.c_int),
    "mf_get_stream_state": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_set_stream_state": ([_vp, ctypes.c_int, _vp, ctypes.c_int], ctypes.c_int),
    "mf_push_chunks_device": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "mf_stream_peaks": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                         ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, _ip], ctypes.c_int),
    "mf_stream_peaks_all": ([_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                             ctypes.c_int, _ip], ctypes.c_int),
    "mf_get_stream_peaks_state": ([_vp, ctypes.c_int, _vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)],
                                  ctypes.c_int),
    "mf_set_stream_peaks_state": ([_vp, ctypes.c_int, _vp, ctypes.c_longlong], ctypes.c_int),
    "mf_stream_peaks_reset": ([_vp, ctypes.c_int], ctypes.c_int),
    "mf_detect_templates": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int, _ip],
                            ctypes.c_int),
    "mf_detect_template": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int, _ip],
                           ctypes.c_int),
    "mf_amplitude_factors": ([ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _vp, ctypes.c_longlong, _dp, _dp,
                              _ip], ctypes.c_int),
    "mf_detection_products": ([_vp, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _dp, _dp, _dp, _dp, _ip], ctypes.c_int),
    "mf_get_timings": ([_vp, _dp, _dp], ctypes.c_int),
    "mf_get_window_stats": ([_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 4, ctypes.c_int),
    "mf_server_window_stats": ([_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 4, ctypes.c_int),
    "mf_autotune": ([ctypes.c_int, _dp, ctypes.c_int, _ip], ctypes.c_int),
    "mf_server_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int],
                         ctypes.c_int),
    "mf_server_destroy": ([_vp], ctypes.c_int),
    "mf_server_add_templates": ([_vp, ctypes.c_int, _ip, ctypes.POINTER(_dp), _ip], ctypes.c_int),
    "mf_server_submit": ([_vp, ctypes.c_int, _ip, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _ip,
                          ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "mf_server_wait": ([_vp, ctypes.c_longlong, _ip], ctypes.c_int),
    "mf_server_poll": ([_vp, ctypes.c_longlong, _ip], ctypes.c_int),
    "mf_server_get_rho": ([_vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _ip],
                          ctypes.c_int),
    "mf_server_get_rho_device": ([_vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp), _ip],
                                 ctypes.c_int),
    "mf_server_latency": ([_vp, ctypes.c_longlong, _dp, _dp], ctypes.c_int),
    "mf_server_release": ([_vp, ctypes.c_longlong], ctypes.c_int),
    "mf_server_cancel": ([_vp, ctypes.c_longlong, _ip], ctypes.c_int),
    "mf_debug_fail_allocations": ([ctypes.c_int, _ip], ctypes.c_int),
    "mf_server_stats": ([_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _dp],
                        ctypes.c_int),
}


def header_symbols() -> list:
    """Every function declared in include/mflib_gpu.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mf_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m mflib_amd.build` "
                               "(the engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().mf_last_error() or b"").decode()


def check(rc: int, what: str = "") -> None:
    if rc == MF_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == MF_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(f"[{rc}] {msg}")


def debug_fail_allocations(n: int) -> int:
    """Arms n injected device-allocation failures (0 disarms); returns the
    count that was still armed."""
    rem = ctypes.c_int()
    check(lib().mf_debug_fail_allocations(int(n), ctypes.byref(rem)), "mf_debug_fail_allocations")
    return rem.value


def compute_optimal_fft_and_block_length(nb: int, nx: int):
    a, b = ctypes.c_int(), ctypes.c_int()
    check(lib().mf_compute_optimal_fft_and_block_length(nb, nx, ctypes.byref(a), ctypes.byref(b)),
          "computeOptimalFFTAndBlockLength")
    return a.value, b.value


def autotune(dtype=np.float32) -> list:
    """Measures the per-block cost of every FFT size and kernel variant on
    the running GPU and installs it for this process's FFT plans."""
    buf = np.zeros(64 * 5, dtype=np.float64)
    n = ctypes.c_int()
    check(lib().mf_autotune(PRECISION[np.dtype(dtype)], buf.ctypes.data_as(_dp), 64, ctypes.byref(n)),
          "mf_autotune")
    return [{"nfft": 2 << int(r[0]), "variant": int(r[1]), "a_us": r[2], "b_shared_us": r[3],
             "b_solo_us": r[4]} for r in buf[:n.value * 5].reshape(-1, 5)]


def peaks(x, threshold: float, min_dist: int):
    """PeakFinder on the GPU for a host numpy array (float32 or float64)."""
    x = np.ascontiguousarray(x)
    if x.dtype not in PRECISION:
        x = x.astype(np.float64)
    idx = np.empty(max(len(x), 1), dtype=np.int32)
    n = ctypes.c_int()
    check(lib().mf_peaks(x.ctypes.data_as(_vp), len(x), PRECISION[x.dtype], float(threshold),
                         int(min_dist), idx.ctypes.data_as(_ip), len(idx), ctypes.byref(n)), "peaks")
    return idx[:n.value].copy()


def peaks_device(ptr: int, n: int, dtype, threshold: float, min_dist: int, cap: int | None = None):
    cap = cap or max(n, 1)
    idx = np.empty(cap, dtype=np.int32)
    cnt = ctypes.c_int()
    check(lib().mf_peaks_device(_vp(ptr), n, PRECISION[np.dtype(dtype)], float(threshold),
                                int(min_dist), idx.ctypes.data_as(_ip), cap, ctypes.byref(cnt)),
          "peaks_device")
    return idx[:cnt.value].copy()


NOALLOC_MAX_SAMPLES = 1 << 22


def peaks_device_noalloc(ptr: int, n: int, dtype, threshold: float, min_dist: int, cap: int | None = None):
    """peaks_device with the candidates in the library's fixed device
    scratch array (no device allocation); n <= NOALLOC_MAX_SAMPLES."""
    cap = cap or max(n, 1)
    idx = np.empty(cap, dtype=np.int32)
    cnt = ctypes.c_int()
    check(lib().mf_peaks_device_noalloc(_vp(ptr), n, PRECISION[np.dtype(dtype)], float(threshold),
                                        int(min_dist), idx.ctypes.data_as(_ip), cap, ctypes.byref(cnt)),
          "peaks_device_noalloc")
    return idx[:cnt.value].copy()


SCAN_PRECISION = {"f64": 0, "f32c": 1}

# DetectionTimeInterpolationType values and mf_detection_products flag bits
INTERP_LANCZOS, INTERP_QUADRATIC, INTERP_NONE = 0, 1, 2
PRODUCT_AMPLITUDE, PRODUCT_TRUNCATED = 1, 2


def window_scale_device(ptr: int, n: int, lt: int, scan: str, out_ptr: int) -> float:
    """Pearson window scales of a device FP32 signal into a device buffer
    (n - lt + 1 floats) from FP64 ("f64") or compensated FP32 ("f32c")
    prefix scans; returns the kernel time in ms."""
    ms = ctypes.c_double()
    check(lib().mf_window_scale_device(_vp(ptr), int(n), int(lt), SCAN_PRECISION[scan], _vp(out_ptr),
                                       ctypes.byref(ms)), "window_scale_device")
    return ms.value


def amplitude_factors(template, windows, dtype=np.float64):
    """RelativeMagnitude's scaling factors of each row of `windows` (2-D,
    one detected window per row) against `template`, on the GPU: (alpha_gr,
    alpha_sr, computed)."""
    t = np.ascontiguousarray(template, dtype=np.float64)
    w = np.ascontiguousarray(np.atleast_2d(windows), dtype=dtype)
    n, lt = w.shape
    if lt != len(t):
        raise ValueError("windows must have the template's length")
    gr, sr = np.empty(max(n, 1)), np.empty(max(n, 1))
    ok = np.empty(max(n, 1), dtype=np.int32)
    check(lib().mf_amplitude_factors(PRECISION[w.dtype], lt, t.ctypes.data_as(_dp), n, w.ctypes.data_as(_vp), lt,
                                     gr.ctypes.data_as(_dp), sr.ctypes.data_as(_dp), ok.ctypes.data_as(_ip)),
          "amplitude_factors")
    return gr[:n].copy(), sr[:n].copy(), ok[:n].astype(bool)


class Handle:
    """One engine instance (one mf_handle)."""

    def __init__(self, dtype=np.float32, mode: str = "multi"):
        self.dtype = np.dtype(dtype)
        self._h = _vp()
        check(lib().mf_create(ctypes.byref(self._h), PRECISION[self.dtype], MODES[mode]), "mf_create")

    def close(self):
        if self._h:
            lib().mf_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- configuration
    def set_signal_size(self, n: int):
        check(lib().mf_set_signal_size(self._h, int(n)), "set_signal_size")

    def set_templates(self, templates, shifts=None, weights=None):
        self._tmpl = [np.ascontiguousarray(t, dtype=np.float64) for t in templates]
        nT = len(self._tmpl)
        lens = np.array([len(t) for t in self._tmpl], dtype=np.int32)
        ptrs = (_dp * nT)(*[t.ctypes.data_as(_dp) for t in self._tmpl])
        sh = None if shifts is None else np.ascontiguousarray(shifts, dtype=np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        check(lib().mf_set_templates(self._h, nT, lens.ctypes.data_as(_ip), ptrs,
                                     None if sh is None else sh.ctypes.data_as(_ip),
                                     None if w is None else w.ctypes.data_as(_dp)), "set_templates")

    def set_stack_absolute_values(self, absval: bool):
        check(lib().mf_set_stack_absolute_values(self._h, 1 if absval else 0))

    def set_batch_layout(self, channel_lengths, template_channel, template_group, n_groups: int):
        cl = np.ascontiguousarray(channel_lengths, dtype=np.int32)
        tc = np.ascontiguousarray(template_channel, dtype=np.int32)
        tg = np.ascontiguousarray(template_group, dtype=np.int32)
        check(lib().mf_set_batch_layout(self._h, len(cl), cl.ctypes.data_as(_ip), tc.ctypes.data_as(_ip),
                                        tg.ctypes.data_as(_ip), int(n_groups)), "set_batch_layout")

    def set_fft_length(self, nfft: int):
        check(lib().mf_set_fft_length(self._h, int(nfft)), "set_fft_length")

    def set_graph_launch(self, on: bool = True):
        check(lib().mf_set_graph_launch(self._h, int(bool(on))), "set_graph_launch")

    def set_kernel_variant(self, variant: int):
        check(lib().mf_set_kernel_variant(self._h, int(variant)), "set_kernel_variant")

    def set_streaming(self, on: bool = True):
        """Streaming state (before initialize): channels keep their last
        carry samples ahead of each pushed chunk."""
        check(lib().mf_set_streaming(self._h, int(bool(on))), "set_streaming")

    def stream_carry(self) -> int:
        return self._int(lib().mf_get_stream_carry)

    def stream_state(self, channel: int) -> np.ndarray:
        """The carry samples channel `channel` continues from (a checkpoint)."""
        out = np.empty(max(self.stream_carry(), 1), dtype=self.dtype)
        ln = ctypes.c_int()
        check(lib().mf_get_stream_state(self._h, int(channel), out.ctypes.data_as(_vp), len(out), ctypes.byref(ln)),
              "stream_state")
        return out[:ln.value].copy()

    def set_stream_state(self, channel: int, state):
        st = np.ascontiguousarray(state, dtype=self.dtype)
        check(lib().mf_set_stream_state(self._h, int(channel), st.ctypes.data_as(_vp), len(st)), "set_stream_state")

    def push_chunks_device(self, ptrs, n: int):
        """Streaming push of every channel from device pointers ptrs[c]."""
        arr = (_vp * len(ptrs))(*[_vp(int(p)) for p in ptrs])
        check(lib().mf_push_chunks_device(self._h, len(ptrs), int(n), arr), "push_chunks_device")

    def stream_peaks(self, group: int, threshold: float, min_dist: int, last: bool = False):
        """PeakFinder over the stream of group's stacks: the lags that became
        final peaks with the latest push (int64, increasing across calls);
        last=True ends the stream.  Call once after every apply."""
        cap = 1024
        while True:
            out = np.empty(cap, dtype=np.int64)
            n = ctypes.c_int()
            rc = lib().mf_stream_peaks(self._h, int(group), float(threshold), int(min_dist), 1 if last else 0,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), cap, ctypes.byref(n))
            if rc == MF_ERR_INVALID_ARGUMENT and n.value > cap:
                cap = n.value
                continue
            check(rc, "stream_peaks")
            return out[:n.value].copy()

    def stream_peaks_all(self, threshold: float, min_dist: int, last: bool = False):
        """stream_peaks for every group in one call: a list of int64 arrays,
        one per group."""
        ng = self.number_of_groups()
        off = np.zeros(ng + 1, dtype=np.int32)
        cap = 4096
        while True:
            out = np.empty(cap, dtype=np.int64)
            rc = lib().mf_stream_peaks_all(self._h, float(threshold), int(min_dist), 1 if last else 0,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), cap,
                                           off.ctypes.data_as(_ip))
            if rc == MF_ERR_INVALID_ARGUMENT and off[ng] > cap:
                cap = int(off[ng])
                continue
            check(rc, "stream_peaks_all")
            return [out[off[g]:off[g + 1]].copy() for g in range(ng)]

    def stream_peaks_state(self, group: int) -> bytes:
        """Checkpoint of a group's stream-peak finder (opaque bytes)."""
        n = ctypes.c_longlong()
        lib().mf_get_stream_peaks_state(self._h, int(group), None, 0, ctypes.byref(n))
        if n.value <= 0:
            check(lib().mf_get_stream_peaks_state(self._h, int(group), None, 0, ctypes.byref(n)),
                  "stream_peaks_state")
        buf = ctypes.create_string_buffer(n.value)
        check(lib().mf_get_stream_peaks_state(self._h, int(group), buf, n.value, ctypes.byref(n)),
              "stream_peaks_state")
        return buf.raw[:n.value]

    def set_stream_peaks_state(self, group: int, state: bytes):
        check(lib().mf_set_stream_peaks_state(self._h, int(group), ctypes.c_char_p(bytes(state)), len(state)),
              "set_stream_peaks_state")

    def stream_peaks_reset(self, group: int = -1):
        check(lib().mf_stream_peaks_reset(self._h, int(group)), "stream_peaks_reset")

    def initialize(self):
        check(lib().mf_initialize(self._h), "initialize")

    # ---- queries
    def _int(self, fn, *args):
        v = ctypes.c_int()
        check(fn(self._h, *args, ctypes.byref(v)))
        return v.value

    def is_initialized(self) -> bool:
        return bool(self._int(lib().mf_is_initialized))

    def number_of_templates(self) -> int:
        return self._int(lib().mf_get_number_of_templates)

    def signal_length(self) -> int:
        return self._int(lib().mf_get_signal_length)

    def filtered_length(self, row: int = 0) -> int:
        return self._int(lib().mf_get_filtered_length, int(row))

    def plan(self, row: int = 0):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().mf_get_plan(self._h, int(row), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # ---- signals
    def set_signal(self, channel: int, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        check(lib().mf_set_signal(self._h, int(channel), len(x), x.ctypes.data_as(_vp)), "set_signal")

    def set_signal_device(self, channel: int, ptr: int, n: int):
        check(lib().mf_set_signal_device(self._h, int(channel), int(n), _vp(ptr)), "set_signal_device")

    def zero_signal(self, channel: int):
        check(lib().mf_zero_signal(self._h, int(channel)), "zero_signal")

    def get_signal(self, channel: int, n: int):
        out = np.empty(n, dtype=self.dtype)
        check(lib().mf_get_signal(self._h, int(channel), out.ctypes.data_as(_vp), n), "get_signal")
        return out

    # ---- compute
    def apply(self):
        check(lib().mf_apply(self._h), "apply")

    def synchronize(self):
        check(lib().mf_synchronize(self._h), "synchronize")

    def have_matched_filtered_signals(self) -> bool:
        return bool(self._int(lib().mf_have_matched_filtered_signals))

    def rho(self, row: int):
        n = self.filtered_length(row)
        out = np.empty(n, dtype=self.dtype)
        check(lib().mf_get_rho(self._h, int(row), out.ctypes.data_as(_vp), n), "get_rho")
        return out

    def rho_device_ptr(self, row: int) -> int:
        p = _vp()
        check(lib().mf_get_rho_device(self._h, int(row), ctypes.byref(p)))
        return p.value

    def shift_and_stack(self, group: int = 0):
        ln = ctypes.c_int()
        cap = self.filtered_length(0) + 1
        out = np.empty(max(cap, 1), dtype=self.dtype)
        # first call obtains the length when the group's rows are longer
        rc = lib().mf_shift_and_stack(self._h, int(group), out.ctypes.data_as(_vp), len(out), ctypes.byref(ln))
        if rc == MF_ERR_INVALID_ARGUMENT and ln.value > len(out):
            out = np.empty(ln.value, dtype=self.dtype)
            rc = lib().mf_shift_and_stack(self._h, int(group), out.ctypes.data_as(_vp), len(out),
                                          ctypes.byref(ln))
        check(rc, "shift_and_stack")
        return out[:ln.value].copy()

    def shift_and_stack_device(self, group: int, dptr: int, cap: int) -> int:
        ln = ctypes.c_int()
        check(lib().mf_shift_and_stack_device(self._h, int(group), _vp(dptr), int(cap), ctypes.byref(ln)),
              "shift_and_stack_device")
        return ln.value

    def number_of_groups(self) -> int:
        return self._int(lib().mf_get_number_of_groups)

    def detect(self, threshold: float, min_dist: int, cap: int | None = None):
        """Stack every group and run PeakFinder on every stack; returns a
        list of index arrays, one per group."""
        ng = self.number_of_groups()
        cap = cap or max(1, 1 << 20)
        idx = np.empty(cap, dtype=np.int32)
        off = np.empty(ng + 1, dtype=np.int32)
        check(lib().mf_detect(self._h, float(threshold), int(min_dist), idx.ctypes.data_as(_ip), cap,
                              off.ctypes.data_as(_ip)), "detect")
        return [idx[off[g]:off[g + 1]].copy() for g in range(ng)]

    def reduce(self, group: int = 0, absval: bool = False):
        """Detector reduce + clamp of a group's rows: (det, template id)."""
        n = ctypes.c_int()
        rc = lib().mf_reduce_group(self._h, int(group), int(absval), None, None, 0, ctypes.byref(n))
        if rc != MF_ERR_INVALID_ARGUMENT:
            check(rc, "reduce")
        det = np.empty(max(n.value, 1), dtype=self.dtype)
        ids = np.empty(max(n.value, 1), dtype=np.int32)
        check(lib().mf_reduce_group(self._h, int(group), int(absval), det.ctypes.data_as(_vp),
                                    ids.ctypes.data_as(_ip), det.size, ctypes.byref(n)), "reduce")
        return det[:n.value], ids[:n.value]

    def reduce_device(self, group: int, absval: bool, det_ptr: int, id_ptr: int, cap: int) -> int:
        n = ctypes.c_int()
        check(lib().mf_reduce_group_device(self._h, int(group), int(absval), _vp(det_ptr),
                                           ctypes.cast(_vp(id_ptr), _ip), int(cap), ctypes.byref(n)),
              "reduce_device")
        return n.value

    def detect_reduced(self, threshold: float, min_dist: int, group: int = 0, absval: bool = False,
                       cap: int | None = None):
        """Detector REDUCED policy up to the peak list: (peak indices,
        template index of each peak)."""
        cap = cap or max(1, 1 << 20)
        idx = np.empty(cap, dtype=np.int32)
        tid = np.empty(cap, dtype=np.int32)
        n = ctypes.c_int()
        check(lib().mf_detect_reduced(self._h, int(group), int(absval), float(threshold), int(min_dist),
                                      idx.ctypes.data_as(_ip), tid.ctypes.data_as(_ip), cap, ctypes.byref(n)),
              "detect_reduced")
        return idx[:n.value].copy(), tid[:n.value].copy()

    def detect_template(self, it: int, threshold: float, min_dist: int, absval: bool = False,
                        cap: int | None = None):
        """Detector SINGLE policy for one template: its peak list."""
        cap = cap or max(1, 1 << 20)
        idx = np.empty(cap, dtype=np.int32)
        n = ctypes.c_int()
        check(lib().mf_detect_template(self._h, int(it), int(absval), float(threshold), int(min_dist),
                                       idx.ctypes.data_as(_ip), cap, ctypes.byref(n)), "detect_template")
        return idx[:n.value].copy()

    def detect_templates(self, threshold: float, min_dist: int, absval: bool = False, cap: int | None = None):
        """Detector SINGLE policy for every template: a list of peak arrays."""
        nt = self.number_of_templates()
        cap = cap or max(1, sum(self.filtered_length(r) // 2 + 1 for r in range(nt)))
        idx = np.empty(cap, dtype=np.int32)
        off = np.empty(nt + 1, dtype=np.int32)
        check(lib().mf_detect_templates(self._h, int(absval), float(threshold), int(min_dist),
                                        idx.ctypes.data_as(_ip), cap, off.ctypes.data_as(_ip)), "detect_templates")
        return [idx[off[r]:off[r + 1]].copy() for r in range(nt)]

    def detection_products(self, peaks, template_ids, sampling_periods=None, absval: bool = False,
                           interpolation: int = INTERP_QUADRATIC, lanczos_alpha: int = 51,
                           lanczos_points: int = 100):
        """Per-detection products of Detector::detect for the peaks
        (lags) on the template rows template_ids: a dict of rho (clamped),
        shift (seconds, or samples when sampling_periods is None), alpha_gr,
        alpha_sr and flags (PRODUCT_AMPLITUDE, PRODUCT_TRUNCATED)."""
        pk = np.ascontiguousarray(peaks, dtype=np.int32)
        ti = np.ascontiguousarray(template_ids, dtype=np.int32)
        if pk.shape != ti.shape or pk.ndim != 1:
            raise ValueError("peaks and template_ids must be 1-D arrays of one length")
        n = len(pk)
        dts = None
        if sampling_periods is not None:
            dts = np.ascontiguousarray(sampling_periods, dtype=np.float64)
            if dts.ndim != 1 or len(dts) != self.number_of_templates():
                raise ValueError("sampling_periods needs one entry per template")
        out = {k: np.empty(max(n, 1)) for k in ("rho", "shift", "alpha_gr", "alpha_sr")}
        flags = np.empty(max(n, 1), dtype=np.int32)
        check(lib().mf_detection_products(
            self._h, n, pk.ctypes.data_as(_ip), ti.ctypes.data_as(_ip),
            dts.ctypes.data_as(_dp) if dts is not None else None, int(absval), int(interpolation),
            int(lanczos_alpha), int(lanczos_points), out["rho"].ctypes.data_as(_dp),
            out["shift"].ctypes.data_as(_dp), out["alpha_gr"].ctypes.data_as(_dp),
            out["alpha_sr"].ctypes.data_as(_dp), flags.ctypes.data_as(_ip)), "detection_products")
        res = {k: v[:n].copy() for k, v in out.items()}
        res["flags"] = flags[:n].copy()
        return res

    def stack_device(self, group: int):
        p, n = _vp(), ctypes.c_int()
        check(lib().mf_get_stack_device(self._h, int(group), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def template_spectrum(self, row: int):
        """getSpectraOfTemplate(row): nfft/2 + 1 complex bins at the FFT
        length the engine exports for the row (the size comes from the
        call itself, so time-domain rows work too)."""
        ln = ctypes.c_int(-1)
        rc = lib().mf_template_spectrum(self._h, int(row), None, 0, ctypes.byref(ln))
        if ln.value < 1:
            check(rc, "template_spectrum")
        out = np.empty(2 * ln.value, dtype=self.dtype)
        check(lib().mf_template_spectrum(self._h, int(row), out.ctypes.data_as(_vp), ln.value,
                                         ctypes.byref(ln)), "template_spectrum")
        return out[0::2] + 1j * out[1::2]

    def timings(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().mf_get_timings(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def window_stats(self):
        """Window statistics summed over the applies since initialize:
        windows (outputs computed), flagged (outputs recomputed exactly in
        FP64), full (the flagged ones recomputed as a two-pass Pearson
        value), solo (outputs of rows that computed their window scale
        alone)."""
        v = [ctypes.c_longlong() for _ in range(4)]
        check(lib().mf_get_window_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("windows", "flagged", "full", "solo"), (x.value for x in v)))


class Server:
    """Request server (mf_server_*): resident template library, continuous
    batching of matched-filter requests, per-request latency."""

    def __init__(self, dtype=np.float32, budget_bytes: float = 8 << 30, max_batch: int = 256,
                 max_inflight: int = 2):
        self.dtype = np.dtype(dtype)
        self._h = _vp()
        check(lib().mf_server_create(ctypes.byref(self._h), PRECISION[self.dtype], float(budget_bytes),
                                     int(max_batch), int(max_inflight)), "mf_server_create")
        self._keep = {}

    def close(self):
        if self._h:
            lib().mf_server_destroy(self._h)
            self._h = _vp()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_templates(self, templates) -> list:
        arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in templates]
        lens = (ctypes.c_int * len(arrs))(*[a.size for a in arrs])
        ptrs = (_dp * len(arrs))(*[a.ctypes.data_as(_dp) for a in arrs])
        first = ctypes.c_int()
        check(lib().mf_server_add_templates(self._h, len(arrs), lens, ptrs, ctypes.byref(first)),
              "mf_server_add_templates")
        return list(range(first.value, first.value + len(arrs)))

    def submit(self, channels, template_ids, on_device: bool = False, lengths=None) -> int:
        """channels: host numpy arrays of the server dtype, or (with
        on_device) device addresses with `lengths`."""
        n = len(channels)
        if on_device:
            lens = list(lengths)
            ptrs = (_vp * n)(*[int(p) for p in channels])
            keep = None
        else:
            arrs = [np.ascontiguousarray(c, dtype=self.dtype) for c in channels]
            lens = [a.size for a in arrs]
            ptrs = (_vp * n)(*[a.ctypes.data for a in arrs])
            keep = arrs
        ids = (ctypes.c_int * len(template_ids))(*[int(t) for t in template_ids])
        rid = ctypes.c_longlong()
        check(lib().mf_server_submit(self._h, n, (ctypes.c_int * n)(*lens), ptrs, int(bool(on_device)),
                                     len(template_ids), ids, ctypes.byref(rid)), "mf_server_submit")
        if keep is not None:
            self._keep[rid.value] = keep  # host inputs must outlive the request
        return rid.value

    def wait(self, rid: int) -> None:
        st = ctypes.c_int()
        check(lib().mf_server_wait(self._h, rid, ctypes.byref(st)), "mf_server_wait")
        check(st.value, f"request {rid}")

    def done(self, rid: int) -> bool:
        d = ctypes.c_int()
        check(lib().mf_server_poll(self._h, rid, ctypes.byref(d)), "mf_server_poll")
        return bool(d.value)

    def rho(self, rid: int, template_index: int, channel: int) -> np.ndarray:
        p, n = _vp(), ctypes.c_int()
        check(lib().mf_server_get_rho_device(self._h, rid, template_index, channel, ctypes.byref(p),
                                             ctypes.byref(n)), "mf_server_get_rho_device")
        out = np.empty(n.value, dtype=self.dtype)
        check(lib().mf_server_get_rho(self._h, rid, template_index, channel, out.ctypes.data, out.size,
                                      ctypes.byref(n)), "mf_server_get_rho")
        return out

    def rho_device(self, rid: int, template_index: int, channel: int):
        p, n = _vp(), ctypes.c_int()
        check(lib().mf_server_get_rho_device(self._h, rid, template_index, channel, ctypes.byref(p),
                                             ctypes.byref(n)), "mf_server_get_rho_device")
        return p.value, n.value

    def latency(self, rid: int):
        q, t = ctypes.c_double(), ctypes.c_double()
        check(lib().mf_server_latency(self._h, rid, ctypes.byref(q), ctypes.byref(t)), "mf_server_latency")
        return q.value, t.value

    def release(self, rid: int) -> None:
        check(lib().mf_server_release(self._h, rid), "mf_server_release")
        self._keep.pop(rid, None)

    def cancel(self, rid: int) -> bool:
        """Cancels a request not yet launched; False when it already runs
        (or ran) on the GPU."""
        c = ctypes.c_int()
        check(lib().mf_server_cancel(self._h, rid, ctypes.byref(c)), "mf_server_cancel")
        return bool(c.value)

    def window_stats(self):
        """Window statistics of the launched batches (see
        Handle.window_stats)."""
        v = [ctypes.c_longlong() for _ in range(4)]
        check(lib().mf_server_window_stats(self._h, *[ctypes.byref(x) for x in v]), "mf_server_window_stats")
        return dict(zip(("windows", "flagged", "full", "solo"), (x.value for x in v)))

    def status(self, rid: int) -> int:
        """Waits for the request and returns its status code (MF_OK, or the
        failure code, MF_ERR_CANCELLED for a cancelled request)."""
        st = ctypes.c_int()
        check(lib().mf_server_wait(self._h, rid, ctypes.byref(st)), "mf_server_wait")
        return st.value

    def stats(self):
        b, r, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        check(lib().mf_server_stats(self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(c)),
              "mf_server_stats")
        return b.value, r.value, c.value
