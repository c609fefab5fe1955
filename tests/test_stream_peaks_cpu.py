"""The streaming PeakFinder's host logic (mflib_amd/csrc/stream_peaks.hpp)
on the CPU: a signal fed in segments of many shapes (single samples, a
segment shorter than minDist, one segment for the whole signal) returns,
over all pushes and the final flush, exactly the oracle's PeakFinder list
of the whole signal (src/peakFinder.cpp:468-586), in increasing index.
Signals rounded to one decimal have plateaus and ties everywhere, which is
where the late decisions (a segment's last sample, greedy fates waiting on
higher-priority candidates) matter."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "stream_peaks_check.cpp")
INC = os.path.join(os.path.dirname(HERE), "mflib_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sp") / "stream_peaks_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D_GLIBCXX_ASSERTIONS", "-I", INC, SRC,
                    "-o", exe], check=True)
    return exe


def _stream(driver, tmp_path, x, segs, thr, md, ckpt=0):
    prec = 0 if x.dtype == np.float32 else 1
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([prec, md, len(segs), len(x), ckpt], np.int32).tofile(f)
        np.array([thr], np.float64).tofile(f)
        np.asarray(segs, np.int32).tofile(f)
        x.tofile(f)
    subprocess.run([driver, str(inp), str(outp)], check=True)
    raw = np.fromfile(outp, np.int64)
    cnt = int(raw[0])
    return raw[1:1 + cnt], int(raw[1 + cnt])


def _segments(rng, n, kind):
    if kind == "ones":
        return [1] * n
    if kind == "whole":
        return [n]
    lo, hi = {"short": (1, 6), "mixed": (1, 90), "long": (50, 400)}[kind]
    segs = []
    while sum(segs) < n:
        segs.append(int(min(rng.integers(lo, hi + 1), n - sum(segs))))
    return segs


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("kind", ["ones", "short", "mixed", "long", "whole"])
def test_stream_equals_whole_signal_peak_finder(driver, tmp_path, seed, kind):
    rng = np.random.default_rng(seed)
    n = 700
    x = np.round(rng.uniform(0, 1, n), 1)
    for dtype in (np.float64, np.float32):
        xs = x.astype(dtype)
        segs = _segments(rng, n, kind)
        for md in (0, 1, 3, 25, 150):
            for thr in (0.0, 0.55):
                got, disorder = _stream(driver, tmp_path, xs, segs, thr, md)
                ref = O.peak_finder(xs, thr, md)
                assert list(got) == list(ref), (kind, md, thr, dtype)
                assert disorder == 0


def test_stream_edges(driver, tmp_path):
    for x in ([0.9], [0.9, 0.1], [0.1, 0.9], [0.9, 0.9, 0.1], [0.5, 0.5], [0.2, 0.9, 0.9, 0.9]):
        x = np.array(x)
        for segs in ([len(x)], [1] * len(x)):
            got, _ = _stream(driver, tmp_path, x, segs, 0.5, 0)
            assert list(got) == list(O.peak_finder(x, 0.5, 0)), (x, segs)


def test_stream_long_chain_settles_late(driver, tmp_path):
    """Rising peaks 10 samples apart with minDist 15: every peak's fate waits
    on the next one, so nothing settles until the flush; with a final drop
    the chain resolves into alternate peaks."""
    n = 2000
    x = np.zeros(n)
    idx = np.arange(5, n - 5, 10)
    x[idx] = np.linspace(0.6, 0.99, len(idx))
    got, disorder = _stream(driver, tmp_path, x, [37] * (n // 37) + [n % 37], 0.5, 15)
    assert list(got) == list(O.peak_finder(x, 0.5, 15))
    assert disorder == 0


@pytest.mark.parametrize("ckpt", [1, 3, 7])
def test_stream_checkpoint_and_resume(driver, tmp_path, ckpt):
    """The state saved after every ckpt-th segment and loaded into a new
    finder (which then carries on) gives the uninterrupted list; the saved
    bytes survive a round trip, and truncated or altered bytes are refused."""
    rng = np.random.default_rng(40 + ckpt)
    n = 900
    x = np.round(rng.uniform(0, 1, n), 1)
    for dtype in (np.float64, np.float32):
        xs = x.astype(dtype)
        segs = _segments(rng, n, "mixed")
        for md in (0, 4, 40, 200):
            got, disorder = _stream(driver, tmp_path, xs, segs, 0.3, md, ckpt=ckpt)
            assert list(got) == list(O.peak_finder(xs, 0.3, md)), (md, dtype)
            assert disorder == 0
