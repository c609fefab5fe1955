"""GPU parity of the HIP engine, called through the C ABI (libmflib_gpu.so).

Each test restates a reference test (file:line in the docstring) or checks
the engine against the CPU oracle on seeded inputs.  Tolerances are the
reference's own where it states one, and 1e-5 (the engine's contract for
FP32 correlation against the FP64 oracle) otherwise.
"""
import numpy as np
import pytest

import oracle as O
import refcases as R
from hipmem import read_array
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _multi(templates, signals, dtype, nfft=0, shifts=None, weights=None):
    h = capi.Handle(dtype, "multi")
    h.set_templates(templates, shifts=shifts, weights=weights)
    h.set_signal_size(len(signals[0]))
    if nfft:
        h.set_fft_length(nfft)
    h.initialize()
    for i, s in enumerate(signals):
        h.set_signal(i, s)
    h.apply()
    return h


def test_basic_double():
    """testing/matchedFilter.cpp:58-168: |rho - naive| < 1e-13, 1 then 4 templates."""
    x, x2, b, b2 = R.basic_test_inputs()
    ref = O.naive_pearson(b, x)
    ref2 = O.naive_pearson(b2, x2)
    h = _multi([b], [x], np.float64)
    r = h.rho(0)
    assert len(r) == 1000 - len(b) + 1
    assert np.abs(r - ref).max() < 1e-13
    h = _multi([b, b2, b, b2], [x, x2, x, x2], np.float64)
    for i in range(4):
        assert np.abs(h.rho(i) - (ref if i % 2 == 0 else ref2)).max() < 1e-13


def test_basic_float():
    """testing/matchedFilter.cpp:172-271: float path within 2e-6."""
    x, x2, b, b2 = R.basic_test_inputs()
    xf, x2f = x.astype(np.float32), x2.astype(np.float32)
    bf = b.astype(np.float32).astype(np.float64)    # the reference test's float template
    b2f = b2.astype(np.float32).astype(np.float64)
    ref = O.pearson(bf, xf)
    ref2 = O.pearson(b2f, x2f)
    h = _multi([bf, b2f, bf, b2f], [xf, x2f, xf, x2f], np.float32)
    for i in range(4):
        assert np.abs(h.rho(i) - (ref if i % 2 == 0 else ref2)).max() < 2e-6


def test_single_channel_uneven_lengths():
    """testing/singleChannel.cpp:104-160: three uneven templates, < 1e-12."""
    rng = np.random.default_rng(20240601)
    wave, tpls = R.single_channel_inputs(rng)
    h = capi.Handle(np.float64, "single")
    h.set_templates(tpls)
    h.set_signal_size(len(wave))
    h.initialize()
    h.set_signal(0, wave)
    h.apply()
    nout = len(wave) - 140 + 1
    for k in range(3):
        r = h.rho(k)
        assert len(r) == nout
        assert np.abs(r - O.naive_pearson(tpls[k], wave)[:nout]).max() < 1e-12


def test_shift_and_stack_reference_case():
    """testing/matchedFilter.cpp:275-465: rho within 1e-11, stack max 1 +/- 1e-3
    at the origin time +/- dt/2."""
    tp, sg, tt, toff, df, origin = R.shift_and_stack_inputs()
    shifts = [R.moveout_shift(tt[k], toff[k], df) for k in range(3)]
    h = _multi(tp, sg, np.float64, shifts=shifts, weights=[1, 1, 1])
    for k in range(3):
        r = h.rho(k)
        assert np.abs(r - O.naive_pearson(tp[k], sg[k])[:len(r)]).max() <= 1e-11
    st = h.shift_and_stack()
    i = int(np.argmax(st))
    assert abs(st[i] - 1.0) < 1e-3
    assert abs(i / df - origin) < 1 / df / 2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stack_bit_exact_vs_oracle(dtype):
    """shiftAndStack (src/matchedFilter.cpp:1318-1364) on identical rho:
    bit-identical to the oracle's sequential accumulation."""
    rng = np.random.default_rng(7)
    nT, n, nb = 5, 20000, 300
    tpls = [rng.standard_normal(nb) for _ in range(nT)]
    sigs = [rng.standard_normal(n).astype(dtype) for _ in range(nT)]
    shifts = rng.integers(0, 150, nT)
    weights = rng.uniform(0.1, 1.0, nT)
    h = _multi(tpls, sigs, dtype, shifts=shifts, weights=weights)
    rho = np.stack([h.rho(i) for i in range(nT)])
    for absval in (True, False):
        h.set_stack_absolute_values(absval)
        st = h.shift_and_stack()
        ref = O.shift_and_stack(rho, shifts, weights, absval=absval)
        assert st.dtype == ref.dtype and len(st) == len(ref)
        assert np.array_equal(st, ref)


def test_peak_finder_kats():
    """testing/peakFinder.cpp:16-143: exact 13 and 21 peak indices."""
    sig, thr, md, ref13, ref21 = R.peak_finder_signal()
    assert list(capi.peaks(sig, thr, md)) == ref13
    assert list(capi.peaks(sig, thr, 0)) == ref21
    assert list(capi.peaks(sig.astype(np.float32), thr, md)) == ref13
    assert len(capi.peaks(np.full(4000, 0.4), thr, md)) == 0


def test_peak_finder_many_candidates_vs_oracle():
    """Tens of thousands of candidates: more than the first readback holds
    (8192), so the rest come back in a second copy; host and device inputs."""
    from hipmem import DeviceArray
    rng = np.random.default_rng(77)
    x = np.round(rng.uniform(0, 1, 120_000), 4)         # white noise: a local maximum every ~3 samples
    for md in (0, 3, 40):
        ref = O.peak_finder(x, 0.0, md)
        assert np.array_equal(capi.peaks(x, 0.0, md), ref)
        d = DeviceArray(x)
        try:
            assert np.array_equal(capi.peaks_device(d.ptr, x.size, np.float64, 0.0, md), ref)
        finally:
            d.free()
    assert len(O.peak_finder(x, 0.0, 0)) > 8192


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_peak_finder_device_unaligned_starts_and_short_lengths(dtype):
    """The candidate scan reads whole 16-byte vectors from the first 16-byte
    boundary on and tests the samples before it and after the last whole
    vector one by one: signals starting 0-3 elements past a boundary, of
    lengths 1-9, 63-65 and 4099, with plateaus, against the oracle."""
    from hipmem import DeviceArray
    rng = np.random.default_rng(31)
    base = np.round(rng.uniform(0, 1, 5000), 1).astype(dtype)   # plateaus and ties
    d = DeviceArray(base)
    try:
        step = np.dtype(dtype).itemsize
        for off in range(4 if dtype == np.float32 else 2):
            for n in list(range(1, 10)) + [63, 64, 65, 4099]:
                x = base[off:off + n]
                for thr, md in ((0.0, 0), (0.45, 3)):
                    got = capi.peaks_device(d.ptr + off * step, n, dtype, thr, md)
                    assert np.array_equal(got, O.peak_finder(x, thr, md)), (off, n, thr, md)
    finally:
        d.free()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_peak_finder_random_vs_oracle(seed):
    rng = np.random.default_rng(seed)
    x = np.convolve(rng.standard_normal(50000), np.hanning(31), mode="same")
    x /= np.abs(x).max()
    x = np.round(x, 3)  # plateaus and ties
    for md in (0, 1, 10, 200):
        assert np.array_equal(capi.peaks(x, 0.3, md), O.peak_finder(x, 0.3, md))
        xf = x.astype(np.float32)
        assert np.array_equal(capi.peaks(xf, 0.3, md), O.peak_finder(xf, 0.3, md))


@pytest.mark.parametrize("nfft", [1024, 2048, 4096, 8192])
def test_multiblock_float_vs_fp64_oracle(nfft):
    """Many blocks per channel, DC offsets, uneven template lengths."""
    rng = np.random.default_rng(nfft)
    n = 60000
    lens = [257, 300, 211]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    sigs = [(rng.standard_normal(n) + rng.uniform(-100, 100)).astype(np.float32) for _ in lens]
    h = _multi(tpls, sigs, np.float32, nfft=nfft)
    assert h.plan(0)[0] == nfft
    nout = n - max(lens) + 1
    for k in range(3):
        r = h.rho(k)
        ref = O.pearson(tpls[k], sigs[k])[:nout]
        assert np.abs(r - ref).max() < 1e-5


@pytest.mark.parametrize("nfft", [1024, 4096])
def test_multiblock_double_vs_fp64_oracle(nfft):
    rng = np.random.default_rng(nfft + 1)
    n = 30000
    lens = [400, 333]
    tpls = [rng.standard_normal(L) for L in lens]
    sigs = [rng.standard_normal(n) * 3 + 1e3 for _ in lens]
    h = _multi(tpls, sigs, np.float64, nfft=nfft)
    nout = n - max(lens) + 1
    for k in range(2):
        ref = O.pearson(tpls[k], sigs[k])[:nout]
        assert np.abs(h.rho(k) - ref).max() < 1e-11


@pytest.mark.parametrize("dtype,nfft", [(np.float32, 1024), (np.float32, 2048), (np.float32, 4096),
                                        (np.float64, 1024), (np.float64, 2048)])
def test_kernel_variants_match_oracle(dtype, nfft):
    """Block-wide (1) and wave-level (2) correlation kernels: single channel,
    several template-length groups (one wave per row, more rows than waves
    in a group, a group smaller than the wave count), a DC offset and a
    quiet stretch that forces the exact fix-up."""
    rng = np.random.default_rng(nfft + (7 if dtype == np.float64 else 0))
    n = 50000
    x = rng.standard_normal(n) * 0.5 + 40.0
    x[20000:21500] = 40.0 + 1e-5 * rng.standard_normal(1500)
    x = x.astype(dtype)
    lens = [200, 200, 173, 173, 173, 173, 173, 150, 211]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    nout = n - max(lens) + 1
    refs = [O.pearson(t, x)[:nout] for t in tpls]
    tol = 1e-5 if dtype == np.float32 else 1e-11
    out = {}
    for variant in (1, 2):
        h = capi.Handle(dtype, "single")
        h.set_templates(tpls)
        h.set_signal_size(n)
        h.set_fft_length(nfft)
        h.set_kernel_variant(variant)
        h.initialize()
        h.set_signal(0, x)
        h.apply()
        out[variant] = [h.rho(k) for k in range(len(tpls))]
        for k in range(len(tpls)):
            assert np.abs(out[variant][k] - refs[k]).max() < tol, (variant, k)
        h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("slen", [1, 2, 3, 2047, 2048, 2049, 4097])
def test_detect_stack_lengths_around_the_block(slen, dtype):
    """mf_detect's one launch of stacks and candidates at stack lengths of 1,
    2, 3 and around its 2048-output blocks: the stacks it wrote equal the
    oracle's shiftAndStack bit for bit, and the peaks equal the oracle
    PeakFinder's at threshold 0 (every local maximum, so the block edges'
    recomputed neighbours and the end-point rules all matter)."""
    rng = np.random.default_rng(900 + slen)
    C, L = 3, 40
    shifts = [0, 7, 3]
    n = L - 1 + slen + max(shifts)
    sigs = [(rng.standard_normal(n) + rng.uniform(-5, 5)).astype(dtype) for _ in range(C)]
    tpls = [rng.standard_normal(L) * np.hanning(L) for _ in range(C)]
    weights = [0.5, 0.25, 0.25]
    h = capi.Handle(dtype, "batched")
    h.set_templates(tpls, shifts=shifts, weights=weights)
    h.set_batch_layout([n] * C, list(range(C)), [0] * C, 1)
    h.initialize()
    for c in range(C):
        h.set_signal(c, sigs[c])
    h.apply()
    rows = [h.rho(r) for r in range(C)]
    st_ref = O.shift_and_stack(np.stack(rows), shifts, weights)
    assert st_ref.size == slen
    peaks = h.detect(0.0, 0)
    ptr, got_len = h.stack_device(0)
    assert got_len == slen
    assert np.array_equal(read_array(ptr, dtype, slen), st_ref)
    assert np.array_equal(peaks[0], O.peak_finder(st_ref, 0.0, 0))
    h.close()


def test_batched_layout_detect_matches_oracle():
    """Batched layout (templates of two groups share three channels): each
    row equals the single-instance result, mf_detect's stacks equal the
    oracle's shiftAndStack of the GPU rows bit for bit, and its peaks equal
    the oracle PeakFinder's."""
    rng = np.random.default_rng(99)
    C, n = 3, 40000
    sigs = [(np.convolve(rng.standard_normal(n), np.hanning(9), "same") + rng.uniform(-20, 20))
            .astype(np.float32) for _ in range(C)]
    G = 2
    tpls, chans, groups, shifts, weights = [], [], [], [], []
    for g in range(G):
        for c in range(C):
            L = 150 + 10 * c + 5 * g
            i0 = 1000 + 7000 * g + 300 * c
            tpls.append(sigs[c][i0:i0 + L].astype(np.float64) + 0.01 * rng.standard_normal(L))
            chans.append(c)
            groups.append(g)
            shifts.append(300 * c)
            weights.append(0.5 + 0.25 * c)
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=shifts, weights=weights)
    h.set_batch_layout([n] * C, chans, groups, G)
    h.initialize()
    for c in range(C):
        h.set_signal(c, sigs[c])
    h.apply()
    rows = [h.rho(r) for r in range(len(tpls))]
    for r, t in enumerate(tpls):
        ref = O.pearson(t, sigs[chans[r]])[:len(rows[r])]
        assert np.abs(rows[r] - ref).max() < 1e-5
    peaks = h.detect(0.3, 25)
    for g in range(G):
        members = [r for r in range(len(tpls)) if groups[r] == g]
        st_ref = O.shift_and_stack(np.stack([rows[r] for r in members]), [shifts[r] for r in members],
                                   [weights[r] for r in members])
        # the stack mf_detect wrote in the same launch as its candidates
        ptr, slen = h.stack_device(g)
        assert np.array_equal(read_array(ptr, np.float32, slen), st_ref)
        st = h.shift_and_stack(g)
        assert np.array_equal(st, st_ref)
        assert np.array_equal(peaks[g], O.peak_finder(st_ref, 0.3, 25))
        assert len(peaks[g]) >= 1
    # threshold 0 with no minimum distance: every local maximum above 0 is a
    # peak (tens of thousands over both groups), more than the detect
    # readback's first batch, so the remainder is read in a second copy
    peaks0 = h.detect(0.0, 0)
    total = 0
    for g in range(G):
        members = [r for r in range(len(tpls)) if groups[r] == g]
        st_ref = O.shift_and_stack(np.stack([rows[r] for r in members]), [shifts[r] for r in members],
                                   [weights[r] for r in members])
        ref0 = O.peak_finder(st_ref, 0.0, 0)
        assert np.array_equal(peaks0[g], ref0)
        total += len(ref0)
    assert total > 8192


def test_dead_and_zeroed_channels():
    """Dead trace (|min| == |max|, src/matchedFilter.cpp:890-894) and
    zeroSignal (823-845) give rho == 0."""
    rng = np.random.default_rng(5)
    n = 5000
    tpls = [rng.standard_normal(100) for _ in range(3)]
    sigs = [np.full(n, 3.25, np.float32), rng.standard_normal(n).astype(np.float32),
            rng.standard_normal(n).astype(np.float32)]
    h = _multi(tpls, sigs, np.float32)
    assert np.all(h.rho(0) == 0)
    assert np.abs(h.rho(1)).max() > 0.1
    h.zero_signal(2)
    h.apply()
    assert np.all(h.rho(2) == 0)


def test_gap_windows_recomputed_exactly():
    """Constant and near-constant stretches inside ordinary blocks: the FFT
    numerator is unreliable there, the engine recomputes those windows
    exactly (FP64 two-pass); must match the FP64 oracle to 1e-5."""
    rng = np.random.default_rng(11)
    n = 40000
    x = rng.standard_normal(n) * 0.4 + 25.0
    x[5000:8000] = 25.0                                 # exact constant
    x[15000:17000] = 3.0 + 1e-4 * rng.standard_normal(2000)  # quiet, offset level
    x = x.astype(np.float32)
    tpls = [rng.standard_normal(200)]
    h = _multi(tpls, [x], np.float32, nfft=4096)
    r = h.rho(0)
    ref = O.pearson(tpls[0], x)[:len(r)]
    assert np.abs(r - ref).max() < 1e-5


def test_errors_match_reference_conventions():
    h = capi.Handle(np.float32, "multi")
    with pytest.raises(RuntimeError):
        h.apply()                       # not initialised: runtime_error
    h.set_templates([np.arange(10.0)])
    h.set_signal_size(100)
    h.initialize()
    with pytest.raises(ValueError):
        h.set_signal(1, np.zeros(100))  # index out of range: invalid_argument
    with pytest.raises(ValueError):
        h.set_signal(0, np.zeros(99))   # wrong length: invalid_argument
    with pytest.raises(RuntimeError):
        h.shift_and_stack()             # not applied: runtime_error
    with pytest.raises(ValueError):
        capi.peaks(np.zeros(10), 1.5, 0)  # threshold outside [0,1]


def test_length_buckets_get_their_own_fft_size():
    """Rows of very different template lengths on one channel are split into
    length buckets, each with its own FFT size; every row still matches the
    FP64 oracle."""
    rng = np.random.default_rng(606)
    n = 200_000
    x = (rng.standard_normal(n) + 30).astype(np.float32)
    lens = [200, 210, 220, 230, 2900, 3000]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(lens), weights=[1.0] * len(lens))
    h.set_batch_layout([n], [0] * len(lens), [0] * len(lens), 1)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    short, long_ = h.plan(0)[0], h.plan(len(lens) - 1)[0]
    assert short < long_, (short, long_)
    nout = n - max(lens) + 1
    lags = rng.integers(0, nout, 3000)
    for i, t in enumerate(tpls):
        r = h.rho(i)
        assert r.size == nout
        assert np.abs(r[lags].astype(np.float64) - O.pearson(t, x, lags)).max() < 1e-5


def test_autotune_then_uneven_batch_matches_oracle():
    """The measured cost table installs and plans built from it stay exact."""
    table = capi.autotune(np.float32)
    assert len(table) >= 6
    assert all(r["a_us"] > 0 and r["b_shared_us"] > 0 and r["b_solo_us"] > 0 for r in table)
    assert {r["nfft"] for r in table} == {1024, 2048, 4096, 8192}
    rng = np.random.default_rng(607)
    C, lens_c = 3, [150_000, 90_000, 120_000]
    xs = [(rng.standard_normal(L) * 2 + rng.uniform(-50, 50)).astype(np.float32) for L in lens_c]
    tl = [300, 1200, 2500]
    tpls, rch, rgr = [], [], []
    for g, L in enumerate(tl):
        for c in range(C):
            tpls.append(rng.standard_normal(L) * np.hanning(L))
            rch.append(c)
            rgr.append(g)
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout(lens_c, rch, rgr, len(tl))
    h.initialize()
    for c in range(C):
        h.set_signal(c, xs[c])
    h.apply()
    for r, (t, c, g) in enumerate(zip(tpls, rch, rgr)):
        nout = min(lens_c) - tl[g] + 1
        row = h.rho(r)
        assert row.size == nout
        lags = rng.integers(0, nout, 1500)
        assert np.abs(row[lags].astype(np.float64) - O.pearson(t, xs[c], lags)).max() < 1e-5


def test_graph_launch_replays_match_eager():
    """mf_apply replayed from a HIP graph gives the eager result on every
    replay, with new signals and host synchronisations between replays."""
    rng = np.random.default_rng(608)
    n, C = 60_000, 3
    lens = [300, 700, 1500]
    tpls, rch = [], []
    for c in range(C):
        for L in lens:
            tpls.append(rng.standard_normal(L) * np.hanning(L))
            rch.append(c)

    def handle(graph):
        h = capi.Handle(np.float32, "batched")
        h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
        h.set_batch_layout([n] * C, rch, [0] * len(tpls), 1)
        h.set_graph_launch(graph)
        h.initialize()
        return h

    eager, graphed = handle(False), handle(True)
    for replay in range(3):
        xs = [(rng.standard_normal(n) + rng.uniform(-20, 20)).astype(np.float32) for _ in range(C)]
        for h in (eager, graphed):
            for c in range(C):
                h.set_signal(c, xs[c])
            h.apply()
            h.synchronize()
        for r in range(len(tpls)):
            assert np.array_equal(graphed.rho(r), eager.rho(r)), (replay, r)


@pytest.mark.parametrize("variant,tol", [(3, 1e-5), (4, 1e-5)])
def test_time_domain_mfma_path_matches_oracle(variant, tol):
    """Short templates (64-256 taps, several lengths) on the MFMA
    time-domain path, FP16 two-part or BF16 three-part operands: every row within the
    FP32 contract of the FP64 oracle, including a constant gap with a DC
    offset (windows recomputed exactly) and a dead channel."""
    rng = np.random.default_rng(700 + variant)
    n = 40_000
    x0 = (rng.standard_normal(n) * 3 + 1e4).astype(np.float32)
    x0[12000:14500] = np.float32(1e4 + 17.0) + (rng.standard_normal(2500) * 1e-6).astype(np.float32)
    x1 = np.full(n, 5.0, dtype=np.float32)   # dead channel
    lens = [64, 100, 128, 129, 200, 256, 256, 300]   # 300 stays on the FFT path
    tpls, rch = [], []
    for c in range(2):
        for L in lens:
            tpls.append(rng.standard_normal(L) * np.hanning(L))
            rch.append(c)
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([n, n], rch, [0] * len(tpls), 1)
    h.set_kernel_variant(variant)
    h.initialize()
    h.set_signal(0, x0)
    h.set_signal(1, x1)
    h.apply()
    nout = n - max(lens) + 1
    assert h.plan(0)[0] == 0 and h.plan(len(lens) - 1)[0] > 0  # time-domain vs FFT row
    for r, (t, c) in enumerate(zip(tpls, rch)):
        row = h.rho(r)
        assert row.size == nout
        if c == 1:
            assert not np.any(row)
            continue
        ref = O.pearson(t, x0)[:nout]
        assert np.abs(row.astype(np.float64) - ref).max() < tol, (r, len(t))


@pytest.mark.parametrize("variant", [3, 4])
def test_time_domain_short_tap_lanes_match_oracle(variant):
    """Lanes whose longest template pads to 64 and to 128 taps (the kernel's
    KT = 64 and 128 forms: 2 and 4 k-steps of 32), with a DC offset and a
    near-flat stretch: every row within the FP32 contract of the FP64
    oracle."""
    rng = np.random.default_rng(780 + variant)
    n = 30_000
    sigs = []
    for c in range(2):
        x = (rng.standard_normal(n) * 2 + 500.0 * (c + 1)).astype(np.float32)
        x[9000:9600] = np.float32(500.0 * (c + 1)) + (rng.standard_normal(600) * 1e-3).astype(np.float32)
        sigs.append(x)
    lens = [[16, 33, 50, 64, 64], [65, 80, 100, 127, 128]]   # channel 0: KT 64; channel 1: KT 128
    tpls, rch = [], []
    for c in range(2):
        for L in lens[c]:
            tpls.append(rng.standard_normal(L) * np.hanning(L))
            rch.append(c)
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([n, n], rch, [0] * len(tpls), 1)
    h.set_kernel_variant(variant)
    h.initialize()
    for c in range(2):
        h.set_signal(c, sigs[c])
    h.apply()
    for r, (t, c) in enumerate(zip(tpls, rch)):
        assert h.plan(r)[0] == 0
        row = h.rho(r)
        ref = O.pearson(t, sigs[c])[:row.size]
        assert np.abs(row.astype(np.float64) - ref).max() < 1e-5, (r, len(t))


@pytest.mark.parametrize("variant", [3, 4])
def test_time_domain_many_rows_per_lane_match_oracle(variant):
    """150 short templates of 20-256 taps on one channel: work items of 128
    and 22 rows, so every wave of a block holds a partial or empty row set
    and the 16-row tiles take two row groups; every row within the FP32
    contract of the FP64 oracle."""
    rng = np.random.default_rng(740 + variant)
    n = 6_000
    x = (rng.standard_normal(n) * 2 + 300.0).astype(np.float32)
    lens = [int(v) for v in rng.integers(20, 257, size=150)]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([n], [0] * len(tpls), [0] * len(tpls), 1)
    h.set_kernel_variant(variant)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    nout = n - max(lens) + 1
    for r, t in enumerate(tpls):
        assert h.plan(r)[0] == 0
        row = h.rho(r)
        assert row.size == nout
        ref = O.pearson(t, x)[:nout]
        assert np.abs(row.astype(np.float64) - ref).max() < 1e-5, (r, len(t))


@pytest.mark.parametrize("variant", [3, 4])
def test_time_domain_flag_queue_overflow_matches_oracle(variant):
    """More flagged windows than the time-domain flag queue holds: a signal
    alternating loud noise with near-flat stretches (1e4 plus or minus one
    ulp), whose windows lose their variance's digits without being exactly
    constant, so every one is queued for exact recomputation.  The queue
    holds max(windows / 32, 65536) entries; the rest are recomputed by the
    MFMA kernel's own per-wave lists.  Every row within the FP32 contract of
    the FP64 oracle."""
    rng = np.random.default_rng(760 + variant)
    n, period = 140_000, 700
    x = np.empty(n, dtype=np.float32)
    ulp = np.spacing(np.float32(1e4))
    for p0 in range(0, n, period):
        q = min(n, p0 + period // 2)
        x[p0:q] = (rng.standard_normal(q - p0) * 1e3).astype(np.float32)
        r = min(n, p0 + period)
        x[q:r] = np.float32(1e4) + ulp * rng.integers(-1, 2, size=r - q).astype(np.float32)
    lens = [int(v) for v in rng.integers(64, 257, size=16)]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([n], [0] * len(tpls), [0] * len(tpls), 1)
    h.set_kernel_variant(variant)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    nout = n - max(lens) + 1
    ws = h.window_stats()
    # far more than the queue holds (16 rows x 139,745 windows / 32 = 69,872)
    assert ws["full"] > 300_000, ws
    for r, t in enumerate(tpls):
        assert h.plan(r)[0] == 0
        row = h.rho(r)
        ref = O.pearson(t, x)[:nout]
        assert np.abs(row.astype(np.float64) - ref).max() < 1e-5, (r, len(t))
