"""Streaming state (mf_set_streaming): a continuous signal pushed in
fixed-size chunks.  The tails of the rows and stacks of consecutive pushes
tile the lag axis, and equal the FP64 oracle over the whole concatenated
signal (Pearson definition, testing/utilities.hpp:33-69; shiftAndStack,
src/matchedFilter.cpp:1318-1364) at the accuracy contract: 1e-5 for FP32
instances, 1e-11 for FP64.  Lags whose windows start before the stream read
the zero carry and are not compared."""
import numpy as np
import pytest

import oracle as O
from hipmem import DeviceArray
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _signal(rng, n, dc):
    x = rng.standard_normal(n) * 0.5 + dc
    x[n // 3:n // 3 + 700] = dc + 0.25          # a constant stretch inside one chunk
    return x


def test_multichannel_stream_rows_and_stacks_match_oracle():
    rng = np.random.default_rng(41)
    lens, shifts, weights = [300, 250, 300], [0, 7, 19], [1.0, 0.5, 0.75]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    chunk, pushes = 10_000, 6
    total = chunk * pushes
    sigs = [_signal(rng, total, 50.0 * c) for c in range(3)]
    for c, t in enumerate(tpls):                  # events straddling chunk boundaries
        for pos in (chunk - 120, 3 * chunk - 5, 4 * chunk + 999):
            sigs[c][pos + shifts[c]:pos + shifts[c] + len(t)] += 3 * t
    h = capi.Handle(np.float32, "multi")
    h.set_templates(tpls, shifts=shifts, weights=weights)
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    nb, smax = max(lens), max(shifts)
    assert h.stream_carry() == nb - 1 + smax
    assert h.filtered_length(0) == chunk + smax
    assert h.signal_length() == chunk
    rows_tail = [[] for _ in tpls]
    stack_tail = []
    for p in range(pushes):
        for c in range(3):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk].astype(np.float32))
        h.apply()
        for r in range(3):
            rows_tail[r].append(h.rho(r)[-chunk:])
        st = h.shift_and_stack()
        assert len(st) == chunk
        stack_tail.append(st)
        assert np.array_equal(h.get_signal(0, chunk), sigs[0][p * chunk:(p + 1) * chunk].astype(np.float32))
    # the oracle over the whole signal (rows of the common length total - nb + 1)
    nout = total - nb + 1
    full = np.stack([O.pearson(tpls[r], sigs[r].astype(np.float32).astype(np.float64))[:nout] for r in range(3)])
    ref_stack = O.shift_and_stack(full, shifts, weights, absval=True)
    for r in range(3):
        got = np.concatenate(rows_tail[r])        # lag j of push p: window start p*chunk - nb + 1 + j
        starts = np.arange(total) - nb + 1
        ok = (starts >= 0) & (starts < nout)
        assert np.abs(got[ok] - full[r][starts[ok]]).max() < 1e-5
    got = np.concatenate(stack_tail)              # stack k of push p: start p*chunk - nb + 1 - smax + k
    starts = np.arange(total) - nb + 1 - smax
    ok = (starts >= 0) & (starts < len(ref_stack))
    assert ok.sum() > total - 2 * nb
    assert np.abs(got[ok] - ref_stack[starts[ok]]).max() < 1e-5
    h.close()


def test_single_channel_stream_chunk_shorter_than_carry_fp64():
    """FP64, uneven templates, chunks shorter than the carry (the carry
    moves through the scratch buffer), pushed from device memory."""
    rng = np.random.default_rng(43)
    lens = [400, 333, 120]
    tpls = [rng.standard_normal(L) for L in lens]
    chunk, pushes = 300, 12
    total = chunk * pushes
    x = _signal(rng, total, 7.0)
    h = capi.Handle(np.float64, "single")
    h.set_templates(tpls)
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    nb = max(lens)
    assert h.stream_carry() == nb - 1 > chunk
    tails = [[] for _ in tpls]
    for p in range(pushes):
        d = DeviceArray(x[p * chunk:(p + 1) * chunk].copy())
        h.set_signal_device(0, d.ptr, chunk)
        d.free()
        h.apply()
        for r in range(len(tpls)):
            tails[r].append(h.rho(r)[-chunk:])
    nout = total - nb + 1
    starts = np.arange(total) - nb + 1
    ok = (starts >= 0) & (starts < nout)
    for r, t in enumerate(tpls):
        ref = O.pearson(t, x)[:nout]
        got = np.concatenate(tails[r])
        assert np.abs(got[ok] - ref[starts[ok]]).max() < 1e-11
    h.close()


def test_stream_restart_and_errors():
    rng = np.random.default_rng(47)
    tpls = [rng.standard_normal(200)]
    chunk = 5000
    a, b = rng.standard_normal(chunk), rng.standard_normal(chunk)

    def fresh():
        h = capi.Handle(np.float32, "multi")
        h.set_templates(tpls)
        h.set_signal_size(chunk)
        h.set_streaming(True)
        h.initialize()
        return h

    h1, h2 = fresh(), fresh()
    h1.set_signal(0, a.astype(np.float32))
    h1.apply()
    h1.zero_signal(0)                             # restart: the carry is zeros again
    h1.set_signal(0, b.astype(np.float32))
    h1.apply()
    h2.set_signal(0, b.astype(np.float32))
    h2.apply()
    assert np.array_equal(h1.rho(0), h2.rho(0))
    # dead-trace rule over carry + chunk, evaluated on the device at apply
    h1.zero_signal(0)
    h1.set_signal(0, np.full(chunk, 0.0, np.float32))
    h1.apply()
    assert not np.any(h1.rho(0))                  # |min| == |max|: every row of the channel is 0
    h1.set_signal(0, b.astype(np.float32))        # live again: the carry of zeros no longer makes it dead
    h1.apply()
    assert np.array_equal(h1.rho(0), h2.rho(0))
    with pytest.raises(ValueError):
        h1.set_signal(0, np.zeros(chunk + 1, np.float32))
    with pytest.raises(RuntimeError):
        h1.set_streaming(False)                   # only before initialize
    h1.close()
    h2.close()


@pytest.mark.parametrize("variant", [0, 3])
def test_batched_stream_with_graph_launch_and_mfma_rows(variant):
    """Streaming in the batched layout (two channels, two stacking groups)
    with graph launch on; under variant 3 the short templates (<= 256 taps)
    take the time-domain MFMA path.  Row tails against the FP64 oracle at
    1e-5."""
    rng = np.random.default_rng(53 + variant)
    lens = [200, 150, 256, 200]
    tpls = [rng.standard_normal(L) for L in lens]
    row_ch, row_grp, shifts = [0, 0, 1, 1], [0, 0, 1, 1], [0, 11, 0, 5]
    chunk, pushes = 8000, 5
    total = chunk * pushes
    sigs = [_signal(rng, total, 3.0 + 20 * c) for c in range(2)]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=shifts, weights=[1.0] * 4)
    h.set_batch_layout([chunk, chunk], row_ch, row_grp, 2)
    h.set_kernel_variant(variant)
    h.set_graph_launch(True)
    h.set_streaming(True)
    h.initialize()
    carry = h.stream_carry()
    assert carry == max(lens) - 1 + max(shifts)
    tails = [[] for _ in tpls]
    for p in range(pushes):
        for c in range(2):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk].astype(np.float32))
        h.apply()
        for r in range(4):
            tails[r].append(h.rho(r)[-chunk:])
    for r in range(4):
        nb = max(lens[k] for k in range(4) if row_grp[k] == row_grp[r])
        ref = O.pearson(tpls[r], sigs[row_ch[r]].astype(np.float32).astype(np.float64))
        got = np.concatenate(tails[r])
        starts = np.arange(total) - nb + 1
        ok = (starts >= 0) & (starts < total - nb + 1)
        assert np.abs(got[ok] - ref[starts[ok]]).max() < 1e-5, r
    h.close()


@pytest.mark.parametrize("dtype,chunk", [(np.float32, 4000), (np.float64, 150)])
def test_batched_push_equals_per_channel_pushes(dtype, chunk):
    """mf_push_chunks_device (every channel in one call: the carries saved,
    then every buffer rewritten, one wait) gives the rows and stacks of one
    mf_set_signal_device per channel, bit for bit; with a chunk of 150
    samples under a carry of 248 the saved carry overlaps the new one."""
    rng = np.random.default_rng(59)
    tpls = [rng.standard_normal(L) for L in (240, 180, 240)]
    pushes = 4 if chunk > 1000 else 9
    sigs = [_signal(rng, chunk * pushes, 5.0 * c) for c in range(3)]

    def fresh():
        h = capi.Handle(dtype, "multi")
        h.set_templates(tpls, shifts=[0, 4, 9], weights=[1.0, 1.0, 1.0])
        h.set_signal_size(chunk)
        h.set_streaming(True)
        h.initialize()
        return h

    a, b = fresh(), fresh()
    for p in range(pushes):
        chunks = [DeviceArray(sigs[c][p * chunk:(p + 1) * chunk].astype(dtype)) for c in range(3)]
        for c in range(3):
            a.set_signal_device(c, chunks[c].ptr, chunk)
        b.push_chunks_device([d.ptr for d in chunks], chunk)
        for d in chunks:
            d.free()
        a.apply()
        b.apply()
        for r in range(3):
            assert np.array_equal(a.rho(r), b.rho(r)), (p, r)
        assert np.array_equal(a.shift_and_stack(), b.shift_and_stack()), p
    with pytest.raises(ValueError):
        b.push_chunks_device([0, 0], chunk)            # not every channel
    a.close()
    b.close()
    h = capi.Handle(dtype, "multi")
    h.set_templates(tpls)
    h.set_signal_size(4000)
    h.initialize()
    with pytest.raises(RuntimeError):
        h.push_chunks_device([0, 0, 0], 4000)          # not a streaming handle
    h.close()


def test_stream_checkpoint_and_resume():
    """The carry of every channel, saved after three pushes and restored into
    a fresh handle, continues the stream: the next pushes' rows and stacks
    are bit-identical to the original handle's."""
    rng = np.random.default_rng(71)
    tpls = [rng.standard_normal(L) for L in (210, 150, 210)]
    chunk = 3000
    sigs = [_signal(rng, chunk * 5, 4.0 * c) for c in range(3)]

    def fresh():
        h = capi.Handle(np.float32, "multi")
        h.set_templates(tpls, shifts=[0, 6, 13], weights=[1.0, 1.0, 1.0])
        h.set_signal_size(chunk)
        h.set_streaming(True)
        h.initialize()
        return h

    def push(h, p):
        for c in range(3):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk].astype(np.float32))
        h.apply()
        return [h.rho(r) for r in range(3)], h.shift_and_stack()

    a = fresh()
    for p in range(3):
        push(a, p)
    states = [a.stream_state(c) for c in range(3)]
    assert all(len(st) == a.stream_carry() == 209 + 13 for st in states)
    b = fresh()
    for c in range(3):
        b.set_stream_state(c, states[c])
    for p in (3, 4):
        rows_a, st_a = push(a, p)
        rows_b, st_b = push(b, p)
        for r in range(3):
            assert np.array_equal(rows_a[r], rows_b[r]), (p, r)
        assert np.array_equal(st_a, st_b), p
    with pytest.raises(ValueError):
        b.set_stream_state(0, states[0][:-1])
    a.close()
    b.close()



def _stream_stack_and_peaks(h, groups, sigs, chunk, pushes, thr, md):
    """Pushes the channels chunk by chunk; per group, the stream's stack
    (lags K >= 0) and the peaks mf_stream_peaks returned over the pushes."""
    carry = h.stream_carry()
    tails = {g: [] for g in groups}
    got = {g: [] for g in groups}
    lead = {}
    for p in range(pushes):
        for c, s in enumerate(sigs):
            h.set_signal(c, s[p * chunk:(p + 1) * chunk])
        h.apply()
        for g in groups:
            st = h.shift_and_stack(g)
            lead[g] = carry - (len(st) - chunk)          # the group's nb - 1 + max shift
            tails[g].append(st[-chunk:])
            got[g].append(h.stream_peaks(g, thr, md, last=(p == pushes - 1)))
    out = {}
    for g in groups:
        stack = np.concatenate(tails[g])[lead[g]:]        # lag K = p * chunk - lead + k, kept from K = 0
        out[g] = (stack, np.concatenate(got[g]))
    return out


@pytest.mark.parametrize("dtype,thr,md", [(np.float32, 0.25, 150), (np.float32, 0.2, 2500),
                                          (np.float32, 0.5, 0), (np.float64, 0.3, 400)])
def test_stream_peaks_equal_whole_signal_peak_finder(dtype, thr, md):
    """mf_stream_peaks over every push equals the oracle's PeakFinder
    (src/peakFinder.cpp:468-586) on the stream's whole stack, index for
    index, in increasing order; a minDist longer than a push makes fates
    wait several pushes."""
    rng = np.random.default_rng(83)
    lens, shifts = [240, 180, 240], [0, 9, 23]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    chunk, pushes = 1500, 8
    sigs = [_signal(rng, chunk * pushes, 3.0 * c) for c in range(3)]
    for c, t in enumerate(tpls):
        for pos in (1400, 5990, 8950, 11450):               # events across push boundaries
            sigs[c][pos + shifts[c]:pos + shifts[c] + len(t)] += 2.5 * t
    sigs = [s.astype(dtype) for s in sigs]
    h = capi.Handle(dtype, "multi")
    h.set_templates(tpls, shifts=shifts, weights=[1.0, 0.8, 0.6])
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    stack, got = _stream_stack_and_peaks(h, [0], sigs, chunk, pushes, thr, md)[0]
    assert len(stack) == chunk * pushes - max(lens) + 1 - max(shifts)
    ref = O.peak_finder(stack, thr, md)
    assert len(ref) >= 3                                 # the injected events at least
    assert list(got) == list(ref)
    assert np.all(np.diff(got) > 0)
    h.close()


def test_stream_peaks_batched_groups_and_errors():
    """Two stacking groups of different lengths and moveouts (so each
    group's stream starts at its own lag of the stack), then the call
    conventions."""
    rng = np.random.default_rng(89)
    lens = [200, 150, 256, 200]
    tpls = [rng.standard_normal(L) for L in lens]
    row_ch, row_grp, shifts = [0, 0, 1, 1], [0, 0, 1, 1], [0, 11, 0, 5]
    chunk, pushes = 2500, 6
    sigs = [_signal(rng, chunk * pushes, 2.0 + 10 * c).astype(np.float32) for c in range(2)]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=shifts, weights=[1.0] * 4)
    h.set_batch_layout([chunk, chunk], row_ch, row_grp, 2)
    h.set_streaming(True)
    h.initialize()
    res = _stream_stack_and_peaks(h, [0, 1], sigs, chunk, pushes, 0.12, 120)
    for g in (0, 1):
        stack, got = res[g]
        ref = O.peak_finder(stack, 0.12, 120)
        assert len(ref) > 3, g
        assert list(got) == list(ref), g
    # a new stream after the last push; one call per apply; fixed parameters
    h.set_signal(0, sigs[0][:chunk])
    h.set_signal(1, sigs[1][:chunk])
    h.apply()
    h.stream_peaks(0, 0.12, 120)
    with pytest.raises(RuntimeError):
        h.stream_peaks(0, 0.12, 120)                   # twice after one apply
    h.apply()
    with pytest.raises(ValueError):
        h.stream_peaks(0, 0.3, 120)                    # another threshold mid-stream
    h.stream_peaks(0, 0.12, 120, last=True)
    h.close()
    m = capi.Handle(np.float32, "multi")
    m.set_templates(tpls[:2], shifts=[0, 3], weights=[1.0, 1.0])
    m.set_signal_size(chunk)
    m.initialize()
    m.set_signal(0, sigs[0][:chunk])
    m.set_signal(1, sigs[1][:chunk])
    m.apply()
    with pytest.raises(RuntimeError):
        m.stream_peaks(0, 0.5, 10)                     # not a streaming handle
    m.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stream_peaks_all_groups_equals_per_group_calls(dtype):
    """mf_stream_peaks_all (every group, one scan and one wait) returns, push
    by push, the lists of one mf_stream_peaks call per group."""
    rng = np.random.default_rng(97)
    lens = [200, 150, 256, 200, 120, 180]
    tpls = [rng.standard_normal(L) for L in lens]
    row_ch, row_grp, shifts = [0, 1, 0, 1, 0, 1], [0, 0, 1, 1, 2, 2], [0, 11, 0, 5, 3, 40]
    chunk, pushes = 1800, 7

    def fresh():
        h = capi.Handle(dtype, "batched")
        h.set_templates(tpls, shifts=shifts, weights=[1.0, 0.5, 1.0, 1.0, 0.7, 0.7])
        h.set_batch_layout([chunk, chunk], row_ch, row_grp, 3)
        h.set_streaming(True)
        h.initialize()
        return h

    sigs = [_signal(rng, chunk * pushes, 1.0 + 6 * c).astype(dtype) for c in range(2)]
    a, b = fresh(), fresh()
    seen = 0
    for p in range(pushes):
        last = p == pushes - 1
        for h in (a, b):
            for c in range(2):
                h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk])
            h.apply()
        per = [a.stream_peaks(g, 0.1, 90, last=last) for g in range(3)]
        allg = b.stream_peaks_all(0.1, 90, last=last)
        for g in range(3):
            assert np.array_equal(per[g], allg[g]), (p, g)
            seen += len(per[g])
    assert seen > 10
    a.close()
    b.close()


def test_stream_peaks_all_with_chunks_shorter_than_the_carry():
    """Chunks of 100 samples under a carry of 269: the first pushes hold no
    lag of the stream yet (their segments are empty) and every later peak's
    pruning window spans several pushes.  The lists equal the oracle's
    PeakFinder over the stream's whole stack."""
    rng = np.random.default_rng(101)
    lens, shifts = [240, 180], [0, 30]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    chunk, pushes = 100, 40
    sigs = [_signal(rng, chunk * pushes, 2.0 * c) for c in range(2)]
    for c, t in enumerate(tpls):
        for pos in (700, 2150, 3300):
            sigs[c][pos + shifts[c]:pos + shifts[c] + len(t)] += 3 * t
    sigs = [s.astype(np.float32) for s in sigs]
    h = capi.Handle(np.float32, "multi")
    h.set_templates(tpls, shifts=shifts, weights=[1.0, 1.0])
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    carry = h.stream_carry()
    assert carry == 239 + 30 > chunk
    tails, got = [], []
    for p in range(pushes):
        for c in range(2):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk])
        h.apply()
        got.append(h.stream_peaks_all(0.15, 250, last=(p == pushes - 1))[0])
        tails.append(h.shift_and_stack(0)[-chunk:])
    stack = np.concatenate(tails)[carry:]
    ref = O.peak_finder(stack, 0.15, 250)
    assert len(ref) >= 3
    assert list(np.concatenate(got)) == list(ref)
    h.close()


def test_stream_peaks_retry_after_a_full_output_keeps_the_state():
    """Threshold 0 and no minimum distance: about a third of all lags are
    peaks, more than the first output buffer holds, so every call fails once
    with the count and is repeated with room (Handle.stream_peaks does
    this).  The failed calls leave the state unchanged: the lists still
    equal PeakFinder over the whole stream stack."""
    rng = np.random.default_rng(103)
    tpls = [rng.standard_normal(L) for L in (150, 120)]
    chunk, pushes = 5000, 4
    sigs = [_signal(rng, chunk * pushes, 1.0).astype(np.float32) for _ in range(2)]
    h = capi.Handle(np.float32, "multi")
    h.set_templates(tpls, shifts=[0, 2], weights=[1.0, 1.0])
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    carry = h.stream_carry()
    tails, got = [], []
    for p in range(pushes):
        for c in range(2):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk])
        h.apply()
        pk = h.stream_peaks(0, 0.0, 0, last=(p == pushes - 1))
        assert len(pk) > 1024
        got.append(pk)
        tails.append(h.shift_and_stack(0)[-chunk:])
    stack = np.concatenate(tails)[carry:]
    assert list(np.concatenate(got)) == list(O.peak_finder(stack, 0.0, 0))
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stream_peaks_checkpoint_and_resume(dtype):
    """A checkpoint after push 3 of 8 (every channel's carry, then every
    group's stream-peak state) restored into a fresh handle: the restored
    handle's peak lists for pushes 4..8 equal the original's, and the
    original's lists over all pushes equal the oracle's PeakFinder on the
    stream's whole stack.  A minDist longer than a push keeps candidates
    pending across the checkpoint."""
    rng = np.random.default_rng(107)
    lens, shifts = [200, 150, 256, 200], [0, 11, 0, 5]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    chunk, pushes, cut = 1500, 8, 3
    sigs = [_signal(rng, chunk * pushes, 2.0 + 7 * c) for c in range(2)]
    for c in range(2):
        for pos in (1300, 4450, 4600, 7400, 10100):          # events around the checkpoint
            sigs[c][pos:pos + 256] += 2.0 * np.hanning(256) * rng.standard_normal(256)
    sigs = [s.astype(dtype) for s in sigs]

    def fresh():
        h = capi.Handle(dtype, "batched")
        h.set_templates(tpls, shifts=shifts, weights=[1.0, 0.5, 1.0, 1.0])
        h.set_batch_layout([chunk, chunk], [0, 1, 0, 1], [0, 0, 1, 1], 2)
        h.set_streaming(True)
        h.initialize()
        return h

    def push(h, p):
        for c in range(2):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk])
        h.apply()
        return h.stream_peaks_all(0.1, 1700, last=(p == pushes - 1))

    a = fresh()
    carry = a.stream_carry()
    got_a = [[] for _ in range(2)]
    tails = [[] for _ in range(2)]
    leads = [0, 0]
    for p in range(pushes):
        for g, pk in enumerate(push(a, p)):
            got_a[g].append(pk)
            st = a.shift_and_stack(g)
            leads[g] = carry - (len(st) - chunk)
            tails[g].append(st[-chunk:])
        if p == cut - 1:
            carries = [a.stream_state(c) for c in range(2)]
            peak_states = [a.stream_peaks_state(g) for g in range(2)]
    b = fresh()
    for c in range(2):
        b.set_stream_state(c, carries[c])
    for g in range(2):
        b.set_stream_peaks_state(g, peak_states[g])
    for p in range(cut, pushes):
        for g, pk in enumerate(push(b, p)):
            assert np.array_equal(pk, got_a[g][p]), (p, g)
    for g in range(2):
        stack = np.concatenate(tails[g])[leads[g]:]
        ref = O.peak_finder(stack, 0.1, 1700)
        assert len(ref) >= 3
        assert list(np.concatenate(got_a[g])) == list(ref), g
    # bytes that are not a stream-peak state change nothing
    with pytest.raises(ValueError):
        b.set_stream_peaks_state(0, peak_states[0][:-3])
    with pytest.raises(ValueError):
        b.set_stream_peaks_state(0, b"\0" * len(peak_states[0]))
    a.close()
    b.close()


def test_stream_peaks_reset_after_a_missed_call():
    """A push whose mf_stream_peaks call was missed leaves the group's stream
    stuck (the next call fails with BAD_STATE, last=1 included);
    mf_stream_peaks_reset drops the stream and the next call starts a new
    one, whose peaks equal PeakFinder over the stack from that push on.  A
    carry restore (mf_set_stream_state) also ends every group's stream."""
    rng = np.random.default_rng(109)
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in (220, 180)]
    chunk = 2000
    sigs = [_signal(rng, chunk * 9, 1.0 + 5 * c).astype(np.float32) for c in range(2)]
    for c in range(2):
        for pos in (9100, 12500, 15800):
            sigs[c][pos:pos + len(tpls[c])] += 3.0 * tpls[c]
    h = capi.Handle(np.float32, "multi")
    h.set_templates(tpls, shifts=[0, 4], weights=[1.0, 1.0])
    h.set_signal_size(chunk)
    h.set_streaming(True)
    h.initialize()
    carry = h.stream_carry()

    def push(p):
        for c in range(2):
            h.set_signal(c, sigs[c][p * chunk:(p + 1) * chunk])
        h.apply()

    push(0)
    h.stream_peaks(0, 0.2, 300)
    push(1)                                            # no stream_peaks call for this push
    push(2)
    with pytest.raises(RuntimeError):
        h.stream_peaks(0, 0.2, 300)
    with pytest.raises(RuntimeError):
        h.stream_peaks(0, 0.2, 300, last=True)
    h.stream_peaks_reset(0)
    got, tails = [], []
    for p in range(3, 9):
        push(p)
        got.append(h.stream_peaks(0, 0.2, 300, last=(p == 8)))
        tails.append(h.shift_and_stack(0)[-chunk:])
    stack = np.concatenate(tails)[carry:]               # lag 0 of the new stream: push 3's first window
    ref = O.peak_finder(stack, 0.2, 300)
    assert len(ref) >= 2
    assert list(np.concatenate(got)) == list(ref)
    # a carry restore mid-stream ends the group's stream: the next call opens a new one
    push(0)
    h.stream_peaks(0, 0.2, 300)
    state = h.stream_state(0)
    h.set_stream_state(0, state)
    push(1)
    push(2)
    h.stream_peaks(0, 0.2, 300)                        # no BAD_STATE: a new stream
    h.stream_peaks_reset(-1)
    h.close()


def test_batched_push_from_host_memory_equals_device_push():
    """mf_push_chunks_device with host pointers goes through the copy engine
    and gives the same rows as the two-kernel push of device chunks."""
    rng = np.random.default_rng(113)
    tpls = [rng.standard_normal(L) for L in (240, 180)]
    chunk = 3000
    sigs = [_signal(rng, chunk * 3, 3.0 * c).astype(np.float32) for c in range(2)]

    def fresh():
        h = capi.Handle(np.float32, "multi")
        h.set_templates(tpls, shifts=[0, 5], weights=[1.0, 1.0])
        h.set_signal_size(chunk)
        h.set_streaming(True)
        h.initialize()
        return h

    a, b = fresh(), fresh()
    for p in range(3):
        host = [np.ascontiguousarray(sigs[c][p * chunk:(p + 1) * chunk]) for c in range(2)]
        dev = [DeviceArray(x) for x in host]
        a.push_chunks_device([d.ptr for d in dev], chunk)
        b.push_chunks_device([x.ctypes.data for x in host], chunk)
        for d in dev:
            d.free()
        a.apply()
        b.apply()
        for r in range(2):
            assert np.array_equal(a.rho(r), b.rho(r)), (p, r)
    a.close()
    b.close()
