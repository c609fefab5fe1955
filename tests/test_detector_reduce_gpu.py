"""GPU Detector reduce step (mf_reduce_group / mf_detect_reduced) against
the oracle's restatement of reduceAllCorrelations + cleanReducedTemplate
(src/singleChannel/detector.cpp:190-250) on the same GPU correlation rows:
det and id are bit-identical (max and argmax are exact; ties keep the first
template), and the peaks equal the oracle PeakFinder's on the oracle det."""
import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu


def _single_channel(dtype, seed):
    rng = np.random.default_rng(seed)
    n = 40_000
    x = rng.standard_normal(n) * 0.3 + 5.0
    lens = [180, 180, 240, 131, 200, 240]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    for k, t in enumerate(tpls):                      # each template injected twice
        for pos in rng.integers(0, n - 300, 2):
            x[pos:pos + len(t)] += (2.0 + k) * t / np.linalg.norm(t)
    h = capi.Handle(dtype, "single")
    h.set_templates(tpls)
    h.set_signal_size(n)
    h.initialize()
    h.set_signal(0, x.astype(dtype))
    h.apply()
    return h, len(tpls)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("absval", [False, True])
def test_reduce_matches_oracle_bit_exact(dtype, absval):
    h, nt = _single_channel(dtype, 31 + absval)
    rows = np.stack([h.rho(k) for k in range(nt)])
    det, ids = h.reduce(0, absval)
    rdet, rids = O.reduce_correlations(rows, absval)
    assert det.dtype == rows.dtype and det.size == rows.shape[1]
    assert np.array_equal(det, rdet)
    assert np.array_equal(ids, rids)
    assert ids.min() >= 0 and ids.max() < nt
    pk, tid = h.detect_reduced(0.55, 150, 0, absval)
    ref_pk = O.peak_finder(rdet, 0.55, 150)
    assert np.array_equal(np.sort(pk), np.sort(ref_pk))
    assert pk.size > 0
    assert np.array_equal(tid, rids[pk])
    h.close()


def test_reduce_batched_groups_report_template_indices():
    """In the batched layout a group's rows are any subset of templates:
    the ids are the handle's template indices, not positions in the group."""
    rng = np.random.default_rng(5)
    n = 20_000
    x = [rng.standard_normal(n).astype(np.float32) for _ in range(2)]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in (100, 150, 120, 100, 90)]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * 5, weights=[1.0] * 5)
    # templates 0, 2, 4 on channel 0 (group 0); 1, 3 on channel 1 (group 1)
    h.set_batch_layout([n, n], [0, 1, 0, 1, 0], [0, 1, 0, 1, 0], 2)
    h.initialize()
    for c in range(2):
        h.set_signal(c, x[c])
    h.apply()
    for g, members in ((0, [0, 2, 4]), (1, [1, 3])):
        det, ids = h.reduce(g, True)
        rows = np.stack([h.rho(k)[:det.size] for k in members])
        rdet, rids = O.reduce_correlations(rows, True)
        assert np.array_equal(det, rdet)
        assert np.array_equal(ids, np.asarray(members)[rids])
    with pytest.raises(ValueError):
        h.reduce(2)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("absval", [False, True])
def test_detect_templates_batch_matches_per_template_and_oracle(dtype, absval):
    """mf_detect_templates (every template's SINGLE-policy peak list in one
    batch) equals mf_detect_template row by row, and the oracle PeakFinder on
    the clamped rows (setCorrelations, src/singleChannel/detector.cpp:252-284)."""
    h, nt = _single_channel(dtype, 61 + absval)
    lists = h.detect_templates(0.25, 30, absval=absval)
    assert len(lists) == nt
    for it in range(nt):
        one = h.detect_template(it, 0.25, 30, absval=absval)
        assert np.array_equal(lists[it], one), it
        row = h.rho(it)
        det, _ = O.reduce_correlations(row[None, :], absval)   # one row: clamp(rho) or clamp(|rho|)
        assert np.array_equal(lists[it], O.peak_finder(det, 0.25, 30)), it
    assert sum(len(x) for x in lists) > 0
    with pytest.raises(ValueError):
        h.detect_templates(1.5, 0)


def test_detect_templates_rows_beyond_the_candidate_room():
    """Threshold 0 on 400,000-lag rows of noise: ~130,000 candidates per row,
    more than the batch's per-row room (65,536), so every row is redone alone
    with full capacity; the lists still equal the per-template path and the
    oracle."""
    rng = np.random.default_rng(67)
    n = 400_000
    x = rng.standard_normal(n).astype(np.float32)
    tpls = [rng.standard_normal(L) for L in (64, 80)]
    h = capi.Handle(np.float32, "single")
    h.set_templates(tpls)
    h.set_signal_size(n)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    lists = h.detect_templates(0.0, 0)
    for it in range(len(tpls)):
        row = h.rho(it)
        det, _ = O.reduce_correlations(row[None, :], False)
        ref = O.peak_finder(det, 0.0, 0)
        assert len(ref) > 65536
        assert np.array_equal(lists[it], ref), it
        assert np.array_equal(lists[it], h.detect_template(it, 0.0, 0)), it
    h.close()
