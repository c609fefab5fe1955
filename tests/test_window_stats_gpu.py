"""mf_get_window_stats: windows computed, windows recomputed exactly, and
windows of solo rows, against what the data and the row grouping fix.

- A stationary signal (white noise on a DC offset) gives no flagged window
  on any path: the fast value's error bound holds everywhere.
- An exactly constant stretch of G samples flags at least the G - Lt + 1
  windows of each row that lie inside it (zero variance, recomputed in
  full), and nothing
  outside the items whose samples overlap it (the block anchor elsewhere
  is clean), on the FFT path and on both MFMA time-domain paths; the
  recomputed rows match the FP64 oracle (the flagged windows are where
  the recomputation runs).  On the same data the BF16 three-part path
  recomputes the same cancelled windows as the FP16 two-part path and has
  no more error-bound flags (its bound is the tighter one).
- Solo rows: rows whose template length fewer than four rows of the block
  share compute their scale alone (every window solo); eight rows of one
  length share the block's scale (no solo window) where the kernel keeps
  the shared buffer (FP32 nfft <= 2048).
- Counts add up over applies.
"""
import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu

TD_ITEM, TD_SPAN_MAX = 512, 512 + 256


def _handle(tpls, x, variant):
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([x.size], [0] * len(tpls), [0] * len(tpls), 1)
    h.set_kernel_variant(variant)
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    return h


def _stationary(rng, n):
    return (rng.standard_normal(n) * 3 + 1e4).astype(np.float32)


@pytest.mark.parametrize("variant", [0, 3, 4])
def test_stationary_signal_flags_nothing(variant):
    rng = np.random.default_rng(900 + variant)
    n = 60_000
    x = _stationary(rng, n)
    lens = [64, 128, 200, 256]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = _handle(tpls, x, variant)
    ws = h.window_stats()
    nout = n - max(lens) + 1
    assert ws["windows"] == len(tpls) * nout
    assert ws["flagged"] == 0 and ws["full"] == 0, (variant, ws)
    h.close()


@pytest.mark.parametrize("variant", [0, 3, 4])
def test_constant_stretch_flags_its_windows(variant):
    rng = np.random.default_rng(910)
    n = 60_000
    x = _stationary(rng, n)
    g0, G = 20_000, 2_500
    x[g0:g0 + G] = np.float32(1e4 + 17.0)
    lens = [64, 100, 128, 200, 256]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = _handle(tpls, x, variant)
    ws = h.window_stats()
    f = ws["flagged"]
    nout = n - max(lens) + 1
    assert ws["windows"] == len(tpls) * nout
    inside = sum(G - L + 1 for L in lens)
    # zero variance: recomputed in full
    assert ws["full"] >= inside and f >= ws["full"], (variant, ws, inside)
    if variant in (3, 4):
        # only items whose sample span overlaps the stretch see its anchor
        first = (g0 - TD_SPAN_MAX) // TD_ITEM * TD_ITEM
        last = (g0 + G) // TD_ITEM * TD_ITEM + TD_ITEM
        assert f <= len(tpls) * (last - first), (f, last - first)
    for r, t in enumerate(tpls):
        ref = O.pearson(t, x)[:nout]
        assert np.abs(h.rho(r).astype(np.float64) - ref).max() < 1e-5, (variant, r)
    h.close()


def test_bf16_three_part_flags_no_more_than_fp16():
    rng = np.random.default_rng(920)
    n = 80_000
    x = _stationary(rng, n)
    for s0, L in ((9_000, 700), (30_000, 3000), (61_000, 1500)):
        x[s0:s0 + L] = np.float32(1e4 - 5.0)
    x[50_000:52_000] = (1e4 + rng.standard_normal(2000) * 0.01).astype(np.float32)   # quiet stretch
    lens = [96, 160, 256, 256]
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    st = {}
    for variant in (3, 4):
        h = _handle(tpls, x, variant)
        st[variant] = h.window_stats()
        h.close()
    # the same cancellation test on both; the BF16 triple's error bound is
    # the tighter one, so its bound flags are a subset of the FP16 pair's
    assert st[4]["full"] == st[3]["full"] > 0, st
    bound = {v: st[v]["flagged"] - st[v]["full"] for v in st}
    assert bound[4] <= bound[3], st
    assert st[4]["flagged"] < 0.1 * st[4]["windows"], st


def test_solo_windows_follow_the_row_groups():
    rng = np.random.default_rng(930)
    n = 50_000
    x = _stationary(rng, n)
    # eight rows of one length: one shared scale per block
    tpls = [rng.standard_normal(300) for _ in range(8)]
    h = _handle(tpls, x, 2)
    nfft = h.plan(0)[0]
    ws = h.window_stats()
    w, s = ws["windows"], ws["solo"]
    assert w == 8 * (n - 300 + 1)
    assert s == (0 if nfft <= 2048 else w), (nfft, s, w)
    h.apply()
    ws2 = h.window_stats()
    assert (ws2["windows"], ws2["solo"]) == (2 * w, 2 * s)
    h.close()
    # three rows of one length: a group below four, every window solo
    h = _handle(tpls[:3], x, 2)
    ws = h.window_stats()
    assert ws["windows"] == 3 * (n - 300 + 1) and ws["solo"] == ws["windows"]
    h.close()


def test_window_stats_before_initialize_raises():
    h = capi.Handle(np.float32, "batched")
    h.set_templates([np.hanning(64)], shifts=[0], weights=[1.0])
    h.set_batch_layout([1000], [0], [0], 1)
    with pytest.raises(RuntimeError):
        h.window_stats()
    h.close()
