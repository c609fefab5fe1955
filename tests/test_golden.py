"""The committed golden vectors (tests/golden/golden_vectors.npz, written by
tests/golden/make_golden.py) on the CPU:
  - the oracle still reproduces every frozen output (a change to the
    oracle's arithmetic shows up here before it can move a GPU parity bar);
  - the frozen outputs themselves obey what the mathematics fixes, so the
    fixture is not merely a copy of whatever the oracle printed: rho in
    [-1, 1], rho = 1 where a template was cut from the signal, rho = 0 in a
    constant stretch, the stack as the weighted mean of |rho| and its
    peaks at the injected events, the reference's PeakFinder lists
    (testing/peakFinder.cpp:16-59) and sizing known answers;
  - the host-only sizing entry point of the C ABI agrees with the fixture.
"""
import os

import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_vectors.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _multi(g):
    tpls = [g[f"multi_template_{i}"] for i in range(3)]
    sigs = [g[f"multi_signal_{i}"] for i in range(3)]
    return tpls, sigs


def _single(g):
    return [g[f"single_template_{i}"] for i in range(3)], g["single_signal"]


def test_oracle_reproduces_multichannel_vectors(g):
    tpls, sigs = _multi(g)
    nout = g["multi_rho_f64"].shape[1]
    for i in range(3):
        assert np.abs(O.pearson(tpls[i], sigs[i])[:nout] - g["multi_rho_f64"][i]).max() <= 1e-14
        x32 = sigs[i].astype(np.float32)
        assert np.abs(O.pearson(tpls[i], x32)[:nout] - g["multi_rho_f32in"][i]).max() <= 1e-14
    sh, w = g["multi_shifts"], g["multi_weights"]
    assert np.array_equal(O.shift_and_stack(g["multi_rho_f64"], sh, w, True), g["multi_stack_abs"])
    assert np.array_equal(O.shift_and_stack(g["multi_rho_f64"], sh, w, False), g["multi_stack_raw"])
    assert np.array_equal(O.peak_finder(g["multi_stack_abs"], 0.3, 200), g["multi_stack_peaks"])


def test_multichannel_vectors_obey_closed_forms(g):
    rho = g["multi_rho_f64"]
    assert np.all(np.abs(rho) <= 1 + 1e-12)
    # channel 1's constant stretch: every window inside it has rho 0
    Lt = len(g["multi_template_1"])
    assert not np.any(rho[1][2000:2600 - Lt + 1])
    # stack = weighted mean of |rho| (or rho) at the moveouts (src/matchedFilter.cpp:1341-1361)
    sh, w = g["multi_shifts"], g["multi_weights"] / g["multi_weights"].sum()
    n = len(g["multi_stack_abs"])
    assert n == rho.shape[1] - sh.max()
    ref = sum(w[c] * np.abs(rho[c][sh[c]:sh[c] + n]) for c in range(3))
    assert np.abs(ref - g["multi_stack_abs"]).max() < 1e-15
    # the three injected events are the stack's peaks
    assert list(g["multi_stack_peaks"]) == [1000, 3500, 5000]


def test_oracle_reproduces_single_channel_vectors(g):
    tpls, x = _single(g)
    rho = g["single_rho_f64"]
    nout = rho.shape[1]
    assert nout == len(x) - max(len(t) for t in tpls) + 1
    for i, t in enumerate(tpls):
        assert np.abs(O.pearson(t, x)[:nout] - rho[i]).max() <= 1e-14
    for a in (0, 1):
        det, who = O.reduce_correlations(rho, absval=bool(a))
        assert np.array_equal(det, g[f"single_reduce_det_{a}"])
        assert np.array_equal(who, g[f"single_reduce_id_{a}"])
    dt = float(g["single_dt"])
    for k, (p, i) in enumerate(zip(g["single_peaks"], g["single_peak_template"])):
        assert O.quadratic_shift(rho[i], p, dt) == g["single_quadratic_shift"][k]
        s, prof = O.lanczos_shift(rho[i], p, dt)
        assert s == g["single_lanczos_shift"][k]
        assert np.abs(prof - g["single_lanczos_profile"][k]).max() <= 1e-14
        f = O.amplitude_factors(tpls[i], x[p:p + len(tpls[i])])
        assert (f is not None) == bool(g["single_amplitude_ok"][k])
        assert f[0] == pytest.approx(g["single_alpha_gr"][k], rel=1e-14)
        assert f[1] == pytest.approx(g["single_alpha_sr"][k], rel=1e-14)


def test_single_channel_vectors_obey_closed_forms(g):
    tpls, x = _single(g)
    rho = g["single_rho_f64"]
    cuts = [p for p, i in zip(g["single_peaks"], g["single_peak_template"]) if rho[i][p] > 1 - 1e-12]
    assert cuts == [40, 1700, 3900]                        # each template cut from the signal: rho = 1
    for i, c in enumerate(cuts):
        k = list(g["single_peaks"]).index(c)
        assert g["single_alpha_gr"][k] == pytest.approx(1.0, rel=1e-12)   # a copy scales by 1
        assert g["single_alpha_sr"][k] == pytest.approx(1.0, rel=1e-12)
        assert abs(g["single_lanczos_shift"][k]) <= 0.5 * float(g["single_dt"])
    assert not np.any(rho[:, 2000:2600 - max(len(t) for t in tpls) + 1])   # constant stretch
    det, who = g["single_reduce_det_0"], g["single_reduce_id_0"]
    assert np.array_equal(det, np.clip(rho.max(axis=0), -1, 1))
    assert np.array_equal(who, rho.argmax(axis=0))        # argmax: the first row reaching the maximum
    det1, who1 = g["single_reduce_det_1"], g["single_reduce_id_1"]
    assert np.array_equal(det1, np.clip(np.abs(rho).max(axis=0), -1, 1))
    assert np.array_equal(who1, np.abs(rho).argmax(axis=0))


def test_peak_finder_vectors_are_the_reference_lists(g):
    sig, thr, md = g["pf_signal"], float(g["pf_threshold"]), int(g["pf_min_dist"])
    assert list(O.peak_finder(sig, thr, md)) == list(g["pf_expected_min_dist"])
    assert list(O.peak_finder(sig, thr, 0)) == list(g["pf_expected_no_min_dist"])
    assert len(g["pf_expected_min_dist"]) == 13 and len(g["pf_expected_no_min_dist"]) == 21


def test_sizing_vectors(g):
    for nb, nx, nfft, L in g["sizing"]:
        assert O.window_length(int(nb), int(nx)) == (nfft, L)
        assert capi.compute_optimal_fft_and_block_length(int(nb), int(nx)) == (nfft, L)
        assert L == nfft - nb + 1
    # the workloads' sizes quoted for the reference (one hour at nb 400, one day at nb 512)
    rows = {(int(r[0]), int(r[1])): (int(r[2]), int(r[3])) for r in g["sizing"]}
    assert rows[(400, 360399)] == (4096, 3697)
    assert rows[(512, 8640511)] == (8192, 7681)
