"""The engine, through the C ABI, against the committed golden vectors
(tests/golden/golden_vectors.npz; the CPU suite pins them to the oracle and
to closed forms):
  - correlation rows at the engine's contract, 1e-11 for FP64 instances and
    1e-5 for FP32 (inputs rounded to FP32 first, as the fixture's
    multi_rho_f32in was computed);
  - stacks (abs and raw) to the same bars, the device PeakFinder on the
    frozen stack and on the reference's Gaussian-sine train index for index;
  - the Detector reduce, and the per-detection products at the frozen
    peaks: quadratic shifts to 1e-9 dt (the GPU's own rows differ from the
    oracle's by ~1e-13), Lanczos picks at a grid point within 1e-9 of the
    frozen profile's maximum, amplitude factors to 1e-11 relative.
"""
import os

import numpy as np
import pytest

from mflib_amd import capi

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_vectors.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
def test_golden_multichannel_rows_stacks_peaks(g, dtype, tol):
    tpls = [g[f"multi_template_{i}"] for i in range(3)]
    sigs = [g[f"multi_signal_{i}"].astype(dtype) for i in range(3)]
    ref = g["multi_rho_f64"] if dtype == np.float64 else g["multi_rho_f32in"]
    h = capi.Handle(dtype, "multi")
    h.set_templates(tpls, shifts=[int(s) for s in g["multi_shifts"]], weights=list(g["multi_weights"]))
    h.set_signal_size(len(sigs[0]))
    h.initialize()
    for c, s in enumerate(sigs):
        h.set_signal(c, s)
    h.apply()
    for i in range(3):
        r = h.rho(i)
        assert len(r) == ref.shape[1]
        assert np.abs(r - ref[i]).max() < tol, i
    st = h.shift_and_stack()
    assert len(st) == len(g["multi_stack_abs"])
    assert np.abs(st - g["multi_stack_abs"]).max() < tol
    assert list(h.detect(0.3, 200)[0]) == list(g["multi_stack_peaks"])
    h.set_stack_absolute_values(False)
    assert np.abs(h.shift_and_stack() - g["multi_stack_raw"]).max() < tol
    h.close()
    # the device PeakFinder on the frozen stack itself: index for index
    assert np.array_equal(capi.peaks(g["multi_stack_abs"], 0.3, 200), g["multi_stack_peaks"])


def test_golden_peak_finder_reference_lists(g):
    sig, thr, md = g["pf_signal"], float(g["pf_threshold"]), int(g["pf_min_dist"])
    for x in (sig, sig.astype(np.float32)):
        assert list(capi.peaks(x, thr, md)) == list(g["pf_expected_min_dist"])
        assert list(capi.peaks(x, thr, 0)) == list(g["pf_expected_no_min_dist"])


def test_golden_single_channel_rows_reduce_products(g):
    tpls = [g[f"single_template_{i}"] for i in range(3)]
    x = g["single_signal"]
    rho = g["single_rho_f64"]
    h = capi.Handle(np.float64, "single")
    h.set_templates(tpls)
    h.set_signal_size(len(x))
    h.initialize()
    h.set_signal(0, x)
    h.apply()
    for i in range(3):
        assert np.abs(h.rho(i) - rho[i]).max() < 1e-12, i
    for a in (0, 1):
        det, who = h.reduce(0, absval=bool(a))
        assert np.abs(det - g[f"single_reduce_det_{a}"]).max() < 1e-12
        v = np.sort(np.abs(rho) if a else rho, axis=0)
        clear = v[-1] - v[-2] > 1e-10                   # no near-tie between the two best rows
        assert np.array_equal(who[clear], g[f"single_reduce_id_{a}"][clear])
    peaks, tids = g["single_peaks"], g["single_peak_template"]
    dt = float(g["single_dt"])
    dts = [dt] * 3
    q = h.detection_products(peaks, tids, dts, interpolation=capi.INTERP_QUADRATIC)
    lz = h.detection_products(peaks, tids, dts, interpolation=capi.INTERP_LANCZOS)
    prof = g["single_lanczos_profile"]
    npts = prof.shape[1] + 1
    for k in range(len(peaks)):
        ctx = f"k={k} lag={peaks[k]} template={tids[k]}"
        assert q["rho"][k] == pytest.approx(min(1.0, rho[tids[k]][peaks[k]]), abs=1e-12), ctx
        assert q["shift"][k] == pytest.approx(g["single_quadratic_shift"][k], abs=1e-9 * dt), ctx
        j = int(round((lz["shift"][k] / dt + 0.5) * npts)) - 1
        assert 0 <= j < npts - 1, ctx
        assert prof[k][j] >= prof[k].max() - 1e-9, ctx
        assert bool(q["flags"][k] & capi.PRODUCT_AMPLITUDE) == bool(g["single_amplitude_ok"][k]), ctx
        assert q["alpha_gr"][k] == pytest.approx(g["single_alpha_gr"][k], rel=1e-11), ctx
        assert q["alpha_sr"][k] == pytest.approx(g["single_alpha_sr"][k], rel=1e-11), ctx
    h.close()
