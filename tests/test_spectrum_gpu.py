"""getSpectraOfTemplate (src/matchedFilter.cpp:1411-1431) through the C ABI.

The reference returns row `it` of mB: the r2c transform of the demeaned,
unit-norm, time-reversed template zero-padded to nfft
(src/matchedFilter.cpp:556-620).  The engine keeps conj(T^)/(2 nfft) of the
unreversed template and rebuilds the export with a phase rotation and a
rescale; this checks that rebuild against the oracle's own FFT of the
oracle's reversed template at the engine's nfft.  A wrong sign, offset or
scale fails it: the phase of bin k carries the template length, so a
non-power-of-two length and an off-by-one in Lt both show."""
import numpy as np
import pytest

import oracle as O
from mflib_amd import capi

pytestmark = pytest.mark.gpu

LENS = [300, 257, 211, 128, 1000]


def _expected(t, nfft):
    b = O.demean_normalize_reverse(t)
    pad = np.zeros(nfft)
    pad[:len(b)] = b
    return O.rfft(pad)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_exported_spectrum_matches_oracle_rfft(dtype, tol):
    rng = np.random.default_rng(31)
    n = 12_000
    tpls = [rng.standard_normal(L) * np.hanning(L) + 0.3 for L in LENS]   # nonzero mean: demean matters
    h = capi.Handle(dtype, "single")
    h.set_templates(tpls)
    h.set_signal_size(n)
    h.initialize()
    for r, t in enumerate(tpls):
        got = h.template_spectrum(r)
        nfft = 2 * (len(got) - 1)
        assert nfft == h.plan(r)[0]
        want = _expected(t, nfft)
        scale = np.abs(want).max()
        err = np.abs(got.astype(np.complex128) - want).max() / scale
        assert err < tol, (r, len(t), nfft, err)
        # the DC bin of a demeaned template is zero; the energy is Parseval's
        assert abs(got[0]) < 10 * tol * scale
        energy = (2 * np.sum(np.abs(want[1:-1]) ** 2) + abs(want[0]) ** 2 + abs(want[-1]) ** 2) / nfft
        assert abs(energy - 1.0) < 1e-10
    h.close()


def test_exported_spectrum_of_time_domain_rows():
    """Rows on the MFMA time-domain path have no FFT plan (nfft 0); their
    export is still the reference's spectrum, at the FFT length the engine
    reports through the length of the returned array."""
    rng = np.random.default_rng(32)
    n = 20_000
    lens = [100, 129, 256, 300]    # the first three run in the time domain
    tpls = [rng.standard_normal(L) * np.hanning(L) for L in lens]
    h = capi.Handle(np.float32, "batched")
    h.set_templates(tpls, shifts=[0] * len(tpls), weights=[1.0] * len(tpls))
    h.set_batch_layout([n], [0] * len(tpls), [0] * len(tpls), 1)
    h.set_kernel_variant(3)
    h.initialize()
    assert h.plan(0)[0] == 0 and h.plan(3)[0] > 0
    for r, t in enumerate(tpls):
        got = h.template_spectrum(r)
        nfft = 2 * (len(got) - 1)
        assert nfft >= 2 * len(t)
        want = _expected(t, nfft)
        err = np.abs(got.astype(np.complex128) - want).max() / np.abs(want).max()
        assert err < 2e-6, (r, len(t), nfft, err)
    h.close()


def test_spectrum_errors_follow_reference():
    """Uninitialized: runtime_error (1419-1421); bad index: invalid_argument (1423-1428)."""
    h = capi.Handle(np.float64, "single")
    h.set_templates([np.hanning(64)])
    h.set_signal_size(1000)
    with pytest.raises(RuntimeError):
        h.template_spectrum(0)
    h.initialize()
    with pytest.raises(ValueError):
        h.template_spectrum(1)
    h.close()
