"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function include/mflib_gpu.h declares (no compute calls here)."""
import ctypes
import os

from mflib_amd import capi


def test_library_exports_every_declared_symbol():
    syms = capi.header_symbols()
    assert len(syms) >= 30
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    # the ctypes table covers the whole header
    assert sorted(capi._SIGNATURES) == syms


def test_sizing_kats_through_the_c_abi():
    """testing/matchedFilterParameters.cpp:19-31, 82-89: host-only entry point."""
    assert capi.compute_optimal_fft_and_block_length(4500, 75000) == (32768, 28269)
    assert capi.compute_optimal_fft_and_block_length(4500, 8640000) == (65536, 61037)
    assert capi.compute_optimal_fft_and_block_length(101, 1024) == (512, 412)


def test_no_gpu_means_loud_failure():
    """Without a device the engine refuses to create a handle (no CPU fallback)."""
    import pytest
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    import numpy as np
    with pytest.raises(RuntimeError):
        capi.Handle(np.float32, "multi")


def test_server_needs_a_device():
    """The request server has no CPU fallback either."""
    import pytest
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    import numpy as np
    with pytest.raises(RuntimeError):
        capi.Server(np.float32, budget_bytes=1 << 26)
