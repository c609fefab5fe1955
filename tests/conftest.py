import os
import sys

import pytest

# torch ships its own HIP runtime.  Loading it before the engine library
# (which the test modules load at collection) leaves one HIP runtime in the
# process, so tests that pass torch GPU tensors through DLPack can share the
# device with the engine.  Importing torch does not touch the GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
