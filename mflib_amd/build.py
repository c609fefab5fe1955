"""Builds the native libraries in-tree (no GPU needed; hipcc cross-compiles).

    python -m mflib_amd.build          # everything
Outputs (git-ignored, shipped with the tree):
    mflib_amd/lib/libmflib_gpu.so      HIP kernels + the C ABI (include/mflib_gpu.h)
    mflib_amd/lib/libmflib.so          C++ API (include/mflib/*.hpp) over the C ABI
    mflib_amd/pymflib*.so              pybind11 module `pymflib`
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def build_gpu_lib() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmflib_gpu.so")
    srcs = [os.path.join(CSRC, f) for f in ("mf_kernels.hip", "mf_tdcorr.hip", "mf_scan.hip", "mf_products.hip", "mf_engine.hip", "mf_server.hip")]
    objs = []
    for s in srcs:
        o = os.path.join(LIB, os.path.basename(s) + ".o")
        _run([hipcc(), *HIP_FLAGS, "-I", INC, "-I", CSRC, "-c", s, "-o", o])
        objs.append(o)
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out + ".tmp", *objs,
          "-Wl,-soname,libmflib_gpu.so"])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


FACADE_SRCS = ["waveformTemplate.cpp", "networkStationPhase.cpp", "matchedFilterParameters.cpp",
               "matchedFilter.cpp", "singleChannelMatchedFilterParameters.cpp",
               "singleChannelMatchedFilter.cpp", "peakFinder.cpp", "detection.cpp", "detector.cpp",
               "relativeMagnitude.cpp"]


def build_facade() -> str:
    out = os.path.join(LIB, "libmflib.so")
    srcs = [os.path.join(CSRC, "facade", f) for f in FACADE_SRCS]
    _run(["g++", *CXX_FLAGS, "-shared", "-I", INC, "-o", out + ".tmp", *srcs,
          "-L", LIB, "-lmflib_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libmflib.so"])
    os.replace(out + ".tmp", out)
    return out


def build_pymflib() -> str:
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "pymflib" + ext)
    pyinc = sysconfig.get_paths()["include"]
    src = os.path.join(CSRC, "pymflib.cpp")
    _run(["g++", *CXX_FLAGS, "-shared", "-fvisibility=hidden", "-I", INC, "-I", pyinc,
          "-I", pybind11.get_include(), "-o", out + ".tmp", src,
          "-L", LIB, "-lmflib", "-lmflib_gpu", "-Wl,-rpath,$ORIGIN/lib"])
    os.replace(out + ".tmp", out)
    return out


TESTS_CPP = os.path.join(ROOT, "tests", "cpp")


def build_cpp_tests() -> str:
    """The C++ API test program (tests/cpp/test_facade.cpp), linked against
    libmflib.so; written to tests/cpp/bin/ (git-ignored)."""
    outdir = os.path.join(TESTS_CPP, "bin")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "test_facade")
    _run(["g++", *CXX_FLAGS, "-I", INC, "-o", out + ".tmp", os.path.join(TESTS_CPP, "test_facade.cpp"),
          "-L", LIB, "-lmflib", "-lmflib_gpu",
          "-Wl,-rpath,$ORIGIN/../../../mflib_amd/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose: bool = True) -> list:
    built = [build_gpu_lib(), build_facade(), build_pymflib(), build_cpp_tests()]
    if verbose:
        for b in built:
            print("built", os.path.relpath(b, ROOT))
    return built


if __name__ == "__main__":
    build_all()
