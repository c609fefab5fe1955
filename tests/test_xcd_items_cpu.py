"""The order of a launch's work items over the 8 XCDs
(mfh::appendXcdItems in mflib_amd/csrc/mf_host.hpp) on the CPU: every block
of every lane is issued exactly once with its row count; a launch of many
lanes keeps each lane on one XCD (its template spectra in one L2); a launch
of few lanes gives every XCD about the same work.  The check is a C++
program over 300 generated lane sets (tests/cpp/xcd_items_check.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "xcd_items_check.cpp")


def test_xcd_item_order(tmp_path):
    exe = str(tmp_path / "xcd_items_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D_GLIBCXX_ASSERTIONS",
                    "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "mflib_amd", "csrc"), SRC, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
