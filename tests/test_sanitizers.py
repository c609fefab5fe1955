"""Host-side sanitizer run of the oracle: tests/cpp/oracle_sanitize.cpp built with g++
-fsanitize=address,undefined and run; any sanitizer finding aborts it.
CPU only: the oracle is the test suite's checker, so its memory safety
underwrites every parity claim."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "oracle"), os.path.join(HERE, "cpp", "oracle_sanitize.cpp"),
           os.path.join(ROOT, "oracle", "mf_oracle.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    res = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120, env=env)
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-4000:]
    assert "0 failure(s)" in res.stdout
