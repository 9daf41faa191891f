"""The host-only C++ (the CPU oracle, the configuration reader, the output
writer) built with AddressSanitizer and UndefinedBehaviorSanitizer by g++
and run on a few steps of every case family (tests/sanitize_driver.cpp).
Any report is fatal (-fno-sanitize-recover): the run must end cleanly."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_and_ubsan(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    exe = tmp_path / "sanitize_driver"
    srcs = [os.path.join(ROOT, "tests", "sanitize_driver.cpp"),
            os.path.join(ROOT, "oracle", "hydro_oracle.cpp"),
            os.path.join(ROOT, "oracle", "hydro_oracle3d.cpp"),
            os.path.join(ROOT, "hydro_amd", "csrc", "config.cpp"),
            os.path.join(ROOT, "hydro_amd", "csrc", "output.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-o", str(exe)] + srcs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env,
                       cwd=str(tmp_path))
    assert r.returncode == 0 and "SANITIZE_OK" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
