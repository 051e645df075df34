"""The host FFT engine behind mifft_cpu_* (CpuFftPlan restated in C++,
src/cpu.rs:52-866) built with g++ under AddressSanitizer and
UndefinedBehaviorSanitizer, host code only, and run over the shapes of
tests/cpu_sanitize/engine_driver.cpp: any out-of-bounds access, leak or
undefined operation in the engine ends the driver with a sanitizer report,
and its own checks (round trips, direct DFT of short lines, in place equal to
out of place) give a non-zero exit status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wgpufft_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "cpu_sanitize", "engine_driver.cpp")
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-I", CSRC]


def test_engine_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    units = [("cpu_fft_base.cpp", []), ("cpu_fft_avx2.cpp", ["-mavx2", "-mfma"])]
    jobs = [subprocess.Popen([gxx, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", str(tmp_path / (src + ".o"))],
                             stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for src, extra in units]
    for j in jobs:
        out, _ = j.communicate(timeout=600)
        assert j.returncode == 0, out[-4000:]
    exe = tmp_path / "engine_driver"
    r = subprocess.run([gxx, *FLAGS, DRIVER, *[str(tmp_path / (s + ".o")) for s, _ in units], "-o", str(exe),
                        "-lpthread"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-6000:]
    assert "fails=0" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
