"""The CPU oracle, the CPU baseline and the argument checks of validation.h built with
AddressSanitizer and UndefinedBehaviorSanitizer (host code only, gcc/g++) and run over small
inputs of every element type, channel count and edge shape (tests/sanitize_driver.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]
CFLAGS = ["-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=gnu11"]


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                    reason="needs gcc and g++")
def test_oracle_baseline_and_checks_under_sanitizers(tmp_path):
    o = str(tmp_path)
    base = os.path.join(ROOT, "oracle", "baseline.c")
    cmds = [
        ["gcc", *SAN, *CFLAGS, "-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", f"{o}/oracle.o"],
        ["gcc", *SAN, *CFLAGS, "-DBASE_ISA=avx512", "-mavx512f", "-mavx512vl", "-mfma", "-c", base,
         "-o", f"{o}/b512.o"],
        ["gcc", *SAN, *CFLAGS, "-DBASE_ISA=avx2", "-mavx2", "-mfma", "-c", base, "-o", f"{o}/b2.o"],
        ["gcc", *SAN, *CFLAGS, "-DBASE_ISA=generic", "-c", base, "-o", f"{o}/bg.o"],
        ["gcc", *SAN, *CFLAGS, "-c", base, "-o", f"{o}/bmain.o"],
        ["g++", *SAN, "-std=c++17", "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "openclipp_amd", "csrc"),
         os.path.join(ROOT, "tests", "sanitize_driver.cpp"), f"{o}/oracle.o", f"{o}/b512.o",
         f"{o}/b2.o", f"{o}/bg.o", f"{o}/bmain.o", "-lm", "-o", f"{o}/driver"],
    ]
    for c in cmds:
        subprocess.run(c, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([f"{o}/driver"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert "sanitized run: ok" in r.stdout
