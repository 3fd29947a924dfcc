"""The core kernel's emission quotient (correctly rounded reciprocal + one
fused residual correction) must equal IEEE single-precision division, which
is what the reference computes (GPU/align_cuda.cl:527).  Checked on the CPU
with glibc's correctly rounded fmaf."""
import os
import subprocess

import numpy as np

from conftest import GOLDEN, ROOT


def test_reciprocal_fma_quotient_equals_ieee_division(tmp_path):
    src = os.path.join(ROOT, "tests", "native", "division_check.c")
    exe = str(tmp_path / "division_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    model = tmp_path / "model.bin"
    np.load(GOLDEN)["model"].astype(np.float32).tofile(model)
    r = subprocess.run([exe, str(model)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout


def test_division_by_three_and_six_equals_ieee_division(tmp_path):
    """Event detection divides the window sums of tame reads by w = 3 and 6
    with a correctly rounded reciprocal and one fused correction: in double
    where the C source divides doubles, in float (or in double, rounded to
    float, for the variance over 6) where it divides floats; the check
    compares every route with IEEE division (tests/native/div3_check.c)."""
    src = os.path.join(ROOT, "tests", "native", "div3_check.c")
    exe = str(tmp_path / "div3_check")
    with open("/proc/cpuinfo") as f:
        fma_flag = [" -mfma"] if " fma " in f.read() else []
    flags = ["-O2", "-ffp-contract=off"] + [x.strip() for x in fma_flag]
    subprocess.run(["gcc"] + flags + ["-o", exe, src, "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout
