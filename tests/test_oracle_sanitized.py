"""The oracle's own checks under AddressSanitizer and UBSan.

The oracle is plain C driven through ctypes; an out-of-bounds band index or
an uninitialised cell would not crash it and could still pass a golden
comparison by luck.  This test builds the sanitizer flavour of the same
sources (oracle/Makefile, liborc_asan.so) and re-runs the golden, brute
force, closed-form and edge-case checks of the oracle tests against it in a
child interpreter, with libasan preloaded (Python itself is not built with
it).  Any sanitizer report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ASAN_LIB = os.path.join(ORACLE, "_build", "liborc_asan.so")
SUITES = ["tests/test_oracle.py", "tests/test_getevents_oracle.py",
          "tests/test_postalign_oracle.py"]

# the planted-mistake tests load their own mutant builds, and the fills test
# only repeats (slowly, under the sanitizer) walks the other tests make
DESELECT = "not planted and not fills_from_moves"

PROBE = r"""
import sys
sys.path.insert(0, sys.argv[1])
import oracle
L = oracle.lib()
maps = open("/proc/self/maps").read()
assert "liborc_asan.so" in maps and "libasan" in maps, "sanitizer build not loaded"
print("loaded", oracle.LIB_PATH)
"""


def _runtime(name):
    try:
        out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                             text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_oracle_checks_under_asan_ubsan():
    libasan = _runtime("libasan.so")
    if libasan is None:
        pytest.skip("gcc's libasan is not installed")
    subprocess.run(["make", "-s", "-C", ORACLE, "_build/liborc_asan.so"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=libasan, ORC_LIB=ASAN_LIB,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    probe = subprocess.run([sys.executable, "-c", PROBE, ORACLE], env=env, cwd=ROOT,
                           capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0, probe.stdout + probe.stderr
    run = subprocess.run([sys.executable, "-m", "pytest", "-p", "no:cacheprovider", "-q",
                          "-m", "not gpu", "-k", DESELECT, *SUITES], env=env, cwd=ROOT,
                         capture_output=True, text=True, timeout=600)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    for marker in ("ERROR: AddressSanitizer", "runtime error:"):
        assert marker not in out, out[-4000:]
    assert " passed" in out and " failed" not in out, out[-2000:]
