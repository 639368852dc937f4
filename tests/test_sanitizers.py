"""The C oracle (fdtd_oracle.c, fdtd2d_oracle.c) under AddressSanitizer and
UndefinedBehaviorSanitizer: its golden-vector, module and pin tests run
against the sanitizer build (make -C oracle asan), loaded in a child
process with libasan preloaded; any report aborts the child."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT as REPO

TESTS = ["tests/test_oracle_golden.py", "tests/test_oracle2d_mod.py", "tests/test_oracle_pins.py",
         "tests/test_rasteriser.py"]

CHILD = r"""
import os, sys
sys.path.insert(0, os.path.join(os.getcwd(), "tests"))
import oracle
assert oracle.LIB.endswith("liboracle_asan.so"), oracle.LIB
oracle.binding()
maps = open("/proc/self/maps").read()
assert "libasan" in maps and "liboracle_asan.so" in maps
print("[sanitizers] oracle loaded from", oracle.LIB, flush=True)
import pytest
sys.exit(pytest.main(["-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + sys.argv[1:]))
"""


def test_oracle_is_clean_under_asan_and_ubsan():
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("no gcc")
    res = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "asan"], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    asan = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(asan) and os.path.exists(asan), asan
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + ((" " + env["LD_PRELOAD"]) if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OFDTD_ORACLE_LIB"] = os.path.join(REPO, "oracle", "_ref", "liboracle_asan.so")
    env["OMP_NUM_THREADS"] = "4"
    res = subprocess.run([sys.executable, "-c", CHILD, *TESTS], cwd=REPO, env=env, capture_output=True, text=True,
                         timeout=1200)
    out = res.stdout + res.stderr
    assert "[sanitizers] oracle loaded from" in out, out[-3000:]
    assert res.returncode == 0, out[-3000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-3000:]
