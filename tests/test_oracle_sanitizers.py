"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's C sources and a driver that calls each of its entry points on small
inputs (tests/oracle_asan_driver.c) are built with -fsanitize=address,undefined
and run; any out-of-bounds access, leak or undefined operation fails the
test.  CPU only."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc is not available")
    sources = sorted(str(p) for p in (ROOT / "oracle").glob("*.c"))
    exe = tmp_path / "oracle_driver"
    cmd = [cc, "-O1", "-g", "-std=gnu11", "-ffp-contract=off", "-fopenmp",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", *sources, str(ROOT / "tests" / "oracle_asan_driver.c"),
           "-lm", "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "oracle driver: ok" in run.stdout
