"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

Every oracle function runs once on small seeded inputs from a C driver
(tests/native/oracle_sanitized_driver.c) built with gcc, the oracle sources
compiled in; any out-of-bounds access, use after free, leak, signed overflow
or misaligned access aborts the run.  Host code only: nothing here touches a
GPU.
"""
import os
import pathlib
import shutil
import subprocess

import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
ORACLE = REPO / "oracle"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_runs_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitized"
    srcs = sorted(str(p) for p in ORACLE.glob("*.c"))
    cmd = ["gcc", "-O1", "-g", "-std=c99", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", f"-I{ORACLE}",
           str(REPO / "tests" / "native" / "oracle_sanitized_driver.c"), *srcs, "-lm",
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "ok" in r.stdout
