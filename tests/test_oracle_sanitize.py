"""The host-loop oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

Builds oracle/sanitize_main.cpp with oracle/sf_oracle.cpp (host code, g++) and runs it:
every boundary type on every face, odd and one-cell extents, both precisions and
smoothers, zero iteration counts, traces far outside the grid, and every per-stage entry
point. Any invalid access, leak or undefined operation aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-ffp-contract=off", "-Wno-unknown-pragmas"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_oracle_clean_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize")
    src = [os.path.join(ROOT, "oracle", "sanitize_main.cpp"), os.path.join(ROOT, "oracle", "sf_oracle.cpp")]
    build = subprocess.run(["g++", *FLAGS, *src, "-o", exe], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    # (verify_asan_link_order=0: a library preloaded by the environment may come first)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0 and "all runs clean" in run.stdout, (run.stdout + run.stderr)[-4000:]
