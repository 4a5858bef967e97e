"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: oracle/san_driver.cpp runs
every oracle entry point on seeded inputs and edge shapes (1 x 1, single rows and columns, empty
and full planes) and the whole segmentNuclei chain on three synthetic H&E tiles; the sanitizers
abort on the first out-of-bounds access, leak or undefined operation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_clean_under_asan_ubsan():
    build = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                           timeout=300)
    if build.returncode != 0 and "asan" in (build.stderr + build.stdout).lower():
        pytest.skip("the sanitizer runtime is not installed: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([os.path.join(ROOT, "oracle", "_san", "san_driver")], capture_output=True, text=True,
                         timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "sanitizer run clean" in run.stdout
    assert "runtime error" not in run.stderr and "ERROR: AddressSanitizer" not in run.stderr
    # the chain really ran: the 512^2 tile has nuclei
    line = [l for l in run.stdout.splitlines() if l.startswith("chain 512x512")][0]
    assert "status 0" in line and int(line.split(",")[1].split()[0]) > 0, line
