"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

``make -C oracle sanitize`` builds the oracle with a driver that calls every
routine on small problems (all sources and boundary kinds, every limiter,
HLL, the multigrid with multipole faces, the converged cooling integrator,
the network's forward pass and gradient).  The test runs it and fails on any
sanitizer report or non-finite result.  Host code only.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
TABLE = os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-B", "sanitize"], cwd=ORACLE, capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_build", "sanitize_oracle"), TABLE],
                       capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out
    assert "every routine ran" in r.stdout
