"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
routine the parity tests use runs on small sets (oracle/selftest.c), and
any out-of-bounds access, use after free, leak or undefined integer
behaviour fails the run."""
import os
import shutil
import subprocess

import pytest

from fheacc_testlib import ROOT

ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_selftest_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "selftest-asan"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_build", "selftest_asan")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ORACLE SELFTEST OK" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
