"""Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/sanitize/roundtrip.cpp runs the product client (keys and queries in
both generator modes, response unpacking, decryption, argument checks) against
the oracle server end to end; any sanitizer report aborts it.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(os.path.join(SAN, "Makefile")) or not os.path.exists(CLANG)
                    or shutil.which("make") is None, reason="sanitizer build needs make and ROCm's clang")
def test_client_and_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", SAN], check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([os.path.join(SAN, "_build", "roundtrip")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitized round trip ok" in r.stdout
