"""The drop-in C++ API (include/onionpir.hpp): tests/cpp/test_pir.cpp repeats
the reference's end-to-end test (src/tests.cpp:40-153) through PirParams /
PirServer / PirClient, streams included, and must retrieve every entry."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_pir")


@pytest.mark.parametrize("shape", [("256", "32", "13", "16"), ("128", "64", "5", "17")])
def test_cpp_drop_in_end_to_end(shape):
    assert os.path.exists(EXE), "build the C++ tests first (python -m onionpirv2_amd.build)"
    r = subprocess.run([EXE, *shape], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Success rate: {shape[2]}/{shape[2]}" in r.stdout
    assert "response size: 13824 bytes" in r.stdout
