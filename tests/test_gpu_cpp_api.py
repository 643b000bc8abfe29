"""The C++ interface (include/qsketch.hpp) driven like the reference's Test
harness, estimates checked bit-exact against the oracle inside the test
program (tests/native/cpp_api_test.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "native", "cpp_api_test")


def test_cpp_interface_matches_oracle(gpu):
    if not os.path.exists(EXE):
        pytest.fail("tests/native/cpp_api_test is not built (run __graft_entry__.build())")
    print("CASE cpp_api_test", flush=True)
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")
