"""qs_fastmod (every bucket / mask index) equals % for all tested dividends
and divisors (random dividends for many divisors plus edge values)."""
import os
import subprocess

EXE = os.path.join(os.path.dirname(__file__), "native", "fastmod_check")


def test_fastmod_equals_modulo():
    if not os.path.exists(EXE):
        from qsketch_amd import build as b
        b.build_native_tests()
    r = subprocess.run([EXE], stdout=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "mismatches: 0" in r.stdout
