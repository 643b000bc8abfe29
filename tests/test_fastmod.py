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


PACK = os.path.join(os.path.dirname(__file__), "native", "pack_check")


def test_item_packing_fields_and_order_keys():
    """Regression: a 32-bit stream-index field made the allocation-order key
    decode (key >> 32 wraps to key >> 0 on the GPU) return the index instead of
    the sketch."""
    if not os.path.exists(PACK):
        from qsketch_amd import build as b
        b.build_native_tests()
    r = subprocess.run([PACK], stdout=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "pack violations: 0" in r.stdout


def test_fastmod_exhaustive_over_all_dividends():
    """Every 32-bit dividend for the divisors the workloads use (1021 masks and
    the reference sizing rule's tables: 3,145,721, 100,663,291, 201,326,557,
    402,653,171 buckets); about 45 s on one core."""
    if not os.path.exists(EXE):
        from qsketch_amd import build as b
        b.build_native_tests()
    r = subprocess.run([EXE, "full"], stdout=subprocess.PIPE, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    assert "mismatches: 0" in r.stdout
