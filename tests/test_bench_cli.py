"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message and a non-zero status (the count path runs on one
device)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single MI355X" in r.stderr
    assert r.stdout.strip() == ""


def test_tile_prime_crossings_against_the_oracle(oracle):
    """bench.py's closed-form crossings of the pre-sieved primes 29..47 (used
    for the roofline fraction with C over p >= 29) equal the oracle's count,
    and with the crossings by p >= 53 they give the measurement contract's
    totals for [0, 1e10] and [1e13, 1e13 + 1e11]."""
    sys.path.insert(0, ROOT)
    import bench
    for lo, hi in ((7, 10 ** 6), (7, 10 ** 10), (10 ** 13, 10 ** 13 + 10 ** 11),
                   (10 ** 18, 10 ** 18 + 2 ** 32), (1000003, 1000003)):
        assert bench.tile_prime_crossings(lo, hi) == oracle.crossings(lo, hi, 29, 52)
    assert bench.tile_prime_crossings(7, 10 ** 10) + oracle.crossings(7, 10 ** 10, 53, 2 ** 32 - 1) \
        == 3095675604
    assert bench.tile_prime_crossings(10 ** 13, 10 ** 13 + 10 ** 11) + 34823572083 == 39161985516
