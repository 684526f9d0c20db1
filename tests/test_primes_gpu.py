"""GPU check of the device-resident sieving-prime cache: caches reaching
2^24 are generated on the GPU (the segment kernel counts, then lists, the
primes of every segment; psg_prime_magics adds their division magics).
The primes are checked against the oracle's counts over random windows and
known values of pi(x); the magics against the oracle's divisionMagic."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIRST = 53          # the first sieving prime struck in LDS (7..47 are pre-sieved)
PI_SMALL = 15       # primes below 53


def _check_windows(pr, oracle, limit, rng, n=40, width=200000):
    for _ in range(n):
        a = rng.randrange(FIRST, max(FIRST + 1, limit - width))
        b = min(limit, a + rng.randrange(1, width))
        got = int(np.searchsorted(pr, b, "right") - np.searchsorted(pr, a, "left"))
        assert got == oracle.count6(a, b)[0], (a, b)


def _check_magics(pr, mg, oracle, rng, n=4000):
    idx = set(rng.randrange(len(pr)) for _ in range(n)) | set(range(min(200, len(pr))))
    idx |= set(range(max(0, len(pr) - 200), len(pr)))
    for i in sorted(idx):
        assert int(mg[i]) == oracle.magic(int(pr[i])), (i, int(pr[i]))


def test_device_primes_to_2_25(gpu, oracle):
    limit = (1 << 25) + 12345
    pr, mg = gpu.sieving_primes(limit)
    print("primes", len(pr), "last", int(pr[-1]), flush=True)
    assert int(pr[0]) == FIRST and int(pr[-1]) <= limit
    assert np.all(np.diff(pr.astype(np.int64)) > 0)
    assert len(pr) + PI_SMALL == oracle.count6(0, limit)[0]
    rng = random.Random(5)
    _check_windows(pr, oracle, limit, rng)
    _check_magics(pr, mg, oracle, rng)


def test_device_primes_to_1e9(gpu, oracle):
    """The cache the 1e18 queries need: pi(1e9) = 50,847,534."""
    limit = 10 ** 9
    pr, mg = gpu.sieving_primes(limit)
    print("primes", len(pr), "last", int(pr[-1]), flush=True)
    assert len(pr) + PI_SMALL == 50847534
    assert int(pr[-1]) == 999999937
    assert np.all(np.diff(pr.astype(np.int64)) > 0)
    for x, pi in [(10 ** 6, 78498), (10 ** 7, 664579), (10 ** 8, 5761455)]:
        assert int(np.searchsorted(pr, x, "right")) + PI_SMALL == pi
    rng = random.Random(9)
    _check_windows(pr, oracle, limit, rng, n=20)
    _check_magics(pr, mg, oracle, rng)
    # the counts that use these primes: a window at 1e18 against the oracle
    a = 10 ** 18
    assert gpu.count_all(a, a + 10 ** 6) == oracle.count6(a, a + 10 ** 6)


@pytest.mark.parametrize("kb,wg,presieve", [(32, 512, True), (150, 256, True), (128, 1024, False)])
def test_device_primes_other_geometries(gpu, kb, wg, presieve):
    """A new tuning makes a new backend and a new cache: generated with other
    segment and workgroup sizes (and without the pre-sieve, from 7) it holds
    the same primes."""
    limit = (1 << 24) + 999
    want, _ = gpu.sieving_primes(limit)
    gpu.set_tuning(sieve_bytes=kb * 1024, work_group_size=wg, presieve=presieve)
    try:
        pr, mg = gpu.sieving_primes(limit)
        print("geometry", kb, wg, presieve, "primes", len(pr), flush=True)
        if not presieve:
            assert list(pr[:12]) == [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
            pr, mg = pr[12:], mg[12:]
        assert np.array_equal(pr, want)
    finally:
        gpu.set_tuning()
