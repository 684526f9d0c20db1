"""psg_gpu_count_all_noalloc, the count entry point that allocates no device
memory: its descriptors, counters, bucket lists and counts live in the code
object's static psg_scratch array.  Checked through the C ABI against
psg_gpu_count_all (the allocating entry point it is a variant of), bit-exact,
on two calls in a row, the second on a smaller interval; and its error paths
(no backend yet, sieving primes not cached, an interval larger than the
scratch holds) fail cleanly before any launch."""
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_equals_count_all(gpu):
    big = (10 ** 13, 10 ** 13 + 10 ** 10)          # ~34 bucket slabs through the scratch
    small = (10 ** 12 + 12345, 10 ** 12 + 10 ** 8)
    want_big = gpu.count_all(*big)                 # creates the backend and its prime cache
    want_small = gpu.count_all(*small)
    got_big = gpu.count_all_noalloc(*big)
    st = gpu.last_stats()
    print("noalloc big", got_big, st.as_dict(), flush=True)
    got_small = gpu.count_all_noalloc(*small)
    print("noalloc small", got_small, gpu.last_stats().as_dict(), flush=True)
    assert got_big == want_big
    assert got_small == want_small
    assert st.launches > 1
    # tiny and empty ranges, and the primes 2, 3, 5
    assert gpu.count_all_noalloc(0, 30) == [10, 4, 5, 2, 3, 1]
    assert gpu.count_all_noalloc(10, 9) == [0] * 6


def test_noalloc_error_paths(gpu):
    # no backend (a tuning change drops it, and with it the prime cache): refused
    gpu.set_tuning()
    with pytest.raises(gpu.PrimesieveError, match="no GPU backend"):
        gpu.count_all_noalloc(0, 10 ** 6)
    gpu.count_all(10 ** 12, 10 ** 12 + 10 ** 6)     # a backend with a small prime cache
    # sieving primes up to isqrt(1e19) are not cached: refused, nothing launched
    with pytest.raises(gpu.PrimesieveError, match="not cached"):
        gpu.count_all_noalloc(10 ** 19, 10 ** 19 + 10 ** 6)
    # more tables and bucket lists than the scratch holds: refused
    gpu.count_all(10 ** 18, 10 ** 18 + 10 ** 6)     # primes to 1e9 cached
    with pytest.raises(gpu.PrimesieveError, match="scratch"):
        gpu.count_all_noalloc(10 ** 9, 10 ** 9 + 5 * 10 ** 13)
    # the backend still counts correctly afterwards, with and without allocating
    a, b = 10 ** 12, 10 ** 12 + 10 ** 8
    assert gpu.count_all_noalloc(a, b) == gpu.count_all(a, b)
