"""GPU parity: the gfx950 kernels, called through the C ABI, against the
reference's golden vectors and the CPU oracle.  Every count is compared
bit-exact.  Mirrors the reference GPU parity test (test/gpu_count.cpp:73-132):
its whole case list, the known pi(x) values, plus the configs of the
benchmark and seeded random ranges checked against the oracle."""
import random

import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_device_and_geometry(gpu):
    s = gpu.devices()
    assert s and "HIP" in s[0]
    gpu.count_all(0, 10 ** 6)
    st = gpu.last_stats()
    assert st.sieve_bytes % 256 == 0 and st.work_group_size % 64 == 0
    _say("stats", st.as_dict())


def test_reference_case_list_per_type(gpu, golden):
    """test/gpu_count.cpp:86-113 through the per-type entry points
    (gpu_count_primes .. gpu_count_sextuplets)."""
    cases = [c for c in golden("cases.json") if "tag" not in c]
    bad = []
    for c in cases:
        for t, want in c["counts"].items():
            got = gpu.count(c["start"], c["stop"], int(t))
            if got != want:
                bad.append((c["start"], c["stop"], t, got, want))
    assert not bad, bad[:20]


def test_reference_case_list_single_pass(gpu, golden):
    """Same cases through the single-pass six-count call."""
    bad = []
    for c in golden("cases.json"):
        if c.get("tag") == "config":
            continue
        got = gpu.count_all(c["start"], c["stop"])
        for t, want in c["counts"].items():
            if got[int(t)] != want:
                bad.append((c["start"], c["stop"], t, got[int(t)], want))
    assert not bad, bad[:20]


def test_known_pi_and_kats(gpu, golden):
    """Known values from the reference test suite (kats.json), including the
    five 1e11-wide chunks from 1e19 (src/app/stressTest.cpp:95-118): above
    2^63, where the sieving primes reach 3.2e9 and each call runs as several
    bucket slabs."""
    bad = []
    for e in golden("kats.json"):
        _say("kat", e["start"], e["stop"], e["type"])
        got = gpu.count(e["start"], e["stop"], e["type"])
        if got != e["count"]:
            bad.append((e, got))
    assert not bad, bad


def test_config1_pi_1e10(gpu, golden):
    want = [c for c in golden("cases.json") if c.get("tag") == "config"][0]
    got = gpu.count_all(want["start"], want["stop"])
    assert got[0] == 455052511
    assert got == [want["counts"][str(t)] for t in range(6)]


def test_config2_six_counts(gpu, golden):
    """[1e13, 1e13+1e11]: six counts from one pass, against the reference."""
    want = [c for c in golden("cases.json") if c.get("tag") == "config"][1]
    got = gpu.count_all(want["start"], want["stop"])
    _say("config2", got, gpu.last_stats().as_dict())
    assert got == [want["counts"][str(t)] for t in range(6)]
    assert got[0] == 3340141707


def test_random_ranges_golden(gpu, golden):
    bad = []
    for c in golden("cases.json"):
        if c.get("tag") != "random":
            continue
        got = gpu.count_all(c["start"], c["stop"])
        if got != [c["counts"][str(t)] for t in range(6)]:
            bad.append((c["start"], c["stop"], got, c["counts"]))
    assert not bad, bad[:5]


def test_high_ranges_golden(gpu, golden):
    """[2^63, 2^64): six counts of windows at 1e19, across 2^63, up to
    2^64 - 1 (whose last byte holds values past 2^64) and seeded windows,
    against the reference CPU library."""
    bad = []
    for c in golden("cases_high.json"):
        _say("high", c["start"], c["stop"])
        got = gpu.count_all(c["start"], c["stop"])
        if got != [c["counts"][str(t)] for t in range(6)]:
            bad.append((c["start"], c["stop"], got, c["counts"]))
    assert not bad, bad[:5]
    top = golden("cases_high.json")[3]          # [2^64 - 1001, 2^64 - 1]
    for t in range(6):
        assert gpu.count(top["start"], top["stop"], t) == top["counts"][str(t)]


def test_wide_intervals_over_several_bucket_slabs(gpu, golden):
    """1e11-wide intervals at 1e15 and 1e17 expect 6.2e9 and 9.5e9 bucket
    entries, more than one slab's budget: each call is cut into several runs
    (and slabs) and must still give the reference's six counts."""
    for c in golden("cases_wide.json"):
        got = gpu.count_all(c["start"], c["stop"])
        st = gpu.last_stats()
        _say("wide", c["start"], st.as_dict())
        assert got == [c["counts"][str(t)] for t in range(6)], (c["start"], got)
        assert st.launches >= 2 and st.bucket_retries == 0


def test_work_group_size_power_of_two(gpu, oracle):
    """Workgroup sizes are rounded down to a power of two, as the reference
    requires (src/gpu/GpuBackend.hpp:49-50, src/gpu/cuda/CudaBackend.cpp:260-263);
    the counts of 192- and 768-thread requests equal the oracle's."""
    try:
        for req, want in [(192, 128), (768, 512), (960, 512), (100, 64)]:
            gpu.set_work_group_size(req)
            a, b = 10 ** 12 + 17, 10 ** 12 + 3 * 10 ** 7
            got = gpu.count_all(a, b)
            assert gpu.last_stats().work_group_size == want, req
            assert got == oracle.count6(a, b), (req, got)
    finally:
        gpu.set_work_group_size(0)


def test_segments_per_launch_bounds_launches(gpu, golden):
    """set_gpu_segments_per_launch bounds the segments of one launch (the
    reference's meaning, src/gpu/cuda/CudaBackend.cpp:315-319, 466): a
    1e10-wide interval of 2544 segments of 128 KiB takes ceil(2544 / n)
    launches, with unchanged counts."""
    want = [e["count"] for e in golden("kats.json")
            if e["start"] == 10 ** 12 and e["stop"] == 10 ** 12 + 10 ** 10 and e["type"] == 0][0]
    try:
        for n in [1000, 300]:
            gpu.set_segments_per_launch(n)
            assert gpu.count(10 ** 12, 10 ** 12 + 10 ** 10, 0) == want
            st = gpu.last_stats()
            _say("segments per launch", n, st.as_dict())
            assert st.segments == 2544 and st.launches == -(-2544 // n)
    finally:
        gpu.set_segments_per_launch(0)
    gpu.count(10 ** 12, 10 ** 12 + 10 ** 10, 0)
    assert gpu.last_stats().launches == 1


@pytest.mark.parametrize("kb,wg", [(16, 256), (32, 512), (64, 1024), (128, 512), (150, 256), (160, 1024),
                                   (64, 192), (128, 768)])
def test_geometries_against_oracle(gpu, oracle, kb, wg):
    """Every segment size / workgroup size gives the oracle's counts,
    including ranges that straddle segment ends and byte ends."""
    gpu.set_sieve_size(kb)
    gpu.set_work_group_size(wg)
    try:
        rng = random.Random(kb * 1000 + wg)
        for base in [0, 7, 10 ** 6, 10 ** 9, 2 ** 32 + 1, 10 ** 12, 10 ** 14]:
            for _ in range(4):
                a = base + rng.randrange(0, 10 ** 6)
                b = a + rng.choice([0, 1, 29, 30, 31, rng.randrange(10 ** 3),
                                    rng.randrange(10 ** 7), rng.randrange(10 ** 8)])
                got = gpu.count_all(a, b)
                want = oracle.count6(a, b)
                assert got == want, (kb, wg, a, b, got, want)
    finally:
        gpu.set_sieve_size(0)
        gpu.set_work_group_size(0)


def test_presieve_disabled(gpu, oracle):
    gpu.set_presieve(False)
    try:
        for a, b in [(0, 10 ** 7), (10 ** 9, 10 ** 9 + 10 ** 7), (0, 100), (7, 7)]:
            assert gpu.count_all(a, b) == oracle.count6(a, b), (a, b)
    finally:
        gpu.set_presieve(True)


def test_empty_and_tiny_ranges(gpu):
    # expected values computed by the reference CPU library (oracle/_ref)
    assert gpu.count_all(10, 9) == [0] * 6
    assert gpu.count_all(0, 0) == [0] * 6
    assert gpu.count_all(2, 2) == [1, 0, 0, 0, 0, 0]
    assert gpu.count_all(5, 17) == [5, 2, 3, 1, 1, 0]
    assert gpu.count_all(0, 17) == [7, 3, 3, 1, 1, 0]
    assert gpu.count_all(0, 30) == [10, 4, 5, 2, 3, 1]
    assert gpu.count_all(3, 31) == [10, 5, 5, 2, 3, 1]
    assert gpu.count(0, 17, 4) == 1


def test_try_count_policy(gpu):
    gpu.set_enabled(False)
    assert gpu.try_count(0, 10 ** 9) is None          # disabled
    gpu.set_enabled(True)
    try:
        assert gpu.try_count(0, 10 ** 7) is None      # below the 1e8 threshold
        assert gpu.try_count(0, 10 ** 9) == 50847534
        assert gpu.try_count(0, 10 ** 9, 1) == 3424506
    finally:
        gpu.set_enabled(False)


def test_bucket_list_room_retry(gpu, oracle):
    """Bucket lists get a room of E + m*sqrt(E) + 64 entries (E = expected
    crossings of the segment).  With m = 1 many lists of a high interval
    outgrow their rooms: the call must repeat once with the counted rooms and
    still give the oracle's counts."""
    a, b = 10 ** 13, 10 ** 13 + 10 ** 9
    want = oracle.count6(a, b)
    assert gpu.count_all(a, b) == want
    st = gpu.last_stats()
    assert st.bucket_retries == 0 and st.bucket_entries > 0
    entries = st.bucket_entries
    gpu.set_tuning(bucket_margin=1)
    try:
        assert gpu.get_tuning()["bucket_margin"] == 1
        assert gpu.count_all(a, b) == want
        st = gpu.last_stats()
        _say("margin 1:", st.as_dict())
        assert st.bucket_retries == 1
        assert st.bucket_entries == entries
    finally:
        gpu.set_tuning()


def test_wide_tier3_against_oracle(gpu, oracle):
    """A bucket threshold of 24 segments puts tier-3 primes above 30 * 2^16,
    where the class starts take the 32-bit multiply path instead of the packed
    16-bit dot product; the counts must not change."""
    gpu.set_tuning(bucket_factor=24)
    try:
        assert gpu.get_tuning()["bucket_factor"] == 24
        for a, b in [(10 ** 14, 10 ** 14 + 10 ** 7), (10 ** 15 + 12345, 10 ** 15 + 3 * 10 ** 6)]:
            _say("bucket factor 24:", a, b)
            assert gpu.count_all(a, b) == oracle.count6(a, b), (a, b)
    finally:
        gpu.set_tuning()


def test_measured_ceilings(gpu):
    """The roofline microbenchmarks the benchmark divides by: every figure is
    positive, the conflict-free LDS atomics are at least as fast as random
    ones, and the HBM copy is in the range of an MI355X stream."""
    r = gpu.lds_roofline()
    _say("ceilings", r)
    for k in ("and_conflict_free_gbs", "and_random_gbs", "or_conflict_free_gbs", "or_random_gbs",
              "lds_read_gbs", "hbm_copy_gbs"):
        assert r[k] > 0, k
    assert r["or_conflict_free_gbs"] >= 0.9 * r["or_random_gbs"]
    assert r["and_conflict_free_gbs"] >= 0.9 * r["and_random_gbs"]
    assert 500 < r["hbm_copy_gbs"] < 10000
    assert r["compute_units"] >= 1
