"""Pins the CPU oracle (test infrastructure) before any GPU result is
compared with it: reference KATs, the reference CPU library, brute force,
and what the mathematics fixes (pattern enumeration, 128-bit division,
additivity)."""
import random

import pytest

PATTERNS = {   # admissible prime k-tuplet patterns (offsets from the smallest member)
    1: [(0, 2)],
    2: [(0, 2, 6), (0, 4, 6)],
    3: [(0, 2, 6, 8)],
    4: [(0, 2, 6, 8, 12), (0, 4, 6, 10, 12)],
    5: [(0, 4, 6, 10, 12, 16)],
}
OFF = [7, 11, 13, 17, 19, 23, 29, 31]


def test_kats_small(oracle, golden):
    """Reference KATs that the oracle finishes in seconds."""
    for e in golden("kats.json"):
        if e["stop"] - e["start"] <= 2 * 10 ** 9 and e["stop"] <= 10 ** 13:
            assert oracle.count6(e["start"], e["stop"])[e["type"]] == e["count"], e


def test_reference_case_list(oracle, golden):
    """test/gpu_count.cpp's case list, expected values from the reference."""
    for c in golden("cases.json"):
        if c.get("tag") or c["stop"] - c["start"] > 2 * 10 ** 8:
            continue
        got = oracle.count6(c["start"], c["stop"])
        for t, want in c["counts"].items():
            assert got[int(t)] == want, (c, got)


def test_random_golden_small(oracle, golden):
    for c in golden("cases.json"):
        if c.get("tag") == "random" and c["start"] < 10 ** 16:
            got = oracle.count6(c["start"], c["stop"])
            assert got == [c["counts"][str(t)] for t in range(6)], c


def test_brute_force_agrees(oracle):
    """Explicit primality + explicit tuplet patterns vs the byte-mask sieve."""
    rng = random.Random(5)
    for _ in range(150):
        a = rng.choice([0, 1, 2, 3, 4, 5, 6, 7, 29, 30, 31, 1000, 10 ** 6]) + rng.randrange(120)
        b = a + rng.randrange(400)
        got = oracle.count6(a, b, 1)
        for k in range(6):
            assert got[k] == oracle.brute(a, b, k), (a, b, k)


def test_brute_force_special_tuplets(oracle):
    # the tuplets containing 3 and 5 (reference src/PrimeSieveClass.cpp:38-48)
    assert oracle.brute(3, 5, 1) == 1 and oracle.brute(5, 7, 1) == 1
    assert oracle.brute(5, 11, 2) == 1 and oracle.brute(5, 13, 3) == 1
    assert oracle.brute(5, 17, 4) == 1 and oracle.brute(7, 23, 5) == 1
    assert oracle.brute(0, 100, 0) == 25
    assert oracle.brute(4, 17, 1) == 2   # (5,7) (11,13); (3,5) starts below 4


def test_ktable_by_pattern_enumeration(oracle):
    """Per-byte k-tuplet counts equal the number of admissible patterns whose
    members are all set bits of the byte (values 30g + OFF[b])."""
    for k in range(6):
        tab = oracle.ktable(k)
        for j in range(256):
            vals = {OFF[b] for b in range(8) if (j >> b) & 1}
            if k == 0:
                want = len(vals)
            else:
                want = sum(1 for pat in PATTERNS[k] for v in vals
                           if all(v + d in vals for d in pat))
            assert tab[j] == want, (k, j)


def test_division_magic_128bit(oracle):
    rng = random.Random(11)
    primes = [7, 29, 31, 65537, 2147483647, 4294967291] + [rng.randrange(3, 2 ** 32) | 1 for _ in range(200)]
    for p in primes:
        L = p.bit_length() - 1
        M = oracle.magic(p)
        assert M == -(-(1 << (64 + L)) // p)
        for _ in range(50):
            x = rng.randrange(0, 2 ** 63)
            assert ((x * M) >> 64) >> L == x // p


def test_crossings_closed_form(oracle):
    rng = random.Random(3)
    for _ in range(60):
        lo = rng.randrange(7, 10 ** 6)
        hi = lo + rng.randrange(0, 20000)
        pmin = rng.choice([7, 29, 100])
        assert oracle.crossings(lo, hi, pmin, 10 ** 6) == oracle.crossings_brute(lo, hi, pmin, 10 ** 6)


def test_additivity(oracle):
    """count(a, c) = count(a, b) + count(b+1, c) for primes; for k-tuplets
    when b == 1 (mod 30), the last value of a byte (no tuplet is split)."""
    rng = random.Random(9)
    for _ in range(40):
        a = rng.randrange(0, 10 ** 9)
        c = a + rng.randrange(10 ** 4, 10 ** 6)
        b = rng.randrange(a, c)
        assert oracle.count6(a, c)[0] == oracle.count6(a, b)[0] + oracle.count6(b + 1, c)[0]
        b = b - (b % 30) + 1
        if a <= b < c:
            whole = oracle.count6(a, c)
            parts = [x + y for x, y in zip(oracle.count6(a, b), oracle.count6(b + 1, c))]
            assert whole == parts


def test_oracle_against_reference_library(oracle, reference):
    """The restatement agrees with the reference CPU library built from its
    own sources (oracle/_ref), on seeded ranges including segment edges."""
    rng = random.Random(17)
    for base in [0, 7, 30, 10 ** 6, 10 ** 9, 2 ** 32, 10 ** 12, 10 ** 13]:
        for _ in range(6):
            a = base + rng.randrange(10 ** 5)
            b = a + rng.choice([0, 1, 30, rng.randrange(10 ** 4), rng.randrange(10 ** 7)])
            got = oracle.count6(a, b)
            want = [reference.count(a, b, t) for t in range(6)]
            assert got == want, (a, b)
