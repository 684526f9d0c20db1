"""CPU check of the rule the segment kernel's tier 3 uses for primes with
16p > 30S (strike_cands_lane in psg_sieve.hip): in a segment of S bytes at
global byte B >= p, the crossings of p are among K classes, namely the first
K coprime cofactors, in cofactor order, from t mod 30, where
t = ceil((30 (B mod p) + 7) / p).  The segment's values hold at most m
consecutive multiples of p when mp > 30S, and m consecutive integers hold at
most K(m) coprime to 30: K = 2, 3, 4, 5 for m = 6, 8, 12, 16.

The rule is restated here from its derivation (wheel-30 layout: bit j of
byte g is 30 g + {7, 11, 13, 17, 19, 23, 29, 31}[j], the reference layout of
src/gpu/GpuWheel.hpp:125) and compared with brute force: every multiple of p
in the segment's value range whose cofactor is coprime to 30.  Just below
each band one more candidate is needed, which the last test shows, so the
bounds are tight."""
import math
import random

OFFSETS = [7, 11, 13, 17, 19, 23, 29, 31]
COPRIMES = [1, 7, 11, 13, 17, 19, 23, 29]
COPRIME_BITS = sum(1 << q for q in COPRIMES)


def _class_rows():
    """Per residue b = p mod 30: the 8 classes in cofactor order, each as
    (cofactor residue q, byte correction e, class j), with p*q = 30*(a*q + e) + o_j."""
    rows = {}
    for b in COPRIMES:
        row = []
        for q in COPRIMES:
            j = next(j for j, o in enumerate(OFFSETS) if (b * q) % 30 == o % 30)
            row.append((q, (b * q - OFFSETS[j]) // 30, j))
        rows[b] = row
    return rows


ROWS = _class_rows()


def _is_prime(n):
    if n < 2:
        return False
    return all(n % d for d in range(2, math.isqrt(n) + 1))


def crossings_brute(p, B, S):
    """(byte within the segment, class) of every multiple of p with a cofactor
    coprime to 30 in the segment's values [30B + 7, 30(B + S) + 7), p itself excluded."""
    hits = set()
    v = -(-(30 * B + 7) // p) * p
    while v < 30 * (B + S) + 7:
        if math.gcd(v // p, 30) == 1 and v != p:
            g = (v - 7) // 30
            hits.add((g - B, OFFSETS.index(v - 30 * g)))
        v += p
    return hits


def crossings_candidates(p, B, S, k=3):
    """The kernel's rule: k classes from the first coprime cofactor >= t mod 30.
    Its table row holds the 8 classes twice (16 slots), and the first slot is
    the number of coprime residues below t, 0..8 (t = 30 gives 8, which is
    residue 0's slot repeated)."""
    a, b = divmod(p, 30)
    R = B % p
    t = (30 * R + 7 + p - 1) // p
    assert 1 <= t <= 30
    s0 = bin(COPRIME_BITS & ((1 << t) - 1)).count("1")
    assert s0 + k <= 16
    hits = set()
    for i in range(k):
        q, e, j = ROWS[b][(s0 + i) % 8]
        off = (a * q + e - R) % p
        if off < S:
            hits.add((off, j))
    return hits


def _random_prime(rng, lo, hi):
    while True:
        p = rng.randrange(lo, hi + 1)
        if _is_prime(p):
            return p


def test_class_rows_cover_every_class():
    for b, row in ROWS.items():
        assert sorted(j for _, _, j in row) == list(range(8))
        assert [q for q, _, _ in row] == COPRIMES


def test_three_candidates_hold_every_crossing():
    rng = random.Random(7)
    for S in (256, 4096, 131072):
        lo, hi = 30 * S // 8 + 1, 7 * S
        cases = [(_random_prime(rng, lo, hi), None) for _ in range(150 if S < 131072 else 40)]
        cases += [(p, None) for p in [x for x in range(lo, lo + 400) if _is_prime(x)][:6]]
        for p, _ in cases:
            for _ in range(6):
                B = rng.randrange(p, 10 ** 15)
                assert crossings_candidates(p, B, S) == crossings_brute(p, B, S), (p, B, S)


def test_two_candidates_above_five_segments():
    """6p > 30S: at most 6 consecutive multiples in a segment, at most 2 of them
    with a cofactor coprime to 30 (the kernel's two-candidate band)."""
    rng = random.Random(9)
    for S in (256, 4096, 131072):
        lo, hi = 30 * S // 6 + 1, 24 * S
        primes = [_random_prime(rng, lo, hi) for _ in range(100 if S < 131072 else 30)]
        primes += [x for x in range(lo, lo + 400) if _is_prime(x)][:6]
        for p in primes:
            for _ in range(6):
                B = rng.randrange(p, 10 ** 15)
                assert crossings_candidates(p, B, S, k=2) == crossings_brute(p, B, S), (p, B, S)


def test_max_coprimes_in_consecutive_integers():
    """K(m), the most integers coprime to 30 among m consecutive ones, for the
    bands the kernel uses (m = 6, 8, 12, 16 -> K = 2, 3, 4, 5)."""
    cop = [math.gcd(x, 30) == 1 for x in range(30)]
    K = {m: max(sum(cop[(s + i) % 30] for i in range(m)) for s in range(30)) for m in range(1, 31)}
    assert [K[m] for m in (6, 8, 12, 16)] == [2, 3, 4, 5]
    assert [K[m] for m in (7, 9, 13, 17)] == [3, 4, 5, 6]


def test_four_and_five_candidates():
    """12p > 30S: four candidates; 16p > 30S: five (the kernel's bands between
    30S/16 and 30S/8)."""
    rng = random.Random(13)
    for k, m in ((4, 12), (5, 16)):
        for S in (256, 4096, 131072):
            lo, hi = 30 * S // m + 1, 30 * S // (m - 4)
            primes = [_random_prime(rng, lo, hi) for _ in range(100 if S < 131072 else 30)]
            primes += [x for x in range(lo, lo + 400) if _is_prime(x)][:6]
            for p in primes:
                for _ in range(6):
                    B = rng.randrange(p, 10 ** 15)
                    assert crossings_candidates(p, B, S, k=k) == crossings_brute(p, B, S), (k, p, B, S)


def test_bound_is_tight():
    """Below 8p > 30S a segment can hold four crossings of p, so three
    candidates would miss one: the kernel keeps those primes on the
    eight-class path."""
    rng = random.Random(11)
    S = 256
    found = False
    for p in [x for x in range(30 * S // 12 + 1, 30 * S // 8 + 1) if _is_prime(x)]:
        for _ in range(200):
            B = rng.randrange(p, 10 ** 9)
            if len(crossings_brute(p, B, S)) == 4:
                assert crossings_candidates(p, B, S, k=3) != crossings_brute(p, B, S)
                assert crossings_candidates(p, B, S, k=4) == crossings_brute(p, B, S)
                found = True
                break
        if found:
            break
    assert found


def test_four_and_five_candidate_bounds_are_tight():
    """Below 12p > 30S a segment can hold five crossings, below 16p > 30S six."""
    rng = random.Random(17)
    S = 256
    for k, m in ((4, 12), (5, 16)):
        found = False
        for p in [x for x in range(30 * S // (m + 4) + 1, 30 * S // m + 1) if _is_prime(x)]:
            for _ in range(400):
                B = rng.randrange(p, 10 ** 9)
                brute = crossings_brute(p, B, S)
                if len(brute) == k + 1:
                    assert crossings_candidates(p, B, S, k=k) != brute
                    assert crossings_candidates(p, B, S, k=k + 1) == brute
                    found = True
                    break
            if found:
                break
        assert found, k


def _magic(p):
    """The device's division magic of p (psg_prime_magics): ceil(2^(64+L) / p)."""
    L = p.bit_length() - 1
    return -(-(1 << (64 + L)) // p), L


def test_first_cofactor_from_the_magic_high_word():
    """The kernel takes t = ceil((30R + 7) / p) as floor(y Mh / 2^(32+L)) with
    y = 30R + 7 + p and Mh the magic's high 32 bits: exact for tier-3 primes
    below 30 * 2^16, checked at random remainders and at the remainders where
    30R + 7 or 30R + 6 is a multiple of p (the cases a truncated magic would
    get wrong)."""
    rng = random.Random(19)
    primes = [_random_prime(rng, 1 << 17, 30 << 16) for _ in range(300)]
    primes += [131071, 1966079] + [x for x in range((30 << 16) - 200, 30 << 16) if _is_prime(x)]
    for p in primes:
        M, L = _magic(p)
        Mh = M >> 32
        assert Mh == (1 << (32 + L)) // p
        Rs = [0, 1, p - 1, p - 2] + [rng.randrange(p) for _ in range(40)]
        inv30 = pow(30, -1, p)
        Rs += [(-7 * inv30) % p, (-6 * inv30) % p]
        for R in Rs:
            y = 30 * R + 7 + p
            assert y < 1 << 26
            t = ((y * Mh) >> 32) >> L
            assert t == -(-(30 * R + 7) // p), (p, R)
