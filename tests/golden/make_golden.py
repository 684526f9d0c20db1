"""Write the golden fixtures of the count path into tests/golden/.

Two kinds of data end up here:

* known-answer values transcribed from the reference's own test files
  (kats.json; each entry names the file:line it comes from), and
* expected counts computed by the reference CPU library itself, built from
  the reference sources by oracle/Makefile into oracle/_ref/ (cases.json,
  config3_expected.json).  The reference is only run here, in the build
  container; the fixtures travel, the library does not have to.

The config-3 query list (config3_queries.npy) is drawn from a fixed seed and
frozen as data so that expected counts never depend on floating-point pow().

Usage: python tests/golden/make_golden.py [extra|wide]   (needs oracle/_ref built: make -C oracle ref;
       "extra" writes only cases_high.json, config3_decades.json and
       config5_expected.json; "wide" only cases_wide.json)
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libprimesieve_ref.so")

NAMES = ["primes", "twins", "triplets", "quadruplets", "quintuplets", "sextuplets"]

# ---------------------------------------------------------------------------
# Known answers from the reference test suite (value, source).
KATS = [
    # pi(10^k), test/count_primes1.cpp:23-34
    *[{"start": 0, "stop": 10 ** (k + 1), "type": 0, "count": v, "src": "test/count_primes1.cpp:25-33"}
      for k, v in enumerate([4, 25, 168, 1229, 9592, 78498, 664579, 5761455, 50847534])],
    # pi(1e10), test/gpu_count.cpp:116-120 and tools/gpu_check.cpp:111-117
    {"start": 0, "stop": 10 ** 10, "type": 0, "count": 455052511, "src": "test/gpu_count.cpp:119"},
    # pi[10^k, 10^k+10^8], test/count_primes2.cpp:22-30
    *[{"start": 10 ** k, "stop": 10 ** k + 10 ** 8, "type": 0, "count": v, "src": "test/count_primes2.cpp:24-29"}
      for k, v in zip(range(12, 18), [3618282, 3342093, 3102679, 2893937, 2714904, 2555873])],
    # [1e12, 1e12+1e9], test/count_primes3.cpp:57 and test/count_<tuplet>.cpp:26-30
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 0, "count": 36190991, "src": "test/count_primes3.cpp:57"},
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 1, "count": 1730012, "src": "test/count_twins.cpp:30"},
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 2, "count": 271316, "src": "test/count_triplets.cpp:30"},
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 3, "count": 7171, "src": "test/count_quadruplets.cpp:30"},
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 4, "count": 1259, "src": "test/count_quintuplets.cpp:30"},
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 9, "type": 5, "count": 42, "src": "test/count_sextuplets.cpp:30"},
    # k-tuplets over 1e10 windows, src/app/test.cpp:74-81
    {"start": 10 ** 12, "stop": 10 ** 12 + 10 ** 10, "type": 1, "count": 17278660, "src": "src/app/test.cpp:76"},
    {"start": 10 ** 13, "stop": 10 ** 13 + 10 ** 10, "type": 2, "count": 2130571, "src": "src/app/test.cpp:77"},
    {"start": 10 ** 14, "stop": 10 ** 14 + 10 ** 10, "type": 3, "count": 38270, "src": "src/app/test.cpp:78"},
    {"start": 10 ** 15, "stop": 10 ** 15 + 10 ** 10, "type": 4, "count": 4107, "src": "src/app/test.cpp:79"},
    {"start": 10 ** 16, "stop": 10 ** 16 + 10 ** 10, "type": 5, "count": 66, "src": "src/app/test.cpp:80"},
    # pi over [10^k, 10^k+10^10], src/app/test.cpp:108-116
    *[{"start": 10 ** k, "stop": 10 ** k + 10 ** 10, "type": 0, "count": v, "src": "src/app/test.cpp:110-115"}
      for k, v in zip(range(12, 18), [361840208, 334067230, 310208140, 289531946, 271425366, 255481287])],
    # 1e11-wide chunks from 1e13 and from 1e19 (first five each), src/app/stressTest.cpp:56-79, 95-118
    *[{"start": 10 ** 13 + i * 10 ** 11, "stop": 10 ** 13 + (i + 1) * 10 ** 11, "type": 0, "count": v,
       "src": "src/app/stressTest.cpp:59"}
      for i, v in enumerate([3340141707, 3339037770, 3337978139, 3336895789, 3335816088])],
    *[{"start": 10 ** 19 + i * 10 ** 11, "stop": 10 ** 19 + (i + 1) * 10 ** 11, "type": 0, "count": v,
       "src": "src/app/stressTest.cpp:98"}
      for i, v in enumerate([2285693139, 2285771432, 2285721987, 2285796471, 2285730708])],
    # tools/ref_sieve.cpp:146-164
    {"start": 1, "stop": 2, "type": 0, "count": 1, "src": "tools/ref_sieve.cpp:155"},
    {"start": 2, "stop": 2, "type": 0, "count": 1, "src": "tools/ref_sieve.cpp:156"},
    {"start": 3, "stop": 5, "type": 0, "count": 2, "src": "tools/ref_sieve.cpp:157"},
    {"start": 5, "stop": 7, "type": 0, "count": 2, "src": "tools/ref_sieve.cpp:158"},
    {"start": 6, "stop": 6, "type": 0, "count": 0, "src": "tools/ref_sieve.cpp:159"},
    {"start": 7, "stop": 7, "type": 0, "count": 1, "src": "tools/ref_sieve.cpp:160"},
    {"start": 10 ** 9, "stop": 10 ** 9 + 1000, "type": 0, "count": 49, "src": "tools/ref_sieve.cpp:161"},
    {"start": 999999999989, "stop": 1000000000039, "type": 0, "count": 2, "src": "tools/ref_sieve.cpp:162"},
    {"start": 999999999000, "stop": 1000000001000, "type": 0, "count": 75, "src": "tools/ref_sieve.cpp:163"},
]


def ref_lib():
    lib = ctypes.CDLL(REF_LIB)
    fns = []
    for n in NAMES:
        f = getattr(lib, "primesieve_count_" + n)
        f.restype = ctypes.c_uint64
        f.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        fns.append(f)
    return fns


def gpu_count_case_list():
    """The case list of the reference GPU parity test, test/gpu_count.cpp:86-113."""
    cases = []
    small_stops = [0, 1, 2, 3, 5, 6, 7, 8, 30, 31, 49, 100, 163, 164, 241, 1000, 65537, 1000000]
    for s in small_stops:                                   # :88-92
        cases.append((0, s, "all"))
    for base in range(0, 121):                              # :95-97
        for length in range(0, 71, 7):
            cases.append((base, base + length, "primes"))
    for start in range(999999900, 1000000001):              # :100-101
        cases.append((start, start + 331, "primes"))
    for a, b in [(0, 10 ** 7), (0, 10 ** 8), (10 ** 9, 11 * 10 ** 8),       # :104-113
                 (999999999989, 1000000200000), (10 ** 13, 10 ** 13 + 10 ** 7)]:
        cases.append((a, b, "all"))
    return cases


def config3_queries(n=8192, seed=20260, path=None):
    """Batched uneven intervals: start = floor(10^u), u ~ U(9, 18); length =
    floor(2^v), v ~ U(16, 32), rounded down to a multiple of 30; k ~ U{1..6};
    stop = start + length - 1 < 2^64."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(9.0, 18.0, n)
    v = rng.uniform(16.0, 32.0, n)
    k = rng.integers(1, 7, n)
    q = np.zeros((n, 3), dtype=np.uint64)
    for i in range(n):
        start = int(10.0 ** float(u[i]))
        length = int(2.0 ** float(v[i]))
        length -= length % 30
        stop = start + length - 1
        assert stop < 2 ** 64
        q[i] = (start, stop, int(k[i]))
    return q


def high_cases(fns):
    """Six-count fixtures above 2^63, where the reference GPU backend leaves
    its division-magic path (src/gpu/cuda/CudaBackend.cpp:409-411): 1e9-wide
    windows at 1e19, across 2^63 and ending at 2^64 - 1, seeded windows in
    [2^63, 2^64), and short ranges whose last byte holds values past 2^64."""
    top = 2 ** 64 - 1
    spans = [(10 ** 19, 10 ** 19 + 10 ** 9), (2 ** 63 - 5 * 10 ** 8, 2 ** 63 + 5 * 10 ** 8),
             (top - 10 ** 9, top), (top - 1000, top), (top - 31, top), (top - 30, top - 1),
             (top, top), (top - 59, top - 29)]
    rng = np.random.Generator(np.random.PCG64(63))
    for _ in range(8):
        a = 2 ** 63 + int(rng.integers(0, 2 ** 63 - 2 * 10 ** 8))
        spans.append((a, a + int(rng.integers(0, 10 ** 8))))
    return [{"start": a, "stop": b, "counts": {str(t): fns[t](a, b) for t in range(6)}}
            for a, b in spans]


def wide_cases():
    """Six counts of 1e11-wide intervals at 1e15 and 1e17 (several bucket slabs
    per call: their expected bucket entries exceed one slab's budget), from the
    reference's one-pass ParallelSieve (oracle/ref_count6.cpp)."""
    lib = ctypes.CDLL(REF_LIB)
    lib.psg_ref_count6.restype = None
    lib.psg_ref_count6.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    out = []
    for a in (10 ** 15, 10 ** 17):
        c = (ctypes.c_uint64 * 6)()
        lib.psg_ref_count6(a, a + 10 ** 11, c, os.cpu_count() or 1)
        out.append({"start": a, "stop": a + 10 ** 11, "counts": {str(t): c[t] for t in range(6)}})
    return out


def config3_decade_subset(fns, q, per_decade=16, skip=()):
    """Reference counts for the first `per_decade` config-3 queries of every
    start decade 10^9 .. 10^17 (lengths up to 2^32), beyond those in `skip`."""
    out, seen = {}, {}
    for i in range(len(q)):
        if str(i) in skip:
            continue
        d = len(str(int(q[i, 0]))) - 1
        if seen.get(d, 0) >= per_decade:
            continue
        seen[d] = seen.get(d, 0) + 1
        out[str(i)] = fns[int(q[i, 2]) - 1](int(q[i, 0]), int(q[i, 1]))
    return out


def config5_subsets(fns):
    """Reference counts of Poisson serving streams (primesieve-gpu_amd/serve.py
    poisson_workload): all 400 queries of the test stream (seed 11) and every
    40th query of the benchmark's 8000-query stream (seed 5)."""
    sys.path.insert(0, os.path.join(ROOT, "primesieve-gpu_amd"))
    import serve
    out = {}
    for seed, n, step in [(11, 400, 1), (5, 8000, 40)]:
        _, qs = serve.poisson_workload(n, 2000.0, seed=seed)
        out["%d:%d" % (seed, n)] = {str(i): fns[k - 1](a, b)
                                    for i, (a, b, k) in enumerate(qs) if i % step == 0}
    return out


def main():
    if not os.path.exists(REF_LIB):
        sys.exit("build the reference library first: make -C oracle ref")
    fns = ref_lib()
    if len(sys.argv) > 1 and sys.argv[1] == "wide":
        with open(os.path.join(HERE, "cases_wide.json"), "w") as f:
            json.dump(wide_cases(), f)
        print("wrote cases_wide.json", flush=True)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "extra":
        # the fixtures added later, written without rewriting the others
        q = np.load(os.path.join(HERE, "config3_queries.npy"), allow_pickle=False)
        with open(os.path.join(HERE, "cases_high.json"), "w") as f:
            json.dump(high_cases(fns), f)
        print("wrote cases_high.json", flush=True)
        skip = json.load(open(os.path.join(HERE, "config3_expected.json")))
        with open(os.path.join(HERE, "config3_decades.json"), "w") as f:
            json.dump(config3_decade_subset(fns, q, skip=skip), f)
        print("wrote config3_decades.json", flush=True)
        with open(os.path.join(HERE, "config5_expected.json"), "w") as f:
            json.dump(config5_subsets(fns), f)
        print("wrote config5_expected.json", flush=True)
        return

    # the transcribed KATs must agree with the reference library itself
    for e in KATS:
        if e["stop"] - e["start"] <= 2 * 10 ** 11 and e["start"] < 10 ** 18:
            got = fns[e["type"]](e["start"], e["stop"])
            assert got == e["count"], (e, got)
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(KATS, f, indent=0)

    cases = []
    for a, b, which in gpu_count_case_list():
        types = range(6) if which == "all" else [0]
        cases.append({"start": a, "stop": b, "counts": {str(t): fns[t](a, b) for t in types}})
    # config 1 and config 2, all six counts
    for a, b in [(0, 10 ** 10), (10 ** 13, 10 ** 13 + 10 ** 11)]:
        cases.append({"start": a, "stop": b, "counts": {str(t): fns[t](a, b) for t in range(6)},
                      "tag": "config"})
    # seeded random ranges near segment and byte boundaries, all six counts
    rng = np.random.Generator(np.random.PCG64(7))
    for base in [0, 10 ** 6, 10 ** 9, 2 ** 32, 10 ** 12, 10 ** 15, 2 ** 62, 10 ** 18]:
        for _ in range(12 if base < 2 ** 40 else 3):
            a = base + int(rng.integers(0, 10 ** 7))
            b = a + int(rng.integers(0, 3 * 10 ** 7))
            cases.append({"start": a, "stop": b, "counts": {str(t): fns[t](a, b) for t in range(6)},
                          "tag": "random"})
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(cases, f)

    q = config3_queries()
    np.save(os.path.join(HERE, "config3_queries.npy"), q, allow_pickle=False)
    # expected counts for a seeded subset that the CPU reference finishes quickly
    pick = [i for i in range(len(q)) if int(q[i, 0]) <= 10 ** 14][:384]
    exp = {str(i): fns[int(q[i, 2]) - 1](int(q[i, 0]), int(q[i, 1])) for i in pick}
    with open(os.path.join(HERE, "config3_expected.json"), "w") as f:
        json.dump(exp, f)
    print("wrote", len(KATS), "kats,", len(cases), "cases,", len(exp), "config-3 expectations")


if __name__ == "__main__":
    main()
