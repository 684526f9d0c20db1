"""GPU parity of the batch API (persistent work-stealing launches and the
plain one-call-per-query path) against the reference's counts and the
oracle.  Batch queries are (start, stop, k) with k = 1 primes .. 6
sextuplets."""
import json
import os
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _queries():
    q = np.load(os.path.join(GOLDEN, "config3_queries.npy"), allow_pickle=False)
    exp = json.load(open(os.path.join(GOLDEN, "config3_expected.json")))
    idx = sorted(int(i) for i in exp)
    return [(int(q[i, 0]), int(q[i, 1]), int(q[i, 2])) for i in idx], [exp[str(i)] for i in idx]


@pytest.mark.parametrize("mode", [0, 1])
def test_config3_subset_matches_reference(gpu, mode):
    """384 fixture queries (start <= 1e14) with reference-computed counts."""
    qs, want = _queries()
    print("queries", len(qs), "mode", mode, flush=True)
    got = gpu.count_batch(qs, mode=mode)
    bad = [(q, g, w) for q, g, w in zip(qs, got, want) if g != w]
    assert not bad, bad[:5]


def test_config3_every_decade_matches_reference(gpu):
    """16 fixture queries from every start decade 1e9 .. 1e17 (lengths up to
    2^32, so up to 1e18 + 2^32) with reference-computed counts, batched and
    plain."""
    q = np.load(os.path.join(GOLDEN, "config3_queries.npy"), allow_pickle=False)
    exp = json.load(open(os.path.join(GOLDEN, "config3_decades.json")))
    idx = sorted(int(i) for i in exp)
    qs = [(int(q[i, 0]), int(q[i, 1]), int(q[i, 2])) for i in idx]
    print("queries", len(qs), "max start", max(a for a, _, _ in qs), flush=True)
    for mode in (0, 1):
        got = gpu.count_batch(qs, mode=mode)
        bad = [(i, g, exp[str(i)]) for i, g in zip(idx, got) if g != exp[str(i)]]
        assert not bad, (mode, bad[:5])


def test_small_and_edge_queries_against_oracle(gpu, oracle):
    rng = random.Random(23)
    qs = [(10, 9, 1), (0, 0, 1), (0, 1, 2), (2, 2, 1), (3, 5, 2), (5, 17, 5), (0, 30, 6),
          (7, 7, 1), (31, 31, 1)]
    for base in [0, 1000, 10 ** 6, 10 ** 9, 2 ** 32, 10 ** 12]:
        for _ in range(12):
            a = base + rng.randrange(10 ** 5)
            qs.append((a, a + rng.choice([0, 1, 29, 30, rng.randrange(10 ** 4), rng.randrange(10 ** 7)]),
                       rng.randrange(1, 7)))
    got = gpu.count_batch(qs)
    for (a, b, k), g in zip(qs, got):
        assert g == oracle.count6(a, b)[k - 1], (a, b, k, g)
    assert gpu.count_batch(qs, mode=1) == got


def test_batch_equals_plain_on_fixture_prefix(gpu):
    """Size-independent check on the full-size query shape: the persistent
    engine and the plain per-query path agree on the first 256 config-3
    queries (starts up to 1e18, lengths up to 2^32)."""
    q = np.load(os.path.join(GOLDEN, "config3_queries.npy"), allow_pickle=False)[:256]
    qs = [(int(a), int(b), int(k)) for a, b, k in q]
    got = gpu.count_batch(qs)
    plain = gpu.count_batch(qs, mode=1)
    assert got == plain


def test_batch_rejects_bad_k(gpu):
    with pytest.raises(gpu.PrimesieveError):
        gpu.count_batch([(0, 100, 0)])
    with pytest.raises(gpu.PrimesieveError):
        gpu.count_batch([(0, 100, 7)])


def test_empty_batch(gpu):
    assert gpu.count_batch([]) == []


def test_batch_over_several_bucket_slabs(gpu):
    """A batch of about 3 slabs of bucket work: slab k+1 is filled on a second
    stream while slab k is sieved, with the two entry buffers alternating.
    Every query must equal its own single-interval call (one stream, no
    overlap), all six counts."""
    span = 3000 * 131072 * 30                     # ~3000 segments of 128 KiB
    starts = [2 * 10 ** 12, 5 * 10 ** 12, 10 ** 13, 3 * 10 ** 13, 7 * 10 ** 13, 10 ** 14]
    qs = [(s + 7 * i, s + 7 * i + span, k) for i, s in enumerate(starts) for k in range(1, 7)]
    print("queries", len(qs), "span", span, flush=True)
    got = gpu.count_batch(qs)
    st = gpu.last_stats()
    print("batch stats", st.as_dict(), flush=True)
    assert st.launches >= 3
    for i, s in enumerate(starts):
        want = gpu.count_all(s + 7 * i, s + 7 * i + span)
        assert got[6 * i:6 * i + 6] == want, (s, got[6 * i:6 * i + 6], want)


def test_graph_replay_matches_reference_and_eager(gpu, oracle):
    """PSG_BATCH_GRAPH: the batch's launch series (clear, bucket fill,
    segment kernel per slab) is captured once per shape and replayed with
    new descriptors.  Every replay -- with eager batches and count calls run
    between them -- must equal the reference's counts and the eager
    persistent path, for two different shapes."""
    q = np.load(os.path.join(GOLDEN, "config3_queries.npy"), allow_pickle=False)
    exp = json.load(open(os.path.join(GOLDEN, "config3_expected.json")))
    exp.update(json.load(open(os.path.join(GOLDEN, "config3_decades.json"))))
    idx = sorted(int(i) for i in exp)
    qs = [(int(q[i, 0]), int(q[i, 1]), int(q[i, 2])) for i in idx]
    want = [exp[str(i)] for i in idx]
    eager = gpu.count_batch(qs)
    assert eager == want
    rng = random.Random(5)
    small = []
    for _ in range(40):
        a = rng.randrange(10 ** 9, 10 ** 12)
        small.append((a, a + rng.randrange(10 ** 6), rng.randrange(1, 7)))
    small_want = [oracle.count6(a, b)[k - 1] for a, b, k in small]
    seen = set()
    for rep in range(3):
        got = gpu.count_batch(qs, mode=gpu.BATCH_GRAPH)
        st = gpu.last_stats()
        print("replay", rep, st.as_dict(), flush=True)
        seen.add(st.graph)
        assert got == want, rep
        # eager work between replays
        assert gpu.count_batch(small) == small_want
        assert gpu.count_all(10 ** 12, 10 ** 12 + 10 ** 8)[0] == 3618282
        got_small = gpu.count_batch(small, mode=gpu.BATCH_GRAPH)
        assert got_small == small_want, rep
    assert 1 in seen      # at least one pure replay of a captured series
