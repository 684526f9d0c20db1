"""GPU check of the serving loop: a short Poisson stream served by continuous
batching returns, for every query, the reference CPU library's count
(tests/golden/config5_expected.json), the count the plain per-query path
gives, and the oracle's count for the queries it can sieve quickly."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu


def test_serving_loop_counts_are_exact(gpu, oracle):
    import serve
    arrivals, qs = serve.poisson_workload(400, qps=2000.0, seed=11)
    print("serving", len(qs), "queries", flush=True)
    counts, lat, wall, batches = serve.serve(arrivals, qs)
    assert all(c is not None for c in counts)
    exp = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                      "config5_expected.json")))["11:400"]
    assert len(exp) == len(qs)
    bad = [(i, qs[i], counts[i], v) for i, v in ((int(i), v) for i, v in exp.items()) if counts[i] != v]
    assert not bad, bad[:5]
    assert gpu.count_batch(qs, mode=1) == counts
    small = [(i, q) for i, q in enumerate(qs) if q[1] - q[0] < 10 ** 6 and q[0] < 10 ** 13][:40]
    for i, (a, b, k) in small:
        assert counts[i] == oracle.count6(a, b)[k - 1], (a, b, k)
    s = serve.summary(lat, wall, len(qs), batches)
    print("serving summary", s, flush=True)
    assert s["p99_ms"] >= s["p50_ms"] > 0 and batches >= 1


def test_serving_loop_graph_replay(gpu):
    """The same stream served through captured launch series (each batch
    replays the series captured for its shape) gives the reference's counts."""
    import serve
    arrivals, qs = serve.poisson_workload(400, qps=2000.0, seed=11)
    counts, lat, wall, batches = serve.serve(arrivals, qs, mode=gpu.BATCH_GRAPH)
    exp = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                      "config5_expected.json")))["11:400"]
    bad = [(i, qs[i], counts[i], v) for i, v in ((int(i), v) for i, v in exp.items()) if counts[i] != v]
    assert not bad, bad[:5]
    s = serve.summary(lat, wall, len(qs), batches)
    print("graph serving summary", s, flush=True)
