"""GPU check of the serving loop: a short Poisson stream served by continuous
batching returns, for every query, the count the plain per-query path gives,
and the oracle's count for the queries it can sieve quickly."""
import pytest

pytestmark = pytest.mark.gpu


def test_serving_loop_counts_are_exact(gpu, oracle):
    import serve
    arrivals, qs = serve.poisson_workload(400, qps=2000.0, seed=11)
    print("serving", len(qs), "queries", flush=True)
    counts, lat, wall, batches = serve.serve(arrivals, qs)
    assert all(c is not None for c in counts)
    assert gpu.count_batch(qs, mode=1) == counts
    small = [(i, q) for i, q in enumerate(qs) if q[1] - q[0] < 10 ** 6 and q[0] < 10 ** 13][:40]
    for i, (a, b, k) in small:
        assert counts[i] == oracle.count6(a, b)[k - 1], (a, b, k)
    s = serve.summary(lat, wall, len(qs), batches)
    print("serving summary", s, flush=True)
    assert s["p99_ms"] >= s["p50_ms"] > 0 and batches >= 1
