"""Serving loop for prime / k-tuplet count queries.

Queries arrive as a Poisson stream.  The loop keeps the sieving primes cached
on the device (the backend's prime cache is filled once, to the largest stop
the workload allows) and serves queries by continuous batching:

  * every query that has arrived joins the waiting queue;
  * admission: queries leave the queue for the next batch in arrival order
    while the batch's estimated big-prime bucket entries stay within an HBM
    budget (and the batch within a size cap); the rest wait;
  * packing: the admitted queries are ordered by the cost model's predicted
    time, largest first, so the persistent work-stealing launch finishes
    evenly (without a model: by length);
  * the batch runs through psg_gpu_count_batch; each query's latency is its
    completion time minus its arrival time.

Nothing here computes a count; every count comes from the GPU kernels.
"""
import math
import time

import numpy as np

import psgpu


def poisson_workload(n, qps, seed, alpha=1.2, min_len=2 ** 12, max_len=2 ** 36):
    """n queries: Poisson arrivals at `qps`; start = floor(10^u), u ~ U(9, 18);
    length ~ Pareto(alpha, x_m = min_len) clipped to [min_len, max_len];
    k ~ U{1..6}.  Returns (arrival seconds, [(start, stop, k)])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / qps, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    u = rng.uniform(9.0, 18.0, n)
    lengths = np.minimum(min_len * (1.0 - rng.uniform(0.0, 1.0, n)) ** (-1.0 / alpha), max_len)
    ks = rng.integers(1, 7, n)
    qs = []
    for i in range(n):
        start = int(10.0 ** float(u[i]))
        length = max(int(lengths[i]), min_len)
        qs.append((start, start + length - 1, int(ks[i])))
    return arrivals, qs


def bucket_entries_estimate(start, stop, seg_bytes=131072, factor=12):
    """Expected big-prime bucket entries of one query: 8 * bytes * sum 1/p over
    the bucket primes (p > factor * seg_bytes; the backend's default bucket
    threshold for a batch, which is how the loop counts, is 12 segments), by
    Mertens' estimate."""
    lo = float(factor * seg_bytes)
    hi = math.sqrt(max(stop, 4))
    if hi <= lo:
        return 0.0
    nbytes = (stop - start + 1) / 30.0
    return 8.0 * nbytes * (math.log(math.log(hi)) - math.log(math.log(lo)))


def serve(arrivals, queries, model=None, entry_budget=2.0e8, max_batch=4096, log=None,
          mode=psgpu.BATCH_PERSISTENT):
    """Run the loop in real time.  mode: psgpu.BATCH_PERSISTENT launches each
    batch's series one by one, psgpu.BATCH_GRAPH replays a captured series
    (captured once per shape).  Returns (counts, latencies in seconds, wall
    seconds, batches)."""
    n = len(queries)
    counts = [None] * n
    lat = np.zeros(n)
    est = [bucket_entries_estimate(a, b) for a, b, _ in queries]
    nxt = 0                      # next query not yet arrived
    queue = []                   # arrived, not admitted
    batches = 0
    t0 = time.perf_counter()
    while nxt < n or queue:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        if not queue:
            time.sleep(max(0.0, min(arrivals[nxt] - now, 0.001)))
            continue
        batch, total = [], 0.0
        while queue and len(batch) < max_batch:
            i = queue[0]
            if batch and total + est[i] > entry_budget:
                break
            batch.append(queue.pop(0))
            total += est[i]
        qs = [queries[i] for i in batch]
        if model is not None and len(qs) > 1:
            import costmodel
            _, order = costmodel.choose_for_batch(model, [(a, b - a + 1, k) for a, b, k in qs],
                                                  grid=[(131072, 16)])
        else:
            order = np.argsort([-(b - a) for a, b, _ in qs], kind="stable")
        batch = [batch[j] for j in order]
        out = psgpu.count_batch([queries[i] for i in batch], mode=mode)
        done = time.perf_counter() - t0
        for i, c in zip(batch, out):
            counts[i] = c
            lat[i] = done - arrivals[i]
        batches += 1
        if log and batches % 50 == 0:
            log("served %d / %d queries in %d batches" % (sum(c is not None for c in counts), n, batches))
    wall = time.perf_counter() - t0
    return counts, lat, wall, batches


def summary(lat, wall, n, batches):
    return {"queries": n, "wall_s": wall, "throughput_qps": n / wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3), "batches": batches}
