"""Learned cost model and tuner of the count path.

An FP32 MLP (6 -> 64 -> 64 -> 1) trained by the library's own HIP forward,
backward and Adam kernels (psg_mlp_* in include/primesieve_gpu.h) on timings
that the sieve itself produces.  It predicts log2 of the device time of one
count call from

    log2 start, log2 length, k, log2 segment bytes, waves per workgroup,
    log2 number of sieving primes

and is used to choose the tuning (segment bytes, waves) of an interval or of
a whole batch, and to order a batch's queries by predicted cost.  Grid
search is the fallback and the judge of its regret: over the same grid for
the model's training and held-out timings, and at run time over a third
axis as well, the bucket threshold (bucket_factor: primes above that many
segments' bytes go to the HBM bucket lists), timed on a seeded sample of
the batch itself -- the reference's tune() (tools/gpu_check.cpp:164-204)
applied to a batch.  The model never touches a count: it only picks among
geometries that all give bit-exact results.
"""
import ctypes
import json
import math
import time

import numpy as np

import psgpu

# tuning grid the model chooses from: segment bytes x waves per workgroup
SEG_BYTES = (32768, 65536, 131072, 163840)
WAVES = (4, 8, 16)
GRID = [(s, w) for s in SEG_BYTES for w in WAVES]
DEFAULT = (131072, 16)          # the backend's own default geometry
# the run-time grid's third axis: the bucket threshold in segments (only for
# power-of-two segments; 160 KiB segments have no bucket tier)
BUCKET_FACTORS = (4, 7, 12)
GRID3 = [(s, w, b) for s in SEG_BYTES for w in WAVES
         for b in (BUCKET_FACTORS if s & (s - 1) == 0 else (0,))]
DEFAULT3 = (131072, 16, 12)   # the backend's default for a batch (12 segments; 7 for one call)


def prime_count_estimate(x):
    """pi(x) by the logarithmic-integral approximation (a feature only)."""
    x = max(float(x), 3.0)
    return x / (math.log(x) - 1.0)


def features(start, length, k, seg_bytes, waves):
    stop = start + length - 1
    return [math.log2(max(start, 2)), math.log2(max(length, 2)), float(k),
            math.log2(seg_bytes), float(waves),
            math.log2(max(prime_count_estimate(math.isqrt(max(stop, 4))), 1.0))]


def draw_queries(n, seed, min_log2_len=16.0, max_log2_len=32.0):
    """Uneven intervals as in the batched workload: start = floor(10^u),
    u ~ U(9, 18); length = floor(2^v) rounded down to a multiple of 30;
    k ~ U{1..6}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        start = int(10.0 ** rng.uniform(9.0, 18.0))
        length = int(2.0 ** rng.uniform(min_log2_len, max_log2_len))
        length -= length % 30
        out.append((start, max(length, 30), int(rng.integers(1, 7))))
    return out


def draw_sweep_queries(n, seed, min_log10_start=9.0, max_log10_start=15.0, min_log2_len=26.0,
                       max_log2_len=37.0):
    """Single large intervals in the manner of the [1e13, 1e13 + 1e11]
    workload (six counts of one interval per call), swept over starts and
    lengths: start = floor(10^u), u ~ U(9, 15); length = floor(2^v),
    v ~ U(26, 37) (6.7e7 .. 1.4e11), rounded down to a multiple of 30;
    k ~ U{1..6}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        start = int(10.0 ** rng.uniform(min_log10_start, max_log10_start))
        length = int(2.0 ** rng.uniform(min_log2_len, max_log2_len))
        length -= length % 30
        out.append((start, max(length, 30), int(rng.integers(1, 7))))
    return out


def apply_tuning(t):
    """Make the backend use tuning t = (segment bytes, waves[, bucket factor])."""
    sb, waves = t[0], t[1]
    bf = t[2] if len(t) > 2 else 0
    psgpu.set_tuning(sieve_bytes=sb, work_group_size=64 * waves, bucket_factor=bf)


def grid_search_batch(queries, grid=GRID3, sample=512, seed=0, reps=2, log=None):
    """Run-time grid search for a batch: time the persistent batch of a
    seeded sample of `queries` (start, stop, k) under every tuning of `grid`
    (each a new backend; its sieving-prime cache is filled by an untimed call
    first) and return (fastest tuning, {tuning: seconds}).  The backend is
    left with the default tuning."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = np.sort(rng.choice(len(queries), size=min(sample, len(queries)), replace=False))
    qs = [queries[i] for i in idx]
    top = max(b for _, b, _ in qs)
    times = {}
    try:
        for t in grid:
            apply_tuning(t)
            psgpu.count_all(max(top - 29, 0), top)
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                psgpu.count_batch(qs)
                best = min(best, time.perf_counter() - t0)
            times[t] = best
            if log:
                log("grid %s: %.4f s" % (t, best))
    finally:
        psgpu.set_tuning()
    return min(times, key=times.get), times


def collect(queries, grid=GRID, reps=1, log=None, done=None):
    """Time one count call per (query, tuning) on the GPU.  Returns an
    (n_queries, len(grid)) array of device milliseconds (min over reps).
    done(times, gi) is called after each tuning (for partial saves)."""
    times = np.zeros((len(queries), len(grid)), dtype=np.float64)
    top = max(queries, key=lambda q: q[0] + q[1])
    for gi, (sb, waves) in enumerate(grid):
        psgpu.set_tuning(sieve_bytes=sb, work_group_size=64 * waves)
        # a new geometry is a new backend: fill its sieving-prime cache once,
        # up to the largest stop, so that no timed call regenerates it
        psgpu.count(top[0], top[0] + 29, 0)
        t0 = time.perf_counter()
        for qi, (a, length, k) in enumerate(queries):
            best = float("inf")
            for _ in range(reps):
                psgpu.count(a, a + length - 1, k - 1)
                best = min(best, psgpu.last_stats().device_ms)
            times[qi, gi] = best
            if log and qi % 1000 == 999:
                log("config %s: %d queries, %.1f s" % ((sb, waves), qi + 1, time.perf_counter() - t0))
        if log:
            log("config %s: %d queries in %.1f s" % ((sb, waves), len(queries), time.perf_counter() - t0))
        if done:
            done(times, gi)
    psgpu.set_tuning()
    return times


def design_matrix(queries, grid=GRID):
    X = [features(a, length, k, sb, w) for (a, length, k) in queries for (sb, w) in grid]
    return np.asarray(X, dtype=np.float64)


class MLP:
    """The FP32 network on the GPU (psg_mlp_* kernels)."""

    def __init__(self, seed=0):
        self.lib = psgpu.lib()
        self.h = self.lib.psg_mlp_create()
        if not self.h:
            raise psgpu.PrimesieveError(psgpu.last_error())
        self.n = self.lib.psg_mlp_param_count()
        self.mean = np.zeros(6)
        self.std = np.ones(6)
        self.set_params(_he_init(seed, self.n))

    def __del__(self):
        try:
            if self.h:
                self.lib.psg_mlp_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def _fp(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def set_params(self, p):
        a, ptr = self._fp(p)
        assert a.size == self.n
        if self.lib.psg_mlp_set_params(self.h, ptr) != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())

    def get_params(self):
        a = np.zeros(self.n, dtype=np.float32)
        if self.lib.psg_mlp_get_params(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())
        return a

    def forward_raw(self, X):
        x, px = self._fp(X)
        y = np.zeros(x.shape[0], dtype=np.float32)
        if self.lib.psg_mlp_forward(self.h, px, x.shape[0], y.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())
        return y

    def gradients_raw(self, X, T):
        x, px = self._fp(X)
        t, pt = self._fp(T)
        g = np.zeros(self.n, dtype=np.float32)
        l = ctypes.c_float(0)
        if self.lib.psg_mlp_gradients(self.h, px, pt, x.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      ctypes.byref(l)) != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())
        return g, l.value

    def adam_step(self, grads, lr):
        g, pg = self._fp(grads)
        if self.lib.psg_mlp_adam_step(self.h, pg, ctypes.c_float(lr)) != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())

    def fit(self, X, y, epochs=40, batch=256, lr=2e-3, seed=0):
        """Normalise the features (training-set mean / std) and train on
        log2 times.  Returns the per-epoch training loss."""
        X = np.asarray(X, dtype=np.float64)
        self.mean = X.mean(0)
        self.std = np.maximum(X.std(0), 1e-6)
        x, px = self._fp((X - self.mean) / self.std)
        t, pt = self._fp(y)
        losses = np.zeros(epochs, dtype=np.float32)
        rc = self.lib.psg_mlp_train(self.h, px, pt, x.shape[0], epochs, batch, ctypes.c_float(lr), seed,
                                    losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        if rc != 0:
            raise psgpu.PrimesieveError(psgpu.last_error())
        return losses

    def predict(self, X):
        X = np.asarray(X, dtype=np.float64)
        return self.forward_raw((X - self.mean) / self.std).astype(np.float64)

    def save(self, path):
        with open(path, "w") as f:
            json.dump({"params": self.get_params().tolist(), "mean": self.mean.tolist(),
                       "std": self.std.tolist()}, f)

    def load(self, path):
        with open(path) as f:
            d = json.load(f)
        self.set_params(np.asarray(d["params"], dtype=np.float32))
        self.mean = np.asarray(d["mean"])
        self.std = np.asarray(d["std"])


def _he_init(seed, n):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, math.sqrt(2 / 6), 384), np.zeros(64), rng.normal(0, math.sqrt(2 / 64), 4096),
             np.zeros(64), rng.normal(0, math.sqrt(1 / 64), 64), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == n
    return p


def choose(model, queries, grid=GRID):
    """Per query, the grid index with the smallest predicted time."""
    pred = model.predict(design_matrix(queries, grid)).reshape(len(queries), len(grid))
    return pred.argmin(1), pred


def choose_for_batch(model, queries, grid=GRID, min_gain=0.05):
    """One tuning for a whole batch, and the packing order of its queries.

    The model predicts the device time of one count call, which includes a
    per-call cost (launches, the sieving-prime set-up of the call) that a
    persistent batch pays once.  So a query adds to the batch its predicted
    time minus the predicted time of the shortest interval (2^16) at the same
    start; the grid point with the smallest sum wins, but the default
    geometry is kept unless the winner is predicted min_gain faster.  The
    queries are ordered by predicted time, largest first (the order handed to
    the persistent engine)."""
    n, G = len(queries), len(grid)
    pred = model.predict(design_matrix(queries, grid)).reshape(n, G)
    base = model.predict(design_matrix([(a, 1 << 16, k) for a, _, k in queries], grid)).reshape(n, G)
    marginal = np.maximum(np.exp2(pred) - np.exp2(base), 0.0).sum(0)
    g = int(marginal.argmin())
    if DEFAULT in grid:
        d = grid.index(DEFAULT)
        if marginal[g] > (1.0 - min_gain) * marginal[d]:
            g = d
    order = np.argsort(-pred[:, g], kind="stable")
    return grid[g], order


def plan_batch(model, queries, sample=512, seed=0, reps=2, min_gain=0.05):
    """The tuning and packing order of a batch of (start, stop, k).

    The model proposes (choose_for_batch).  It is trained on single large
    intervals, while a batch of short ones spends its time on each query's
    sweep over the sieving primes, so a proposal other than the default
    geometry is confirmed by measurement: both are timed on a seeded sample
    of the batch (grid_search_batch over the two) and the faster is kept.
    Returns ((segment bytes, waves), order, {tuning: sample seconds} or {})."""
    (sb, waves), order = choose_for_batch(model, [(a, b - a + 1, k) for a, b, k in queries],
                                          min_gain=min_gain)
    pick = (sb, waves, DEFAULT3[2])
    times = {}
    if pick != DEFAULT3:
        best, times = grid_search_batch(queries, grid=[DEFAULT3, pick], sample=sample, seed=seed,
                                        reps=reps)
        sb, waves = best[0], best[1]
    return (sb, waves), order, times


def regret(chosen, times, default_index):
    """Regret of choosing `chosen` (per query) against grid search on the
    measured times (n x grid), and the gain over the fixed default."""
    n = times.shape[0]
    t_model = times[np.arange(n), chosen]
    t_best = times.min(1)
    t_default = times[:, default_index]
    return {
        "mean_regret": float(np.mean(t_model / t_best - 1.0)),
        "total_regret": float(t_model.sum() / t_best.sum() - 1.0),
        "default_total_regret": float(t_default.sum() / t_best.sum() - 1.0),
        "gain_over_default": float(t_default.sum() / t_model.sum()),
        "queries": int(n),
    }
