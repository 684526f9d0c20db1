"""Benchmark of the gfx950 counting backend on one MI355X.

Headline workload (a "step"): the single large interval [1e13, 1e13 + 1e11]
-- all six counts (primes, twins, triplets, quadruplets, quintuplets,
sextuplets) from one sieve pass, through the C ABI (psg_gpu_count_all).  The
sieving primes (<= isqrt(1.01e13)) are generated and uploaded during warm-up,
so the timed steps start with every input resident in HBM.  Every count of
every step is checked against the reference's values.

Prints ONE JSON line:
  value         integers sieved per second, (stop - start + 1) * steps / seconds
  roofline      the segment kernel (psg_sieve_count) against the measured LDS
                roofline: achieved = algorithmic LDS bytes per launch / average
                launch duration (HIP events on the backend's stream), with
                algorithmic LDS bytes = 8*C + 8*W, the measurement
                contract's fixed convention for this path (C = crossings by
                the sieving primes >= 29: one 4-byte read + one 4-byte write
                each; W = 32-bit words of the bit array: one fill write +
                one count read); peak = the same 8-bytes-per-ds_or_b32
                figure (the sieve's strike) measured by the conflict-free
                LDS microbenchmark on the same device;
                traffic = HBM bytes per launch from the rocprofv3 FETCH_SIZE /
                WRITE_SIZE summary in profiles/ when present;
                frac_struck_p53 = the stricter fraction with C over the
                primes this kernel strikes in LDS (p >= 53: 7..47 are
                pre-sieve tiles, whose crossings cost no atomics)
  cpu_baseline  the reference CPU sieve (oracle/_ref, built from the
                reference sources) on the host's CPU share (16 threads),
                counting all six types over the same interval in one pass;
                the host's core count, quota and CPU model are recorded
  workloads     the other configurations: pi(1e10); 8192 uneven intervals
                batched vs one call per interval (intervals/s), with and
                without the cost model; the cost model's training and regret;
                the Poisson serving loop
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "primesieve-gpu_amd")
sys.path.insert(0, PKG)

START, STOP = 10 ** 13, 10 ** 13 + 10 ** 11
GOLDEN = [3340141707, 147308319, 21308636, 517256, 84202, 2390]   # reference CPU sieve
PI_1E10 = 455052511
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "psg_sieve_count_interval_1e13_pmc.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """What the host offers: logical CPUs, the CPUs this process may run on,
    the cgroup CPU quota and the CPU model."""
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        q = open("/sys/fs/cgroup/cpu.max").read().split()
        if q and q[0] != "max":
            info["cgroup_cpu_quota"] = float(q[0]) / float(q[1])
    except (OSError, ValueError, IndexError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return info


# Worker threads of the CPU baseline: the host's CPU share for one GPU on
# these machines is 16 cores, whatever os.cpu_count() reports.
CPU_SHARE = 16


def cpu_baseline():
    """The reference CPU library (oracle/_ref, built from the reference
    sources) counting all six types over the whole interval in ONE pass --
    its ParallelSieve with the six COUNT_* flags, as `primesieve 1e13
    --dist=1e11 --count=123456` runs -- on the host's CPU share."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libprimesieve_ref.so")
    host = host_cpu()
    threads = max(1, min(CPU_SHARE, host.get("affinity_cpus") or host["os_cpu_count"] or 1))
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        L.psg_ref_count6.restype = None
        L.psg_ref_count6.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        out = (ctypes.c_uint64 * 6)()
        t0 = time.perf_counter()
        L.psg_ref_count6(START, STOP, out, threads)
        dt = time.perf_counter() - t0
        if list(out) != GOLDEN:
            raise RuntimeError("reference CPU counts mismatch: %s" % list(out))
        return {"value": (STOP - START + 1) / dt, "unit": "integers/s", "cores": threads,
                "kind": "reference", "host": host, "seconds": dt,
                "sample": "reference ParallelSieve, all six counts of [1e13, 1e13+1e11] (the whole "
                          "interval) in one pass, %d threads, %.2f s" % (threads, dt)}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    a, b = START, START + 10 ** 10
    t0 = time.perf_counter()
    Oracle().count6(a, b, threads)
    dt = time.perf_counter() - t0
    return {"value": (b - a + 1) / dt, "unit": "integers/s", "cores": threads, "kind": "port",
            "host": host, "seconds": dt,
            "sample": "oracle count6 over [1e13, 1e13+1e10], %d threads, %.2f s" % (threads, dt)}


def profiled_traffic():
    """HBM bytes per psg_sieve_count launch from the committed PMC summary."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def tile_prime_crossings(lo, hi):
    """Crossings of the pre-sieved primes 29..47 in [lo, hi]: for each p the
    cofactors q coprime to 30 with max(p, ceil(lo/p)) <= q <= floor(hi/p),
    counted in closed form (cop(n) = 8 floor(n/30) + #coprime residues <= n mod 30).
    Added to the crossings of the primes struck in LDS (p >= 53) this gives C
    over p >= 29, the convention of the measurement contract for this path."""
    res, c = [], 0
    for r in range(30):
        c += 1 if r and all(r % d for d in (2, 3, 5)) else 0
        res.append(c)

    def cop(n):
        return 0 if n <= 0 else 8 * (n // 30) + res[n % 30]
    total = 0
    for p in (29, 31, 37, 41, 43, 47):
        a, b = max(p, -(-lo // p)), hi // p
        total += cop(b) - cop(a - 1) if b >= a else 0
    return total


def headline(m, steps, warmup):
    for _ in range(max(1, warmup)):
        if m.count_all(START, STOP) != GOLDEN:
            raise SystemExit("count mismatch in warm-up")
    launch_ms, launches = [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        got = m.count_all(START, STOP)
        st = m.last_stats()
        launch_ms.append(st.kernel_ms / max(1, st.launches))
        launches = st.launches
        if got != GOLDEN:
            raise SystemExit("count mismatch: %s" % got)
    dt = time.perf_counter() - t0
    return got, dt, m.last_stats(), sum(launch_ms) / len(launch_ms), launches


def workload_pi(m):
    t0 = time.perf_counter()
    for _ in range(3):
        n = m.count(0, 10 ** 10, 0)
        if n != PI_1E10:
            raise SystemExit("pi(1e10) mismatch: %d" % n)
    dt = (time.perf_counter() - t0) / 3
    return {"workload": "pi(1e10), primes in [0, 1e10]", "count": n, "ms": dt * 1e3,
            "integers_per_s": (10 ** 10 + 1) / dt, "device_ms": m.last_stats().device_ms}


def workload_batch(m, model):
    q = np.load(os.path.join(ROOT, "tests", "golden", "config3_queries.npy"), allow_pickle=False)
    qs = [(int(a), int(b), int(k)) for a, b, k in q]
    res = {"workload": "8192 uneven intervals, start 1e9..1e18, length 2^16..2^32, k 1..6"}
    # cold: a fresh backend builds its sieving-prime cache (to ~1e9, on the
    # GPU) inside the timed call
    m.set_tuning()
    t0 = time.perf_counter()
    cold = m.count_batch(qs)
    res["cold_batched_s"] = time.perf_counter() - t0
    res["cold_batched_intervals_per_s"] = len(qs) / res["cold_batched_s"]
    # warm: the cache stays resident between calls
    t0 = time.perf_counter()
    batched = m.count_batch(qs)
    res["batched_s"] = time.perf_counter() - t0
    if cold != batched:
        raise SystemExit("cold and warm batched counts differ")
    res["batched_stats"] = {k: v for k, v in m.last_stats().as_dict().items()
                            if k in ("kernel_ms", "fill_ms", "device_ms", "launches", "segments")}
    # the same batch through a captured launch series: the first call
    # captures it, the second replays it
    t0 = time.perf_counter()
    g1 = m.count_batch(qs, mode=m.BATCH_GRAPH)
    res["graph_capture_batched_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    g2 = m.count_batch(qs, mode=m.BATCH_GRAPH)
    res["graph_batched_s"] = time.perf_counter() - t0
    if g1 != batched or g2 != batched or m.last_stats().graph != 1:
        raise SystemExit("graph-replayed batch differs from the eager batch")
    res["graph_batched_intervals_per_s"] = len(qs) / res["graph_batched_s"]
    t0 = time.perf_counter()
    plain = m.count_batch(qs, mode=m.BATCH_PLAIN)
    res["plain_s"] = time.perf_counter() - t0
    if plain != batched:
        raise SystemExit("batched and plain counts differ")
    exp = json.load(open(os.path.join(ROOT, "tests", "golden", "config3_expected.json")))
    exp.update(json.load(open(os.path.join(ROOT, "tests", "golden", "config3_decades.json"))))
    bad = [i for i, v in exp.items() if batched[int(i)] != v]
    if bad:
        raise SystemExit("batched counts differ from the reference on %d queries" % len(bad))
    res["checked_against_reference"] = len(exp)
    res["batched_intervals_per_s"] = len(qs) / res["batched_s"]
    res["plain_intervals_per_s"] = len(qs) / res["plain_s"]
    res["speedup"] = res["plain_s"] / res["batched_s"]
    # run-time grid search (the fallback tuner): segment bytes x waves x bucket
    # threshold, timed on a seeded 512-query sample of the batch
    import costmodel
    t0 = time.perf_counter()
    gt, gtimes = costmodel.grid_search_batch(qs, log=log)
    tune_s = time.perf_counter() - t0
    costmodel.apply_tuning(gt)
    m.count_all(10 ** 18, 10 ** 18 + 2 ** 33)
    t0 = time.perf_counter()
    gb = m.count_batch(qs)
    grid_s = time.perf_counter() - t0
    m.set_tuning()
    if gb != batched:
        raise SystemExit("grid-tuned batch counts differ")
    res["grid_search"] = {"choice": {"sieve_bytes": gt[0], "waves": gt[1], "bucket_factor": gt[2]},
                          "tune_s": tune_s, "sample_queries": 512, "grid_points": len(gtimes),
                          "batched_s": grid_s, "batched_intervals_per_s": len(qs) / grid_s,
                          "sample_s_default": gtimes[costmodel.DEFAULT3], "sample_s_best": gtimes[gt]}
    if model is not None:
        t0 = time.perf_counter()
        (sb, waves), order, confirm = costmodel.plan_batch(model, qs)
        m.set_tuning(sieve_bytes=sb, work_group_size=64 * waves)
        m.count_all(10 ** 18, 10 ** 18 + 10 ** 6)
        plan_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        tuned = m.count_batch([qs[i] for i in order])
        res["model_batched_s"] = time.perf_counter() - t0
        m.set_tuning()
        back = [0] * len(qs)
        for j, i in enumerate(order):
            back[i] = tuned[j]
        if back != batched:
            raise SystemExit("model-packed counts differ")
        res["model_choice"] = {"sieve_bytes": sb, "waves": waves, "plan_s": plan_s,
                               "confirmed_on_sample": {"%d/%d/%d" % t: s for t, s in confirm.items()}}
        res["model_batched_intervals_per_s"] = len(qs) / res["model_batched_s"]
        res["model_speedup_vs_plain"] = res["plain_s"] / res["model_batched_s"]
    return res


def workload_batched_once(m):
    """One warm persistent batch of the 8192 uneven intervals (for profilers:
    the sieving-prime cache is built by an untimed call first)."""
    q = np.load(os.path.join(ROOT, "tests", "golden", "config3_queries.npy"), allow_pickle=False)
    qs = [(int(a), int(b), int(k)) for a, b, k in q]
    m.count_all(10 ** 18, 10 ** 18 + 2 ** 33)
    t0 = time.perf_counter()
    got = m.count_batch(qs)
    dt = time.perf_counter() - t0
    exp = json.load(open(os.path.join(ROOT, "tests", "golden", "config3_expected.json")))
    if any(got[int(i)] != v for i, v in exp.items()):
        raise SystemExit("batched counts differ from the reference")
    st = m.last_stats().as_dict()
    return {"workload": "8192 uneven intervals, one warm persistent batch", "batched_s": dt,
            "batched_intervals_per_s": len(qs) / dt,
            "stats": {k: st[k] for k in ("kernel_ms", "fill_ms", "device_ms", "launches", "segments",
                                         "bucket_entries")}}


def workload_costmodel():
    path = os.path.join(PKG, "data", "costmodel_timings.npz")
    if not os.path.exists(path):
        return None, None
    import costmodel
    d = np.load(path, allow_pickle=False)
    grid = [tuple(int(v) for v in g) for g in d["grid"]]
    qtr = [tuple(int(v) for v in q) for q in d["q_train"]]
    qte = [tuple(int(v) for v in q) for q in d["q_test"]]
    X = costmodel.design_matrix(qtr, grid)
    y = np.log2(d["t_train"].reshape(-1))
    net = costmodel.MLP(seed=0)
    t0 = time.perf_counter()
    losses = net.fit(X, y, epochs=60, batch=256, lr=2e-3, seed=0)
    train_s = time.perf_counter() - t0
    chosen, pred = costmodel.choose(net, qte, grid)
    r = costmodel.regret(chosen, d["t_test"], grid.index(costmodel.DEFAULT))
    test_mse = float(np.mean((pred.reshape(-1) - np.log2(d["t_test"].reshape(-1))) ** 2))
    return net, {"workload": "MLP 6-64-64-1 FP32, %d training timings, %d held-out queries x %d geometries"
                 % (X.shape[0], len(qte), len(grid)),
                 "train_s": train_s, "epochs": len(losses), "final_train_loss": float(losses[-1]),
                 "heldout_mse_log2": test_mse, **r}


def workload_serving(m, model, seconds=4.0, qps=2000.0):
    import serve
    n = int(seconds * qps)
    arrivals, qs = serve.poisson_workload(n, qps, seed=5)
    m.count_all(10 ** 18, 10 ** 18 + 10 ** 6)
    counts, lat, wall, batches = serve.serve(arrivals, qs, model=model)
    gcounts, glat, gwall, gbatches = serve.serve(arrivals, qs, model=model, mode=m.BATCH_GRAPH)
    check = m.count_batch(qs, mode=m.BATCH_PLAIN)
    if check != counts or gcounts != counts:
        raise SystemExit("served counts differ from the plain path")
    exp = json.load(open(os.path.join(ROOT, "tests", "golden", "config5_expected.json")))
    exp = exp.get("5:%d" % n, {})
    bad = [i for i, v in exp.items() if counts[int(i)] != v]
    if bad:
        raise SystemExit("served counts differ from the reference on %d queries" % len(bad))
    s = serve.summary(lat, wall, n, batches)
    s["workload"] = "Poisson %.0f QPS for %.0f s, Pareto(1.2) lengths in [2^12, 2^36], k 1..6" % (qps, seconds)
    s["counts_checked"] = ("all %d against the plain per-query path, %d against the reference CPU library"
                           % (n, len(exp)))
    g = serve.summary(glat, gwall, n, gbatches)
    s["graph_replay"] = {k: g[k] for k in ("throughput_qps", "p50_ms", "p99_ms", "max_ms", "batches")}
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads")
    ap.add_argument("--only", choices=["pi_1e10", "cost_model", "batched_intervals", "batched_once", "serving"],
                    help="run just this workload (no headline) and print its JSON, e.g. under a profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the count path runs on a single MI355X; "
              "sharding an interval over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import psgpu as m
    if not m.available():
        raise SystemExit("no usable GPU: " + m.last_error())
    if args.only:
        model = None
        if args.only in ("cost_model", "batched_intervals", "serving"):
            model, cm = workload_costmodel()
        res = {"pi_1e10": lambda: workload_pi(m), "cost_model": lambda: cm,
               "batched_intervals": lambda: workload_batch(m, model),
               "batched_once": lambda: workload_batched_once(m),
               "serving": lambda: workload_serving(m, model)}[args.only]()
        print(json.dumps({"workload": args.only, "result": res}), flush=True)
        return

    got, dt, st, avg_launch_ms, launches = headline(m, args.steps, args.warmup)
    value = (STOP - START + 1) * args.steps / dt
    work = m.algorithmic_work(START, STOP, st.sieve_bytes)
    seconds = avg_launch_ms / 1e3
    # C over the sieving primes >= 29 (the convention), and over the primes
    # struck in LDS (>= 53; 29..47 are pre-sieve tiles)
    crossings = work.crossings + tile_prime_crossings(max(START, 7), STOP)
    lds_bytes_launch = (8 * crossings + 8 * work.words32) / max(1, launches)
    struck_bytes_launch = (8 * work.crossings + 8 * work.words32) / max(1, launches)
    achieved = lds_bytes_launch / seconds / 1e9
    achieved_struck = struck_bytes_launch / seconds / 1e9
    lds = m.lds_roofline()
    peak = lds["or_conflict_free_gbs"]          # the sieve strikes with ds_or_b32
    roofline = {"bound": "lds", "achieved": achieved, "peak": peak, "unit": "GB/s",
                "frac": achieved / peak, "traffic": profiled_traffic(),
                "kernel": "psg_sieve_count", "avg_launch_ms": avg_launch_ms,
                "launches_per_step": launches, "lds_bytes_per_launch": lds_bytes_launch,
                "crossings": crossings, "crossings_primes_from": 29, "words32": work.words32,
                "peak_source": "measured ds_or_b32 microbenchmark (psg_lds_and mode 2, conflict-free)",
                "crossings_struck_p53": work.crossings,
                "lds_bytes_struck_per_launch": struck_bytes_launch,
                "achieved_struck": achieved_struck,
                "frac_struck_p53": achieved_struck / peak,
                # the same strike with random word addresses, the access pattern of a scattered sieve
                "peak_random_address": lds["or_random_gbs"],
                "frac_struck_vs_random_address_peak": achieved_struck / lds["or_random_gbs"],
                "measured": lds}
    out = {
        "metric": "integers sieved/s on 1 MI355X, exact counts",
        "value": value,
        "unit": "integers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic (fixed interval)",
        "config": {"workload": "single interval [1e13, 1e13+1e11], six counts in one pass",
                   "sieve_bytes": st.sieve_bytes, "work_group_size": st.work_group_size,
                   "sieving_primes": st.sieving_primes, "segments": st.segments,
                   "bucket_entries": st.bucket_entries,
                   "segment_kernel_ms": st.kernel_ms, "bucket_kernels_ms": st.fill_ms},
        "counts": got,
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    if not args.headline_only:
        wl = {}
        log("workload: pi(1e10)")
        wl["pi_1e10"] = workload_pi(m)
        log("workload: cost model")
        model, wl["cost_model"] = workload_costmodel()
        log("workload: batched intervals")
        wl["batched_intervals"] = workload_batch(m, model)
        log("workload: serving")
        wl["serving"] = workload_serving(m, model)
        out["workloads"] = wl
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
