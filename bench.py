"""Benchmark of the gfx950 counting backend on one MI355X.

Workload (a "step"): the single large interval [1e13, 1e13 + 1e11] -- all six
counts (primes, twins, triplets, quadruplets, quintuplets, sextuplets) from
one sieve pass, through the C ABI (psg_gpu_count_all).  The sieving primes
(<= isqrt(1.01e13)) are generated and uploaded during warm-up, so the timed
steps start with every input resident in HBM.

Prints ONE JSON line:
  value      integers sieved per second, (stop - start + 1) * steps / seconds
  roofline   the segment kernel against the measured LDS roofline:
             achieved = algorithmic LDS bytes per launch / average launch
             duration (HIP events on the backend's stream);
             algorithmic LDS bytes = 8 * C + 8 * W, C = crossings by sieving
             primes >= 29, W = 32-bit words of the bit array (4 B fill write +
             4 B count read each); peak = the same 8-byte-per-atomic figure
             measured by the LDS ds_and_b32 microbenchmark
  cpu_baseline  the reference CPU sieve (oracle/_ref, built from the
             reference sources) on the host cores, counting primes over the
             same interval
Every count of every step is checked against the reference's golden values.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "primesieve-gpu_amd"))

START, STOP = 10 ** 13, 10 ** 13 + 10 ** 11
GOLDEN = [3340141707, 147308319, 21308636, 517256, 84202, 2390]   # reference CPU sieve
HBM_PEAK_GBS = 8000.0


def cpu_baseline():
    """The reference CPU library on the host's cores (counting primes over
    the whole interval; the CPU's sieving work is the same for six counts)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libprimesieve_ref.so")
    threads = min(16, os.cpu_count() or 1)
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        L.primesieve_count_primes.restype = ctypes.c_uint64
        L.primesieve_count_primes.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.primesieve_set_num_threads.argtypes = [ctypes.c_int]
        L.primesieve_set_num_threads(threads)
        a, b = START, STOP
        t0 = time.perf_counter()
        n = L.primesieve_count_primes(a, b)
        dt = time.perf_counter() - t0
        if n != GOLDEN[0]:
            raise RuntimeError("reference CPU count mismatch: %d" % n)
        return {"value": (b - a + 1) / dt, "unit": "integers/s", "cores": threads,
                "kind": "reference",
                "sample": "primesieve_count_primes over [1e13, 1e13+1e11], %d threads, %.2f s"
                          % (threads, dt)}
    # fall back to the oracle restatement on a bounded sample
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    o = Oracle()
    a, b = START, START + 10 ** 10
    t0 = time.perf_counter()
    o.count6(a, b, threads)
    dt = time.perf_counter() - t0
    return {"value": (b - a + 1) / dt, "unit": "integers/s", "cores": threads, "kind": "port",
            "sample": "oracle count6 over [1e13, 1e13+1e10], %d threads, %.2f s" % (threads, dt)}


def lds_peak_gbs(m):
    """Measured LDS roofline in the 8-bytes-per-ds_and convention (GB/s)."""
    try:
        r = m.lds_roofline()
        return r["and_conflict_free_gbs"], r
    except Exception as e:  # microbenchmark not available
        return None, {"error": str(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the count path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    import psgpu as m
    if not m.available():
        raise SystemExit("no usable GPU: " + m.last_error())

    for _ in range(max(1, args.warmup)):
        got = m.count_all(START, STOP)
        if got != GOLDEN:
            raise SystemExit("count mismatch in warm-up: %s" % got)

    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got = m.count_all(START, STOP)
        st = m.last_stats()
        kernel_ms.append(st.kernel_ms / max(1, st.launches))
        if got != GOLDEN:
            raise SystemExit("count mismatch: %s" % got)
    dt = time.perf_counter() - t0

    n_int = (STOP - START + 1) * args.steps
    value = n_int / dt
    st = m.last_stats()
    work = m.algorithmic_work(START, STOP, st.sieve_bytes)
    lds_bytes_launch = (8 * work.crossings + 8 * work.words32) / max(1, st.launches)
    avg_launch_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    achieved = lds_bytes_launch / avg_launch_s / 1e9
    peak, lds_info = lds_peak_gbs(m)
    roofline = {"bound": "lds", "achieved": achieved, "peak": peak, "unit": "GB/s",
                "frac": (achieved / peak) if peak else None, "traffic": None,
                "kernel": "psg_sieve_count", "avg_launch_ms": avg_launch_s * 1e3,
                "launches_per_step": st.launches,
                "lds_bytes_per_launch": lds_bytes_launch,
                "crossings": work.crossings, "words32": work.words32}

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
                   "sieving_primes": st.sieving_primes, "segments": st.segments},
        "counts": got,
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
        "lds_microbench": lds_info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
