"""Collect the cost model's training and held-out timings on the GPU.

Draws seeded uneven intervals (costmodel.draw_queries), times one count call
per (interval, geometry) over the tuning grid with the backend's own HIP
events, and writes data/costmodel_timings.npz:

  q_train, q_test   (n, 3) uint64: start, length, k
  t_train, t_test   (n, len(grid)) float64 device milliseconds
  grid              (len(grid), 2): segment bytes, waves per workgroup

Usage: python primesieve-gpu_amd/collect_timings.py [--train 4096] [--test 2048]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import costmodel  # noqa: E402
import psgpu      # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--test", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2027)
    ap.add_argument("--out", default=os.path.join(HERE, "data", "costmodel_timings.npz"))
    args = ap.parse_args()
    if not psgpu.available():
        raise SystemExit("no usable GPU: " + psgpu.last_error())
    qs = costmodel.draw_queries(args.train + args.test, args.seed)
    # warm the sieving-prime cache up to the largest stop
    a, length, _ = max(qs, key=lambda q: q[0] + q[1])
    psgpu.count_all(a, a + length - 1)
    t0 = time.perf_counter()
    times = costmodel.collect(qs, log=lambda s: print(s, flush=True))
    print("collected %d timings in %.1f s" % (times.size, time.perf_counter() - t0), flush=True)
    q = np.asarray(qs, dtype=np.uint64)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.savez_compressed(args.out, q_train=q[:args.train], q_test=q[args.train:],
                        t_train=times[:args.train], t_test=times[args.train:],
                        grid=np.asarray(costmodel.GRID, dtype=np.int64))
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
