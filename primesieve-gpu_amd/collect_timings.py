"""Collect the cost model's training and held-out timings on the GPU.

Draws seeded intervals -- by default single large intervals swept over
starts 1e9..1e15 and lengths 2^26..2^37 (costmodel.draw_sweep_queries), or
the uneven batch intervals (--style batch) -- times one count call per
(interval, geometry) over the tuning grid with the backend's own HIP events,
and writes data/costmodel_timings.npz:

  q_train, q_test   (n, 3) uint64: start, length, k
  t_train, t_test   (n, len(grid)) float64 device milliseconds
  grid              (len(grid), 2): segment bytes, waves per workgroup

Usage: python primesieve-gpu_amd/collect_timings.py [--train 4096] [--test 2048] [--style sweep|batch]
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
    ap.add_argument("--style", choices=["sweep", "batch"], default="sweep",
                    help="sweep: single large intervals (costmodel.draw_sweep_queries); "
                         "batch: the uneven batch intervals (costmodel.draw_queries)")
    args = ap.parse_args()
    if not psgpu.available():
        raise SystemExit("no usable GPU: " + psgpu.last_error())
    draw = costmodel.draw_sweep_queries if args.style == "sweep" else costmodel.draw_queries
    qs = draw(args.train + args.test, args.seed)
    # warm the sieving-prime cache up to the largest stop
    a, length, _ = max(qs, key=lambda q: q[0] + q[1])
    psgpu.count_all(a, a + length - 1)
    t0 = time.perf_counter()
    q = np.asarray(qs, dtype=np.uint64)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)

    def save(times, gi, path=args.out + ".partial.npz"):
        np.savez_compressed(path, q=q, times=times, done=gi + 1,
                            grid=np.asarray(costmodel.GRID, dtype=np.int64))

    times = costmodel.collect(qs, log=lambda s: print(s, flush=True), done=save)
    print("collected %d timings in %.1f s" % (times.size, time.perf_counter() - t0), flush=True)
    np.savez_compressed(args.out, q_train=q[:args.train], q_test=q[args.train:],
                        t_train=times[:args.train], t_test=times[args.train:],
                        grid=np.asarray(costmodel.GRID, dtype=np.int64))
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
