"""Benchmark: the spin_glass_solver local search (single flips, double flips,
Gray seeds, Gray walkers, reheat) on one MI355X.

Workload (one step = one whole local search from a fixed start, until no
phase improves): random 3-regular graph, n = 2048, couplings uniform +-1
(int8), spins +-1; 4096 Gray seeds per Gray phase (block_size 128 x 32
blocks) scored by the int8 MFMA contraction; gray_iterations = n^2.

Prints one JSON line:
  value = flip evaluations per second of the whole local search (single-flip
  gains, pair gains and Gray steps; each is one exact gain evaluation),
  roofline = the dominant kernel (the Gray walker) against HBM bandwidth,
  cpu_baseline = the CPU oracle (a C restatement of the reference CPU path with
  the reference's O(n)-per-evaluation cost) on a bounded sample of the same
  instance, timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def c2_instance(n=2048, seed=1002):
    from pyqrackising_amd import graphs

    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, 3, rng)
    W = graphs.dense_from_edges(n, lo, hi, graphs.pm1(lo.size, rng), dtype=np.int8)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    return W, x0


def cpu_baseline(W, x0, walkers, seed_multiple, gray_iterations):
    """Bounded sample: the oracle's local search on the same instance and start
    with the Gray phase capped; flip evaluations per second."""
    from oracle import oracle as orc

    threads = os.cpu_count() or 1
    t = time.perf_counter()
    _, r = orc.dense_local_search(W.astype(np.float64), x0, is_spin_glass=1, walkers=walkers,
                                  seed_multiple=seed_multiple, gray_iterations=gray_iterations)
    dt = time.perf_counter() - t
    evals = r["single_evals"] + r["double_evals"] + r["gray_evals"]  # the oracle computes every step
    return {
        "value": evals / dt,
        "unit": "flip evals/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle dense_local_search on the same instance and start, gray_iterations={gray_iterations} "
                   f"(GPU run: n^2), walkers={walkers}, seed_multiple={seed_multiple}; {evals} flip evals in "
                   f"{dt:.2f} s; final cut {r['cut_value']:.0f}"),
        "seconds": dt,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--walkers", type=int, default=128)
    ap.add_argument("--seed-multiple", type=int, default=1)
    ap.add_argument("--gray-iterations", type=int, default=None)
    ap.add_argument("--cpu-gray-iterations", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from pyqrackising_amd import _lib

    n = args.n
    W, x0 = c2_instance(n)
    gi = n * n if args.gray_iterations is None else args.gray_iterations
    dg = _lib.DeviceGraph.dense(W)
    params = _lib.make_params(n, is_spin_glass=1, gray_iterations=gi, walkers=args.walkers,
                              seed_multiple=args.seed_multiple, seed_engine=_lib.SEEDS_MFMA,
                              flags=_lib.FLAG_TIME_KERNELS)
    for i in range(args.warmup):
        dg.local_search(x0, params)
        print(f"# warmup {i} done", file=sys.stderr, flush=True)

    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = dg.local_search(x0, params)
        stats.append(st)
        print(f"# step {i}: cut={st['cut_value']:.0f} rounds={st['rounds']}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0

    st = stats[-1]
    # computed gain evaluations only; Gray steps that repeat an already rejected
    # step of an unchanged walker are resolved without computing and not counted
    evals = st["single_evals"] + st["double_evals"] + st["gray_evals_computed"]
    total_evals = sum(s["single_evals"] + s["double_evals"] + s["gray_evals_computed"] for s in stats)
    total_ref = sum(s["single_evals"] + s["double_evals"] + s["gray_evals"] for s in stats)
    value = total_evals / dt
    ms_per_step = dt / args.steps * 1e3

    # dominant kernel: the Gray walker.  Algorithmic bytes per Gray step: the
    # gain's field entry (4 B int32) and spin (1 B); per accepted flip the
    # walker also reads an adjacency row, which is not counted.
    walk_s = sum(s["walk_kernel_seconds"] for s in stats)
    walk_launches = sum(s["walk_launches"] for s in stats)
    gray_evals = sum(s["gray_evals_computed"] for s in stats)
    avg_launch_s = walk_s / max(1, walk_launches)
    bytes_per_launch = 5.0 * gray_evals / max(1, walk_launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0

    out = {
        "metric": "time-to-equal-cut & flip evals/s, learned+batched vs plain HIP, 1 MI355X",
        "value": value,
        "unit": "flip evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8",
        "data": "synthetic",
        "config": {
            "workload": (f"spin_glass_solver local search, random 3-regular n={n}, +-1 int8 couplings, "
                         f"{args.walkers * args.seed_multiple * ((n + 63) // 64)} Gray seeds per phase on int8 MFMA, "
                         f"walkers={args.walkers}, gray_iterations={gi}"),
            "n": n,
            "walkers": args.walkers,
            "seed_multiple": args.seed_multiple,
            "gray_iterations": gi,
            "time_to_cut_s": ms_per_step / 1e3,
            "final_cut": st["cut_value"],
            "min_energy": st["min_energy"],
            "flip_evals_per_step": evals,
            "reference_flip_steps_per_step": st["single_evals"] + st["double_evals"] + st["gray_evals"],
            "reference_flip_steps_per_s": total_ref / dt,
            "seed_candidates_per_step": st["seed_evals"],
            "phase_seconds": {k: st[k] for k in ("seconds_single", "seconds_double", "seconds_seeds",
                                                  "seconds_walk", "seconds_reheat")},
        },
        "roofline": {
            "kernel": "k_gray_walk",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_launch_s * 1e3,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cgi = n * n // 4 if args.cpu_gray_iterations is None else args.cpu_gray_iterations
        cb = cpu_baseline(W, x0, 16, 8, cgi)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
