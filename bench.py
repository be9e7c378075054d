"""Benchmark: the spin_glass_solver local search (single flips, double flips,
Gray seeds, Gray walkers, reheat) on one MI355X.

Workload (one step = one whole local search from a fixed start, until no
phase improves): random 3-regular graph, n = 2048, couplings uniform +-1
(int8), spins +-1; 4096 Gray seeds per Gray phase (block_size 128 x 32
blocks) scored by the int8 MFMA contraction; gray_iterations = n^2.

Prints one JSON line:
  value = flip evaluations per second of the whole local search (single-flip
  gains, pair gains and Gray steps; each is one exact gain evaluation),
  roofline = the dominant kernel (the single-flip descent, k_descent_reg,
  packed keys in registers) against HBM bandwidth; its duration comes from
  HIP events the library records around each descent launch on its own
  stream,
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
# HBM bytes per k_descent_reg launch on this workload, from rocprofv3 FETCH_SIZE
# and WRITE_SIZE passes scaled by a plain copy of known size (see the file)
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_descent_reg_c2_n2048.md")


def measured_traffic():
    """Bytes per launch recorded in the committed profile summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def device_sync():
    """Barrier for the timed region: a device-wide synchronize (one process,
    one GPU, so no inter-rank barrier is needed)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def copy_kernel(mib):
    """A plain device-to-device copy of known size: `mib` MiB read and written
    per launch.  Its FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic
    figures of the other kernels."""
    import torch

    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_mib": mib, "launches": 5}))


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


def time_to_equal_cut(log):
    """64 Erdos-Renyi graphs, n = 8192, mean degree 8, couplings U[-1, 1] (fp32),
    spin-glass mode, from fixed random starts: the plain HIP local search one
    instance at a time (the reference loop per instance) against the batched
    search (all instances per launch, identical moves) and the learned+batched
    search (network-ordered proposals with exact gains, stopping at the plain
    search's final energy; instances that end short are re-solved by the
    batched plain search from their start, so every instance reaches it)."""
    from pyqrackising_amd import _lib, learned

    P = _lib.make_params(0, is_spin_glass=1, walkers=16, seed_multiple=8)
    P.gray_iterations = -1
    solver = learned.LearnedSolver(seed=11)
    t = time.perf_counter()
    solver.fit(learned.regular_training_graphs(48, seed=3001), P, seed=5, steps=300, lr=2e-3)
    train_s = time.perf_counter() - t
    G5 = learned.er_graphs(64, 8192, 8.0, seed=5005)
    dgs = [_lib.DeviceGraph.csr(G) for G in G5]
    rng = np.random.default_rng(77)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in G5]
    solver.bind(dgs)
    # warm up every path once
    dgs[0].local_search(starts[0], _lib.make_params(8192, is_spin_glass=1, walkers=16, seed_multiple=8))
    _lib.batch_local_search(dgs[:2], starts[:2], P)
    t = time.perf_counter()
    plain = [dg.local_search(x, _lib.make_params(dg.n, is_spin_glass=1, walkers=16, seed_multiple=8))
             for dg, x in zip(dgs, starts)]
    t_plain = time.perf_counter() - t
    targets = np.array([st["max_energy"] for _, st in plain])
    log("# time-to-equal-cut: plain sequential done")
    t = time.perf_counter()
    bs, _ = _lib.batch_local_search(dgs, starts, P)
    t_batch = time.perf_counter() - t
    identical = all(np.array_equal(a, b[0]) for a, b in zip(bs, plain))
    t = time.perf_counter()
    _, lst = solver.solve(dgs, starts, P, targets=targets, passes=4, min_score=1.0, bind=False)
    short = [i for i, s in enumerate(lst) if s["max_energy"] < targets[i]]
    if short:
        _lib.batch_local_search([dgs[i] for i in short], [starts[i] for i in short], P)
    t_learned = time.perf_counter() - t
    return {
        "workload": "64 x ER(n=8192, mean degree 8), U[-1,1] fp32 couplings, spin glass, gray_iterations=n^2, "
                    "walkers=16, seed_multiple=8",
        "plain_sequential_s": t_plain,
        "batched_s": t_batch,
        "batched_identical_to_plain": identical,
        "learned_batched_s": t_learned,
        "learned_reached_without_fallback": 64 - len(short),
        "speedup_batched": t_plain / t_batch,
        "speedup_learned_batched": t_plain / t_learned,
        "network_training_s": train_s,
    }


def batched_eval(log):
    """256 instances, n ~ U[256, 4096], d ~ U{3..10}, +-1 int8, 1024 candidates
    each, one graph-captured persistent launch with the autotuned variant."""
    from pyqrackising_amd import _lib, graphs

    Ws = graphs.batched_regular_instances(256, seed=1003)
    dgs = [_lib.DeviceGraph.dense(W) for W in Ws]
    rng = np.random.default_rng(53)
    bits = [rng.integers(0, 1 << 32, (1024, (W.shape[0] + 31) // 32), dtype=np.uint64).astype(np.uint32)
            for W in Ws]
    batch = _lib.Batch(dgs, 1024)
    batch.set_candidates(bits)
    chosen, ms = batch.autotune(reps=3)
    info = batch.info()
    reps = 10
    batch.eval(1)
    t = time.perf_counter()
    for _ in range(reps):
        batch.eval(1)
    dt = (time.perf_counter() - t) / reps
    best_ms = float(ms[chosen])
    log("# batched evaluation done")
    return {
        "workload": "256 instances, n~U[256,4096], d~U{3..10}, +-1 int8, 1024 candidates each",
        "autotune_ms_per_variant": [round(float(v), 4) for v in ms],
        "variant": chosen,
        "kernel_ms": best_ms,
        "call_ms": dt * 1e3,
        "algorithmic_ops": info["algorithmic_ops"],
        "executed_ops": info["executed_ops"],
        "weight_bytes": info["weight_bytes"],
        "mfma_int8_tops": info["algorithmic_ops"] / (best_ms * 1e-3) / 1e12,
        "mfma_frac_of_5000_tops": info["algorithmic_ops"] / (best_ms * 1e-3) / 5.0e15,
    }


def sk_bf16(log):
    """SK couplings w_ij ~ N(0, 1/n), n = 16384, fp32; 1024 random candidates
    scored on split-BF16 MFMA with an FP64 re-score of the best 16."""
    from pyqrackising_amd import _lib

    n, B = 16384, 1024
    rng = np.random.default_rng(43)
    A = rng.standard_normal((n, n), dtype=np.float32) / np.float32(np.sqrt(n))
    W = np.triu(A, 1)
    del A
    W += W.T
    dg = _lib.DeviceGraph.dense(W)
    del W
    bits = _lib.pack_bits(rng.integers(0, 2, (B, n)).astype(bool))
    dg.eval_batch(bits, 1, engine=_lib.ENGINE_BF16)
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        bi, bv, _, err = dg.eval_batch(bits, 1, engine=_lib.ENGINE_BF16, return_error=True)
    dt = (time.perf_counter() - t) / reps
    log("# SK split-BF16 done")
    return {
        "workload": "SK n=16384 fp32, 1024 candidates, split hi/lo BF16 MFMA + FP64 re-score of the best 16",
        "call_ms": dt * 1e3,
        "algorithmic_flop": 2.0 * B * n * n,
        "best_energy_fp64": bv,
        "bf16_max_rel_err_top16": err,
    }


def sampler_round(log, with_cpu):
    """One maxcut_tfim sampling round at the same shape (random 3-regular
    n = 2048, quality 6): roundup32((n << 6) / 2) = 65536 candidates drawn with
    neighbour repulsion (base 5), scored and reduced to the best on the device.
    CPU baseline: the numpy restatement of the reference sampler (vectorised
    over shots) on a bounded sample of 512 shots."""
    from pyqrackising_amd import _lib
    from pyqrackising_amd import sampler as smp

    n = 2048
    W, _ = c2_instance(n)
    Wd = W.astype(np.float64)
    J, z = smp.init_J_and_z_dense(Wd)
    cum = smp.maxcut_hamming_cdf(n, J, z, 6, 8.0, 8.0)
    shots = ((((n << 6) >> 1) + 31) >> 5) << 5
    dg = _lib.DeviceGraph.dense(W)
    dg.sample_round(cum, 5.0, 1, 0, shots, 1)
    reps = 3
    device_sync()
    t = time.perf_counter()
    for r in range(reps):
        _, bv, _ = dg.sample_round(cum, 5.0, 2 + r, 0, shots, 1)
    device_sync()
    dt = (time.perf_counter() - t) / reps
    dg.close()
    log("# sampler round done")
    res = {
        "workload": f"maxcut_tfim sampling round, random 3-regular n={n}, +-1 int8, {shots} shots, repulsion base 5",
        "round_ms": dt * 1e3,
        "shots_per_s": shots / dt,
        "best_energy": bv,
    }
    if with_cpu:
        from oracle.sampler_ref import RepulsionSampler

        cs = 512
        rs = RepulsionSampler(Wd, cum, 5.0, np.random.default_rng(1))
        t = time.perf_counter()
        rs.sample(cs)
        ct = time.perf_counter() - t
        res["cpu_baseline"] = {"value": cs / ct, "unit": "shots/s", "cores": 1, "kind": "port",
                               "sample": f"numpy restatement of the reference sampler, {cs} shots, same prior"}
        res["gpu_over_cpu"] = res["shots_per_s"] / res["cpu_baseline"]["value"]
    return res


def bnb_exact(log, with_cpu):
    """Branch and bound (solve_maxcut_bnb's loop) certifying MAXCUT on random
    +-1 graphs (ER p = 0.5) from the all-zeros incumbent, batch_cap 256: one
    device evaluation per batch of popped nodes.  CPU baseline: the numpy
    restatement of the reference loop on the n = 20 instance (same counts)."""
    from pyqrackising_amd import _lib

    def inst(n):
        rng = np.random.default_rng(n)
        U = np.triu(rng.choice([-1.0, 1.0], (n, n)) * (rng.random((n, n)) < 0.5), 1)
        return U + U.T

    res = {"workload": "MAXCUT branch and bound, ER(n, 0.5) +-1 weights, zero incumbent, batch_cap 256"}
    for n in (20, 24):
        W = inst(n)
        dg = _lib.DeviceGraph.dense(W)
        x0 = np.zeros(n, np.uint8)
        dg.bnb(x0, 0.0, False, 256)
        t = time.perf_counter()
        _, v, st = dg.bnb(x0, 0.0, False, 256)
        dt = time.perf_counter() - t
        dg.close()
        res[f"n{n}"] = {"seconds": dt, "optimum": v, "certified": bool(st["certified"]),
                        "nodes_explored": st["nodes_explored"], "nodes_per_s": st["nodes_explored"] / dt,
                        "batches": st["batches"], "kernel_seconds": st["kernel_seconds"]}
    log("# branch and bound done")
    if with_cpu:
        from oracle import bnb_ref

        W = inst(20)
        t = time.perf_counter()
        _, v, cert, ex, _ = bnb_ref.branch_and_bound(W, np.zeros(20, bool), 0.0, False, 256)
        ct = time.perf_counter() - t
        res["cpu_baseline"] = {"value": ex / ct, "unit": "nodes/s", "cores": 1, "kind": "port",
                               "sample": "numpy restatement of the reference loop, the n=20 instance",
                               "same_result": bool(v == res["n20"]["optimum"] and ex == res["n20"]["nodes_explored"])}
        res["gpu_over_cpu_n20"] = res["n20"]["nodes_per_s"] / res["cpu_baseline"]["value"]
    return res


# the extra legs, each (log, with_cpu) -> dict
LEGS = {
    "time_to_equal_cut": lambda lg, cpu: time_to_equal_cut(lg),
    "batched_eval": lambda lg, cpu: batched_eval(lg),
    "sk_bf16": lambda lg, cpu: sk_bf16(lg),
    "sampler": lambda lg, cpu: sampler_round(lg, cpu),
    "branch_and_bound": lambda lg, cpu: bnb_exact(lg, cpu),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--walkers", type=int, default=128)
    ap.add_argument("--seed-multiple", type=int, default=1)
    ap.add_argument("--gray-iterations", type=int, default=None)
    ap.add_argument("--cpu-gray-iterations", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched, learned and SK legs")
    ap.add_argument("--copy-kernel", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    ap.add_argument("--only", choices=sorted(LEGS), default=None,
                    help="run one extra leg alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.copy_kernel:
        copy_kernel(args.copy_kernel)
        return
    if args.only:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        print(json.dumps({args.only: LEGS[args.only](log, not args.no_cpu)}))
        return
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
    device_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = dg.local_search(x0, params)
        stats.append(st)
    device_sync()
    dt = time.perf_counter() - t0
    for i, st in enumerate(stats):
        print(f"# step {i}: cut={st['cut_value']:.0f} rounds={st['rounds']}", file=sys.stderr, flush=True)

    st = stats[-1]
    # computed gain evaluations only; Gray steps that repeat an already rejected
    # step of an unchanged walker are resolved without computing and not counted
    evals = st["single_evals"] + st["double_evals"] + st["gray_evals_computed"]
    total_evals = sum(s["single_evals"] + s["double_evals"] + s["gray_evals_computed"] for s in stats)
    total_ref = sum(s["single_evals"] + s["double_evals"] + s["gray_evals"] for s in stats)
    value = total_evals / dt
    ms_per_step = dt / args.steps * 1e3

    # dominant kernel: the single-flip descent (k_descent_reg).  Algorithmic
    # bytes per flip evaluation: the gain (4 B int32) and the spin (1 B); per
    # accepted flip, 16 B per neighbour (index and weight read, gain read and
    # written) -- degree 3 here.
    desc_s = sum(s["descent_kernel_seconds"] for s in stats)
    desc_launches = sum(s["descent_launches"] for s in stats)
    desc_evals = sum(s["descent_evals"] for s in stats)
    desc_flips = sum(s["single_accepts"] for s in stats)
    avg_launch_s = desc_s / max(1, desc_launches)
    bytes_per_launch = (5.0 * desc_evals + 16.0 * 3 * desc_flips) / max(1, desc_launches)
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
            "kernel": "k_descent_reg<32, 4>",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "traffic_source": "profiles/k_descent_reg_c2_n2048.md",
            "avg_launch_ms": avg_launch_s * 1e3,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    if not args.no_extra:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        for key, fn in LEGS.items():
            try:
                out[key] = fn(log, not args.no_cpu)
            except Exception as e:  # a failing leg is reported, not fatal to the headline
                out[key] = {"error": repr(e)}
    if not args.no_cpu:
        cgi = n * n if args.cpu_gray_iterations is None else args.cpu_gray_iterations
        cb = cpu_baseline(W, x0, 16, 8, cgi)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
