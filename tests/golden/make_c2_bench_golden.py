"""Writes tests/golden/c2_bench_config.json: the oracle's local search at the
benchmark's configuration (bench.py's c2_instance(): random 3-regular graph,
n = 2048, +-1 couplings, seed 1002; spin-glass mode, 128 walkers x 1, i.e.
4096 Gray seeds per phase, gray_iterations = n^2).  The oracle evaluates every
Gray step (about 4e8 of them), which takes minutes, so the GPU test compares
against this record instead of re-running it.

Run from the repository root: python tests/golden/make_c2_bench_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from bench import c2_instance  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def instance_digest(W, x0):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(W, dtype=np.int8).tobytes())
    h.update(np.ascontiguousarray(x0, dtype=np.uint8).tobytes())
    return h.hexdigest()


def main():
    W, x0 = c2_instance()
    n = W.shape[0]
    t = time.perf_counter()
    x, r = orc.dense_local_search(W.astype(np.float64), x0, is_spin_glass=1, walkers=128, seed_multiple=1,
                                  gray_iterations=n * n)
    dt = time.perf_counter() - t
    out = {
        "workload": "c2_instance(): random 3-regular n=2048, +-1, seed 1002; spin glass, walkers=128, "
                    "seed_multiple=1, gray_iterations=n^2",
        "n": n,
        "instance_sha256": instance_digest(W, x0),
        "state_packed_hex": np.packbits(x.astype(np.uint8), bitorder="little").tobytes().hex(),
        "result": {k: (float(v) if isinstance(v, float) else int(v)) for k, v in r.items()},
        "oracle_seconds": dt,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_bench_config.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path} ({dt:.1f} s of oracle time)")


if __name__ == "__main__":
    main()
