"""Public paths against the oracle:

* the exact small-n path of maxcut_tfim / maxcut_tfim_sparse in spin-glass
  mode (pyqrackising/maxcut_tfim.py:315-339, 364-372): the 2^n brute force of
  -E with a strict '>' from -1, against the oracle's brute force;
* maxcut_tfim_sparse above its threshold (pyqrackising/maxcut_tfim_sparse.py:
  159-201, 373-472): the device sampling rounds replayed with the same seeded
  candidates, scored and reduced by the oracle;
* the local search at exactly the benchmark's configuration: random 3-regular
  n = 2048, +-1 int8, 128 walkers x 1 (4096 Gray seeds per phase on int8
  MFMA), gray_iterations = n^2, through the default device-decided fast path.
"""
import os
import sys

import numpy as np
import pytest

from pyqrackising_amd import _lib, graphs

pytestmark = pytest.mark.gpu

REL = 1e-9
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert _lib.device_count() > 0, "no GPU visible"
    return _lib.default_context()


def er_dense(n, p, seed, kind):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges(n, p, rng)
    w = rng.integers(-3, 4, lo.size).astype(np.float64) if kind == "int" else rng.uniform(-1.0, 1.0, lo.size)
    return graphs.dense_from_edges(n, lo, hi, w, dtype=np.float64)


@pytest.mark.parametrize("n,kind,dtype", [(8, "int", np.int32), (14, "int", np.float32), (18, "int", np.int8),
                                          (12, "real", np.float64), (16, "real", np.float32),
                                          (20, "real", np.float32)])
def test_exact_spin_glass_matches_oracle(ctx, orc, n, kind, dtype):
    from pyqrackising_amd import maxcut_tfim, maxcut_tfim_sparse

    log(f"exact spin glass n={n} {kind} {np.dtype(dtype).name}")
    W = er_dense(n, 0.5, 40 + n, kind).astype(dtype).astype(np.float64)
    obits, ov = orc.exact(W, 1)
    expect = "".join("1" if b else "0" for b in obits)
    iu, ju = np.nonzero(np.triu(W, 1))
    G = graphs.upper_csr(n, iu, ju, W[iu, ju], dtype=np.float64)
    for fn, arg in ((maxcut_tfim, W.astype(dtype)), (maxcut_tfim_sparse, G)):
        bits, value, (l, r) = fn(arg, is_spin_glass=True)
        assert bits == expect, fn.__name__
        if kind == "int":
            assert value == ov
        else:
            assert value == pytest.approx(ov, rel=REL)
        assert sorted(l + r) == list(range(n))
        # -E of the returned state is the value (pyqrackising/maxcut_tfim.py:328-329)
        x = np.array([c == "1" for c in bits], np.uint8)
        assert value == pytest.approx(-orc.compute_energy(W, x), rel=REL, abs=1e-12)


def test_exact_guard_on_an_empty_graph(ctx, orc):
    """The guard `best_value < 0 -> ("0" * n, 0.0, (nodes, []))` cannot fire on
    the exact path: the all-zeros mask scores 0 for MAXCUT and -E averages 0
    over all masks, so both maxima are >= 0 > -1.  On a graph without edges
    every mask scores 0 and the first one (all zeros) is kept, as with the
    reference's strict '>' from -1."""
    from pyqrackising_amd import maxcut_tfim

    W = np.zeros((10, 10), np.float32)
    for sg in (False, True):
        bits, value, (l, r) = maxcut_tfim(W, is_spin_glass=sg)
        obits, ov = orc.exact(W.astype(np.float64), int(sg))
        assert bits == "0" * 10 and value == 0.0 == ov and not obits.any()
        assert l == list(range(10)) and r == []


def sparse_instance(n, kind, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges(n, 6.0 / (n - 1), rng)
    if kind == "int":
        w = graphs.pm1(lo.size, rng).astype(np.float64)
    else:
        w = rng.uniform(-1.0, 1.0, lo.size)
    return graphs.upper_csr(n, lo, hi, w, dtype=np.float32)


@pytest.mark.parametrize("kind,sg", [("int", 0), ("int", 1), ("real", 0), ("real", 1)])
def test_maxcut_tfim_sparse_rounds_replay(ctx, orc, kind, sg):
    """Above the threshold: rounds of roundup32(shots/2) candidates until a
    round fails to improve.  The same seeded device candidates are scored by
    the oracle's CSR scorer and reduced with the first maximum; the result
    must be the one maxcut_tfim_sparse returned."""
    from pyqrackising_amd import maxcut_tfim_sparse
    from pyqrackising_amd import sampler as smp

    n, q, seed = (60, 4, 11) if kind == "int" else (90, 4, 12)
    log(f"maxcut_tfim_sparse rounds n={n} {kind} sg={sg}")
    G = sparse_instance(n, kind, 900 + n + sg)
    bits, value, (l, r) = maxcut_tfim_sparse(G, quality=q, is_spin_glass=bool(sg), seed=seed)
    J, z = smp.init_J_and_z_sparse(G)
    cum = smp.maxcut_hamming_cdf(n, J, z, q, 8.0, 8.0)
    shots = ((max(1, (n << q) >> 1) + 31) >> 5) << 5
    dg = _lib.DeviceGraph.csr(G)
    s0 = int(np.random.default_rng(seed).integers(0, 2 ** 63))
    best = np.zeros(n, bool)
    best_e = orc.compute_energy_sparse(best.astype(np.uint8), G) if sg else 0.0
    rnd = 0
    while True:
        X = dg.sample_candidates(cum, 5.0, s0 + rnd, 0, shots)
        rnd += 1
        bi, vals = orc.eval_batch_sparse(G, X, sg)
        if vals[bi] > best_e:
            best_e, best = vals[bi], X[bi].copy()
        else:
            break
    # spin-glass mode returns the cut of the best state
    # (pyqrackising/maxcut_tfim_sparse.py:199-200), then the negative guard
    final = orc.compute_cut_sparse(best.astype(np.uint8), G) if sg else best_e
    log(f"  {rnd} rounds, best {best_e}, returned value {final}")
    assert rnd >= 2
    if final < 0.0:
        assert bits == "0" * n and value == 0.0
        return
    assert bits == "".join("1" if b else "0" for b in best)
    if kind == "int":
        assert value == final
    else:
        assert value == pytest.approx(final, rel=REL)


def load_c2_golden():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "c2_bench_config.json")) as f:
        return json.load(f)


def test_local_search_at_the_benchmark_configuration(ctx):
    """bench.py's step exactly: c2_instance() (3-regular n = 2048, +-1, seed
    1002), spin-glass mode, 128 walkers x 1 (4096 seeds per Gray phase, MFMA
    seed scoring), gray_iterations = n^2 (32768 walk iterations per walker,
    most of them resolved by the repeat rule), the default fast path.  State,
    every accept counter, Gray steps, cut and energy equal the oracle's record
    of the same run (tests/golden/c2_bench_config.json; the oracle evaluates
    every Gray step and needs minutes, so it is recorded once by
    tests/golden/make_c2_bench_golden.py)."""
    import hashlib

    sys.path.insert(0, ROOT)
    from bench import c2_instance

    gold = load_c2_golden()
    W, x0 = c2_instance()
    n = W.shape[0]
    digest = hashlib.sha256(np.ascontiguousarray(W, np.int8).tobytes() + x0.astype(np.uint8).tobytes())
    assert digest.hexdigest() == gold["instance_sha256"]
    xc = np.unpackbits(np.frombuffer(bytes.fromhex(gold["state_packed_hex"]), np.uint8), bitorder="little")[:n]
    r = gold["result"]
    log(f"benchmark configuration n={n}")
    dg = _lib.DeviceGraph.dense(W)
    p = _lib.make_params(n, is_spin_glass=1, gray_iterations=n * n, walkers=128, seed_multiple=1,
                         seed_engine=_lib.SEEDS_MFMA)
    x, st = dg.local_search(x0, p)
    log(f"  GPU cut {st['cut_value']}, oracle cut {r['cut_value']}, rounds {st['rounds']}")
    assert np.array_equal(x, xc)
    for k in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts",
              "gray_evals", "seed_evals", "double_evals"):
        assert st[k] == r[k], k
    assert st["cut_value"] == r["cut_value"] and st["min_energy"] == r["min_energy"]
    assert st["max_energy"] == -r["min_energy"]


def test_back_to_back_searches_read_their_own_reports(ctx):
    """Searches on one context, back to back, alternating a short search (a
    state that stops after its first report) with the benchmark's search:
    every benchmark search must reproduce the oracle's record, whatever the
    report slots still hold from the call before (a report is taken only when
    its check word matches its number, record and state)."""
    sys.path.insert(0, ROOT)
    from bench import c2_instance

    gold = load_c2_golden()
    r = gold["result"]
    W, x0 = c2_instance()
    n = W.shape[0]
    xc = np.unpackbits(np.frombuffer(bytes.fromhex(gold["state_packed_hex"]), np.uint8), bitorder="little")[:n]
    dg = _lib.DeviceGraph.dense(W)
    p = _lib.make_params(n, is_spin_glass=1, gray_iterations=n * n, walkers=128, seed_multiple=1,
                         seed_engine=_lib.SEEDS_MFMA)
    rng = np.random.default_rng(77)
    m = 96
    lo, hi = graphs.random_regular_edges(m, 3, rng)
    Ws = graphs.dense_from_edges(m, lo, hi, graphs.pm1(lo.size, rng), dtype=np.int8)
    ds = _lib.DeviceGraph.dense(Ws)
    ps = _lib.make_params(m, is_spin_glass=1, gray_iterations=64, walkers=4, seed_multiple=4)
    ys, _ = ds.local_search(rng.integers(0, 2, m).astype(np.uint8), ps)
    ys_first = None
    for k in range(40):
        # the short search starts from its own optimum: it stops after one iteration
        y, sts = ds.local_search(ys.copy(), ps)
        if ys_first is None:
            ys_first = (y.copy(), sts["cut_value"], sts["rounds"])
        assert np.array_equal(y, ys_first[0]) and sts["cut_value"] == ys_first[1], k
        x, st = dg.local_search(x0, p)
        if not (np.array_equal(x, xc) and st["cut_value"] == r["cut_value"]):
            log(f"  call {k}: GPU cut {st['cut_value']}, oracle cut {r['cut_value']}, rounds {st['rounds']}")
        assert np.array_equal(x, xc) and st["cut_value"] == r["cut_value"], k
        assert st["single_accepts"] == r["single_accepts"] and st["gray_evals"] == r["gray_evals"], k
    log(f"  40 pairs of back-to-back searches, short search rounds {ys_first[2]}")
