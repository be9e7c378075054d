"""GPU parity: every libpqi entry point against the CPU oracle on the same
seeded inputs.  Integer weights must match bit for bit (values, selected
indices, accepted moves, final states); real weights within 1e-9 relative of
the FP64 oracle."""
import itertools
import sys

import numpy as np
import pytest

from pyqrackising_amd import _lib, graphs

pytestmark = pytest.mark.gpu

REL = 1e-9  # FP64 fields and re-scores: summation-order differences only


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def regular_pm1(n, d, seed, dtype=np.int8):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, d, rng)
    return graphs.dense_from_edges(n, lo, hi, graphs.pm1(lo.size, rng), dtype=dtype)


def er_weights(n, p, seed, kind):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges(n, p, rng)
    if kind == "int":
        w = rng.integers(-5, 6, lo.size).astype(np.float64)
    elif kind == "unit":
        w = np.ones(lo.size)
    else:
        w = rng.normal(size=lo.size)
    return graphs.dense_from_edges(n, lo, hi, w, dtype=np.float64)


def upper(W):
    n = W.shape[0]
    iu, ju = np.nonzero(np.triu(W, 1))
    return graphs.upper_csr(n, iu, ju, W[iu, ju], dtype=np.float64)


@pytest.fixture(scope="module")
def ctx():
    assert _lib.device_count() > 0, "no GPU visible"
    return _lib.default_context()


# ---------------------------------------------------------------------------
# candidate scoring
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,B", [(24, 7), (100, 129), (130, 300), (257, 1000), (2048, 256), (300, 1025),
                                 (1000, 2500), (512, 4096)])
@pytest.mark.parametrize("sg", [0, 1])
def test_eval_batch_int8_mfma(ctx, orc, n, B, sg):
    log(f"eval_batch int8 n={n} B={B} sg={sg}")
    W = regular_pm1(n, 3, seed=n) if n % 2 == 0 else er_weights(n, 0.1, n, "int")
    rng = np.random.default_rng(n + B)
    X = rng.integers(0, 2, (B, n)).astype(bool)
    X[B // 2] = X[0]  # a tie: first index wins
    dg = _lib.DeviceGraph.dense(W.astype(np.float32))
    assert dg.info()["weights"] == _lib.WEIGHTS_I8
    bi, bv, vals = dg.eval_batch(_lib.pack_bits(X), sg)
    obi, ovals = orc.eval_batch(W, X, sg)
    assert np.array_equal(vals, ovals)
    assert bi == obi and bv == ovals[obi]


@pytest.mark.parametrize("n,B", [(384, 777), (129, 256)])
def test_eval_batch_int8_full_range(ctx, orc, n, B):
    """Dense weights over the whole int8 range: the int32 MFMA accumulation and
    the int64 row totals stay exact."""
    log(f"eval_batch int8 full range n={n} B={B}")
    rng = np.random.default_rng(n * B)
    U = np.triu(rng.integers(-127, 128, (n, n)), 1)
    W = (U + U.T).astype(np.float64)
    X = rng.integers(0, 2, (B, n)).astype(bool)
    dg = _lib.DeviceGraph.dense(W.astype(np.int8))
    bi, bv, vals = dg.eval_batch(_lib.pack_bits(X), 1)
    obi, ovals = orc.eval_batch(W, X, 1)
    assert np.array_equal(vals, ovals)
    assert bi == obi


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("sg", [0, 1])
def test_eval_batch_real(ctx, orc, dtype, sg):
    n, B = 150, 200
    log(f"eval_batch real n={n} dtype={dtype.__name__} sg={sg}")
    W = er_weights(n, 0.3, 5, "real").astype(dtype).astype(np.float64)
    X = np.random.default_rng(6).integers(0, 2, (B, n)).astype(bool)
    dg = _lib.DeviceGraph.dense(W.astype(dtype))
    bi, bv, vals = dg.eval_batch(_lib.pack_bits(X), sg, engine=_lib.ENGINE_EXACT)
    obi, ovals = orc.eval_batch(W, X, sg)
    np.testing.assert_allclose(vals, ovals, rtol=REL, atol=REL * np.abs(ovals).max())
    assert bi == obi
    # default engine: fp32 weights use split BF16 + FP64 re-score of the best 16
    bi2, bv2, _ = dg.eval_batch(_lib.pack_bits(X), sg)
    assert bi2 == obi and bv2 == pytest.approx(ovals[obi], rel=REL)


def sk_matrix(n, seed):
    """Sherrington-Kirkpatrick couplings: w_ij ~ N(0, 1/n) i.i.d. for i < j, fp32."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n), dtype=np.float32) / np.float32(np.sqrt(n))
    W = np.triu(A, 1)
    del A
    W += W.T
    return W


# split-BF16 tolerance: |approx - FP64| <= BF16_TOL * sum_{i<j} |w_ij|.  The
# split leaves |r| <= 2^-18 |w| per weight; the rest is FP32 accumulation.
BF16_TOL = 2e-5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,B", [(100, 1), (130, 5), (257, 16), (1000, 16), (2050, 9)])
def test_eval_batch_real_few_candidates(ctx, orc, dtype, n, B):
    """Up to 16 candidates take the streaming FP64 re-score kernel (W read
    once; 256-column groups, so n_pad that is not a multiple of 256 and the
    row splits are covered): values within 1e-9 relative of the oracle."""
    log(f"few-candidate re-score n={n} B={B} {dtype.__name__}")
    rng = np.random.default_rng(n + B)
    W = er_weights(n, 0.3, n + B, "real").astype(dtype)
    X = rng.integers(0, 2, (B, n)).astype(bool)
    dg = _lib.DeviceGraph.dense(W)
    for sg in (0, 1):
        bi, bv, vals = dg.eval_batch(_lib.pack_bits(X), sg, engine=_lib.ENGINE_EXACT)
        ob, ov = orc.eval_batch(W.astype(np.float64), X, sg)
        scale = np.abs(np.triu(W.astype(np.float64), 1)).sum()
        np.testing.assert_allclose(vals, ov, rtol=0, atol=1e-9 * scale)
        assert vals[bi] == bv == vals.max()
    dg.close()


@pytest.mark.parametrize("sg", [0, 1])
@pytest.mark.parametrize("n,B", [(1000, 300), (1000, 777), (1100, 600)])
def test_eval_batch_bf16_split(ctx, orc, sg, n, B):
    """B >= 512 takes the wide 256 x 256 tile (n = 1100: n_pad 1152, a last
    column tile half past n_pad); smaller batches the 128 x 128 tile."""
    log(f"eval_batch bf16 split n={n} B={B} sg={sg}")
    W = sk_matrix(n, 41)
    X = np.random.default_rng(42).integers(0, 2, (B, n)).astype(bool)
    dg = _lib.DeviceGraph.dense(W)
    assert dg.info()["weights"] == _lib.WEIGHTS_F32
    bi, bv, vals, err = dg.eval_batch(_lib.pack_bits(X), sg, engine=_lib.ENGINE_BF16, return_error=True)
    obi, ovals = orc.eval_batch(W.astype(np.float64), X, sg)
    scale = np.abs(np.triu(W.astype(np.float64), 1)).sum()
    approx = np.ones(B, bool)
    top = np.argsort(-vals, kind="stable")[:16]
    approx[top] = False
    assert np.all(np.abs(vals[approx] - ovals[approx]) <= BF16_TOL * scale)
    np.testing.assert_allclose(vals[top], ovals[top], rtol=1e-9)  # FP64 re-scored
    assert bi == obi and bv == pytest.approx(ovals[obi], rel=1e-9)
    log(f"  max rel err among re-scored: {err:.3e}")


def test_eval_batch_sk_full_size(ctx, orc):
    """C4 shape: SK n=16384, fp32, 1024 random candidates.  The split-BF16
    energies against the FP64 engine on the device (size-independent
    agreement), the FP64 re-score of the selected best against the oracle."""
    n, B = 16384, 1024
    log("eval_batch SK n=16384 B=1024")
    W = sk_matrix(n, 43)
    X = np.random.default_rng(44).integers(0, 2, (B, n)).astype(bool)
    bits = _lib.pack_bits(X)
    dg = _lib.DeviceGraph.dense(W)
    bi, bv, vals, err = dg.eval_batch(bits, 1, engine=_lib.ENGINE_BF16, return_error=True)
    ei, ev, evals = dg.eval_batch(bits, 1, engine=_lib.ENGINE_EXACT)
    scale = float(np.abs(np.triu(W, 1)).sum(dtype=np.float64))
    assert np.all(np.abs(vals - evals) <= BF16_TOL * scale)
    assert bi == ei and bv == pytest.approx(ev, rel=1e-9)
    e64 = orc.compute_energy(W, X[bi].astype(np.uint8))
    assert bv == pytest.approx(e64, rel=1e-9)
    log(f"  best={bi} E={bv:.6f} oracle={e64:.6f} bf16 max rel err (top16)={err:.3e}")


@pytest.mark.parametrize("kind", ["int", "real"])
def test_eval_batch_csr(ctx, orc, kind):
    n, B = 300, 64
    log(f"eval_batch csr kind={kind}")
    W = er_weights(n, 0.05, 7, kind)
    G = upper(W)
    X = np.random.default_rng(8).integers(0, 2, (B, n)).astype(bool)
    dg = _lib.DeviceGraph.csr(G)
    for sg in (0, 1):
        bi, bv, vals = dg.eval_batch(_lib.pack_bits(X), sg)
        obi, ovals = orc.eval_batch_sparse(G, X, sg)
        if kind == "int":
            assert np.array_equal(vals, ovals) and bi == obi
        else:
            np.testing.assert_allclose(vals, ovals, rtol=REL, atol=1e-9)
            assert bi == obi


# ---------------------------------------------------------------------------
# single / double flips
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["int", "real"])
@pytest.mark.parametrize("sg", [0, 1])
def test_single_flip(ctx, orc, kind, sg):
    n = 90
    log(f"single flip kind={kind} sg={sg}")
    W = er_weights(n, 0.2, 9, kind)
    dg = _lib.DeviceGraph.dense(W)
    rng = np.random.default_rng(10)
    for _ in range(5):
        x = rng.integers(0, 2, n).astype(np.uint8)
        g, k = dg.single_flip_best(x, sg)
        og, ok = orc.single_bit_flips(x, sg, W)
        assert k == ok
        assert g == pytest.approx(og, rel=REL, abs=1e-9) if kind == "real" else g == og


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("kind", ["int", "real"])
def test_double_flip(ctx, orc, mode, kind):
    n = 70
    log(f"double flip mode={mode} kind={kind}")
    W = er_weights(n, 0.2, 11, kind)
    dg = _lib.DeviceGraph.dense(W)
    dcsr = _lib.DeviceGraph.csr(upper(W))
    rng = np.random.default_rng(12)
    for T in (3, 8, 64):
        x = rng.integers(0, 2, n).astype(np.uint8)
        og, ok, ol = orc.double_bit_flips(x, 1, W, T, mode)
        for d in (dg, dcsr):
            g, k, l = d.double_flip_best(x, 1, T, mode)
            assert (k, l) == (ok, ol)
            assert g == pytest.approx(og, rel=REL, abs=1e-9) if kind == "real" else g == og


# ---------------------------------------------------------------------------
# Gray phase
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("engine", [_lib.SEEDS_MFMA, _lib.SEEDS_LOWRANK])
@pytest.mark.parametrize("n", [64, 70, 200])
def test_gray_search(ctx, orc, engine, n):
    log(f"gray search engine={engine} n={n}")
    W = er_weights(n, 0.08, 13 + n, "int")
    dg = _lib.DeviceGraph.dense(W)
    rng = np.random.default_rng(14)
    walkers, mult, iters = 6, 3, 600
    # start from a single-flip local optimum so the walk phase is reached
    x = rng.integers(0, 2, n).astype(np.uint8)
    x, _ = orc.dense_local_search(W, x, walkers=walkers, seed_multiple=mult, gray_iterations=0, max_rounds=200)
    p = _lib.make_params(n, is_spin_glass=1, gray_iterations=iters, walkers=walkers, seed_multiple=mult,
                         seed_engine=engine)
    sg_gain, wg_gain, state = dg.gray_search(x, p)
    seeds, en, idx = orc.pick_gray_seeds(x, walkers, mult, W)
    assert sg_gain == 2 * en[0]
    if en[0] > 0:
        assert np.array_equal(state, seeds[0])
        return
    bw, gain, it, en2, _ = orc.gray_optimization(seeds, en, iters, walkers, 1, W)
    assert wg_gain == gain
    assert np.array_equal(state, it[bw])


def _walk_start(orc, W, seed, count=4):
    """A single-flip local optimum with `count` non-adjacent nodes flipped away
    from it.  The nodes sit at block positions 0-5 (reached by a short walk) and
    away from the nodes the Gray seeds mutate and their neighbours, so no seed
    improves and the walkers must find the moves back."""
    n = W.shape[0]
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2, n).astype(np.uint8)
    while True:
        g, k = orc.single_bit_flips(x, 1, W)
        if not g > 0:
            break
        x[k] ^= 1
    seeded = {(n - 1 - j - 64 * o) % n for o in range((n + 63) // 64) for j in range(4)}
    near = set(seeded)
    for v in seeded:
        near.update(np.nonzero(W[v])[0].tolist())
    chosen = []
    for v in (int(v) for v in rng.permutation(n) if v not in near and (v & 63) < 6):
        if len(chosen) == count:
            break
        if all(W[v, u] == 0 for u in chosen):
            chosen.append(v)
    for v in chosen:
        x[v] ^= 1
    return x


@pytest.mark.parametrize("n,d", [(64, 3), (1030, 6), (2048, 3), (2048, 13), (500, 16)])
def test_gray_search_register_walkers(ctx, orc, n, d):
    """Walkers that climb: regular +-1 graphs with neighbour rows of 4, 8 and 16
    entries and n <= 2048 (register-resident walkers; a partial last block at
    n=1030).  The start is a perturbed local optimum that no seed improves, so
    the walk phase itself must take the moves: gains and the best walker's
    state against the oracle."""
    log(f"gray search register walkers n={n} d={d}")
    W = regular_pm1(n, d, seed=90 + d).astype(np.float64)
    x = _walk_start(orc, W, 91)
    walkers, mult, iters = 4, 1, 4000
    dg = _lib.DeviceGraph.dense(W.astype(np.int8))
    p = _lib.make_params(n, is_spin_glass=1, gray_iterations=iters, walkers=walkers, seed_multiple=mult,
                         seed_engine=_lib.SEEDS_MFMA)
    sg_gain, wg_gain, state = dg.gray_search(x, p)
    seeds, en, idx = orc.pick_gray_seeds(x, walkers, mult, W)
    assert en[0] <= 0 and sg_gain == 2 * en[0]
    bw, gain, it, en2, _ = orc.gray_optimization(seeds, en, iters, walkers, 1, W)
    assert gain > 0
    assert wg_gain == gain
    assert np.array_equal(state, it[bw])


# ---------------------------------------------------------------------------
# the whole local search
# ---------------------------------------------------------------------------

def _cmp_search(dg, orc_fn, W_or_G, x0, sg, real, **kw):
    p = _lib.make_params(len(x0), is_spin_glass=sg, **kw)
    xg, st = dg.local_search(x0, p)
    okw = {k: v for k, v in kw.items() if k != "seed_engine"}
    xc, r = orc_fn(W_or_G, x0, is_spin_glass=sg, **okw)
    assert np.array_equal(xg, xc), "final states differ"
    for key in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts", "gray_evals"):
        assert st[key] == r[key], key
    if real:
        for key in ("cut_value", "min_energy", "max_energy"):
            assert st[key] == pytest.approx(r[key], rel=REL, abs=1e-9), key
    else:
        for key in ("cut_value", "min_energy", "max_energy"):
            assert st[key] == r[key], key
    return st


@pytest.mark.parametrize("sg", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_local_search_int_dense(ctx, orc, sg, mode):
    n = 128
    log(f"local search int dense sg={sg} mode={mode}")
    W = regular_pm1(n, 3, seed=15)
    dg = _lib.DeviceGraph.dense(W)
    x0 = np.random.default_rng(16).integers(0, 2, n).astype(np.uint8)
    _cmp_search(dg, orc.dense_local_search, W.astype(np.float64), x0, sg, False, walkers=8, seed_multiple=4,
                gray_iterations=2000, double_mode=mode)


@pytest.mark.parametrize("engine", [_lib.SEEDS_MFMA, _lib.SEEDS_LOWRANK])
def test_local_search_seed_engines_agree(ctx, orc, engine):
    n = 200
    log(f"local search engine={engine}")
    W = er_weights(n, 0.05, 17, "int")
    dg = _lib.DeviceGraph.dense(W)
    x0 = np.random.default_rng(18).integers(0, 2, n).astype(np.uint8)
    _cmp_search(dg, orc.dense_local_search, W, x0, 1, False, walkers=5, seed_multiple=2, gray_iterations=1500,
                seed_engine=engine)


@pytest.mark.parametrize("sg", [0, 1])
def test_local_search_real_dense(ctx, orc, sg):
    n = 100
    log(f"local search real dense sg={sg}")
    W = er_weights(n, 0.3, 19, "real")
    dg = _lib.DeviceGraph.dense(W)
    x0 = np.random.default_rng(20).integers(0, 2, n).astype(np.uint8)
    _cmp_search(dg, orc.dense_local_search, W, x0, sg, True, walkers=4, seed_multiple=4, gray_iterations=1000)


@pytest.mark.parametrize("kind", ["int", "real"])
def test_local_search_sparse(ctx, orc, kind):
    n = 300
    log(f"local search sparse kind={kind}")
    W = er_weights(n, 8.0 / n, 21, kind)
    G = upper(W)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(22).integers(0, 2, n).astype(np.uint8)
    p = _lib.make_params(n, is_spin_glass=1, walkers=8, seed_multiple=2, gray_iterations=3000)
    xg, st = dg.local_search(x0, p)
    xc, r = orc.sparse_local_search(G, x0, is_spin_glass=1, walkers=8, seed_multiple=2, gray_iterations=3000)
    assert np.array_equal(xg, xc)
    if kind == "int":
        assert st["cut_value"] == r["cut_value"] and st["min_energy"] == r["min_energy"]
    else:
        assert st["cut_value"] == pytest.approx(r["cut_value"], rel=REL)
        assert st["min_energy"] == pytest.approx(r["min_energy"], rel=REL)


def test_local_search_c2_shape(ctx, orc):
    """The benchmark's instance shape (3-regular, n=2048, +-1) with the Gray
    phase capped so the oracle finishes in seconds."""
    n = 2048
    log("local search C2 shape")
    W = regular_pm1(n, 3, seed=23)
    dg = _lib.DeviceGraph.dense(W)
    x0 = np.random.default_rng(24).integers(0, 2, n).astype(np.uint8)
    st = _cmp_search(dg, orc.dense_local_search, W.astype(np.float64), x0, 1, False, walkers=16, seed_multiple=8,
                     gray_iterations=4096, seed_engine=_lib.SEEDS_MFMA)
    assert st["seed_evals"] > 0


def test_determinism(ctx):
    n = 256
    W = regular_pm1(n, 5 if False else 4, seed=25)
    dg = _lib.DeviceGraph.dense(W)
    x0 = np.random.default_rng(26).integers(0, 2, n).astype(np.uint8)
    p = _lib.make_params(n, walkers=8, seed_multiple=2, gray_iterations=2048)
    a, sa = dg.local_search(x0, p)
    b, sb = dg.local_search(x0, p)
    assert np.array_equal(a, b) and sa["cut_value"] == sb["cut_value"]


# ---------------------------------------------------------------------------
# Python entry points
# ---------------------------------------------------------------------------

def test_exact_small_graphs_known_answers(ctx):
    """maxcut_tfim below the heuristic threshold is the 2^n brute force
    (pyqrackising/maxcut_tfim.py:364-372); known answers from the reference's
    scripts (Petersen 12, icosahedral 20, K_{m,n} m*n)."""
    import json
    import os

    import networkx as nx

    from pyqrackising_amd import maxcut_tfim, maxcut_tfim_sparse

    with open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")) as f:
        cases = json.load(f)
    for c in cases:
        G = nx.Graph()
        G.add_nodes_from(range(c["n"]))
        for (u, v), w in zip(c["edges"], c["weights"]):
            G.add_edge(u, v, weight=w)
        for fn in (maxcut_tfim, maxcut_tfim_sparse):
            bits, value, (l, r) = fn(G)
            assert value == c["maxcut"], (c["name"], fn.__name__)
            x = [b == "1" for b in bits]
            assert nx.cut_size(G, [i for i in range(c["n"]) if x[i]], weight="weight") == value


def test_spin_glass_solver_matches_oracle(ctx, orc):
    from pyqrackising_amd import spin_glass_solver, spin_glass_solver_sparse

    n = 96
    W = regular_pm1(n, 3, seed=27).astype(np.float32)
    guess = np.random.default_rng(28).integers(0, 2, n).astype(bool).tolist()
    bits, cut, (l, r), e = spin_glass_solver(W, best_guess=guess, gray_iterations=900, gray_seed_multiple=4)
    xc, res = orc.dense_local_search(W, np.array(guess, np.uint8), is_spin_glass=1, walkers=4, seed_multiple=4,
                                     gray_iterations=900)
    assert bits == "".join("1" if b else "0" for b in xc)
    assert cut == res["cut_value"] and e == res["min_energy"]
    assert sorted(l + r) == list(range(n))
    G = upper(W.astype(np.float64))
    bits2, cut2, _, e2 = spin_glass_solver_sparse(G, best_guess=guess, gray_iterations=900, gray_seed_multiple=4)
    assert (bits2, cut2, e2) == (bits, cut, e)


def test_maxcut_tfim_heuristic_path(ctx, orc):
    """Above the threshold: sampled candidates scored on the GPU; the reported
    cut equals the cut of the returned bitstring (the reference's own check,
    scripts/maxcut_benchmarks.py:127-129)."""
    from pyqrackising_amd import maxcut_tfim

    n = 40
    W = er_weights(n, 0.5, 29, "unit").astype(np.float32)
    bits, value, (l, r) = maxcut_tfim(W, quality=3, seed=1)
    x = np.array([b == "1" for b in bits], np.uint8)
    assert value == orc.compute_cut(W, x)
    assert value > 0


def test_c1_pipeline_matches_oracle(ctx, orc):
    """The CPU-runnable configuration: ER(64, 0.5) unit weights, int32 dense,
    quality 6 -> maxcut_tfim then spin_glass_solver.  The candidate rounds are
    replayed with the same seeded sampler and scored by the oracle; the local
    search is compared on the resulting guess."""
    from pyqrackising_amd import maxcut_tfim, spin_glass_solver
    from pyqrackising_amd import sampler as smp

    n, seed = 64, 31
    W = er_weights(n, 0.5, 1001, "unit").astype(np.int32)
    bits, value, _ = maxcut_tfim(W, quality=6, seed=seed)
    # replay (pyqrackising/maxcut_tfim.py:148-186 semantics): the same device
    # candidates (same seed schedule), scored and reduced by the oracle
    Wd = W.astype(np.float64)
    J, z = smp.init_J_and_z_dense(Wd)
    cum = smp.maxcut_hamming_cdf(n, J, z, 6, 8.0, 8.0)
    shots = ((max(1, (n << 6) >> 1) + 31) >> 5) << 5
    dg = _lib.DeviceGraph.dense(W)
    s0 = int(np.random.default_rng(seed).integers(0, 2 ** 63))
    best, best_e, rnd = np.zeros(n, bool), 0.0, 0
    while True:
        X = dg.sample_candidates(cum, 5.0, s0 + rnd, 0, shots)
        rnd += 1
        bi, vals = orc.eval_batch(Wd, X, 0)
        if vals[bi] > best_e:
            best_e, best = vals[bi], X[bi].copy()
        else:
            break
    assert bits == "".join("1" if b else "0" for b in best)
    assert value == best_e
    out = spin_glass_solver(W, best_guess=bits, is_spin_glass=False, gray_seed_multiple=8, gray_iterations=4096)
    xc, r = orc.dense_local_search(Wd, best.astype(np.uint8), is_spin_glass=0, walkers=8, seed_multiple=8,
                                   gray_iterations=4096)
    assert out[0] == "".join("1" if b else "0" for b in xc)
    assert out[1] == r["cut_value"] and out[3] == r["min_energy"]


# ---------------------------------------------------------------------------
# many instances in one persistent, graph-captured launch
# ---------------------------------------------------------------------------

def test_batch_eval_uneven_instances(ctx, orc):
    log("batch eval uneven")
    Ws = graphs.batched_regular_instances(12, seed=51, n_lo=24, n_hi=600)
    cands = 300
    rng = np.random.default_rng(52)
    dgs = [_lib.DeviceGraph.dense(W) for W in Ws]
    batch = _lib.Batch(dgs, cands)
    Xs = [rng.integers(0, 2, (cands, W.shape[0])).astype(bool) for W in Ws]
    batch.set_candidates([_lib.pack_bits(X) for X in Xs])
    expect = [orc.eval_batch(W.astype(np.float64), X, sg) for sg in (0, 1) for W, X in zip(Ws, Xs)]
    for v in range(_lib.Batch.N_VARIANTS):
        batch.set_variant(v)
        for sg in (0, 1):
            bi, bv, vals = batch.eval(sg, want_values=True)
            for i in range(len(Ws)):
                obi, ovals = expect[sg * len(Ws) + i]
                assert np.array_equal(vals[i], ovals), (v, sg, i)
                assert bi[i] == obi and bv[i] == ovals[obi]
    chosen, ms = batch.autotune(reps=2)
    assert 0 <= chosen < _lib.Batch.N_VARIANTS and ms[chosen] == ms.min()
    # graph replays with new candidates, eager single-instance work in between
    for rep in range(3):
        Xs = [rng.integers(0, 2, (cands, W.shape[0])).astype(bool) for W in Ws]
        batch.set_candidates([_lib.pack_bits(X) for X in Xs])
        bi, bv, vals = batch.eval(0, want_values=True)
        for i, (dg, X) in enumerate(zip(dgs, Xs)):
            ebi, ebv, evals = dg.eval_batch(_lib.pack_bits(X), 0)
            assert np.array_equal(vals[i], evals) and bi[i] == ebi, (rep, i)


def test_batch_eval_c3_shape(ctx, orc):
    """The batched configuration at full size: 256 instances, n in [256, 4096],
    d in {3..10}, 1024 candidates each.  Every instance against the
    single-instance path (itself pinned to the oracle); a few candidates of the
    largest instances against the oracle directly."""
    log("batch eval C3 shape")
    Ws = graphs.batched_regular_instances(256, seed=1003)
    cands = 1024
    rng = np.random.default_rng(53)
    dgs = [_lib.DeviceGraph.dense(W) for W in Ws]
    bits = [rng.integers(0, 1 << 32, (cands, (W.shape[0] + 31) // 32), dtype=np.uint64).astype(np.uint32)
            for W in Ws]
    for b, W in zip(bits, Ws):  # clear bits beyond n
        n = W.shape[0]
        if n % 32:
            b[:, -1] &= np.uint32((1 << (n % 32)) - 1)
    batch = _lib.Batch(dgs, cands)
    batch.set_candidates(bits)
    chosen, ms = batch.autotune(reps=2)
    log(f"  autotune chose {chosen}: {np.round(ms, 3).tolist()} ms")
    bi, bv, vals = batch.eval(1, want_values=True)
    for i in range(0, 256, 17):
        ebi, ebv, evals = dgs[i].eval_batch(bits[i], 1)
        assert np.array_equal(vals[i], evals) and bi[i] == ebi and bv[i] == ebv, i
    big = np.argsort([W.shape[0] for W in Ws])[-2:]
    for i in big:
        n = Ws[i].shape[0]
        X = ((bits[i][:4, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(4, -1)[:, :n].astype(bool)
        _, ov = orc.eval_batch(Ws[i].astype(np.float64), X, 1)
        assert np.array_equal(vals[i][:4], ov), i


# ---------------------------------------------------------------------------
# batched local search
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["int", "real"])
def test_batch_local_search_matches_single(ctx, orc, kind):
    """Each instance of the batched search makes exactly the moves of the
    single-instance search (and of the oracle)."""
    log(f"batch local search kind={kind}")
    rng = np.random.default_rng(61)
    Ws, x0s = [], []
    for t in range(7):
        n = int(rng.integers(40, 260))
        if kind == "int":
            W = er_weights(n, 4.0 / n, 100 + t, "int")
        else:
            W = er_weights(n, 0.1, 100 + t, "real")
        Ws.append(W)
        x0s.append(rng.integers(0, 2, n).astype(np.uint8))
    dgs = [_lib.DeviceGraph.dense(W) for W in Ws]
    p = _lib.make_params(0, is_spin_glass=1, gray_iterations=1500, walkers=6, seed_multiple=3)
    states, stats = _lib.batch_local_search(dgs, x0s, p)
    for i, (dg, W, x0) in enumerate(zip(dgs, Ws, x0s)):
        xs, st = dg.local_search(x0, _lib.make_params(len(x0), is_spin_glass=1, gray_iterations=1500, walkers=6,
                                                      seed_multiple=3))
        assert np.array_equal(states[i], xs), i
        for key in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts",
                    "gray_evals", "rounds"):
            assert stats[i][key] == st[key], (i, key)
        assert stats[i]["cut_value"] == st["cut_value"] and stats[i]["min_energy"] == st["min_energy"]
        if i < 2:
            xc, r = orc.dense_local_search(W, x0, is_spin_glass=1, walkers=6, seed_multiple=3, gray_iterations=1500)
            assert np.array_equal(states[i], xc)


def test_batch_local_search_csr_default_iterations(ctx):
    """CSR instances, gray_iterations = n^2 per instance (negative value)."""
    log("batch local search csr")
    rng = np.random.default_rng(62)
    Gs, x0s = [], []
    for t in range(5):
        n = int(rng.integers(64, 400))
        W = er_weights(n, 6.0 / n, 200 + t, "real")
        Gs.append(upper(W))
        x0s.append(rng.integers(0, 2, n).astype(np.uint8))
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    p = _lib.make_params(0, is_spin_glass=1, walkers=8, seed_multiple=2)
    p.gray_iterations = -1
    states, stats = _lib.batch_local_search(dgs, x0s, p)
    for i, (dg, x0) in enumerate(zip(dgs, x0s)):
        xs, st = dg.local_search(x0, _lib.make_params(len(x0), is_spin_glass=1, walkers=8, seed_multiple=2))
        assert np.array_equal(states[i], xs), i
        assert stats[i]["cut_value"] == st["cut_value"]
