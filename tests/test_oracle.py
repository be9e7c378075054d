"""Pins the CPU oracle (oracle/pqi_oracle.c) against checks that do not restate it:
networkx.cut_size, brute force over bitstrings and pairs, closed forms
(E = 2 cut - W_up = -1/2 s^T W s), textbook Gray-code facts, hand-derived
known-answer vectors of the reference formulas, and the known MAXCUT values the
reference's scripts quote (scripts/maxcut_approx.py:47-58,
scripts/maxcut_exact.py:104-115).
"""
import itertools
import json
import os

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import graphs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rand_graph(n, rng, p=0.5, kind="int"):
    W = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                w = rng.integers(-3, 4) if kind == "int" else rng.normal()
                W[i, j] = W[j, i] = w
    return W


def nx_cut(W, x):
    G = nx.from_numpy_array(W)
    S = [i for i in range(len(x)) if x[i]]
    return nx.cut_size(G, S, weight="weight") if S else 0.0


def w_up(W):
    return np.triu(W, 1).sum()


@pytest.mark.parametrize("kind", ["int", "real"])
def test_cut_energy_closed_forms(orc, kind):
    rng = np.random.default_rng(1)
    for n in (2, 5, 17, 40):
        W = rand_graph(n, rng, kind=kind)
        for _ in range(5):
            x = rng.integers(0, 2, n).astype(np.uint8)
            s = 1.0 - 2.0 * x
            cut = orc.compute_cut(W, x)
            assert cut == pytest.approx(nx_cut(W, x), abs=1e-9)
            e = orc.compute_energy(W, x)
            assert e == pytest.approx(2 * cut - w_up(W), abs=1e-9)
            assert e == pytest.approx(-0.5 * s @ W @ s, abs=1e-9)


def test_cut_diff_family(orc):
    rng = np.random.default_rng(2)
    n = 23
    W = rand_graph(n, rng)
    x = rng.integers(0, 2, n).astype(np.uint8)
    base = nx_cut(W, x)
    for u in range(n):
        y = x.copy()
        y[u] ^= 1
        assert orc.cut_diff(u, y, W) == nx_cut(W, y) - base
    for k, l in [(0, 1), (3, 17), (22, 5), (10, 11)]:
        y = x.copy()
        y[k] ^= 1
        y[l] ^= 1
        assert orc.cut_diff_2(k, l, y, W) == nx_cut(W, y) - base
    for _ in range(10):
        y = rng.integers(0, 2, n).astype(np.uint8)
        assert orc.cut_diff_between(x, y, W) == nx_cut(W, y) - base


def test_sparse_values_match_dense(orc):
    rng = np.random.default_rng(3)
    n = 31
    W = rand_graph(n, rng, kind="real")
    iu, ju = np.nonzero(np.triu(W, 1))
    G = graphs.upper_csr(n, iu, ju, W[iu, ju], dtype=np.float64)
    for _ in range(5):
        x = rng.integers(0, 2, n).astype(np.uint8)
        assert orc.compute_cut_sparse(x, G) == pytest.approx(nx_cut(W, x), abs=1e-9)
        s = 1.0 - 2.0 * x
        assert orc.compute_energy_sparse(x, G) == pytest.approx(-0.5 * s @ W @ s, abs=1e-9)


def test_eval_batch_first_max(orc):
    rng = np.random.default_rng(4)
    n = 37
    W = rand_graph(n, rng)
    X = rng.integers(0, 2, (50, n)).astype(bool)
    X[7] = X[3]  # a tie: first index wins
    X[3] = np.zeros(n, bool)
    X[3][: n // 2] = True
    for sg in (0, 1):
        best, vals = orc.eval_batch(W, X, sg)
        ref = np.array([nx_cut(W, x) if not sg else 2 * nx_cut(W, x) - w_up(W) for x in X])
        assert np.array_equal(vals, ref)
        assert best == int(np.argmax(ref))
    iu, ju = np.nonzero(np.triu(W, 1))
    G = graphs.upper_csr(n, iu, ju, W[iu, ju], dtype=np.float64)
    for sg in (0, 1):
        bd, vd = orc.eval_batch(W, X, sg)
        bs, vs = orc.eval_batch_sparse(G, X, sg)
        assert bd == bs and np.array_equal(vd, vs)


def test_pack_bits_layout(orc):
    X = np.zeros((2, 40), bool)
    X[0, 0] = X[0, 33] = X[1, 31] = True
    P = orc.pack_bits(X)
    assert P.dtype == np.uint32 and P.shape == (2, 2)
    assert P[0, 0] == 1 and P[0, 1] == 2 and P[1, 0] == 1 << 31 and P[1, 1] == 0


def test_unrank_pair_enumeration(orc):
    for n in (2, 3, 7, 30):
        for s, (i, j) in enumerate(itertools.combinations(range(n), 2)):
            assert orc.unrank_pair(s, n) == (i, j)


def test_gray_flip_position(orc):
    for i in range(5000):
        g0, g1 = i ^ (i >> 1), (i + 1) ^ ((i + 1) >> 1)
        d = g0 ^ g1
        assert d & (d - 1) == 0
        p = orc.gray_flip_pos(i)
        assert 1 << p == d
        assert p == ((i + 1) & -(i + 1)).bit_length() - 1  # ctz(i+1)


def test_gray_mutation_known_answers(orc):
    n = 70
    base = np.zeros(n, np.uint8)

    def flipped(idx, off):
        return sorted(np.flatnonzero(orc.gray_mutation(idx, base, off)).tolist())

    # gray(1)=1 -> position n-off-1; gray(2)=3 -> n-off-1, n-off-2; gray(3)=2 -> n-off-2
    assert flipped(0, 0) == []
    assert flipped(1, 0) == [69]
    assert flipped(2, 0) == [68, 69]
    assert flipped(3, 0) == [68]
    assert flipped(1, 64) == [5]
    # gray(64) = 96: bits 5 and 6 -> positions 0 and -1 (wraps to n-1)
    assert flipped(64, 64) == [0, 69]
    # XOR into a non-zero seed
    seed = np.ones(n, np.uint8)
    out = orc.gray_mutation(2, seed, 0)
    assert out[69] == 0 and out[68] == 0 and out[:68].all()


@pytest.mark.parametrize("sg", [0, 1])
def test_single_bit_flips_brute_force(orc, sg):
    rng = np.random.default_rng(5)
    n = 26
    W = rand_graph(n, rng)
    for _ in range(4):
        x = rng.integers(0, 2, n).astype(np.uint8)
        base = nx_cut(W, x)
        gains = []
        for k in range(n):
            y = x.copy()
            y[k] ^= 1
            gains.append(nx_cut(W, y) - base)
        e, k = orc.single_bit_flips(x, sg, W)
        assert k == int(np.argmax(gains))
        assert e == (2 if sg else 1) * max(gains)


def _last_pairs_by_loop(n, thread_count):
    """sgs:64-84's thread loop run literally: the pair index each thread ends on."""
    C = n * (n - 1) // 2
    tb = (C + thread_count - 1) // thread_count
    last = []
    for t in range(thread_count):
        s, got = t, -1
        while s < C:
            got = s
            s += tb
        last.append(got)
    return last


def test_double_last_index_closed_form(orc):
    for n in (3, 5, 24, 50):
        C = n * (n - 1) // 2
        for T in (1, 2, 3, 8, 16, C, C + 3):
            assert [orc.double_last_index(t, C, T) for t in range(T)] == _last_pairs_by_loop(n, T)


@pytest.mark.parametrize("sg", [0, 1])
def test_double_bit_flips_both_modes(orc, sg):
    rng = np.random.default_rng(6)
    n = 24
    W = rand_graph(n, rng)
    x = rng.integers(0, 2, n).astype(np.uint8)
    base = nx_cut(W, x)
    pairs = list(itertools.combinations(range(n), 2))

    def gain(i, j):
        y = x.copy()
        y[i] ^= 1
        y[j] ^= 1
        return nx_cut(W, y) - base

    scale = 2 if sg else 1
    all_g = [gain(i, j) for i, j in pairs]
    e, k, l = orc.double_bit_flips(x, sg, W, 8, mode=1)
    b = int(np.argmax(all_g))
    assert (k, l) == pairs[b] and e == scale * all_g[b]
    for T in (3, 8, 16):
        last = [s for s in _last_pairs_by_loop(n, T) if s >= 0]
        g = [all_g[s] for s in last]
        b = int(np.argmax(g))
        e, k, l = orc.double_bit_flips(x, sg, W, T, mode=0)
        assert (k, l) == pairs[last[b]] and e == scale * g[b]


def test_pick_gray_seeds(orc):
    rng = np.random.default_rng(7)
    n = 70  # two 64-bit blocks, the second one partial
    W = rand_graph(n, rng, p=0.2)
    x = rng.integers(0, 2, n).astype(np.uint8)
    walkers, mult = 4, 2
    seeds, en, idx, alle = orc.pick_gray_seeds(x, walkers, mult, W, want_all=True)
    bs = walkers * mult
    assert alle.size == bs * 2
    base = nx_cut(W, x)
    for s in range(alle.size):
        y = orc.gray_mutation(s % bs, x, (s // bs) * 64)
        assert alle[s] == nx_cut(W, y) - base
    order = sorted(range(alle.size), key=lambda s: (-alle[s], s))[:walkers]
    assert idx.tolist() == order
    assert en.tolist() == [alle[s] for s in order]
    for w, s in enumerate(order):
        assert np.array_equal(seeds[w], orc.gray_mutation(s % bs, x, (s // bs) * 64))


def test_gray_optimization_invariants(orc):
    rng = np.random.default_rng(8)
    n = 100
    W = rand_graph(n, rng, p=0.1)
    x = rng.integers(0, 2, n).astype(np.uint8)
    walkers = 4
    seeds, en, _ = orc.pick_gray_seeds(x, walkers, 2, W)
    base = nx_cut(W, x)
    for sg in (0, 1):
        bw, gain, it, en2, evals = orc.gray_optimization(seeds, en, 256, walkers, sg, W)
        # each walker's total (seed difference + accepted gains) is its true cut difference
        for w in range(walkers):
            assert en2[w] == nx_cut(W, it[w]) - base
            assert en2[w] >= en[w]
        assert bw == int(np.argmax(en2))
        assert gain == (2 if sg else 1) * en2[bw]
        # ceil(256/4)=64 iterations x 2 blocks; flips at ctz(i+1)+64 >= 100 are skipped
        expect = sum(1 for i in range(64) for b in range(2) if orc.gray_flip_pos(i) + 64 * b < n)
        assert evals == walkers * expect


@pytest.mark.parametrize("double_mode", [0, 1])
@pytest.mark.parametrize("sg", [0, 1])
def test_dense_local_search_invariants(orc, sg, double_mode):
    rng = np.random.default_rng(9 + sg + 2 * double_mode)
    n = 40
    W = rand_graph(n, rng, p=0.3)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    best, r = orc.dense_local_search(W, x0, is_spin_glass=sg, walkers=4, double_mode=double_mode,
                                     gray_iterations=400)
    cut = nx_cut(W, best)
    assert r["cut_value"] == cut
    en = 2 * cut - w_up(W)
    assert r["max_energy"] == (en if sg else cut)  # bookkeeping is exact (D4)
    assert r["min_energy"] == (-en if sg else en)
    # local optimality at exit: no improving single flip
    for k in range(n):
        y = best.copy()
        y[k] ^= 1
        assert nx_cut(W, y) <= cut
    if double_mode == 1:
        for i, j in itertools.combinations(range(n), 2):
            y = best.copy()
            y[i] ^= 1
            y[j] ^= 1
            assert nx_cut(W, y) <= cut
    assert cut >= nx_cut(W, x0)


@pytest.mark.parametrize("sg", [0, 1])
def test_sparse_local_search_matches_dense(orc, sg):
    """On the same graph the sparse loop (absolute values) and the dense loop
    (gains) accept the same moves for integer weights."""
    rng = np.random.default_rng(11)
    n = 72
    W = rand_graph(n, rng, p=0.1)
    iu, ju = np.nonzero(np.triu(W, 1))
    G = graphs.upper_csr(n, iu, ju, W[iu, ju], dtype=np.float64)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    bd, rd = orc.dense_local_search(W, x0, is_spin_glass=sg, walkers=4, gray_iterations=300)
    bs, rs = orc.sparse_local_search(G, x0, is_spin_glass=sg, walkers=4, gray_iterations=300)
    assert rs["cut_value"] == nx_cut(W, bs)
    # the first phases coincide; the final cuts are both local optima
    e1, k1 = orc.single_bit_flips(x0, sg, W)
    e2, k2 = orc.sparse_single(x0, sg, G)
    assert k1 == k2
    base = orc.compute_energy_sparse(x0, G) if sg else orc.compute_cut_sparse(x0, G)
    assert e2 == base + e1


def _brute(W, sg):
    n = W.shape[0]
    masks = np.arange(1 << n, dtype=np.int64)
    bits = ((masks[:, None] >> np.arange(n)) & 1).astype(np.int8)
    s = 1 - 2 * bits
    quad = np.einsum("bi,ij,bj->b", s.astype(float), W, s.astype(float)) / 2  # sum_{i<j} w s_i s_j
    vals = quad if sg else (w_up(W) - quad) / 2
    return int(np.argmax(vals)), vals.max()


@pytest.mark.parametrize("sg", [0, 1])
def test_exact_solver_brute_force(orc, sg):
    rng = np.random.default_rng(12)
    W = rand_graph(12, rng)
    bits, v = orc.exact(W, sg)
    m, vmax = _brute(W, sg)
    assert v == vmax
    assert int(sum(int(b) << i for i, b in enumerate(bits))) == m


def _load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_known_answers():
    """scripts/maxcut_approx.py:47-58: Petersen 12, icosahedral 20, K_{m,n} m*n."""
    from oracle import oracle as orc

    cases = _load_golden("known_answers.json")
    assert {c["name"] for c in cases} >= {"petersen", "icosahedral", "K_8_8"}
    for c in cases:
        n = c["n"]
        e = np.array(c["edges"], dtype=np.int64)
        W = graphs.dense_from_edges(n, e[:, 0], e[:, 1], np.array(c["weights"], float), dtype=np.float64)
        bits, v = orc.exact(W, 0)
        assert v == c["maxcut"], c["name"]
        assert nx_cut(W, bits) == c["maxcut"]


def test_first_improvement_passes(orc):
    """Each pass visits the order once and flips only improving nodes; the
    result is single-flip optimal when the last pass flips nothing."""
    rng = np.random.default_rng(13)
    n = 45
    W = rand_graph(n, rng, p=0.2)
    x = rng.integers(0, 2, n).astype(np.uint8)
    order = rng.permutation(n)
    xs, flips, passes, gain = orc.first_improvement(W, x, order, 100)
    assert gain == nx_cut(W, xs) - nx_cut(W, x)
    assert flips > 0 and passes >= 2
    for k in range(n):
        y = xs.copy()
        y[k] ^= 1
        assert nx_cut(W, y) <= nx_cut(W, xs)
    # one pass, literally: visit in order, keep a flip iff the cut rises
    y = x.copy()
    for v in order:
        z = y.copy()
        z[v] ^= 1
        if nx_cut(W, z) > nx_cut(W, y):
            y = z
    x1, _, p1, _ = orc.first_improvement(W, x, order, 1)
    assert p1 == 1 and np.array_equal(x1, y)


def test_benchmark_configuration_record():
    """tests/golden/c2_bench_config.json (the oracle's search at the bench
    configuration, recorded by tests/golden/make_c2_bench_golden.py): the
    instance is the one bench.c2_instance() builds, and the recorded state's
    cut and energy, computed here in closed form (cut = sum over cut edges,
    E = 2 cut - W_up), are the recorded values."""
    import hashlib
    import json
    import os

    from bench import c2_instance

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "tests", "golden", "c2_bench_config.json")) as f:
        gold = json.load(f)
    W, x0 = c2_instance()
    n = W.shape[0]
    digest = hashlib.sha256(np.ascontiguousarray(W, np.int8).tobytes() + x0.astype(np.uint8).tobytes())
    assert digest.hexdigest() == gold["instance_sha256"]
    x = np.unpackbits(np.frombuffer(bytes.fromhex(gold["state_packed_hex"]), np.uint8), bitorder="little")[:n]
    Wd = W.astype(np.float64)
    iu, ju = np.triu_indices(n, 1)
    w = Wd[iu, ju]
    cut = float(w[x[iu] != x[ju]].sum())
    r = gold["result"]
    assert r["cut_value"] == cut
    assert r["min_energy"] == -(2 * cut - w.sum()) and r["max_energy"] == -r["min_energy"]
    # the run accepted moves in the walk phase, and every phase ran
    assert r["single_accepts"] > 0 and r["walk_accepts"] > 0 and r["seed_evals"] > 0


def _trace_cases():
    import json
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "tests", "golden", "loop_traces.json")) as f:
        return json.load(f)["cases"]


def _bits(s):
    return np.array([c == "1" for c in s], np.uint8)


def _dense_from_edges(n, edges):
    W = np.zeros((n, n))
    for u, v, w in edges:
        W[u, v] = W[v, u] = w
    return W


def _closed_cut(W, x):
    iu, ju = np.triu_indices(W.shape[0], 1)
    return float(W[iu, ju][x[iu] != x[ju]].sum())


@pytest.mark.parametrize("case", _trace_cases(), ids=lambda c: c["name"][:1] + ("sg" if c["is_spin_glass"] else ""))
def test_loop_traces_derived_by_hand(orc, case):
    """tests/golden/loop_traces.json: tiny graphs on which the reference loop
    (pyqrackising/spin_glass_solver.py:603-717) takes, in order, a reference-
    mode double flip after the singles stall (D2 last pairs), a Gray-seed
    accept, a walk accept whose total starts from the seed's difference, and a
    walk that must reject zero-gain steps.  Each case's moves, phase counters,
    bookkeeping and final state were worked out by hand (the derivation is in
    the file).  Here every derived step is re-checked against the closed-form
    cut, the per-phase oracle functions against the derived intermediate
    results, and both oracle loops (dense and sparse) against the final
    outcome."""
    n, sg = case["n"], case["is_spin_glass"]
    W = _dense_from_edges(n, case["edges"])
    x = _bits(case["x0"])
    cut = _closed_cut(W, x)
    for st in case["steps"]:
        y = x.copy()
        y[st["flip"]] ^= 1
        assert _closed_cut(W, y) - cut == st["gain"] and _closed_cut(W, y) == st["cut"], st
        x, cut = y, _closed_cut(W, y)
    ex = case["expect"]
    assert np.array_equal(x, _bits(ex["x"])) and cut == ex["cut_value"]
    scale = 2 if sg else 1
    for chk in case["checks"]:
        xc = _bits(chk["x"])
        e = chk["expect"]
        if chk["fn"] == "single_bit_flips":
            assert orc.single_bit_flips(xc, sg, W) == (scale * e["gain"], e["k"])
        elif chk["fn"] == "double_bit_flips":
            assert orc.double_bit_flips(xc, sg, W, 2, mode=0) == (e["gain"], e["k"], e["l"])
        elif chk["fn"] == "pick_gray_seeds":
            seeds, en, idx, alle = orc.pick_gray_seeds(xc, 2, 2, W, want_all=True)
            assert idx.tolist() == e["order"] and en.tolist() == e["energies"] and alle.tolist() == e["all"]
        else:
            seeds, en, idx = orc.pick_gray_seeds(xc, 2, 2, W)
            bw, g, it, tot, _ = orc.gray_optimization(seeds, en, 4, 2, sg, W)
            assert (bw, g, tot.tolist()) == (e["best"], scale * e["gain"], e["totals"])
            assert np.array_equal(it[bw], _bits(e["state"]))
    x0 = _bits(case["x0"])
    G = sp.csr_matrix(np.triu(W, 1))
    for name, (xf, r) in (("dense", orc.dense_local_search(W, x0, is_spin_glass=sg, walkers=2, seed_multiple=2,
                                                           gray_iterations=4)),
                          ("sparse", orc.sparse_local_search(G, x0, is_spin_glass=sg, walkers=2, seed_multiple=2,
                                                             gray_iterations=4))):
        assert np.array_equal(xf, _bits(ex["x"])), name
        for k in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts",
                  "max_energy", "cut_value"):
            assert r[k] == ex[k], (name, k)
        if "min_energy" in ex:
            assert r["min_energy"] == ex["min_energy"], name


@pytest.mark.parametrize("kind", ["int", "real"])
def test_post_reheat_branch_never_accepts(orc, kind):
    """The post-reheat phase (pyqrackising/spin_glass_solver.py:670-717;
    spin_glass_solver_sparse.py:802-854) cannot accept a move once the walkers
    start from their seeds' differences (the zero-initialised accumulator).
    Seed 0 is gray_mutation(0, theta) = theta itself, with difference 0, the
    largest any seed can have when the seed phase was rejected, and the lowest
    index, so walker 0 starts at theta with total 0.  It only keeps flips of
    positive gain, so if the walk is rejected walker 0 kept none: it is still
    at theta with total 0, every other total is <= 0, and the first argmax
    hands theta itself, with reheat energy 0, to the reheat phase, whose
    single and double flips on theta were rejected at the start of the same
    iteration.  (In the reference the dense accumulator is np.empty, so only
    uninitialised memory can make the branch fire.)  Checked here on random
    small instances for both loops."""
    rng = np.random.default_rng(5 if kind == "int" else 6)
    walks = 0
    for t in range(40):
        n = int(rng.integers(6, 24))
        U = np.triu(rng.random((n, n)) < 0.4, 1)
        w = rng.integers(-3, 4, (n, n)) if kind == "int" else rng.uniform(-1, 1, (n, n))
        W = np.where(U, w, 0).astype(np.float64)
        W = W + W.T
        x0 = rng.integers(0, 2, n).astype(np.uint8)
        G = int(rng.integers(1, 5))
        kw = dict(is_spin_glass=int(t % 2), walkers=G, seed_multiple=int(rng.integers(1, 4)),
                  gray_iterations=int(rng.integers(1, 200)))
        _, r = orc.dense_local_search(W, x0, **kw)
        _, rs = orc.sparse_local_search(sp.csr_matrix(np.triu(W, 1)), x0, **kw)
        assert r["reheat_accepts"] == 0 and rs["reheat_accepts"] == 0
        walks += r["walk_accepts"]
    assert walks > 0  # the walks did accept moves on these instances
