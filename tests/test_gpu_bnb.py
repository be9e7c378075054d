"""Branch and bound on the GPU (pqi_bnb_eval, pqi_bnb, solve_maxcut_bnb(_sparse))
against the restatement in oracle/bnb_ref.py (pyqrackising/solve_maxcut_bnb.py:48-321).

Integer weights: node bounds, leaf values and branch variables are bit-exact,
and so are the whole searches (returned bits and value, explored and pruned
counts, certified flag) for the same batch_cap.  fp32 weights: bounds and leaf
values to 1e-12 relative (the device sums in another association order; every
fp32 value and partial sum here is exact in fp64, so in practice they agree
bit for bit), branch variables exact."""
import sys

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import bnb_ref
from pyqrackising_amd import _lib, solve_maxcut_bnb, solve_maxcut_bnb_sparse

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_graph(n, rng, p=0.6, kind="int"):
    mask = rng.random((n, n)) < p
    if kind == "int":
        U = np.triu(rng.integers(-3, 4, (n, n)) * mask, 1).astype(np.float32)
    else:
        U = np.triu(rng.uniform(-1, 1, (n, n)).astype(np.float32) * mask, 1)
    return U + U.T


def oracle_branch(W, F):
    out = np.empty(F.shape[0], dtype=np.int64)
    for k in range(F.shape[0]):
        if (F[k] < 0).any():
            out[k] = int(np.argmax(bnb_ref.influence_scores(W, F[k])))
        else:
            out[k] = -1
    return out


@pytest.mark.parametrize("n", [3, 33, 64, 65, 130, 300])
@pytest.mark.parametrize("kind", ["int", "real"])
@pytest.mark.parametrize("sg", [False, True])
def test_node_evaluation(n, kind, sg):
    log(f"bnb_eval n={n} kind={kind} sg={sg}")
    rng = np.random.default_rng(n * 7 + (kind == "real") + 2 * sg)
    W = rand_graph(n, rng, kind=kind)
    cnt = 37
    F = rng.integers(-1, 2, (cnt, n)).astype(np.int8)
    F[::4] = rng.integers(0, 2, (len(F[::4]), n))  # leaves
    F[1] = -1  # the root
    dg = _lib.DeviceGraph.dense(W)
    ub, leaf, br = dg.bnb_eval(F, sg)
    dg.close()
    W64 = W.astype(np.float64)
    ub_ref = bnb_ref.upper_bound_batch(W64, F)
    leaf_ref = bnb_ref.eval_leaves(W64, F, sg)
    is_leaf = ~(F < 0).any(axis=1)
    if kind == "int":
        np.testing.assert_array_equal(ub, ub_ref)
        np.testing.assert_array_equal(leaf[is_leaf], leaf_ref[is_leaf])
    else:
        np.testing.assert_allclose(ub, ub_ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(leaf[is_leaf], leaf_ref[is_leaf], rtol=1e-12, atol=1e-12)
    assert np.isnan(leaf[~is_leaf]).all()
    np.testing.assert_array_equal(br, oracle_branch(W64, F))


def test_node_evaluation_edges():
    W = rand_graph(20, np.random.default_rng(1))
    dg = _lib.DeviceGraph.dense(W)
    ub, leaf, br = dg.bnb_eval(np.zeros((0, 20), np.int8), False)
    assert ub.size == 0
    with pytest.raises(_lib.PqiError):
        dg.bnb_eval(np.full((2, 20), 2, np.int8), False)
    with pytest.raises(_lib.PqiError):
        dg.bnb(np.zeros(20, np.uint8), 0.0, False, 0)
    dg.close()


CASES = [(8, 1), (12, 5), (14, 64), (16, 256), (18, 64), (10, 3)]


@pytest.mark.parametrize("n,cap", CASES)
@pytest.mark.parametrize("mode", ["cut", "sg", "sg_energy"])
def test_search_matches_oracle(n, cap, mode):
    log(f"bnb n={n} cap={cap} mode={mode}")
    sg = mode != "cut"
    eb = mode == "sg_energy"
    rng = np.random.default_rng(1000 + n * 3 + cap)
    W = rand_graph(n, rng)
    W64 = W.astype(np.float64)
    x0 = rng.integers(0, 2, n).astype(bool)
    dg = _lib.DeviceGraph.dense(W)
    c0, e0 = dg.score(x0)
    warm = e0 if sg else c0
    bits, v, st = dg.bnb(x0, warm, sg, cap, energy_bound=eb)
    dg.close()
    rb, rv, rc, rex, rpr = bnb_ref.branch_and_bound(W64, x0, warm, sg, cap, energy_bound=eb)
    log(f"  explored={st['nodes_explored']} pruned={st['nodes_pruned']} value={v}")
    assert v == rv
    np.testing.assert_array_equal(bits, rb)
    assert (st["nodes_explored"], st["nodes_pruned"], bool(st["certified"])) == (rex, rpr, rc)


def test_search_real_weights():
    rng = np.random.default_rng(77)
    n = 13
    W = rand_graph(n, rng, kind="real")
    x0 = np.zeros(n, bool)
    dg = _lib.DeviceGraph.dense(W)
    c0, _ = dg.score(x0)
    bits, v, st = dg.bnb(x0, c0, False, 32)
    dg.close()
    rb, rv, rc, rex, rpr = bnb_ref.branch_and_bound(W.astype(np.float64), x0, c0, False, 32)
    assert v == pytest.approx(rv, rel=1e-12)
    assert (st["nodes_explored"], st["nodes_pruned"]) == (rex, rpr)
    np.testing.assert_array_equal(bits, rb)


def test_solve_petersen():
    """pyqrackising's scripts/bnb_exact.py: the Petersen graph, maximum cut 12."""
    G = nx.petersen_graph()
    bs, cut, (l, r), en, cert = solve_maxcut_bnb(G, verbose=False, batch_cap=64, seed=3)
    assert cert and cut == 12.0
    assert len(l) + len(r) == 10
    x = np.array([b == "1" for b in bs])
    edges = sum(1 for u, v in G.edges() if x[u] != x[v])
    assert edges == 12
    assert en == 2 * 12 - 15  # E = 2 cut - sum w
    bs2, cut2, _, en2, cert2 = solve_maxcut_bnb_sparse(G, verbose=False, batch_cap=64, seed=3)
    assert (bs2, cut2, en2, cert2) == (bs, cut, en, cert)


@pytest.mark.parametrize("sg", [False, True])
def test_solve_with_guess_formats_and_sparse(sg):
    rng = np.random.default_rng(5 + sg)
    n = 14
    W = rand_graph(n, rng)
    guess = rng.integers(0, 2, n).astype(bool)
    s = "".join("1" if b else "0" for b in guess)
    i = int(s[::-1], 2)
    outs = [solve_maxcut_bnb(W, best_guess=g, is_spin_glass=sg, verbose=False, batch_cap=16, return_stats=True)
            for g in (s, i, [bool(b) for b in guess])]
    assert outs[0][:5] == outs[1][:5] == outs[2][:5]
    bs, cut, _, en, cert, st = outs[0]
    # the same search through the oracle, from the same incumbent
    W64 = W.astype(np.float64)
    dg = _lib.DeviceGraph.dense(W)
    c0, e0 = dg.score(guess)
    rb, rv, rc, rex, rpr = bnb_ref.branch_and_bound(W64, guess, e0 if sg else c0, sg, 16)
    cb, eb = dg.score(rb)
    dg.close()
    assert bs == "".join("1" if b else "0" for b in rb)
    assert (st["nodes_explored"], st["nodes_pruned"], bool(st["certified"])) == (rex, rpr, rc)
    if sg:
        assert (cut, en) == (cb, -rv)
    else:
        assert (cut, en) == (rv, eb)
    sp_out = solve_maxcut_bnb_sparse(sp.csr_matrix(np.triu(W, 1)), best_guess=s, is_spin_glass=sg, verbose=False,
                                     batch_cap=16)
    assert sp_out == outs[0][:5]


def test_solve_small_and_time_limit():
    assert solve_maxcut_bnb(np.zeros((0, 0)), verbose=False) == ("", 0, ([], []), 0, True)
    assert solve_maxcut_bnb(np.zeros((1, 1)), verbose=False) == ("0", 0, ([0], []), 0, True)
    out = solve_maxcut_bnb(np.array([[0, 2.0], [2.0, 0]]), verbose=False)
    assert out[0] == "01" and out[1] == 2.0 and out[4] is True
    W = rand_graph(16, np.random.default_rng(9))
    g = "0" * 16
    bs, cut, _, _, cert = solve_maxcut_bnb(W, best_guess=g, verbose=False, time_limit=0.0)
    assert not cert and bs == g and cut == 0.0
    with pytest.raises(ValueError):
        solve_maxcut_bnb_sparse(sp.csr_matrix(W), best_guess=g, verbose=False)
