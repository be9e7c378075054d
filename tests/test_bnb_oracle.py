"""Pins the branch-and-bound restatement (oracle/bnb_ref.py) against checks that
do not restate it: brute force over every assignment (the certified optimum,
and the bound's validity over every completion of a partial assignment),
networkx.cut_size through the C oracle's compute_cut, closed forms (all
fixed: the bound is the cut; all free: the sum of the positive weights; E =
2 cut - sum w), the Petersen graph's maximum cut of 12 (the graph the
reference's scripts/bnb_exact.py solves), and the time limit's early exit.
"""
import itertools

import networkx as nx
import numpy as np
import pytest

from oracle import bnb_ref


def rand_int_graph(n, rng, p=0.6, lo=-3, hi=4):
    U = np.triu(rng.integers(lo, hi, (n, n)) * (rng.random((n, n)) < p), 1).astype(np.float64)
    return U + U.T


def all_bits(n):
    return np.array(list(itertools.product((0, 1), repeat=n)), dtype=np.uint8)[:, ::-1]


def brute_max(orc, W, is_spin_glass):
    X = all_bits(W.shape[0])
    f = orc.compute_energy if is_spin_glass else orc.compute_cut
    return max(f(W, x) for x in X)


@pytest.mark.parametrize("n", [5, 8, 11])
@pytest.mark.parametrize("batch_cap", [1, 6, 64])
def test_maxcut_certified_optimum_is_brute_force(orc, n, batch_cap):
    rng = np.random.default_rng(100 + n + batch_cap)
    for _ in range(3):
        W = rand_int_graph(n, rng)
        x0 = rng.integers(0, 2, n).astype(bool)
        bits, v, cert, ex, pr = bnb_ref.branch_and_bound(W, x0, orc.compute_cut(W, x0.astype(np.uint8)), False,
                                                         batch_cap)
        assert cert
        assert v == brute_max(orc, W, False)
        assert orc.compute_cut(W, bits.astype(np.uint8)) == v
        assert pr <= ex


@pytest.mark.parametrize("n", [6, 9])
def test_spin_glass_energy_bound_is_exact(orc, n):
    rng = np.random.default_rng(7 + n)
    for _ in range(4):
        W = rand_int_graph(n, rng)
        x0 = np.zeros(n, bool)
        e0 = orc.compute_energy(W, x0.astype(np.uint8))
        bits, v, cert, _, _ = bnb_ref.branch_and_bound(W, x0, e0, True, 16, energy_bound=True)
        assert cert and v == brute_max(orc, W, True)
        assert orc.compute_energy(W, bits.astype(np.uint8)) == v


def test_spin_glass_reference_bound_can_miss_the_optimum(orc):
    """The reference prunes spin-glass nodes with the cut bound; the result is
    a real assignment's energy, never above the optimum, and on this instance
    below it while still reported as certified."""
    rng = np.random.default_rng(1)
    found_gap = False
    for n in (6, 9, 12):
        W = np.triu(rng.integers(-3, 4, (n, n)).astype(float), 1)
        W = W + W.T
        x0 = np.zeros(n, bool)
        e0 = orc.compute_energy(W, x0.astype(np.uint8))
        bits, v, cert, _, _ = bnb_ref.branch_and_bound(W, x0, e0, True, 64)
        best = brute_max(orc, W, True)
        assert cert and v <= best
        assert orc.compute_energy(W, bits.astype(np.uint8)) == v
        found_gap |= v < best
    assert found_gap


def test_bound_dominates_every_completion(orc):
    rng = np.random.default_rng(3)
    n = 9
    for _ in range(20):
        W = rand_int_graph(n, rng)
        F = rng.integers(-1, 2, (8, n)).astype(np.int8)
        ub = bnb_ref.upper_bound_batch(W, F)
        for k in range(F.shape[0]):
            free = np.flatnonzero(F[k] < 0)
            best = -np.inf
            for m in range(1 << free.size):
                x = np.where(F[k] < 0, 0, F[k]).astype(np.uint8)
                x[free] = [(m >> t) & 1 for t in range(free.size)]
                best = max(best, orc.compute_cut(W, x))
            assert ub[k] >= best


def test_bound_and_leaf_closed_forms(orc):
    rng = np.random.default_rng(4)
    n = 13
    W = rand_int_graph(n, rng)
    X = rng.integers(0, 2, (10, n)).astype(np.int8)
    ub = bnb_ref.upper_bound_batch(W, X)
    cut = bnb_ref.eval_leaves(W, X, False)
    en = bnb_ref.eval_leaves(W, X, True)
    for k in range(X.shape[0]):
        x = X[k].astype(np.uint8)
        G = nx.from_numpy_array(W)
        S = [i for i in range(n) if x[i]]
        ref_cut = nx.cut_size(G, S, weight="weight") if S else 0.0
        assert ub[k] == cut[k] == ref_cut == orc.compute_cut(W, x)
        assert en[k] == 2 * cut[k] - np.triu(W, 1).sum() == orc.compute_energy(W, x)
    root = np.full((1, n), -1, np.int8)
    assert bnb_ref.upper_bound_batch(W, root)[0] == np.triu(np.maximum(W, 0), 1).sum()


def test_influence_scores():
    rng = np.random.default_rng(5)
    n = 12
    W = rand_int_graph(n, rng)
    root = np.full(n, -1, np.int8)
    np.testing.assert_array_equal(bnb_ref.influence_scores(W, root), np.abs(W).sum(axis=1))
    for _ in range(10):
        row = rng.integers(-1, 2, n).astype(np.int8)
        a = bnb_ref.influence_scores(W, row)
        np.testing.assert_array_equal(a, bnb_ref.influence_scores_fast(W, row))
        assert np.all(a[row >= 0] == -1.0)


def test_petersen_max_cut_is_12(orc):
    W = nx.to_numpy_array(nx.petersen_graph(), weight="weight", nonedge=0.0)
    x0 = np.zeros(10, bool)
    bits, v, cert, _, _ = bnb_ref.branch_and_bound(W, x0, 0.0, False, 64)
    assert cert and v == 12.0
    assert orc.compute_cut(W, bits.astype(np.uint8)) == 12.0


def test_time_limit_returns_the_incumbent_uncertified(orc):
    rng = np.random.default_rng(6)
    W = rand_int_graph(10, rng)
    x0 = rng.integers(0, 2, 10).astype(bool)
    c0 = orc.compute_cut(W, x0.astype(np.uint8))
    bits, v, cert, ex, _ = bnb_ref.branch_and_bound(W, x0, c0, False, 64, time_limit=-1.0)
    assert not cert and ex == 0 and v == c0
    np.testing.assert_array_equal(bits, x0)
