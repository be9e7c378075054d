"""Belief propagation on the GPU (pqi_bp_marginals, bp_scale in the solvers)
against the restatement in oracle/bp_ref.py
(pyqrackising/spin_glass_solver.py:311-421; spin_glass_solver_sparse.py:378-514).

From the reference's zero start the marginals are exactly zero after one
sweep (a fixed point).  From nonzero starts (the restatement's `init`) the
device marginals match the restatement to 1e-10 absolute: the device forms a
node's cavity sum as H - m instead of a sum over the other neighbours, and
its tanh/atanh may differ from numpy's in the last bit."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import bp_ref
from pyqrackising_amd import _lib, graphs, spin_glass_solver, spin_glass_solver_sparse
from pyqrackising_amd import bp as pbp

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_sym(n, rng, p=0.3, kind="int"):
    mask = rng.random((n, n)) < p
    if kind == "int":
        U = np.triu(rng.integers(-3, 4, (n, n)) * mask, 1).astype(np.float64)
    else:
        U = np.triu(rng.normal(0, 0.7, (n, n)) * mask, 1)
    return U + U.T


@pytest.mark.parametrize("fmt", ["dense", "csr"])
def test_zero_start_fixed_point(fmt):
    rng = np.random.default_rng(2)
    n = 300
    W = rand_sym(n, rng, p=0.05)
    dg = _lib.DeviceGraph.dense(W) if fmt == "dense" else _lib.DeviceGraph.csr(sp.csr_matrix(np.triu(W, 1)))
    marg, it, arcs = dg.bp_marginals(max(1, int(1.0 * n)), 0.5)
    dg.close()
    assert arcs == np.count_nonzero(W)
    assert it == 1
    assert np.all(marg == 0.0)


@pytest.mark.parametrize("fmt,kind,n", [("dense", "int", 40), ("dense", "real", 33), ("dense", "int", 130),
                                        ("csr", "int", 200), ("csr", "real", 64)])
@pytest.mark.parametrize("iters", [1, 2, 5, 17])
def test_messages_match_restatement(fmt, kind, n, iters):
    log(f"bp fmt={fmt} kind={kind} n={n} iters={iters}")
    rng = np.random.default_rng(n + iters)
    if fmt == "csr" and kind == "int":
        lo, hi = graphs.random_regular_edges(n, 3, rng)
        W = graphs.dense_from_edges(n, lo, hi, rng.integers(1, 4, lo.size) * rng.choice([-1, 1], lo.size),
                                    dtype=np.float64)
    else:
        W = rand_sym(n, rng, p=0.3 if n < 100 else 0.1, kind=kind)
    us, vs, ws = bp_ref.edges_dense(W)
    _, _, m2 = bp_ref.arc_order(us, vs, n)
    init = rng.normal(0, 0.4, m2)
    damping = 0.35
    ref = bp_ref.marginals(us, vs, ws, n, damping=damping, init=init, iterations=iters)
    dg = _lib.DeviceGraph.dense(W) if fmt == "dense" else _lib.DeviceGraph.csr(sp.csr_matrix(np.triu(W, 1)))
    marg, it, arcs = dg.bp_marginals(iters, damping, init=init)
    dg.close()
    assert arcs == m2 and it == iters
    np.testing.assert_allclose(marg, ref, rtol=0, atol=1e-10)


def test_reference_named_functions():
    rng = np.random.default_rng(3)
    n = 50
    W = rand_sym(n, rng)
    assert np.all(pbp.belief_propagation_marginals_dense(W, n, 1.0, 0.5) == 0.0)
    assert not pbp.bp_warm_start_dense(W, n).any()
    M = sp.csr_matrix(np.triu(W, 1))
    assert np.all(pbp.belief_propagation_marginals(M.data, M.indptr, M.indices, n, 2.0, 0.1) == 0.0)
    assert not pbp.bp_warm_start(M.data, M.indptr, M.indices, n).any()
    assert np.all(pbp.belief_propagation_marginals_dense(np.zeros((5, 5)), 5) == 0.0)


@pytest.mark.parametrize("sparse", [False, True])
def test_solver_with_bp_scale(sparse):
    """All-negative couplings in spin-glass mode: the BP partition (all
    zeros, every spin equal) is the optimum E = sum |w|; with bp_scale the
    solver keeps it unless sampling scores as high, and ends at the optimum."""
    rng = np.random.default_rng(8)
    n = 48
    lo, hi = graphs.random_regular_edges(n, 4, rng)
    W = graphs.dense_from_edges(n, lo, hi, -np.ones(lo.size), dtype=np.float32)
    G = sp.csr_matrix(np.triu(W, 1)) if sparse else W
    fn = spin_glass_solver_sparse if sparse else spin_glass_solver
    bs, cut, _, en = fn(G, bp_scale=1.0, seed=5, gray_seed_multiple=4)
    assert cut == 0.0 and en == -float(lo.size)
    assert bs in ("0" * n, "1" * n)
