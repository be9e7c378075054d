"""Pins the belief-propagation restatement (oracle/bp_ref.py) with checks that
do not restate it: the zero start is a fixed point (all marginals exactly 0,
the reference's warm start is all zeros); on a single edge every message
decays as d^t m0 (the cavity field is empty); on a cycle with tanh(w) = 1 and
no damping each sweep passes every message one step along the cycle, so L
sweeps on an L-cycle return the start and one sweep gives m(a->b) = the old
m(c->a) (c the other neighbour of a), which pins the exclusion of the
receiver and the direction of every message; odd symmetry; and the dense and
CSR edge lists give the same result."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import bp_ref


def rand_sym(n, rng, p=0.4):
    U = np.triu(rng.integers(-3, 4, (n, n)) * (rng.random((n, n)) < p), 1).astype(np.float64)
    return U + U.T


@pytest.mark.parametrize("damping", [0.0, 0.5, 0.9])
def test_zero_start_is_a_fixed_point(damping):
    rng = np.random.default_rng(1)
    W = rand_sym(15, rng)
    us, vs, ws = bp_ref.edges_dense(W)
    marg = bp_ref.marginals(us, vs, ws, 15, bp_scale=1.0, damping=damping)
    assert np.all(marg == 0.0)
    assert not np.any(marg > 0.0)


def test_single_edge_decay():
    us, vs, ws = [0], [1], [1.7]
    for d in (0.25, 0.5, 0.8):
        for t in (1, 2, 5):
            fwd, bwd, m2 = bp_ref.arc_order(us, vs, 2)
            init = np.zeros(m2)
            init[fwd[0]], init[bwd[0]] = 0.6, -0.35  # m(0->1), m(1->0)
            marg = bp_ref.marginals(us, vs, ws, 2, damping=d, init=init, iterations=t)
            np.testing.assert_allclose(marg, [np.tanh(-0.35 * d ** t), np.tanh(0.6 * d ** t)], rtol=1e-14)


def cycle(L):
    us = list(range(L - 1)) + [0]
    vs = list(range(1, L)) + [L - 1]
    return us, vs, [40.0] * L  # tanh(40) == 1.0 in double


@pytest.mark.parametrize("L", [3, 5, 8])
def test_cycle_rotation(L):
    us, vs, ws = cycle(L)
    assert np.tanh(40.0) == 1.0
    fwd, bwd, m2 = bp_ref.arc_order(us, vs, L)
    rng = np.random.default_rng(L)
    init = rng.uniform(-0.4, 0.4, m2)
    m0 = bp_ref.marginals(us, vs, ws, L, damping=0.0, init=init, iterations=L)
    # L sweeps: every message is back where it started
    ref = np.zeros(L)
    for k in range(L):
        ref[us[k]] += init[bwd[k]]
        ref[vs[k]] += init[fwd[k]]
    np.testing.assert_allclose(m0, np.tanh(ref), rtol=1e-12, atol=1e-15)
    # one sweep: m(a->b) = old m(c->a); node b's marginal sums the messages into it
    msg = {}
    nb = {a: [] for a in range(L)}
    for k in range(L):
        msg[(us[k], vs[k])] = init[fwd[k]]
        msg[(vs[k], us[k])] = init[bwd[k]]
        nb[us[k]].append(vs[k])
        nb[vs[k]].append(us[k])
    new_into = np.zeros(L)
    for (a, b) in msg:
        c = [x for x in nb[a] if x != b][0]
        new_into[b] += msg[(c, a)]
    m1 = bp_ref.marginals(us, vs, ws, L, damping=0.0, init=init, iterations=1)
    np.testing.assert_allclose(m1, np.tanh(new_into), rtol=1e-12, atol=1e-15)


def test_odd_symmetry_and_formats():
    rng = np.random.default_rng(4)
    n = 12
    W = rand_sym(n, rng)
    us, vs, ws = bp_ref.edges_dense(W)
    fwd, bwd, m2 = bp_ref.arc_order(us, vs, n)
    assert sorted(np.concatenate([fwd, bwd]).tolist()) == list(range(m2))
    init = rng.normal(0, 0.3, m2)
    a = bp_ref.marginals(us, vs, ws, n, damping=0.3, init=init, iterations=4)
    b = bp_ref.marginals(us, vs, ws, n, damping=0.3, init=-init, iterations=4)
    np.testing.assert_array_equal(a, -b)
    M = sp.csr_matrix(np.triu(W, 1))
    cu, cv, cw = bp_ref.edges_csr(M.data, M.indptr, M.indices, n)
    assert (cu, cv, cw) == (us, vs, ws)
