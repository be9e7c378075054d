"""Belief-propagation warm start, under the reference's function names
(pyqrackising/spin_glass_solver.py:311-421 belief_propagation_marginals_dense,
bp_warm_start_dense; spin_glass_solver_sparse.py:378-514
belief_propagation_marginals, bp_warm_start).

The message passing runs in libpqi (pqi_bp_marginals) on the device.  The
reference starts every message at zero, which its update keeps at zero, so
its marginals are all zero and its warm start is the all-zeros partition; the
device loop reaches that fixed point after one sweep and stops there.
"""
import numpy as np
import scipy.sparse as sp

from . import _lib


def _iterations(bp_scale, n):
    return max(1, int(bp_scale * n))


def belief_propagation_marginals_dense(G_m, n, bp_scale=1.0, damping=0.5):
    """Marginals in (-1, 1) from the upper triangle of G_m (|w| > 1e-12)."""
    U = np.triu(np.asarray(G_m, dtype=np.float64)[:n, :n], 1)
    if not np.any(np.abs(U) > 1e-12):
        return np.zeros(n, dtype=np.float64)
    dg = _lib.DeviceGraph.dense(U + U.T)
    try:
        marg, _, _ = dg.bp_marginals(_iterations(bp_scale, n), damping)
    finally:
        dg.close()
    return marg


def bp_warm_start_dense(G_m, n, bp_scale=1.0, damping=0.5):
    return belief_propagation_marginals_dense(G_m, n, bp_scale, damping) > 0.0


def belief_propagation_marginals(G_data, G_rows, G_cols, n, bp_scale=1.0, damping=0.5):
    """Marginals in (-1, 1) from a strictly upper-triangular CSR matrix."""
    M = sp.csr_matrix((np.asarray(G_data), np.asarray(G_cols), np.asarray(G_rows)), shape=(n, n))
    if M.count_nonzero() == 0:
        return np.zeros(n, dtype=np.float64)
    dg = _lib.DeviceGraph.csr(M)
    try:
        marg, _, _ = dg.bp_marginals(_iterations(bp_scale, n), damping)
    finally:
        dg.close()
    return marg


def bp_warm_start(G_data, G_rows, G_cols, n, bp_scale=1.0, damping=0.5):
    return belief_propagation_marginals(G_data, G_rows, G_cols, n, bp_scale, damping) > 0.0
