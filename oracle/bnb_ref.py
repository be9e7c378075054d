"""CPU restatement of warm-started branch and bound
(pyqrackising/solve_maxcut_bnb.py:48-201).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the pyqrackising_amd
package.

Each evaluator reads only the i < j entries of W, as the reference's kernels
do; W may be any square matrix.  The batch evaluators are vectorised over the
batch with numpy, so their float sums may associate differently from the
reference's sequential loops; with integer weights every value is exact.
"""
import time

import numpy as np


def upper_bound_batch(W, F):
    """_upper_bound_batch (solve_maxcut_bnb.py:48-75): per node, the cut of
    the fixed-fixed pairs plus max(w, 0) for every pair with a free end."""
    U = np.triu(np.asarray(W, dtype=np.float64), 1)
    F = np.asarray(F, dtype=np.int8)
    fixed = F >= 0
    both = fixed[:, :, None] & fixed[:, None, :]
    diff = F[:, :, None] != F[:, None, :]
    contrib = np.where(both, np.where(diff, U, 0.0), np.maximum(U, 0.0))
    return contrib.sum(axis=(1, 2))


def eval_leaves(W, F, is_spin_glass):
    """_eval_leaves_cut / _eval_leaves_energy (solve_maxcut_bnb.py:78-103)."""
    U = np.triu(np.asarray(W, dtype=np.float64), 1)
    F = np.asarray(F, dtype=np.int8)
    diff = F[:, :, None] != F[:, None, :]
    if is_spin_glass:
        return np.where(diff, U, -U).sum(axis=(1, 2))
    return np.where(diff, U, 0.0).sum(axis=(1, 2))


def influence_scores(W, row):
    """_influence_scores (solve_maxcut_bnb.py:106-121): -1 for a fixed
    variable, else the sum of |W[min(i,j), max(i,j)]| over the other free j,
    summed in j order."""
    W = np.asarray(W, dtype=np.float64)
    n = W.shape[0]
    scores = np.full(n, -1.0)
    for i in range(n):
        if row[i] >= 0:
            continue
        s = 0.0
        for j in range(n):
            if i == j or row[j] >= 0:
                continue
            s += abs(W[min(i, j), max(i, j)])
        scores[i] = s
    return scores


def influence_scores_fast(W, row):
    """influence_scores, vectorised (sums in numpy's order)."""
    U = np.triu(np.asarray(W, dtype=np.float64), 1)
    A = np.abs(U + U.T)
    free = np.asarray(row) < 0
    return np.where(free, (A * free[None, :]).sum(axis=1), -1.0)


def branch_and_bound(W, warm_bits, warm_value, is_spin_glass, batch_cap, time_limit=None, energy_bound=False):
    """_branch_and_bound (solve_maxcut_bnb.py:128-201) with batch_cap given.
    Returns (bits, value, certified, nodes_explored, nodes_pruned).

    The reference prunes with the cut bound in both modes.  In spin-glass mode
    the leaf value is E = 2 cut - sum_{i<j} w, which that bound does not
    bound, so the optimum can be pruned.  energy_bound=True prunes with
    2 ub - sum_{i<j} w instead, a bound on E."""
    n = W.shape[0]
    total = float(np.triu(np.asarray(W, dtype=np.float64), 1).sum())
    best_bits = np.asarray(warm_bits, dtype=bool).copy()
    best_value = warm_value
    stack = [np.full(n, -1, dtype=np.int8)]
    t_start = time.monotonic()
    explored = pruned = 0
    while stack:
        if time_limit is not None and (time.monotonic() - t_start) > time_limit:
            return best_bits, best_value, False, explored, pruned
        bs = min(batch_cap, len(stack))
        batch = np.array([stack.pop() for _ in range(bs)], dtype=np.int8)
        explored += bs
        ubs = upper_bound_batch(W, batch)
        if is_spin_glass and energy_bound:
            ubs = 2.0 * ubs - total
        leaves, interior = [], []
        for k in range(bs):
            if ubs[k] <= best_value + 1e-9:
                pruned += 1
                continue
            if int(np.sum(batch[k] < 0)) == 0:
                leaves.append(k)
            else:
                interior.append(k)
        if leaves:
            leaf_arr = batch[np.array(leaves, dtype=np.int64)]
            vals = eval_leaves(W, leaf_arr, is_spin_glass)
            bl = int(np.argmax(vals))
            if vals[bl] > best_value:
                best_value = float(vals[bl])
                best_bits = (leaf_arr[bl] >= 1).copy()
        for k in interior:
            row = batch[k]
            v = int(np.argmax(influence_scores_fast(W, row)))
            warm = int(best_bits[v])
            for val in (warm, 1 - warm):
                child = row.copy()
                child[v] = np.int8(val)
                stack.append(child)
    return best_bits, best_value, True, explored, pruned
