"""Exact MAXCUT / spin-glass solvers by warm-started branch and bound:
solve_maxcut_bnb and solve_maxcut_bnb_sparse
(pyqrackising/solve_maxcut_bnb.py:208-321, solve_maxcut_bnb_sparse.py:213-322).

The warm start comes from this package's spin_glass_solver(_sparse), as in
the reference; the branch-and-bound loop and its batched node evaluation
(upper bound, leaf value, branch variable) run in libpqi (pqi_bnb), the
evaluation as one HIP kernel launch per batch of nodes.  The loop pops, prunes,
adopts leaves and pushes children exactly as the reference's
_branch_and_bound does (solve_maxcut_bnb.py:128-201), so with the same
batch_cap (the reference's is os.cpu_count()**2) the explored and pruned
counts, the certified flag and the returned bitstring are the reference's.

Both solvers evaluate nodes over the dense symmetric matrix: the sparse
reference's CSR kernels (solve_maxcut_bnb_sparse.py:32-114) compute the same
bound, leaf values and influence scores from a strictly upper-triangular CSR
matrix, and branch and bound is exponential in n, so n stays small.

The reference prunes spin-glass nodes with the cut bound, which does not bound
the energy it maximises (E = 2 cut - sum_{i<j} w), so its certified spin-glass
optimum can be wrong; that behaviour is kept by default.

Keyword arguments added after the reference's own (all optional):
  batch_cap     nodes evaluated per launch (default os.cpu_count()**2)
  energy_bound  spin-glass mode: prune with 2 ub - sum_{i<j} w, a valid bound,
                so that the certified optimum is exact
  seed          forwarded to the warm-start heuristic's sampler
  return_stats  append the search statistics dict to the result
"""
import os
import time

import networkx as nx
import numpy as np
import scipy.sparse as sp

from . import _lib
from . import graphs
from .solvers import DTYPE, get_cut, int_to_bitstring, spin_glass_solver, spin_glass_solver_sparse


def _small(n, nodes, G_m):
    """n < 3 (solve_maxcut_bnb.py:257-266)."""
    if n == 0:
        return "", 0, ([], []), 0, True
    if n == 1:
        return "0", 0, (nodes, []), 0, True
    weight = G_m[0, 1]
    if weight < 0.0:
        return "00", 0, (nodes, []), weight, True
    return "01", weight, ([nodes[0]], [nodes[1]]), -weight, True


def _guess(best_guess, n):
    """The warm-start formats the reference accepts (solve_maxcut_bnb.py:271-276);
    None means: run the heuristic."""
    if isinstance(best_guess, str):
        return best_guess
    if isinstance(best_guess, int):
        return int_to_bitstring(best_guess, n)
    if isinstance(best_guess, list):
        return "".join(["1" if b else "0" for b in best_guess])
    return None


def _heuristic_kwargs(quality, shots, anneal_t, anneal_h, repulsion_base, gray_iterations, gray_seed_multiple,
                      is_maxcut_gpu, seed):
    """solve_maxcut_bnb.py:281-289: only the arguments the caller set."""
    kwargs = {}
    for k, v in (("quality", quality), ("shots", shots), ("anneal_t", anneal_t), ("anneal_h", anneal_h),
                 ("repulsion_base", repulsion_base), ("gray_iterations", gray_iterations),
                 ("gray_seed_multiple", gray_seed_multiple), ("seed", seed)):
        if v is not None:
            kwargs[k] = v
    kwargs["is_maxcut_gpu"] = is_maxcut_gpu
    return kwargs


def _run(W_bnb, nodes, n, heuristic, G_heur, best_guess, is_spin_glass, verbose, time_limit, batch_cap,
         energy_bound, return_stats, kwargs):
    """solve_maxcut_bnb.py:268-321 from the warm start on."""
    bitstring = _guess(best_guess, n)
    cut_value = None
    energy_value = None
    if bitstring is None:
        if verbose:
            print("Running PyQrackIsing heuristic...")
        t0 = time.monotonic()
        bitstring, cut_value, _, energy_value = heuristic(G_heur, is_spin_glass=is_spin_glass, **kwargs)
        if verbose:
            print(f"Heuristic value: {cut_value:.6f}  ({time.monotonic()-t0:.3f}s)")
    best_theta = np.array([b == "1" for b in list(bitstring)], dtype=np.bool_)
    if best_theta.size != n:
        raise ValueError(f"best_guess has {best_theta.size} bits, graph has {n} nodes")

    dg = _lib.DeviceGraph.dense(W_bnb)
    try:
        if is_spin_glass:
            max_energy = dg.score(best_theta)[1] if energy_value is None else energy_value
        elif cut_value is None:
            max_energy = dg.score(best_theta)[0]
        else:
            max_energy = cut_value

        if verbose:
            print("Starting branch and bound...")
            print(f"Warm-start incumbent: {max_energy:.6f}")
        if batch_cap is None:
            batch_cap = (os.cpu_count() or 1) ** 2
        best_theta, max_energy, stats = dg.bnb(best_theta, max_energy, is_spin_glass, batch_cap, time_limit,
                                               energy_bound)
        certified = bool(stats["certified"])
        if verbose:
            if certified:
                print(f"\nExact optimum: {max_energy:.6f}")
                print(f"Nodes explored: {stats['nodes_explored']}  |  Pruned: {stats['nodes_pruned']}  |  "
                      f"Time: {stats['seconds']:.3f}s")
            else:
                print(f"Time limit reached. Nodes: {stats['nodes_explored']}, Pruned: {stats['nodes_pruned']}")

        bitstring, l, r = get_cut(best_theta, nodes)
        cut, energy = dg.score(best_theta)
    finally:
        dg.close()
    if is_spin_glass:
        cut_value = cut
        min_energy = -max_energy
    else:
        cut_value = max_energy
        min_energy = energy
    out = (bitstring, float(cut_value), (l, r), float(min_energy), certified)
    return out + (stats,) if return_stats else out


def _mirror_upper(W):
    """The upper triangle mirrored: every quantity the reference's bound, leaf
    and influence kernels read is an i < j entry."""
    U = np.triu(W, 1)
    return U + U.T


def solve_maxcut_bnb(G, best_guess=None, quality=None, shots=None, is_spin_glass=False, anneal_t=None, anneal_h=None,
                     repulsion_base=None, is_maxcut_gpu=True, verbose=True, time_limit=None, gray_iterations=None,
                     gray_seed_multiple=None, batch_cap=None, energy_bound=False, seed=None, return_stats=False):
    """pyqrackising/solve_maxcut_bnb.py:208-321.  Returns
    (bitstring, cut_value, (l, r), min_energy, certified)."""
    if isinstance(G, nx.Graph):
        nodes = list(G.nodes())
        n = len(nodes)
        G_m = graphs.dense_from_graph(G, DTYPE)
    else:
        n = len(G)
        nodes = list(range(n))
        G_m = np.asarray(G, dtype=DTYPE)
    if n < 3:
        return _small(n, nodes, G_m)
    kwargs = _heuristic_kwargs(quality, shots, anneal_t, anneal_h, repulsion_base, gray_iterations,
                               gray_seed_multiple, is_maxcut_gpu, seed)
    return _run(_mirror_upper(G_m), nodes, n, spin_glass_solver, G_m, best_guess, is_spin_glass, verbose,
                time_limit, batch_cap, energy_bound, return_stats, kwargs)


def solve_maxcut_bnb_sparse(G, best_guess=None, quality=None, shots=None, is_spin_glass=False, anneal_t=None,
                            anneal_h=None, repulsion_base=None, is_maxcut_gpu=True, verbose=True, time_limit=None,
                            gray_iterations=None, gray_seed_multiple=None, batch_cap=None, energy_bound=False,
                            seed=None, return_stats=False):
    """pyqrackising/solve_maxcut_bnb_sparse.py:213-322 (strictly
    upper-triangular CSR input, as every sparse entry point here).  Returns
    (bitstring, cut_value, (l, r), min_energy, certified)."""
    if isinstance(G, nx.Graph):
        nodes = list(G.nodes())
        n = len(nodes)
        G_m = graphs.upper_csr_from_graph(G, DTYPE)
    else:
        G_m = sp.csr_matrix(G)
        n = G_m.shape[0]
        nodes = list(range(n))
    if n < 3:
        return _small(n, nodes, G_m)
    if sp.tril(G_m, 0).nnz:
        raise ValueError("sparse input must be strictly upper triangular")
    U = G_m.astype(np.float64).toarray()
    kwargs = _heuristic_kwargs(quality, shots, anneal_t, anneal_h, repulsion_base, gray_iterations,
                               gray_seed_multiple, is_maxcut_gpu, seed)
    return _run(U + U.T, nodes, n, spin_glass_solver_sparse, G_m, best_guess, is_spin_glass, verbose, time_limit,
                batch_cap, energy_bound, return_stats, kwargs)
