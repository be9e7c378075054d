"""The four drop-in entry points: maxcut_tfim, maxcut_tfim_sparse,
spin_glass_solver, spin_glass_solver_sparse.

Signatures, defaults, small-n paths and return tuples follow the reference
(pyqrackising/maxcut_tfim.py:342-422, maxcut_tfim_sparse.py:373-472,
spin_glass_solver.py:424-727, spin_glass_solver_sparse.py:517-864).  Every
candidate evaluation, flip evaluation and re-score runs in libpqi on the GPU;
there is no CPU fallback.

Keyword arguments added after the reference's own (all optional):
  seed               seeds the candidate sampler (the reference's is unseedable)
  gray_walkers       number of Gray walkers; the reference ties it to
                     gray_seed_multiple (spin_glass_solver.py:603)
  double_flip_mode   "reference" (spin_glass_solver.py:55-92) or "exhaustive"
                     (kernels.cl:186-231)
  seed_engine        "auto", "mfma" or "lowrank" for scoring Gray seeds
  return_stats       append the search statistics dict to the result
"""
import os

import networkx as nx
import numpy as np
import scipy.sparse as sp

from . import _lib
from . import graphs
from . import sampler as smp

HEURISTIC_THRESHOLD = 24         # pyqrackising/maxcut_tfim_util.py:165
HEURISTIC_THRESHOLD_SPARSE = 23  # pyqrackising/maxcut_tfim_util.py:166
DTYPE = np.float32               # pyqrackising/maxcut_tfim_util.py:86-88

_DOUBLE_MODES = {"reference": _lib.DOUBLE_REFERENCE, "exhaustive": _lib.DOUBLE_EXHAUSTIVE}
_SEED_ENGINES = {"auto": _lib.SEEDS_AUTO, "mfma": _lib.SEEDS_MFMA, "lowrank": _lib.SEEDS_LOWRANK}


def get_cut(solution, nodes):
    """pyqrackising/maxcut_tfim_util.py:482-494."""
    bits = "".join("1" if b else "0" for b in solution)
    l = [nodes[i] for i in range(len(solution)) if not solution[i]]
    r = [nodes[i] for i in range(len(solution)) if solution[i]]
    return bits, l, r


def int_to_bitstring(integer, length):
    """pyqrackising/maxcut_tfim_util.py:510-511 (LSB first)."""
    return (bin(integer)[2:].zfill(length))[::-1]


def _exact(dg, n, is_spin_glass):
    """pyqrackising/maxcut_tfim.py:288-339 on the GPU: all 2^n masks scored in
    one batch; strict '>' from max_cut = -1 keeps the first maximum.  The
    spin-glass variant maximises sum_{i<j} w_ij (+1 same side, -1 otherwise),
    which is -E (pyqrackising/maxcut_tfim.py:328-329).  Every mask is scored
    by the exact engine (FP64 for real weights): the approximate split-BF16
    values would decide the argmax of -E, which is never re-scored."""
    masks = np.arange(1 << n, dtype=np.uint32).reshape(-1, 1)
    best_mask, best_val = 0, -1.0
    chunk = 1 << 20
    for c0 in range(0, masks.shape[0], chunk):
        _, _, vals = dg.eval_batch(masks[c0:c0 + chunk], is_spin_glass, engine=_lib.ENGINE_EXACT)
        if is_spin_glass:
            vals = -vals
        i = int(np.argmax(vals))
        if vals[i] > best_val:
            best_val, best_mask = float(vals[i]), c0 + i
    bits = np.array([(best_mask >> i) & 1 for i in range(n)], dtype=bool)
    return bits, best_val


def _sample_and_score(dg, W_dense, n, shots, quality, is_spin_glass, anneal_t, anneal_h, repulsion_base, rng,
                      prior):
    """pyqrackising/maxcut_tfim.py:148-186 (the accelerator path): rounds of
    shots' = roundup32(shots/2) candidates scored on the device until a round
    fails to improve."""
    J_eff, degrees = prior
    cum_prob = smp.maxcut_hamming_cdf(n, J_eff, degrees, quality, anneal_t, anneal_h)
    shots = ((max(1, shots >> 1) + 31) >> 5) << 5
    best_solution = np.zeros(n, dtype=bool)
    if is_spin_glass:
        _, e0 = dg.score(best_solution)
        best_energy = e0
    else:
        best_energy = 0.0
    improved = True
    seed = int(rng.integers(0, 2 ** 63))
    rnd = 0
    while improved:
        improved = False
        # the reference draws fresh samples for shot indices 0..shots-1 each
        # round; here each round gets its own counter range of the stream
        _, bv, bits = dg.sample_round(cum_prob, repulsion_base, seed + rnd, 0, shots, is_spin_glass)
        rnd += 1
        if bv > best_energy:
            best_energy = bv
            best_solution = bits
            improved = True
    if is_spin_glass:
        best_energy, _ = dg.score(best_solution)
    return best_solution, float(best_energy)


def _ingest_dense(G):
    if isinstance(G, nx.Graph):
        nodes = list(G.nodes())
        W = graphs.dense_from_graph(G, DTYPE)
    else:
        W = np.asarray(G)
        nodes = list(range(len(W)))
    return nodes, W


def _ingest_sparse(G):
    if isinstance(G, nx.Graph):
        nodes = list(G.nodes())
        M = graphs.upper_csr_from_graph(G, DTYPE)
    else:
        M = sp.csr_matrix(G)
        nodes = list(range(M.shape[0]))
    return nodes, M


def maxcut_tfim(G, quality=None, shots=None, is_spin_glass=False, anneal_t=None, anneal_h=None,
                repulsion_base=None, is_maxcut_gpu=True, is_nested=False, seed=None, _dg=None):
    """pyqrackising/maxcut_tfim.py:342-422."""
    nodes, G_m = _ingest_dense(G)
    n = len(G_m)
    dg = _dg or _lib.DeviceGraph.dense(G_m)
    if n < HEURISTIC_THRESHOLD:
        best_solution, best_value = _exact(dg, n, is_spin_glass)
        bit_string, l, r = get_cut(best_solution, nodes)
        if best_value < 0.0:
            return "0" * n, 0.0, (nodes, [])
        return bit_string, best_value, (l, r)
    quality = 6 if quality is None else quality
    shots = (n << quality) if shots is None else shots
    anneal_t = 8.0 if anneal_t is None else anneal_t
    anneal_h = 8.0 if anneal_h is None else anneal_h
    repulsion_base = 5.0 if repulsion_base is None else repulsion_base
    rng = np.random.default_rng(seed)
    Wd = np.asarray(G_m, dtype=np.float64)
    best_solution, best_value = _sample_and_score(dg, Wd, n, shots, quality, is_spin_glass, anneal_t, anneal_h,
                                                  repulsion_base, rng, smp.init_J_and_z_dense(Wd))
    bit_string, l, r = get_cut(best_solution, nodes)
    if best_value < 0.0:
        return "0" * n, 0.0, (nodes, [])
    return bit_string, best_value, (l, r)


def maxcut_tfim_sparse(G, quality=None, shots=None, is_spin_glass=False, anneal_t=None, anneal_h=None,
                       repulsion_base=None, is_maxcut_gpu=True, is_nested=False, seed=None, _dg=None):
    """pyqrackising/maxcut_tfim_sparse.py:373-472."""
    nodes, G_m = _ingest_sparse(G)
    n = G_m.shape[0]
    dg = _dg or _lib.DeviceGraph.csr(G_m)
    if n < HEURISTIC_THRESHOLD_SPARSE:
        best_solution, best_value = _exact(dg, n, is_spin_glass)
        bit_string, l, r = get_cut(best_solution, nodes)
        if best_value < 0.0:
            return "0" * n, 0.0, (nodes, [])
        return bit_string, best_value, (l, r)
    quality = 6 if quality is None else quality
    shots = (n << quality) if shots is None else shots
    anneal_t = 8.0 if anneal_t is None else anneal_t
    anneal_h = 8.0 if anneal_h is None else anneal_h
    repulsion_base = 5.0 if repulsion_base is None else repulsion_base
    rng = np.random.default_rng(seed)
    Wsym = (G_m + G_m.T).toarray().astype(np.float64)
    best_solution, best_value = _sample_and_score(dg, Wsym, n, shots, quality, is_spin_glass, anneal_t, anneal_h,
                                                  repulsion_base, rng, smp.init_J_and_z_sparse(G_m))
    bit_string, l, r = get_cut(best_solution, nodes)
    if best_value < 0.0:
        return "0" * n, 0.0, (nodes, [])
    return bit_string, best_value, (l, r)


def _guess_bits(best_guess, n):
    if isinstance(best_guess, str):
        return best_guess
    if isinstance(best_guess, (int, np.integer)) and not isinstance(best_guess, bool):
        return int_to_bitstring(int(best_guess), n)
    if isinstance(best_guess, (list, np.ndarray)):
        return "".join(["1" if b else "0" for b in best_guess])
    return None


def _solve(G, sparse, quality, shots, best_guess, is_spin_glass, anneal_t, anneal_h, repulsion_base,
           gray_iterations, gray_seed_multiple, bp_scale, bp_damping, seed, gray_walkers, double_flip_mode,
           seed_engine, max_rounds, return_stats):
    if sparse:
        nodes, G_m = _ingest_sparse(G)
        n = G_m.shape[0]
        threshold = HEURISTIC_THRESHOLD_SPARSE
    else:
        nodes, G_m = _ingest_dense(G)
        n = len(G_m)
        threshold = HEURISTIC_THRESHOLD

    if n < 3:
        if n == 0:
            return "", 0, ([], []), 0
        if n == 1:
            return "0", 0, (nodes, []), 0
        weight = G_m[0, 1]
        if weight < 0.0:
            return "00", 0, (nodes, []), weight
        return "01", weight, ([nodes[0]], [nodes[1]]), -weight

    if n < threshold:
        best_guess = None
    dg = _lib.DeviceGraph.csr(G_m) if sparse else _lib.DeviceGraph.dense(G_m)

    bitstring = _guess_bits(best_guess, n)
    cut_value = None
    if bitstring is None:
        use_bp = bp_scale is not None and n >= threshold
        if use_bp:
            marginals, _, _ = dg.bp_marginals(max(1, int(bp_scale * n)), bp_damping)
            bp_theta = marginals > 0.0
        fn = maxcut_tfim_sparse if sparse else maxcut_tfim
        bitstring, cut_value, _ = fn(G_m, quality=quality, shots=shots, is_spin_glass=is_spin_glass,
                                     anneal_t=anneal_t, anneal_h=anneal_h, repulsion_base=repulsion_base,
                                     is_nested=True, seed=seed, _dg=dg)
        if use_bp:
            # spin_glass_solver.py:496-523: keep the BP partition if it scores higher
            k = 1 if is_spin_glass else 0
            bp_energy = dg.score(bp_theta)[k]
            sample_energy = dg.score(np.array([b == "1" for b in bitstring], dtype=bool))[k]
            if bp_energy > sample_energy:
                bitstring = "".join("1" if b else "0" for b in bp_theta)
                cut_value = bp_energy if not is_spin_glass else None
    best_theta = np.array([b == "1" for b in bitstring], dtype=bool)
    if best_theta.size != n:
        raise ValueError(f"best_guess has {best_theta.size} bits, graph has {n} nodes")

    if n < threshold:
        cut, en = dg.score(best_theta)
        bitstring, l, r = get_cut(best_theta, nodes)
        if is_spin_glass:
            return bitstring, float(cut), (l, r), float(-en)
        cv = cut if cut_value is None else cut_value
        return bitstring, float(cv), (l, r), float(en)

    if gray_seed_multiple is None:
        gray_seed_multiple = max(2, os.cpu_count() or 2)
    walkers = gray_seed_multiple if gray_walkers is None else gray_walkers
    params = _lib.make_params(n, is_spin_glass=is_spin_glass, gray_iterations=gray_iterations, walkers=walkers,
                              seed_multiple=gray_seed_multiple, double_mode=_DOUBLE_MODES[double_flip_mode],
                              max_rounds=max_rounds, seed_engine=_SEED_ENGINES[seed_engine])
    state, stats = dg.local_search(best_theta.astype(np.uint8), params)
    state = state.astype(bool)
    bitstring, l, r = get_cut(state, nodes)
    out = (bitstring, float(stats["cut_value"]), (l, r), float(stats["min_energy"]))
    return out + (stats,) if return_stats else out


def spin_glass_solver(G, quality=None, shots=None, best_guess=None, is_maxcut_gpu=True, is_spin_glass=True,
                      anneal_t=None, anneal_h=None, repulsion_base=None, is_log=False, gray_iterations=None,
                      gray_seed_multiple=None, bp_scale=None, bp_damping=0.5, seed=None, gray_walkers=None,
                      double_flip_mode="reference", seed_engine="auto", max_rounds=0, return_stats=False):
    """pyqrackising/spin_glass_solver.py:424-727.  Returns
    (bitstring, cut_value, (l, r), min_energy)."""
    return _solve(G, False, quality, shots, best_guess, is_spin_glass, anneal_t, anneal_h, repulsion_base,
                  gray_iterations, gray_seed_multiple, bp_scale, bp_damping, seed, gray_walkers, double_flip_mode,
                  seed_engine, max_rounds, return_stats)


def spin_glass_solver_sparse(G, quality=None, shots=None, best_guess=None, is_maxcut_gpu=True, is_spin_glass=True,
                             anneal_t=None, anneal_h=None, repulsion_base=None, is_log=False, gray_iterations=None,
                             gray_seed_multiple=None, bp_scale=None, bp_damping=0.5, seed=None, gray_walkers=None,
                             double_flip_mode="reference", seed_engine="auto", max_rounds=0, return_stats=False):
    """pyqrackising/spin_glass_solver_sparse.py:517-864 (absolute-value form of
    the same search; identical moves on the symmetrised matrix)."""
    return _solve(G, True, quality, shots, best_guess, is_spin_glass, anneal_t, anneal_h, repulsion_base,
                  gray_iterations, gray_seed_multiple, bp_scale, bp_damping, seed, gray_walkers, double_flip_mode,
                  seed_engine, max_rounds, return_stats)
