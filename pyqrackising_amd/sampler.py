"""Host side of the candidate sampler: the TFIM Hamming-weight prior
(pyqrackising/maxcut_tfim_util.py:545-687; maxcut_tfim.py:189-209;
maxcut_tfim_sparse.py:204-224), restated in float64 numpy.  Its output, the
cumulative Hamming-weight distribution, drives the device sampler
(pqi_sample_candidates / pqi_sample_round), which draws the candidates with
neighbour repulsion on the GPU."""
import math

import numpy as np

EPSILON = 2.0 ** -23  # pyqrackising/maxcut_tfim_util.py:87-88 (float32 build)


# ---- prior ------------------------------------------------------------------

def init_J_and_z_dense(G_m):
    """pyqrackising/maxcut_tfim.py:189-209."""
    G = np.asarray(G_m, dtype=np.float64)
    G_min = G.min()
    val = G - G_min
    mask = val > EPSILON
    degrees = mask.sum(axis=1).astype(np.uint32)
    J = np.where(mask, val, 0.0).sum(axis=1)
    J_eff = np.where(degrees > 0, -J / np.maximum(degrees, 1), 0.0)
    return J_eff, degrees


def init_J_and_z_sparse(G_m):
    """pyqrackising/maxcut_tfim_sparse.py:204-224 on the upper-triangular CSR:
    every one of the n entries of a row counts, implicit zeros included."""
    n = G_m.shape[0]
    data = np.asarray(G_m.data, dtype=np.float64)
    indptr = np.asarray(G_m.indptr)
    nnz_row = np.diff(indptr)
    G_min = min(data.min() if data.size else 0.0, 0.0 if (nnz_row < n).any() else np.inf)
    rows = np.repeat(np.arange(n), nnz_row)
    val = data - G_min
    m = val > EPSILON
    degrees = np.bincount(rows[m], minlength=n).astype(np.int64)
    J = np.bincount(rows[m], weights=val[m], minlength=n).astype(np.float64)
    zval = 0.0 - G_min
    if zval > EPSILON:
        implicit = n - nnz_row
        degrees += implicit
        J += implicit * zval
    J_eff = np.where(degrees > 0, -J / np.maximum(degrees, 1), 0.0)
    return J_eff, degrees.astype(np.uint32)


def init_thresholds(n_qubits):
    """pyqrackising/maxcut_tfim_util.py:558-575.  The reference allocates with
    np.empty and never writes one entry (index n/2 for even n, (n+1)/2 for odd
    n), so that entry holds whatever the allocator returns; it is 0 here."""
    n_bias = n_qubits + 1
    thresholds = np.zeros(n_bias, dtype=np.float64)
    normalizer = 0
    for q in range(n_qubits >> 1):
        normalizer += math.comb(n_qubits, q) << 1
    if n_qubits & 1:
        normalizer += math.comb(n_qubits, n_qubits >> 1)
    p = 1
    for q in range(n_qubits >> 1):
        val = p / normalizer
        thresholds[q] = val
        thresholds[n_bias - (q + 1)] = val
        p = math.comb(n_qubits, q + 1)
    if n_qubits & 1:
        thresholds[n_qubits >> 1] = p / normalizer
    return thresholds


def _theta_c(J, h, z):
    zJ = z * J
    if abs(zJ) < EPSILON:
        return (np.pi if J > 0 else -np.pi) / 2
    return np.arcsin(max(-1.0, min(1.0, h / zJ)))


def _time_envelope(bias, t, n_qubits):
    """pyqrackising/maxcut_tfim_util.py:578-623 (alpha = 1)."""
    half = n_qubits / 2.0
    t_eff = t if t > 1e-9 else 1e-9
    k = np.arange(n_qubits + 1, dtype=np.float64)
    out = bias * np.exp(-np.abs(k - half) / t_eff)
    tot = out.sum()
    if tot <= 0.0:
        out = np.zeros(n_qubits + 1)
        out[n_qubits // 2] = 1.0
        return out
    return out / tot


def probability_by_hamming_weight(J, h, z, theta, t, n_bias, normalized=True, omega=1.5 * np.pi):
    """pyqrackising/maxcut_tfim_util.py:626-651."""
    theta_c = _theta_c(J, h, z)
    p = (2.0 ** (abs(J / h) - 1.0) * (1.0 + np.sin(theta - theta_c) * np.cos(omega * J * t + theta)
                                       / (1.0 + np.sqrt(t))) - 0.5)
    factor = 2.0 ** -(p / n_bias)
    bias = factor ** np.arange(1, n_bias + 1, dtype=np.float64)
    if J > 0.0:
        return bias[::-1]
    if normalized:
        bias = bias / bias.sum()
    return _time_envelope(bias / bias.sum(), t, n_bias - 1)


def maxcut_hamming_cdf(n_qubits, J_func, degrees, quality, tot_t, h_mult, omega=1.5 * np.pi):
    """pyqrackising/maxcut_tfim_util.py:654-687 (with init_theta, :545-555).
    Nodes sharing (J, degree) share every bias vector, so each distinct pair
    is evaluated once per step and weighted by its multiplicity."""
    hamming_prob = init_thresholds(n_qubits)
    n_steps = 1 << quality
    delta_t = tot_t / n_steps
    n_bias = n_qubits + 1
    h_abs = abs(h_mult)
    pairs = {}
    for q in range(n_qubits):
        J = float(J_func[q])
        if abs(J) < EPSILON:
            continue
        key = (J, int(degrees[q]))
        pairs[key] = pairs.get(key, 0) + 1
    for (J, z), mult in pairs.items():
        theta_eff = _theta_c(J, h_abs, z)
        for step in range(1, n_steps):
            t = step * delta_t
            tm1 = (step - 1) * delta_t
            h_t = h_mult * (tot_t - t)
            bias = probability_by_hamming_weight(J, h_t, z, theta_eff, t, n_bias, False, omega)
            last = probability_by_hamming_weight(J, h_t, z, theta_eff, tm1, n_bias, False, omega)
            hamming_prob += mult * (bias - last)
    hamming_prob /= hamming_prob.sum() - (hamming_prob[0] + hamming_prob[-1])
    cum_prob = np.cumsum(hamming_prob[1:-1])
    cum_prob[-1] = 1.0
    return cum_prob
