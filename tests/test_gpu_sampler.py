"""The device candidate sampler (pqi_sample_candidates / pqi_sample_round)
against the reference's sampling semantics (pyqrackising/maxcut_tfim.py:36-75,
maxcut_tfim_util.py:690-724).  The reference draws from numba's unseedable
generator, so parity is structural (exact properties every sample must have)
and statistical (distributions against the host restatement in
oracle/sampler_ref.py on the same prior); the tolerances are stated per test."""
import sys

import numpy as np
import pytest

from pyqrackising_amd import _lib, graphs
from pyqrackising_amd import sampler as smp

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected_count(m):
    # local_repulsion_choice: m == 1 -> the first node; m == 0 -> first + one pick
    return np.where(m == 1, 1, np.where(m == 0, 2, m))


def regular_dense(n, d, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, d, rng)
    return graphs.dense_from_edges(n, lo, hi, np.ones(lo.size), dtype=np.float64)


@pytest.mark.parametrize("base", [5.0, 1.0])
def test_sampler_structure(base):
    """Every shot selects the first node s mod n, and as many nodes as its
    Hamming-weight draw allows (picks never repeat a node)."""
    n = 200
    W = regular_dense(n, 4, 1)
    J, z = smp.init_J_and_z_dense(W)
    cum = smp.maxcut_hamming_cdf(n, J, z, 4, 8.0, 8.0)
    dg = _lib.DeviceGraph.dense(W)
    shots, off = 3000, 123
    X = dg.sample_candidates(cum, base, 7, off, shots)
    s = off + np.arange(shots)
    assert X[np.arange(shots), s % n].all()
    counts = X.sum(axis=1)
    # the count pins the Hamming-weight draw: m == 1 -> 1, m == 0 -> 2, else m
    assert (counts >= 1).all() and (counts <= cum.shape[0]).all()
    # counter-based stream: the same shot index gives the same sample
    Y = dg.sample_candidates(cum, base, 7, off + 1000, 1000)
    assert np.array_equal(X[1000:2000], Y)
    Z = dg.sample_candidates(cum, base, 8, off, 50)
    assert not np.array_equal(X[:50], Z)


def test_sampler_hamming_distribution():
    """On a monotone CDF the selected counts follow it: chi-square of the count
    histogram against m ~ cum_prob mapped to counts as the reference does
    (m == 1 -> 1, m < 2 -> 2, else m); p > 1e-4."""
    from scipy import stats

    n = 120
    W = regular_dense(n, 3, 2)
    rng = np.random.default_rng(3)
    pm = rng.random(60) ** 3
    cum = np.cumsum(pm / pm.sum())
    cum[-1] = 1.0
    dg = _lib.DeviceGraph.dense(W)
    shots = 40000
    counts = dg.sample_candidates(cum, 5.0, 11, 0, shots).sum(axis=1)
    pmf = np.diff(np.concatenate([[0.0], cum]))
    expect = np.zeros(cum.shape[0] + 2)
    np.add.at(expect, expected_count(np.arange(cum.shape[0])), pmf)
    obs = np.bincount(counts, minlength=expect.shape[0])[:expect.shape[0]]
    keep = expect * shots >= 5
    e = expect[keep] * shots
    o = obs[keep] * (e.sum() / obs[keep].sum())
    chi2 = ((o - e) ** 2 / e).sum()
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    log(f"  hamming chi2={chi2:.1f} dof={keep.sum() - 1} p={p:.3g}")
    assert p > 1e-4


def test_sampler_raw_prior_against_host():
    """The reference's prior for this graph is not monotone, so the Hamming
    weight comes from sample_mag's own search path: the device and the host
    restatement give the same count distribution (two-sample chi-square,
    p > 1e-4)."""
    from scipy import stats
    from oracle import sampler_ref

    n = 120
    W = regular_dense(n, 3, 2)
    J, z = smp.init_J_and_z_dense(W)
    cum = smp.maxcut_hamming_cdf(n, J, z, 3, 8.0, 8.0)
    assert (np.diff(cum) < 0).any()  # the case this test is about
    dg = _lib.DeviceGraph.dense(W)
    shots = 20000
    cd = dg.sample_candidates(cum, 5.0, 12, 0, shots).sum(axis=1)
    ch = sampler_ref.RepulsionSampler(W, cum, 5.0, np.random.default_rng(8)).sample(shots, 0).sum(axis=1)
    top = max(cd.max(), ch.max()) + 1
    a, b = np.bincount(cd, minlength=top), np.bincount(ch, minlength=top)
    keep = (a + b) >= 10
    table = np.stack([a[keep], b[keep]])
    chi2, p, dof, _ = stats.chi2_contingency(table)
    log(f"  raw prior: device mean {cd.mean():.2f} host mean {ch.mean():.2f} chi2={chi2:.1f} dof={dof} p={p:.3g}")
    assert p > 1e-4


def test_sampler_repulsion_against_host():
    """Repulsion statistics match the host restatement: the mean number of
    selected edges (both ends selected) per shot agrees within 5 standard
    errors, and repulsion lowers it against base 1."""
    from oracle import sampler_ref

    n = 150
    W = regular_dense(n, 6, 3)
    J, z = smp.init_J_and_z_dense(W)
    cum = smp.maxcut_hamming_cdf(n, J, z, 4, 8.0, 8.0)
    dg = _lib.DeviceGraph.dense(W)
    shots = 6000
    means = {}
    for base in (5.0, 1.0):
        Xd = dg.sample_candidates(cum, base, 21, 0, shots).astype(np.float64)
        Xh = sampler_ref.RepulsionSampler(W, cum, base, np.random.default_rng(5)).sample(shots, 0).astype(np.float64)
        ed = np.einsum("si,ij,sj->s", Xd, W, Xd) / 2
        eh = np.einsum("si,ij,sj->s", Xh, W, Xh) / 2
        se = np.sqrt(ed.var() / shots + eh.var() / shots)
        log(f"  base {base}: device {ed.mean():.3f} host {eh.mean():.3f} se {se:.3f}")
        assert abs(ed.mean() - eh.mean()) <= 5 * se + 1e-9
        means[base] = ed.mean()
    assert means[5.0] < means[1.0]


def test_sampler_uniform_marginals_without_repulsion():
    """Base 1: apart from the first node, picks are uniform over the remaining
    nodes, so every node's selection frequency, first-node shots excluded, is
    equal within 5 standard errors."""
    n = 64
    W = regular_dense(n, 3, 4)
    cum = np.linspace(0.0, 1.0, 20)[1:]
    dg = _lib.DeviceGraph.dense(W)
    shots = 64 * 600
    X = dg.sample_candidates(cum, 1.0, 33, 0, shots)
    first = np.arange(shots) % n
    X[np.arange(shots), first] = False
    freq = X.sum(axis=0) / shots
    mean = freq.mean()
    se = np.sqrt(mean * (1 - mean) / shots)
    assert np.abs(freq - mean).max() <= 5 * se


def test_sample_round_matches_sample_then_eval(orc):
    """pqi_sample_round = pqi_sample_candidates + exact scoring + first argmax."""
    n = 300
    W = regular_dense(n, 3, 5)
    W[W != 0] *= np.where(np.random.default_rng(6).random((W != 0).sum()) < 0.5, -1, 1)
    W = np.triu(W, 1)
    W = W + W.T
    J, z = smp.init_J_and_z_dense(W)
    cum = smp.maxcut_hamming_cdf(n, J, z, 3, 8.0, 8.0)
    dg = _lib.DeviceGraph.dense(W.astype(np.int8))
    for sg in (0, 1):
        bi, bv, bits = dg.sample_round(cum, 5.0, 44, 10, 1024, sg)
        X = dg.sample_candidates(cum, 5.0, 44, 10, 1024)
        obi, ovals = orc.eval_batch(W, X, sg)
        assert bi == obi and bv == ovals[obi]
        assert np.array_equal(bits, X[obi])


def test_sampler_sparse_graph():
    """CSR graphs sample through their adjacency lists: same structure rules."""
    n = 500
    rng = np.random.default_rng(9)
    lo, hi = graphs.erdos_renyi_edges_sparse(n, 8.0 / (n - 1), rng)
    G = graphs.upper_csr(n, lo, hi, rng.uniform(-1, 1, lo.size), dtype=np.float32)
    J, z = smp.init_J_and_z_sparse(G)
    cum = smp.maxcut_hamming_cdf(n, J, z, 3, 8.0, 8.0)
    dg = _lib.DeviceGraph.csr(G)
    X = dg.sample_candidates(cum, 5.0, 3, 0, 2000)
    assert X[np.arange(2000), np.arange(2000) % n].all()
    assert (X.sum(axis=1) <= cum.shape[0]).all()
