"""Learned flip proposals for the batched sparse search.

A 3-layer message-passing network (libpqi's pqi_mpnn_*, hidden 64, BF16 MFMA
transforms) scores every spin; the batched search visits spins in score order
and flips a spin only when its exact gain is positive at that moment, then
continues with the reference loop (single flips, double flips, Gray seeds and
walkers, reheat) as fallback and judge.  The network is trained on the GPU by
its own forward/backward/Adam kernels, on trajectories of the plain batched
search: the label of a spin is whether the plain search's final state differs
from the start state there.
"""
import numpy as np
import scipy.sparse as sp

from . import _lib, graphs as gr


def node_features(G_upper, x, age=None):
    """Features of the network (same definition as the device kernel): gain /
    strength, degree / 16, spin, min(age, 64) / 64.  G_upper: strictly upper
    CSR; x: 0/1 state."""
    A = (G_upper + G_upper.T).tocsr().astype(np.float64)
    s = 1.0 - 2.0 * np.asarray(x, dtype=np.float64)
    h = A @ s
    d = s * h
    strength = np.asarray(abs(A).sum(axis=1)).ravel()
    deg = np.diff(A.indptr).astype(np.float64)
    if age is None:
        age = np.full(A.shape[0], 64.0)
    f = np.stack([np.where(strength > 0, d / np.where(strength > 0, strength, 1.0), 0.0), deg / 16.0, s,
                  np.minimum(age, 64) / 64.0], axis=1)
    return f.astype(np.float32)


def regular_training_graphs(count, seed, n_lo=256, n_hi=2048, d_lo=3, d_hi=10):
    """Training instances from the single-instance and batched generators:
    random d-regular graphs with +-1 couplings, as upper-triangular CSR."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        d = int(rng.integers(d_lo, d_hi + 1))
        if (n * d) % 2:
            n += 1
        lo, hi = gr.random_regular_edges(n, d, rng)
        out.append(gr.upper_csr(n, lo, hi, gr.pm1(lo.size, rng).astype(np.float64), dtype=np.float32))
    return out


def er_graphs(count, n, mean_degree, seed, weights="uniform"):
    """Erdos-Renyi graphs with the given mean degree; weights U[-1, 1] (fp32)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        lo, hi = gr.erdos_renyi_edges_sparse(n, mean_degree / (n - 1), rng)
        w = rng.uniform(-1.0, 1.0, lo.size) if weights == "uniform" else gr.pm1(lo.size, rng).astype(np.float64)
        out.append(gr.upper_csr(n, lo, hi, w, dtype=np.float32))
    return out


def trajectory_dataset(G_list, starts, params):
    """Run the plain batched search; return (features, labels) over all nodes."""
    dgs = [_lib.DeviceGraph.csr(G) for G in G_list]
    finals, _ = _lib.batch_local_search(dgs, starts, params)
    feats = np.concatenate([node_features(G, x) for G, x in zip(G_list, starts)])
    labels = np.concatenate([(f != x).astype(np.float32) for f, x in zip(finals, starts)])
    return dgs, feats, labels


def train(net, dgs, feats, labels, steps=300, lr=2e-3, log=None):
    """Full-batch Adam on the GPU; returns the loss curve."""
    net.bind(dgs)
    losses = []
    for t in range(steps):
        losses.append(net.train_step(feats, labels, lr=lr))
        if log and (t % 50 == 0 or t == steps - 1):
            log(f"step {t}: loss {losses[-1]:.5f}")
    return losses


class LearnedSolver:
    """Batched solver with network-ordered first-improvement passes."""

    def __init__(self, net=None, seed=0):
        self.net = net or _lib.Mpnn(seed=seed)

    def fit(self, train_graphs, params, seed=0, steps=300, lr=2e-3, log=None):
        rng = np.random.default_rng(seed)
        starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in train_graphs]
        dgs, feats, labels = trajectory_dataset(train_graphs, starts, params)
        return train(self.net, dgs, feats, labels, steps=steps, lr=lr, log=log)

    def bind(self, dgs):
        return self.net.bind(dgs)

    def solve(self, dgs, starts, params, targets=None, passes=8, min_score=-np.inf, bind=True, fallback=False):
        if bind:
            self.net.bind(dgs)
        return _lib.learned_search(self.net, dgs, starts, params, targets=targets, passes=passes,
                                   min_score=min_score, fallback=fallback)
