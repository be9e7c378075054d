"""Host restatement of the reference's neighbour-repulsion sampler, used as the
statistical reference for the device sampler.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the pyqrackising_amd
package.  It follows pyqrackising/maxcut_tfim.py:36-75 (local_repulsion_choice)
and maxcut_tfim_util.py:690-724 (sample_mag, bit_pick) in float64 numpy, drawing
from a seeded numpy generator, vectorised over shots.
"""
import numpy as np

EPSILON = 2.0 ** -23  # pyqrackising/maxcut_tfim_util.py:87-88 (float32 build)


def sample_mag(cum, p):
    """pyqrackising/maxcut_tfim_util.py:690-708 step by step (negative indices
    wrap, as in numba); after 4096 steps (the reference would not end) the
    lower bound of p."""
    left, right = 0, len(cum) - 1
    for _ in range(4096):
        m = (left + right) >> 1
        if m < -len(cum) or m >= len(cum) or m - 1 < -len(cum):
            break
        if cum[m] >= p and (m == 0 or cum[m - 1] < p):
            return m
        if cum[m] < p:
            left = m + 1
        else:
            right = m - 1
    return int(np.searchsorted(cum, p, side="left"))


class RepulsionSampler:
    """pyqrackising/maxcut_tfim.py:36-75 (local_repulsion_choice) with
    sample_mag / bit_pick (maxcut_tfim_util.py:690-724), vectorised over shots.

    Semantics kept: the first node of shot s is s mod n; a sampled m <= 1 picks
    as the reference does (m == 1 gives one node, m == 0 gives two); a node is
    "selected" when its weight is exactly 0 at the end (weights == 0.0)."""

    def __init__(self, W, cum_prob, repulsion_base, rng, chunk=None):
        self.W = np.asarray(W, dtype=np.float64)  # symmetric dense
        self.n = self.W.shape[0]
        self.cum_prob = np.asarray(cum_prob, dtype=np.float64)
        self.base = float(repulsion_base)
        self.rng = rng
        self.no_repulsion = abs(1.0 - self.base) <= EPSILON
        if not self.no_repulsion:
            self.factor = self.base ** (-self.W)  # row k: multiplier after picking k
        self.chunk = chunk or max(1, min(4096, (1 << 24) // max(1, self.n)))

    def sample(self, shots, start=0):
        n = self.n
        out = np.empty((shots, n), dtype=bool)
        for c0 in range(0, shots, self.chunk):
            c1 = min(shots, c0 + self.chunk)
            out[c0:c1] = self._chunk(np.arange(start + c0, start + c1))
        return out

    def _pick(self, weights):
        p = weights.sum(axis=1) * self.rng.random(weights.shape[0])
        cum = np.cumsum(weights, axis=1)
        hit = p[:, None] < cum
        node = np.argmax(hit, axis=1)
        node[~hit.any(axis=1)] = 0
        return node

    def _update(self, weights, rows, node):
        weights[rows, node] = 0.0
        if not self.no_repulsion:
            weights[rows] *= self.factor[node]

    def _chunk(self, s_idx):
        n = self.n
        S = s_idx.size
        m = np.array([sample_mag(self.cum_prob, p) for p in self.rng.random(S)], dtype=np.int64)
        first = s_idx % n
        weights = np.ones((S, n), dtype=np.float64)
        rows = np.arange(S)
        one = m == 1
        self._update(weights, rows, first)
        # picks 1 .. m-2 update, pick m-1 (or 1 when m == 0) only zeroes
        last_pick = np.maximum(m - 1, 1)
        tmax = int(last_pick.max()) if S else 0
        for t in range(1, tmax + 1):
            act = (t <= last_pick) & ~one
            if not act.any():
                break
            r = rows[act]
            node = self._pick(weights[r])
            upd = t < last_pick[r]
            self._update(weights, r[upd], node[upd])
            weights[r[~upd], node[~upd]] = 0.0
        sel = weights == 0.0
        sel[one] = False
        sel[rows[one], first[one]] = True
        return sel
