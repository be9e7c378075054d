"""The flip-proposal network's forward and backward kernels against a float64
torch CPU restatement of the same function (the gradient oracle), and a
training check."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import _lib, graphs

pytestmark = pytest.mark.gpu

HID, F0 = 64, 4


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def unpack(p):
    o = 0
    out = {}
    for name, shape in (("T1", (2 * F0, HID)), ("B1", (HID,)), ("T2", (2 * HID, HID)), ("B2", (HID,)),
                        ("T3", (2 * HID, HID)), ("B3", (HID,)), ("TO", (HID,)), ("BO", (1,))):
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    assert o == p.shape[0]
    return out


def torch_loss_grad(p, feats, labels, A):
    import torch

    P = torch.tensor(p.astype(np.float64), requires_grad=True)
    q = unpack(P)
    At = torch.sparse_csr_tensor(torch.tensor(A.indptr, dtype=torch.int64), torch.tensor(A.indices, dtype=torch.int64),
                                 torch.tensor(A.data, dtype=torch.float64), size=A.shape)
    H = torch.tensor(feats.astype(np.float64))
    for l in (1, 2, 3):
        M = At @ H
        Z = torch.cat([H, M], dim=1) @ q[f"T{l}"] + q[f"B{l}"]
        H = torch.relu(Z)
    s = H @ q["TO"] + q["BO"]
    y = torch.tensor(labels.astype(np.float64))
    loss = (torch.nn.functional.softplus(s) - y * s).mean()
    loss.backward()
    return loss.item(), P.grad.numpy(), s.detach().numpy()


def make_batch(seed):
    rng = np.random.default_rng(seed)
    Gs, blocks = [], []
    for n in (60, 140, 257):
        lo, hi = graphs.erdos_renyi_edges(n, 6.0 / n, rng)
        w = rng.uniform(-1, 1, lo.size)
        G = graphs.upper_csr(n, lo, hi, w, dtype=np.float32)
        Gs.append(G)
        S = (G + G.T).tocsr()
        S.sort_indices()
        blocks.append(S.astype(np.float64))
    A = sp.block_diag(blocks, format="csr")
    A.sort_indices()
    N = A.shape[0]
    feats = np.stack([rng.normal(size=N), rng.integers(1, 12, N) / 16.0, rng.choice([-1.0, 1.0], N),
                      rng.integers(0, 65, N) / 64.0], axis=1).astype(np.float32)
    labels = (rng.random(N) < 0.3).astype(np.float32)
    return Gs, A, feats, labels


@pytest.mark.parametrize("precision", [0, 1])
def test_mpnn_gradients_against_float64(precision):
    log(f"mpnn gradient check precision={precision}")
    Gs, A, feats, labels = make_batch(71)
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=5, precision=precision)
    N = net.bind(dgs)
    assert N == A.shape[0]
    p = net.params()
    loss, g = net.loss_grad(feats, labels)
    rl, rg, rs = torch_loss_grad(p, feats, labels, A)
    scores = net.forward(feats)
    if precision == 0:
        assert loss == pytest.approx(rl, rel=1e-5)
        np.testing.assert_allclose(scores, rs, rtol=1e-4, atol=1e-4)
        assert np.max(np.abs(g - rg)) <= 1e-4 * np.max(np.abs(rg)) + 1e-7
    else:
        # BF16 operands in layers 2 and 3: about 3 significant digits
        assert loss == pytest.approx(rl, rel=2e-2)
        assert np.max(np.abs(scores - rs)) <= 3e-2 * (1 + np.max(np.abs(rs)))
        cos = float(g @ rg / (np.linalg.norm(g) * np.linalg.norm(rg)))
        assert cos > 0.99, cos


def test_mpnn_training_reduces_loss():
    """Adam on a target the features determine (spins with positive gain)."""
    log("mpnn training")
    Gs, A, feats, _ = make_batch(72)
    labels = (feats[:, 0] > 0).astype(np.float32)
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=6)
    net.bind(dgs)
    first = net.train_step(feats, labels, lr=3e-3)
    for _ in range(150):
        last = net.train_step(feats, labels, lr=3e-3)
    log(f"  loss {first:.4f} -> {last:.4f}")
    assert last < 0.5 * first
    pred = net.forward(feats) > 0
    assert (pred == (labels > 0)).mean() > 0.9


# ---------------------------------------------------------------------------
# learned-proposal search
# ---------------------------------------------------------------------------

def _er_int(n, deg, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges(n, deg / (n - 1), rng)
    w = rng.integers(-3, 4, lo.size).astype(np.float64)
    keep = w != 0
    G = graphs.upper_csr(n, lo[keep], hi[keep], w[keep], dtype=np.float32)
    W = (G + G.T).toarray().astype(np.float64)
    return G, W


def test_learned_passes_in_index_order_match_oracle(orc):
    """With all scores equal the score order is the index order: the learned
    phase is index-order first-improvement passes, then the reference loop."""
    log("learned search, index order vs oracle")
    rng = np.random.default_rng(81)
    Gs, Ws, x0s = [], [], []
    for t in range(5):
        G, W = _er_int(int(rng.integers(60, 200)), 5.0, 300 + t)
        Gs.append(G)
        Ws.append(W)
        x0s.append(rng.integers(0, 2, W.shape[0]).astype(np.uint8))
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=1)
    net.set_params(np.zeros(net.n_params, np.float32))
    net.bind(dgs)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=800)
    states, stats = _lib.learned_search(net, dgs, x0s, p, passes=50)
    for i, (W, x0) in enumerate(zip(Ws, x0s)):
        x1, flips, passes, gain = orc.first_improvement(W, x0, np.arange(W.shape[0]), 50)
        assert stats[i]["learned_accepts"] == flips and stats[i]["learned_passes"] == passes, i
        xc, r = orc.dense_local_search(W, x1, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=800)
        assert np.array_equal(states[i], xc), i
        assert stats[i]["min_energy"] == r["min_energy"]
        assert stats[i]["max_energy"] == -r["min_energy"]  # every accepted move was a true gain


def test_learned_search_moves_are_exact(orc):
    """Arbitrary network scores: every accepted learned move is a true
    improvement (exact bookkeeping) and the search ends single-flip optimal."""
    log("learned search exactness")
    rng = np.random.default_rng(82)
    Gs, Ws, x0s = [], [], []
    for t in range(4):
        G, W = _er_int(int(rng.integers(80, 300)), 6.0, 400 + t)
        Gs.append(G)
        Ws.append(W)
        x0s.append(rng.integers(0, 2, W.shape[0]).astype(np.uint8))
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=9)
    net.bind(dgs)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=500)
    states, stats = _lib.learned_search(net, dgs, x0s, p, passes=6, min_score=0.0)
    for i, W in enumerate(Ws):
        e = orc.compute_energy(W, states[i])
        assert stats[i]["max_energy"] == e and stats[i]["min_energy"] == -e
        g, _ = orc.single_bit_flips(states[i], 1, W)
        assert g <= 0


def test_adversarial_network_reaches_the_plain_result():
    """Adversarial weights: the score is 10 relu(-f0), f0 = gain / strength, so
    only spins whose flip LOSES are proposed (min_score 1), worst first.  The
    exact gain check rejects every one of them, and the reference loop that
    follows as fallback and judge ends exactly where the plain batched search
    from the same starts ends (same states, same counters)."""
    log("adversarial network")
    rng = np.random.default_rng(83)
    Gs, x0s = [], []
    for t in range(6):
        G, W = _er_int(int(rng.integers(100, 400)), 6.0, 500 + t)
        Gs.append(G)
        x0s.append(rng.integers(0, 2, W.shape[0]).astype(np.uint8))
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=1)
    q = unpack(np.zeros(net.n_params, np.float32))
    q["T1"][0, 0] = -1.0  # h0 = relu(-f0)
    q["T2"][0, 0] = 1.0
    q["T3"][0, 0] = 1.0
    q["TO"][0] = 10.0
    net.set_params(np.concatenate([q[k].ravel() for k in ("T1", "B1", "T2", "B2", "T3", "B3", "TO", "BO")]))
    net.bind(dgs)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=600)
    states, stats = _lib.learned_search(net, dgs, x0s, p, passes=4, min_score=1.0)
    plain_states, plain = _lib.batch_local_search(dgs, x0s, p)
    for i in range(len(dgs)):
        assert stats[i]["learned_passes"] >= 1 and stats[i]["learned_accepts"] == 0, i
        assert np.array_equal(states[i], plain_states[i]), i
        for k in ("rounds", "single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "max_energy",
                  "min_energy", "cut_value"):
            assert stats[i][k] == plain[i][k], (i, k)
