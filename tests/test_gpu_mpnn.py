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
