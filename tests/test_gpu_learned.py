"""The learned-proposal search on the C5 workload's graph family: Erdos-Renyi
CSR graphs with couplings U[-1, 1] stored as fp32.

* One learned pass in the network's own score order, then the reference loop,
  against the oracle (first_improvement over the same order, then
  dense_local_search): identical states and counters, energies within 1e-9
  relative.  The order is rebuilt on the host from the network's scores of
  host-computed features (the device's feature definition and summation order,
  so the features are bit-identical) and the device's 256-bucket stable
  counting sort.
* The full C5 shape, 64 x ER(8192, mean degree 8), with a trained network:
  per instance the bookkeeping equals the exact re-score, the final state is
  single-flip optimal and every accepted learned move had a positive gain.
* What training buys: after four learned passes over every spin in score
  order (no loop), the trained network reaches higher energies than the same
  passes in index order or with the untrained network.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import _lib, graphs, learned

pytestmark = pytest.mark.gpu

REL = 1e-9


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def er_uniform(n, mean_degree, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges_sparse(n, mean_degree / (n - 1), rng)
    return graphs.upper_csr(n, lo, hi, rng.uniform(-1.0, 1.0, lo.size), dtype=np.float32)


def device_features(G, x):
    """Node features exactly as kb_features forms them from a fresh field:
    h_v summed in double over the sorted symmetric row (kb_field's order),
    f0 = float(s_v h_v / sum |w|), f1 = deg / 16, f2 = s_v, f3 = 1 (age 64)."""
    A = (G + G.T).tocsr()
    A.sort_indices()
    ap, ai, aw = A.indptr, A.indices, A.data.astype(np.float32)
    n = A.shape[0]
    f = np.zeros((n, 4), np.float32)
    for v in range(n):
        acc = 0.0
        st = 0.0
        for e in range(ap[v], ap[v + 1]):
            w = float(aw[e])
            acc += -w if x[ai[e]] else w
            st += abs(w)
        sv = -1.0 if x[v] else 1.0
        f[v, 0] = np.float32((sv * acc) / st) if st > 0 else np.float32(0)
        f[v, 1] = np.float32(ap[v + 1] - ap[v]) * np.float32(1.0 / 16.0)
        f[v, 2] = np.float32(sv)
        f[v, 3] = np.float32(1.0)
    return f


def device_order(scores):
    """kb_order: 256 buckets over [lo, hi] of the float32 scores, highest
    first, index order inside a bucket."""
    sc = scores.astype(np.float32)
    lo, hi = sc.min(), sc.max()
    span = np.float32(hi - lo) if hi > lo else np.float32(1.0)
    b = ((np.float32(hi) - sc) / span * np.float32(255.0)).astype(np.int64)
    b = np.clip(b, 0, 255)
    return np.argsort(b, kind="stable")


@pytest.mark.parametrize("min_score", [-np.inf, 0.0])
def test_learned_pass_in_network_order_matches_oracle(orc, min_score):
    log(f"learned pass in network order vs oracle, U[-1,1] fp32, min_score={min_score}")
    rng = np.random.default_rng(91)
    Gs, Ws, x0s = [], [], []
    for t in range(5):
        G = er_uniform(int(rng.integers(120, 320)), 8.0, 700 + t)
        Gs.append(G)
        Ws.append((G + G.T).toarray().astype(np.float64))
        x0s.append(rng.integers(0, 2, G.shape[0]).astype(np.uint8))
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    net = _lib.Mpnn(seed=17)
    net.bind(dgs)
    feats = np.concatenate([device_features(G, x) for G, x in zip(Gs, x0s)])
    scores = net.forward(feats)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=800)
    states, stats = _lib.learned_search(net, dgs, x0s, p, passes=1, min_score=min_score)
    off = 0
    for i, (W, x0) in enumerate(zip(Ws, x0s)):
        n = W.shape[0]
        sc = scores[off:off + n]
        off += n
        order = device_order(sc)
        order = order[sc[order] > np.float32(min_score)]
        x1, flips, passes, gain = orc.first_improvement(W, x0, order, 1)
        log(f"  instance {i}: n={n}, {order.size} spins visited, {flips} learned flips")
        assert stats[i]["learned_passes"] == 1 and stats[i]["learned_accepts"] == flips, i
        assert flips > 0
        xc, r = orc.dense_local_search(W, x1, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=800)
        assert np.array_equal(states[i], xc), i
        for k in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts"):
            assert stats[i][k] == r[k], (i, k)
        assert stats[i]["min_energy"] == pytest.approx(r["min_energy"], rel=REL)
        e0 = orc.compute_energy(W, x0)
        assert stats[i]["max_energy"] == pytest.approx(-r["min_energy"], rel=REL)
        assert stats[i]["learned_min_gain"] > 0
        assert e0 + 2.0 * gain <= -r["min_energy"] + REL * abs(r["min_energy"])


def test_learned_search_c5_full_size_properties():
    """64 x ER(8192, 8), U[-1, 1] fp32, the trained network, four passes of
    confident proposals, then the loop with n^2 Gray iterations (16 walkers x
    8): per instance, max_energy equals the exact FP64 re-score, the final
    state is single-flip optimal, every accepted learned move gained."""
    log("learned search at the full C5 shape")
    P = _lib.make_params(0, is_spin_glass=1, walkers=16, seed_multiple=8)
    P.gray_iterations = -1
    solver = learned.LearnedSolver(seed=11)
    solver.fit(learned.regular_training_graphs(24, seed=3001, n_hi=1024), P, seed=5, steps=150, lr=2e-3)
    G5 = learned.er_graphs(64, 8192, 8.0, seed=5005)
    dgs = [_lib.DeviceGraph.csr(G) for G in G5]
    rng = np.random.default_rng(77)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in G5]
    states, st = solver.solve(dgs, starts, P, passes=4, min_score=1.0)
    total_learned = 0
    for i, (G, x) in enumerate(zip(G5, states)):
        A = (G + G.T).tocsr().astype(np.float64)
        s = 1.0 - 2.0 * x.astype(np.float64)
        h = A @ s
        e = -0.5 * float(s @ h)
        assert st[i]["max_energy"] == pytest.approx(e, rel=REL), i
        assert st[i]["min_energy"] == pytest.approx(-e, rel=REL), i
        gains = s * h  # cut gain of flipping each spin
        assert gains.max() <= 1e-9 * max(1.0, np.abs(h).max()), i
        assert st[i]["learned_accepts"] > 0 and st[i]["learned_min_gain"] > 0, i
        total_learned += st[i]["learned_accepts"]
    log(f"  {total_learned} learned moves over 64 instances, all with positive exact gain")


def test_trained_network_gain():
    """What the trained network gains over the same learned passes without it:
    four first-improvement passes over every spin, in the trained network's
    score order, against index order (all scores equal) and the untrained
    network's order, from the same starts on held-out ER(2048, 8) graphs with
    U[-1, 1] couplings; the measure is the energy the passes reach (no loop)."""
    log("trained vs untrained network")
    P = _lib.make_params(0, is_spin_glass=1, walkers=16, seed_multiple=8)
    P.gray_iterations = -1
    only = _lib.make_params(0, is_spin_glass=1, walkers=16, seed_multiple=8, max_rounds=-1)
    solver = learned.LearnedSolver(seed=11)
    losses = solver.fit(learned.regular_training_graphs(48, seed=3001), P, seed=5, steps=300, lr=2e-3)
    assert losses[-1] < 0.6 * losses[0]
    trained = solver.net.params()
    G5 = learned.er_graphs(32, 2048, 8.0, seed=5006)
    dgs = [_lib.DeviceGraph.csr(G) for G in G5]
    rng = np.random.default_rng(78)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in G5]
    solver.bind(dgs)
    energies = {}
    for name, prm in (("trained", trained), ("index order", np.zeros_like(trained)),
                      ("untrained", _lib.Mpnn(seed=11).params())):
        solver.net.set_params(prm)
        _, st = solver.solve(dgs, starts, only, passes=4, min_score=-np.inf, bind=False)
        assert all(s["learned_min_gain"] > 0 and s["rounds"] == 0 for s in st)
        energies[name] = np.array([s["max_energy"] for s in st])
        log(f"  {name:11s}: mean energy after 4 passes {energies[name].mean():.3f}")
    tr = energies["trained"]
    for other in ("index order", "untrained"):
        o = energies[other]
        wins = int((tr > o).sum())
        log(f"  trained over {other}: +{(tr.mean() / o.mean() - 1) * 100:.2f} % mean energy, higher on {wins}/32")
        assert tr.mean() > 1.005 * o.mean()
        assert wins >= 20


def test_passes_zero_is_the_plain_search_to_targets():
    """passes = 0: the plain batched loop with targets, the network unused (here
    bound to other graphs).  With the plain search's own final energies as
    targets it stops at the same states, with the same accepted moves, having
    skipped only the last, non-improving iteration."""
    log("learned search with passes=0 on graphs the network is not bound to")
    rng = np.random.default_rng(93)
    Gs = [er_uniform(int(rng.integers(200, 600)), 8.0, 800 + t) for t in range(6)]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    net = _lib.Mpnn(seed=3)
    net.bind([_lib.DeviceGraph.csr(er_uniform(100, 4.0, 5))])
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=600)
    xs, plain = _lib.batch_local_search(dgs, starts, p)
    targets = np.array([s["max_energy"] for s in plain])
    ys, st = _lib.learned_search(net, dgs, starts, p, targets=targets, passes=0)
    for i in range(len(Gs)):
        assert np.array_equal(ys[i], xs[i]), i
        assert st[i]["max_energy"] == targets[i] and st[i]["learned_passes"] == 0
        for k in ("single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts"):
            assert st[i][k] == plain[i][k], (i, k)


def test_fallback_in_call_matches_two_calls():
    """PQI_FLAG_FALLBACK: the instances that end below their target are
    searched again by the plain batched loop from their starts inside the same
    call (started on a second stream after the learned search's first descent,
    beside the rest of its round).  The result equals the learned search followed by a separate plain
    search (passes = 0) of the short instances: the same states and counters,
    the learned counters kept on the fallback records, and the replaced
    search's single-flip accepts recorded apart.  With the untrained network
    and two passes, 5 of the 8 instances fall back (asserted non-empty); one
    target is set above what the plain search reaches."""
    log("learned search with the in-call fallback vs two calls")
    rng = np.random.default_rng(97)
    Gs = [er_uniform(int(rng.integers(300, 700)), 8.0, 900 + t) for t in range(8)]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    net = _lib.Mpnn(seed=4)
    net.bind(dgs)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=800)
    _, plain = _lib.batch_local_search(dgs, starts, p)
    targets = np.array([s["max_energy"] for s in plain])
    targets[0] += 1.0
    counters = ("rounds", "single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts",
                "single_evals", "double_evals", "seed_evals", "gray_evals")
    ys, st = _lib.learned_search(net, dgs, starts, p, targets=targets, passes=2, fallback=True)
    xs, ls = _lib.learned_search(net, dgs, starts, p, targets=targets, passes=2)
    short = [i for i in range(len(Gs)) if ls[i]["max_energy"] < targets[i]]
    assert short
    zs, fs = _lib.learned_search(net, [dgs[i] for i in short], [starts[i] for i in short], p,
                                 targets=targets[short], passes=0)
    for i in range(len(Gs)):
        if i in short:
            t = short.index(i)
            assert st[i]["fallback"] == 1, i
            assert np.array_equal(ys[i], zs[t]), i
            assert st[i]["max_energy"] == fs[t]["max_energy"], i
            for k in counters:
                assert st[i][k] == fs[t][k], (i, k)
            assert st[i]["replaced_single_accepts"] == ls[i]["single_accepts"], i
            for k in ("learned_accepts", "learned_passes"):
                assert st[i][k] == ls[i][k], (i, k)
        else:
            assert st[i]["fallback"] == 0 and st[i]["replaced_single_accepts"] == 0, i
            assert np.array_equal(ys[i], xs[i]), i
            assert st[i]["max_energy"] == ls[i]["max_energy"] >= targets[i], i
            for k in counters + ("learned_accepts", "learned_passes"):
                assert st[i][k] == ls[i][k], (i, k)
    # the call's descent pairs cover both searches, the same in every record;
    # the fallback starts after the learned search's first descent, for every
    # instance still below its target then (a superset of the short ones)
    pairs = {s["descent_instance_launches"] for s in st}
    assert len(pairs) == 1
    assert pairs.pop() >= ls[0]["descent_instance_launches"] + fs[0]["descent_instance_launches"]


def test_fallback_after_learned_only_search():
    """PQI_FLAG_FALLBACK with max_rounds < 0 (the learned passes only, no round
    of the loop): no first descent, so the fallback of every short instance
    runs after the learned search, as the plain search (max_rounds 0), with
    the same records as two calls."""
    log("learned-only search with the in-call fallback vs two calls")
    rng = np.random.default_rng(101)
    Gs = [er_uniform(int(rng.integers(300, 600)), 8.0, 950 + t) for t in range(5)]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    net = _lib.Mpnn(seed=6)
    net.bind(dgs)
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=500)
    _, plain = _lib.batch_local_search(dgs, starts, p)
    targets = np.array([s["max_energy"] for s in plain])
    q = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=500, max_rounds=-1)
    ys, st = _lib.learned_search(net, dgs, starts, q, targets=targets, passes=2, fallback=True)
    xs, ls = _lib.learned_search(net, dgs, starts, q, targets=targets, passes=2)
    short = [i for i in range(len(Gs)) if ls[i]["max_energy"] < targets[i]]
    assert short  # learned passes alone stop short of the plain search's energies
    zs, fs = _lib.learned_search(net, [dgs[i] for i in short], [starts[i] for i in short], p,
                                 targets=targets[short], passes=0)
    for i in range(len(Gs)):
        if i in short:
            t = short.index(i)
            assert st[i]["fallback"] == 1 and np.array_equal(ys[i], zs[t]), i
            assert st[i]["max_energy"] == fs[t]["max_energy"] >= targets[i], i
            assert st[i]["single_accepts"] == fs[t]["single_accepts"], i
            assert st[i]["learned_accepts"] == ls[i]["learned_accepts"], i
        else:
            assert st[i]["fallback"] == 0 and np.array_equal(ys[i], xs[i]), i
