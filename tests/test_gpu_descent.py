"""Single-flip descent (the reference's single-flip phase loop,
pyqrackising/spin_glass_solver.py:604-624) on the GPU, against the oracle's
best-single-flip function applied step by step (small graphs) and against the
textbook steepest-ascent rule -- first maximum of Delta_k = s_k h_k, flip while
positive, h_v -= 2 w_vk s_k -- restated with numpy at sizes the oracle does not
reach.  Cases cover every descent kernel: packed keys with the adjacency staged
in LDS, packed keys over global adjacency (too large to stage), packed keys
over dense int8 rows, the unpacked integer scan (row strengths too large for
packed keys) and the FP64 scan for real weights."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import _lib, graphs

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def steepest(A, x):
    """First-max steepest ascent of the cut from x; A symmetric (CSR or dense)."""
    dense = not sp.issparse(A)
    s = 1 - 2 * x.astype(np.int64 if A.dtype.kind in "iu" else np.float64)
    h = A @ s
    flips, gain = 0, 0
    while True:
        d = s * h
        k = int(np.argmax(d))
        if not d[k] > 0:
            break
        if dense:
            h -= 2 * s[k] * A[k]
        else:
            lo, hi = A.indptr[k], A.indptr[k + 1]
            h[A.indices[lo:hi]] -= 2 * s[k] * A.data[lo:hi]
        s[k] = -s[k]
        gain += d[k]
        flips += 1
    return ((1 - s) // 2 if A.dtype.kind in "iu" else (1 - s) / 2).astype(np.uint8), flips, gain


def regular_int(n, d, seed, wmax=1):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, d, rng)
    w = rng.integers(1, wmax + 1, lo.size) * rng.choice([-1, 1], lo.size)
    G = graphs.upper_csr(n, lo, hi, w.astype(np.float64), dtype=np.float32)
    A = (G + G.T).tocsr().astype(np.int64)
    A.sort_indices()
    return G, A


@pytest.mark.parametrize("n,d,wmax,seed", [(2048, 3, 1, 1), (12000, 3, 3, 2), (3000, 10, 5, 3)])
def test_descent_sparse_int(n, d, wmax, seed):
    log(f"descent sparse int n={n} d={d} wmax={wmax}")
    G, A = regular_int(n, d, seed, wmax)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(seed + 10).integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)


@pytest.mark.parametrize("n,wmax,seed", [(1024, 127, 4), (8192, 127, 5), (700, 1, 6)])
def test_descent_dense_int(n, wmax, seed):
    """Dense int8 rows; at n=8192 the row strengths exceed the packed-key range."""
    log(f"descent dense int n={n} wmax={wmax}")
    rng = np.random.default_rng(seed)
    U = np.triu(rng.integers(-wmax, wmax + 1, (n, n), dtype=np.int8), 1)
    W = U + U.T
    dg = _lib.DeviceGraph.dense(W)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, True)
    xr, fr, gr = steepest(W.astype(np.int64), x0)
    assert acc == fr and gain == 2 * gr
    assert np.array_equal(x1, xr)


@pytest.mark.parametrize("n,d,wmax,seed", [(300, 3, 1, 11), (1000, 5, 2, 12), (2000, 7, 1, 13), (4096, 3, 1, 14),
                                            (4000, 12, 1, 15), (4096, 16, 3, 16), (64, 3, 1, 17)])
def test_descent_register_resident(n, d, wmax, seed):
    """Packed keys held in registers (n <= 4096, neighbour rows of at most 16
    entries, the neighbour table in LDS): 8 to 64 registers per lane, rows of
    4, 8 and 16 entries; tables above 160 KiB (n=4000, d=12 and n=4096, d=16)
    take the LDS-key descent instead."""
    log(f"register descent n={n} d={d} wmax={wmax}")
    G, A = regular_int(n, d, seed, wmax)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(seed + 10).integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)


def test_descent_sparse_real():
    log("descent sparse real")
    rng = np.random.default_rng(7)
    n = 5000
    lo, hi = graphs.erdos_renyi_edges_sparse(n, 8.0 / (n - 1), rng)
    G = graphs.upper_csr(n, lo, hi, rng.uniform(-1, 1, lo.size), dtype=np.float32)
    A = (G + G.T).tocsr().astype(np.float64)
    A.sort_indices()
    dg = _lib.DeviceGraph.csr(G)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr
    assert np.array_equal(x1, xr)
    assert gain == pytest.approx(gr, rel=1e-9)


def test_descent_sparse_real_ties():
    """Sparse real weights in {+-0.5, +-1.5} (fp32, so the FP64 gain path): gains
    are exact binary fractions with many exact ties, which the round descent
    must break like the full scan (lowest node among equal gains), including
    ties at its threshold.  Exact equality with the textbook ascent."""
    log("descent sparse real with ties")
    rng = np.random.default_rng(21)
    n = 6000
    lo, hi = graphs.erdos_renyi_edges_sparse(n, 6.0 / (n - 1), rng)
    w = rng.choice([-1.5, -0.5, 0.5, 1.5], lo.size)
    G = graphs.upper_csr(n, lo, hi, w, dtype=np.float32)
    A = (G + G.T).tocsr().astype(np.float64)
    A.sort_indices()
    dg = _lib.DeviceGraph.csr(G)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)


def test_descent_sparse_int_beyond_key_range():
    """CSR int8 weights whose row strengths exceed the packed-key range
    (n = 4096, about 2100 neighbours of weight +-127 per node): the descent
    keeps int32 gains in LDS and runs in rounds on integer keys (ties are
    frequent).  Exact equality with the textbook ascent."""
    log("descent sparse int beyond the key range")
    rng = np.random.default_rng(22)
    n = 4096
    lo, hi = graphs.erdos_renyi_edges_sparse(n, 2100.0 / (n - 1), rng)
    w = 127.0 * rng.choice([-1.0, 1.0], lo.size)
    G = graphs.upper_csr(n, lo, hi, w, dtype=np.float32)
    A = (G + G.T).tocsr().astype(np.int64)
    A.sort_indices()
    assert np.abs(A).sum(axis=1).max() >= 2 ** 18  # beyond 2^(30 - ib), ib = 12
    dg = _lib.DeviceGraph.csr(G)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)


def test_descent_small_against_oracle(orc):
    """Oracle best single flip, applied until no gain, on small int graphs."""
    rng = np.random.default_rng(8)
    for t in range(6):
        n = int(rng.integers(20, 300))
        U = np.triu(rng.integers(-3, 4, (n, n)) * (rng.random((n, n)) < 0.1), 1)
        W = (U + U.T).astype(np.int32)
        x = rng.integers(0, 2, n).astype(np.uint8)
        log(f"descent vs oracle n={n}")
        dg = _lib.DeviceGraph.dense(W)
        x1, acc, gain = dg.single_flip_descent(x, True)
        xo, steps, total = x.copy(), 0, 0.0
        while True:
            g, k = orc.single_bit_flips(xo, 1, W.astype(np.float64))
            if not g > 0:
                break
            xo[k] ^= 1
            total += g
            steps += 1
        assert acc == steps and gain == total
        assert np.array_equal(x1, xo)


@pytest.mark.parametrize("n,d,seed", [(2048, 3, 21), (2046, 3, 22), (4096, 3, 23), (1500, 5, 24), (2500, 7, 25),
                                       (1000, 9, 26), (2000, 15, 27), (2100, 16, 28), (70, 16, 29), (33, 2, 30)])
def test_descent_bit_planes(n, d, seed):
    """+-1 weights: the bit-plane descent (gains as Delta + S in 3 to 6
    planes, one or two words per lane, a partial last word) against the
    textbook steepest ascent: the same state, flip count and total gain."""
    log(f"bit-plane descent n={n} d={d}")
    G, A = regular_int(n, d, seed, 1)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(seed + 10).integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)
    # bounded descents continue where the last one stopped
    x2 = x0
    for cap in (1, 7, 100000):
        x2, a2, _ = dg.single_flip_descent(x2, False, max_steps=cap)
        assert a2 <= cap
    assert np.array_equal(x2, xr)


@pytest.mark.parametrize("n,d,drop,seed", [(2048, 3, 0.1, 31), (1900, 3, 0.3, 32), (3000, 3, 0.1, 33),
                                            (1200, 10, 0.2, 34), (2048, 16, 0.05, 35)])
def test_descent_bit_planes_irregular(n, d, drop, seed):
    """+-1 weights on a regular graph with a share of its edges removed:
    mixed degree parities (full, not halved, plane values) and rows shorter
    than the stride (padding entries in the register rows and the table)."""
    log(f"bit-plane descent irregular n={n} d={d} drop={drop}")
    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, d, rng)
    keep = rng.random(lo.size) >= drop
    lo, hi = lo[keep], hi[keep]
    w = rng.choice([-1.0, 1.0], lo.size)
    G = graphs.upper_csr(n, lo, hi, w, dtype=np.float32)
    A = (G + G.T).tocsr().astype(np.int64)
    A.sort_indices()
    deg = np.diff(A.indptr)
    assert len(set(deg % 2)) == 2  # mixed parities
    dg = _lib.DeviceGraph.csr(G)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    x1, acc, gain = dg.single_flip_descent(x0, False)
    xr, fr, gr = steepest(A, x0)
    assert acc == fr and gain == gr
    assert np.array_equal(x1, xr)


def test_descent_max_steps():
    G, A = regular_int(600, 4, 9)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(19).integers(0, 2, 600).astype(np.uint8)
    x1, acc, _ = dg.single_flip_descent(x0, False, max_steps=5)
    assert acc == 5
    s = 1 - 2 * x0.astype(np.int64)
    h = A @ s
    for _ in range(5):
        k = int(np.argmax(s * h))
        lo, hi = A.indptr[k], A.indptr[k + 1]
        h[A.indices[lo:hi]] -= 2 * s[k] * A.data[lo:hi]
        s[k] = -s[k]
    assert np.array_equal(x1, ((1 - s) // 2).astype(np.uint8))


def test_batched_staging_mix_matches_single():
    """One instance too large to stage its adjacency beside small ones that
    stage theirs: every instance equals its single-instance search."""
    log("batched staging mix")
    Gs = [regular_int(n, 3, 40 + i)[0] for i, n in enumerate((12000, 300, 1000, 258))]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    rng = np.random.default_rng(41)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    p = _lib.make_params(0, is_spin_glass=0, walkers=4, seed_multiple=2, gray_iterations=2000)
    states, stats = _lib.batch_local_search(dgs, starts, p)
    for i, dg in enumerate(dgs):
        xs, st = dg.local_search(starts[i], p)
        assert np.array_equal(states[i], xs), i
        assert stats[i]["cut_value"] == st["cut_value"], i


@pytest.mark.parametrize("sizes", [((300, 3), (1000, 7), (2048, 3), (4000, 12)), ((64, 3), (700, 5), (2048, 13))])
def test_batched_register_descent_matches_single(sizes):
    """Batches of different sizes and row widths: with one table too large for
    LDS the batch takes the LDS-key descent; with every n <= 2048 the
    register-resident descent and walkers.  Every instance equals its
    single-instance search, which the parity tests check against the oracle."""
    log(f"batched register descent {sizes}")
    Gs = [regular_int(n, d, 80 + i)[0] for i, (n, d) in enumerate(sizes)]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    rng = np.random.default_rng(81)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=2000)
    states, stats = _lib.batch_local_search(dgs, starts, p)
    for i, dg in enumerate(dgs):
        xs, st = dg.local_search(starts[i], p)
        assert np.array_equal(states[i], xs), i
        assert stats[i]["cut_value"] == st["cut_value"], i
        assert stats[i]["single_accepts"] == st["single_accepts"], i


def test_consecutive_searches_of_different_sizes():
    """A local search returns while its last (idle) iteration may still be on
    the stream.  Searches of different sizes back to back on one context --
    larger after smaller, with the pinned staging already large -- give the
    host-decided loop's states and counters."""
    log("consecutive searches of different sizes")
    ns = (3000, 700, 2048, 64, 1500)
    Gs = [regular_int(n, 3 + i % 4, 90 + i)[0] for i, n in enumerate(ns)]
    dgs = [_lib.DeviceGraph.csr(G) for G in Gs]
    rng = np.random.default_rng(91)
    starts = [rng.integers(0, 2, G.shape[0]).astype(np.uint8) for G in Gs]
    p = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=2000)
    ph = _lib.make_params(0, is_spin_glass=1, walkers=4, seed_multiple=2, gray_iterations=2000)
    ph.flags = _lib.FLAG_HOST_LOOP
    ref = [dg.local_search(x0, ph) for dg, x0 in zip(dgs, starts)]
    for rep in range(2):
        for i, (dg, x0) in enumerate(zip(dgs, starts)):
            xd, sd = dg.local_search(x0, p)
            assert np.array_equal(xd, ref[i][0]), (rep, i)
            for k in ("cut_value", "max_energy", "rounds", "single_accepts", "walk_accepts"):
                assert sd[k] == ref[i][1][k], (rep, i, k)


@pytest.mark.parametrize("graphs_on", ["0", "1"])
def test_fast_path_direct_and_graph_launches(graphs_on, monkeypatch):
    """The int8 fast path (direct launches by default, iteration graphs with
    PQI_LS_FAST_GRAPH=1): a dense 3-regular instance with MFMA-scored Gray
    seeds and a CSR instance, three searches each, give the host-decided
    loop's states and counters."""
    log(f"fast path PQI_LS_FAST_GRAPH={graphs_on}")
    monkeypatch.setenv("PQI_LS_FAST_GRAPH", graphs_on)
    G1, A1 = regular_int(1024, 3, 95)
    W1 = A1.toarray().astype(np.int8)
    G2, _ = regular_int(1500, 5, 96)
    rng = np.random.default_rng(97)
    for name, dg, n in (("dense n=1024", _lib.DeviceGraph.dense(W1), 1024),
                        ("csr n=1500", _lib.DeviceGraph.csr(G2), 1500)):
        p = _lib.make_params(n, is_spin_glass=1, walkers=16, seed_multiple=2, gray_iterations=n * 4)
        ph = _lib.make_params(n, is_spin_glass=1, walkers=16, seed_multiple=2, gray_iterations=n * 4)
        ph.flags = _lib.FLAG_HOST_LOOP
        for rep in range(3):
            x0 = rng.integers(0, 2, n).astype(np.uint8)
            xr, sr = dg.local_search(x0, ph)
            xd, sd = dg.local_search(x0, p)
            assert np.array_equal(xd, xr), (name, rep)
            for k in ("cut_value", "max_energy", "rounds", "single_accepts", "double_accepts", "seed_accepts",
                      "walk_accepts", "reheat_accepts", "gray_evals_computed"):
                assert sd[k] == sr[k], (name, rep, k)


COUNTERS = ("max_energy", "cut_value", "min_energy", "rounds", "single_accepts", "double_accepts", "seed_accepts",
            "walk_accepts", "reheat_accepts", "single_evals", "double_evals", "seed_evals", "gray_evals",
            "gray_evals_computed", "descent_evals")


def _instances():
    rng = np.random.default_rng(50)
    out = []
    G, _ = regular_int(2048, 3, 51)
    out.append(("csr 3-regular n=2048", _lib.DeviceGraph.csr(G), 2048))
    U = np.triu(rng.integers(-2, 3, (300, 300), dtype=np.int8), 1)
    out.append(("dense int n=300", _lib.DeviceGraph.dense(U + U.T), 300))
    lo, hi = graphs.erdos_renyi_edges_sparse(1500, 8.0 / 1499, rng)
    Gr = graphs.upper_csr(1500, lo, hi, rng.uniform(-1, 1, lo.size), dtype=np.float32)
    out.append(("csr real n=1500", _lib.DeviceGraph.csr(Gr), 1500))
    Ud = np.triu(rng.normal(size=(256, 256)), 1)
    out.append(("dense real n=256", _lib.DeviceGraph.dense(Ud + Ud.T), 256))
    return out


@pytest.mark.parametrize("sg", [0, 1])
def test_device_loop_matches_host_loop(sg):
    """The device-decided loop takes the same moves, in the same order, as the
    host-decided loop: identical states, counters and max_energy."""
    rng = np.random.default_rng(60 + sg)
    for name, dg, n in _instances():
        for mode in (_lib.DOUBLE_REFERENCE, _lib.DOUBLE_EXHAUSTIVE):
            if mode == _lib.DOUBLE_EXHAUSTIVE and n > 1600:
                continue
            log(f"device vs host loop: {name} sg={sg} double_mode={mode}")
            x0 = rng.integers(0, 2, n).astype(np.uint8)
            p = _lib.make_params(n, is_spin_glass=sg, walkers=8, seed_multiple=2, gray_iterations=4 * n)
            p.double_mode = mode
            xd, sd = dg.local_search(x0, p)
            p.flags = _lib.FLAG_HOST_LOOP
            xh, sh = dg.local_search(x0, p)
            assert np.array_equal(xd, xh), name
            for k in COUNTERS:
                assert sd[k] == sh[k], (name, k, sd[k], sh[k])
            # the captured iteration is replayed by later calls with the same buffers
            p.flags = 0
            for _ in range(2):
                x2, s2 = dg.local_search(x0, p)
                assert np.array_equal(x2, xh), name
                assert all(s2[k] == sh[k] for k in COUNTERS), name


def test_device_loop_max_rounds():
    G, _ = regular_int(1024, 3, 70)
    dg = _lib.DeviceGraph.csr(G)
    x0 = np.random.default_rng(71).integers(0, 2, 1024).astype(np.uint8)
    for r in (1, 2, 3):
        p = _lib.make_params(1024, is_spin_glass=1, walkers=8, seed_multiple=2, gray_iterations=2048)
        p.max_rounds = r
        xd, sd = dg.local_search(x0, p)
        p.flags = _lib.FLAG_HOST_LOOP
        xh, sh = dg.local_search(x0, p)
        assert np.array_equal(xd, xh)
        assert sd["rounds"] == sh["rounds"] <= r
        for k in COUNTERS:
            assert sd[k] == sh[k], (r, k)
