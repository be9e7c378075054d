"""Graph ingestion and synthetic instance generators (host side).

Ingestion keeps the reference's matrix conventions:
  * dense: symmetric n x n weights, zero diagonal, built as
    ``nx.to_numpy_array(G, weight="weight", nonedge=0.0)``
    (pyqrackising/maxcut_tfim.py:355-360, spin_glass_solver.py:458-465);
  * sparse: strictly upper-triangular CSR in node order
    (pyqrackising/maxcut_tfim_util.py:533-542).  The reference builds it with an
    O(n^2) ``G.has_edge`` loop; here it is built from the edge list in O(m log m)
    with the same result.

The generators are seeded numpy PCG64 streams so that every configuration is
reproducible without networkx on the GPU host.
"""
import numpy as np
import scipy.sparse as sp


def nodes_of(G):
    return list(G.nodes())


def dense_from_graph(G, dtype=np.float32):
    """nx.Graph -> dense weight matrix, as the reference's nx.to_numpy_array call."""
    import networkx as nx

    return nx.to_numpy_array(G, weight="weight", nonedge=0.0, dtype=dtype)


def upper_csr_from_graph(G, dtype=np.float32):
    """nx.Graph -> strictly upper-triangular CSR in node order (util:533-542)."""
    nodes = list(G.nodes())
    pos = {v: i for i, v in enumerate(nodes)}
    n = len(nodes)
    rows, cols, vals = [], [], []
    for u, v, data in G.edges(data=True):
        i, j = pos[u], pos[v]
        if i == j:
            continue
        if i > j:
            i, j = j, i
        rows.append(i)
        cols.append(j)
        vals.append(data.get("weight", 1.0))
    return upper_csr(n, np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64),
                     np.array(vals, dtype=np.float64), dtype)


def upper_csr(n, rows, cols, vals, dtype=np.float32):
    """Edge arrays (any orientation) -> strictly upper-triangular CSR, sorted columns."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    lo = np.minimum(rows, cols)
    hi = np.maximum(rows, cols)
    keep = lo != hi
    M = sp.csr_matrix((np.asarray(vals, dtype=np.float64)[keep], (lo[keep], hi[keep])), shape=(n, n))
    M.sum_duplicates()
    M.sort_indices()
    M = M.astype(dtype)
    M.indptr = M.indptr.astype(np.int32)
    M.indices = M.indices.astype(np.int32)
    return M


def dense_from_edges(n, rows, cols, vals, dtype=np.float32):
    W = np.zeros((n, n), dtype=np.float64)
    np.add.at(W, (rows, cols), vals)
    np.add.at(W, (cols, rows), vals)
    np.fill_diagonal(W, 0.0)
    return W.astype(dtype)


def random_regular_edges(n, d, rng, max_tries=50):
    """Random d-regular simple graph: pairing model, then degree-preserving
    edge switches that remove every loop and repeated edge.  n*d must be even."""
    if (n * d) % 2:
        raise ValueError("n*d must be even")
    if d >= n:
        raise ValueError("d must be < n")
    for _ in range(max_tries):
        stubs = np.repeat(np.arange(n, dtype=np.int64), d)
        rng.shuffle(stubs)
        a, b = stubs[0::2].copy(), stubs[1::2].copy()
        if _switch_repair(a, b, rng):
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            order = np.lexsort((hi, lo))
            return lo[order], hi[order]
    raise RuntimeError("could not build a simple regular graph")


def _switch_repair(a, b, rng, max_steps=1000000):
    m = a.size
    count = {}
    for u, v in zip(a.tolist(), b.tolist()):
        k = (min(u, v), max(u, v))
        count[k] = count.get(k, 0) + 1

    def bad(e):
        u, v = int(a[e]), int(b[e])
        return u == v or count[(min(u, v), max(u, v))] > 1

    todo = [e for e in range(m) if bad(e)]
    steps = 0
    while todo and steps < max_steps:
        steps += 1
        e = todo[-1]
        if not bad(e):
            todo.pop()
            continue
        f = int(rng.integers(m))
        if f == e:
            continue
        u, v, x, y = int(a[e]), int(b[e]), int(a[f]), int(b[f])
        # rewire (u,v),(x,y) -> (u,x),(v,y)
        if u == x or v == y:
            continue
        k1, k2 = (min(u, x), max(u, x)), (min(v, y), max(v, y))
        if count.get(k1, 0) or count.get(k2, 0) or k1 == k2:
            continue
        for k in ((min(u, v), max(u, v)), (min(x, y), max(x, y))):
            count[k] -= 1
            if not count[k]:
                del count[k]
        count[k1] = 1
        count[k2] = 1
        a[e], b[e], a[f], b[f] = u, x, v, y
        if bad(f):
            todo.append(f)
    return not any(bad(e) for e in range(m))


def erdos_renyi_edges(n, p, rng):
    """G(n, p): every pair i<j independently with probability p."""
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.size) < p
    return iu[keep].astype(np.int64), ju[keep].astype(np.int64)


def erdos_renyi_edges_sparse(n, p, rng):
    """G(n, p) for large n without materialising all pairs (geometric skipping)."""
    total = n * (n - 1) // 2
    m = rng.binomial(total, p)
    idx = np.unique(rng.choice(total, size=m, replace=False) if m < total else np.arange(total))
    # unrank row-major pair index -> (i, j)
    i = (n - 2 - np.floor(np.sqrt(-8.0 * idx + 4.0 * n * (n - 1) - 7) / 2.0 - 0.5)).astype(np.int64)
    j = (idx + i + 1 - n * (n - 1) // 2 + (n - i) * ((n - i) - 1) // 2).astype(np.int64)
    return i, j


def pm1(size, rng):
    """Uniform +-1 couplings (int8)."""
    return (rng.integers(0, 2, size=size, dtype=np.int8) * 2 - 1).astype(np.int8)


def dense_int8_from_edges(n, lo, hi, w):
    """Symmetric int8 weights from an edge list (zero diagonal)."""
    W = np.zeros((n, n), dtype=np.int8)
    W[lo, hi] = w
    W[hi, lo] = w
    return W


def batched_regular_instances(count, seed, n_lo=256, n_hi=4096, d_lo=3, d_hi=10):
    """The batched workload: `count` random d-regular graphs with n uniform in
    [n_lo, n_hi] and d uniform in {d_lo..d_hi} (n += 1 when n*d is odd), +-1
    couplings.  Returns a list of int8 dense matrices."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        d = int(rng.integers(d_lo, d_hi + 1))
        if (n * d) % 2:
            n += 1
        lo, hi = random_regular_edges(n, d, rng)
        out.append(dense_int8_from_edges(n, lo, hi, pm1(lo.size, rng)))
    return out
