"""CPU restatement of the belief-propagation warm start
(pyqrackising/spin_glass_solver.py:311-421, dense;
pyqrackising/spin_glass_solver_sparse.py:378-514, sparse).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the pyqrackising_amd
package.

The reference starts every message at zero.  Its update
m(u->v) <- d m(u->v) + (1 - d) atanh(clip(tanh(w) tanh(h_u\\v))) then keeps
every message at zero (tanh(0) = 0), so its marginals are all zero and the
warm start is the all-zeros partition.  `init` generalises the start so that
the update itself can be checked: init[k] for arc k in the library's arc
order (rows ascending, neighbours ascending within a row; the arc (a, b)
holds the message from b to a).  With init=None this is the reference's
function.
"""
import numpy as np


def edges_dense(W):
    """Undirected edges (u < v) of the upper triangle with |w| > 1e-12, in the
    reference's order (spin_glass_solver.py:341-351)."""
    n = W.shape[0]
    us, vs, ws = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            w = W[i, j]
            if abs(w) > 1e-12:
                us.append(i)
                vs.append(j)
                ws.append(float(w))
    return us, vs, ws


def edges_csr(data, indptr, indices, n):
    """Every stored entry of an upper-triangular CSR matrix, in storage order
    (spin_glass_solver_sparse.py:432-441)."""
    us, vs, ws = [], [], []
    for u in range(n):
        for p in range(indptr[u], indptr[u + 1]):
            us.append(u)
            vs.append(int(indices[p]))
            ws.append(float(data[p]))
    return us, vs, ws


def arc_order(us, vs, n):
    """Position of each directed message in the library's arc order: for edge
    k, (pos of msg u->v, pos of msg v->u)."""
    nbrs = [[] for _ in range(n)]
    for u, v in zip(us, vs):
        nbrs[u].append(v)
        nbrs[v].append(u)
    start = np.zeros(n + 1, dtype=np.int64)
    for a in range(n):
        nbrs[a].sort()
        start[a + 1] = start[a] + len(nbrs[a])
    pos = {}
    for a in range(n):
        for t, b in enumerate(nbrs[a]):
            pos[(a, b)] = start[a] + t  # arc (a, b): message b -> a
    fwd = [pos[(v, u)] for u, v in zip(us, vs)]
    bwd = [pos[(u, v)] for u, v in zip(us, vs)]
    return np.array(fwd, dtype=np.int64), np.array(bwd, dtype=np.int64), int(start[n])


def marginals(us, vs, ws, n, bp_scale=1.0, damping=0.5, init=None, iterations=None):
    """belief_propagation_marginals(_dense) (spin_glass_solver.py:339-415),
    loop for loop.  iterations overrides max(1, int(bp_scale * n))."""
    max_iterations = max(1, int(bp_scale * n)) if iterations is None else iterations
    m = len(us)
    if m == 0:
        return np.zeros(n, dtype=np.float64)
    fwd_pos, bwd_pos, _ = arc_order(us, vs, n)
    msg_fwd = np.zeros(m, dtype=np.float64)
    msg_bwd = np.zeros(m, dtype=np.float64)
    if init is not None:
        init = np.asarray(init, dtype=np.float64)
        msg_fwd = init[fwd_pos].copy()
        msg_bwd = init[bwd_pos].copy()
    node_edges = [[] for _ in range(n)]
    for k in range(m):
        node_edges[us[k]].append((k, 0))
        node_edges[vs[k]].append((k, 1))
    for _ in range(max_iterations):
        new_fwd = np.empty(m, dtype=np.float64)
        new_bwd = np.empty(m, dtype=np.float64)
        for k in range(m):
            u, v, w = us[k], vs[k], ws[k]
            h_u = 0.0
            for ek, d in node_edges[u]:
                if ek != k:
                    h_u += msg_bwd[ek] if d == 0 else msg_fwd[ek]
            tanh_w = np.tanh(w)
            p = np.clip(tanh_w * np.tanh(h_u), -1.0 + 1e-9, 1.0 - 1e-9)
            new_fwd[k] = damping * msg_fwd[k] + (1.0 - damping) * np.arctanh(p)
            h_v = 0.0
            for ek, d in node_edges[v]:
                if ek != k:
                    h_v += msg_bwd[ek] if d == 0 else msg_fwd[ek]
            p = np.clip(tanh_w * np.tanh(h_v), -1.0 + 1e-9, 1.0 - 1e-9)
            new_bwd[k] = damping * msg_bwd[k] + (1.0 - damping) * np.arctanh(p)
        msg_fwd, msg_bwd = new_fwd, new_bwd
    marg = np.zeros(n, dtype=np.float64)
    for k in range(m):
        marg[us[k]] += msg_bwd[k]
        marg[vs[k]] += msg_fwd[k]
    return np.tanh(marg)
