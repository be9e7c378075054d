"""ctypes binding of libpqi.so (the HIP C ABI declared in include/pqi.h).

This module is the replacement for the reference's OpenCL seam
(pyqrackising/maxcut_tfim_util.py:9-301).  There is no fallback: if the
library is missing or no GPU is present, every entry point raises.
"""
import atexit
import ctypes
import os
import threading
import weakref
from collections.abc import Mapping

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpqi.so")

PQI_OK = 0
PQI_ERR_INVALID = -1
PQI_ERR_NO_DEVICE = -2
PQI_ERR_HIP = -3
PQI_ERR_OOM = -4
PQI_ERR_UNSUPPORTED = -5

DTYPE_I8, DTYPE_I32, DTYPE_F32, DTYPE_F64 = 1, 2, 3, 4
WEIGHTS_I8, WEIGHTS_F32, WEIGHTS_F64 = 1, 3, 4
DOUBLE_REFERENCE, DOUBLE_EXHAUSTIVE = 0, 1
SEEDS_AUTO, SEEDS_MFMA, SEEDS_LOWRANK = 0, 1, 2
FLAG_TIME_KERNELS = 1
FLAG_HOST_LOOP = 2
FLAG_TEST_DRIFT = 4
FLAG_FALLBACK = 8
ENGINE_AUTO, ENGINE_EXACT, ENGINE_BF16 = 0, 1, 2

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double


class PqiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpqi error {code}: {msg}")
        self.code = code


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("weights", ctypes.c_int32),
        ("has_dense", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("w_up", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]


class SearchParams(ctypes.Structure):
    _fields_ = [
        ("is_spin_glass", ctypes.c_int32),
        ("double_mode", ctypes.c_int32),
        ("gray_iterations", ctypes.c_int64),
        ("walkers", ctypes.c_int64),
        ("seed_multiple", ctypes.c_int64),
        ("max_rounds", ctypes.c_int64),
        ("seed_engine", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


class SearchStats(ctypes.Structure):
    _fields_ = [
        ("max_energy", ctypes.c_double),
        ("cut_value", ctypes.c_double),
        ("min_energy", ctypes.c_double),
        ("rounds", ctypes.c_int64),
        ("single_accepts", ctypes.c_int64),
        ("double_accepts", ctypes.c_int64),
        ("seed_accepts", ctypes.c_int64),
        ("walk_accepts", ctypes.c_int64),
        ("reheat_accepts", ctypes.c_int64),
        ("single_evals", ctypes.c_int64),
        ("double_evals", ctypes.c_int64),
        ("seed_evals", ctypes.c_int64),
        ("gray_evals", ctypes.c_int64),
        ("gray_evals_computed", ctypes.c_int64),
        ("seconds_single", ctypes.c_double),
        ("seconds_double", ctypes.c_double),
        ("seconds_seeds", ctypes.c_double),
        ("seconds_walk", ctypes.c_double),
        ("seconds_reheat", ctypes.c_double),
        ("walk_kernel_seconds", ctypes.c_double),
        ("descent_kernel_seconds", ctypes.c_double),
        ("walk_launches", ctypes.c_int64),
        ("descent_launches", ctypes.c_int64),
        ("descent_evals", ctypes.c_int64),
        ("learned_accepts", ctypes.c_int64),
        ("learned_passes", ctypes.c_int64),
        ("seconds_learned", ctypes.c_double),
        ("field_checks", ctypes.c_int64),
        ("field_repairs", ctypes.c_int64),
        ("max_field_drift", ctypes.c_double),
        ("learned_min_gain", ctypes.c_double),
        ("descent_instance_launches", ctypes.c_int64),
        ("descent_instance_writes", ctypes.c_int64),
        ("fallback", ctypes.c_int64),
        ("replaced_single_accepts", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_STATS_NAMES = tuple(f for f, _ in SearchStats._fields_)
_STATS_DTYPE = np.dtype([(f, "<f8" if t is ctypes.c_double else "<i8") for f, t in SearchStats._fields_])
assert _STATS_DTYPE.itemsize == ctypes.sizeof(SearchStats)


class StatsRecord(Mapping):
    """One instance's pqi_search_stats of a batched call: a read-only mapping
    from the field names to Python numbers (columns converted once per call)."""

    __slots__ = ("_cols", "_i")

    def __init__(self, cols, i):
        self._cols = cols
        self._i = i

    def __getitem__(self, key):
        return self._cols[key][self._i]

    def __iter__(self):
        return iter(_STATS_NAMES)

    def __len__(self):
        return len(_STATS_NAMES)

    def __repr__(self):
        return repr(dict(self))


def _stats_records(st):
    """The records of a ctypes SearchStats array, one StatsRecord each."""
    arr = np.frombuffer(st, dtype=_STATS_DTYPE)
    cols = {f: arr[f].tolist() for f in _STATS_NAMES}
    return [StatsRecord(cols, i) for i in range(len(st))]


def _batch_states(graphs, thetas):
    """Copies of the start states (one block when all n agree) and the
    pointer array the C calls take."""
    k = len(graphs)
    ns = [g.n for g in graphs]
    if k and all(n == ns[0] for n in ns):
        block = np.array(thetas, dtype=np.uint8)  # a copy
        if block.shape != (k, ns[0]):
            raise ValueError("theta shape does not match its graph")
        base, n = block.ctypes.data, ns[0]
        return list(block), (ctypes.c_void_p * k)(*range(base, base + k * n, n))
    states = [np.ascontiguousarray(t, dtype=np.uint8).copy() for t in thetas]
    for g, t in zip(graphs, states):
        if t.shape != (g.n,):
            raise ValueError("theta shape does not match its graph")
    return states, (ctypes.c_void_p * k)(*[t.ctypes.data for t in states])


class BnbStats(ctypes.Structure):
    _fields_ = [
        ("nodes_explored", ctypes.c_int64),
        ("nodes_pruned", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("max_stack", ctypes.c_int64),
        ("certified", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("seconds", ctypes.c_double),
        ("kernel_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


# name -> (restype, argtypes); every symbol include/pqi.h declares
SIGNATURES = {
    "pqi_last_error": (ctypes.c_char_p, []),
    "pqi_version": (_int, []),
    "pqi_device_count": (_int, [_vp]),
    "pqi_create": (_int, [_int, _vp]),
    "pqi_destroy": (_int, [_vp]),
    "pqi_graph_dense": (_int, [_vp, _vp, _int, _i64, _vp]),
    "pqi_graph_csr": (_int, [_vp, _vp, _vp, _vp, _int, _i64, _i64, _vp]),
    "pqi_graph_free": (_int, [_vp]),
    "pqi_graph_get_info": (_int, [_vp, _vp]),
    "pqi_eval_batch": (_int, [_vp, _vp, _i64, _int, _vp, _vp, _vp]),
    "pqi_eval_batch_ex": (_int, [_vp, _vp, _i64, _int, _int, _vp, _vp, _vp, _vp]),
    "pqi_eval_batch_noalloc": (_int, [_vp, _vp, _i64, _int, _vp, _vp, _vp]),
    "pqi_single_flip_best": (_int, [_vp, _vp, _int, _vp, _vp]),
    "pqi_single_flip_descent": (_int, [_vp, _vp, _int, _i64, _vp, _vp]),
    "pqi_double_flip_best": (_int, [_vp, _vp, _int, _i64, _int, _vp, _vp, _vp]),
    "pqi_gray_search": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pqi_local_search": (_int, [_vp, _vp, _vp, _vp]),
    "pqi_score": (_int, [_vp, _vp, _vp, _vp]),
    "pqi_sample_candidates": (_int, [_vp, _vp, _i64, _dbl, ctypes.c_uint64, _i64, _i64, _vp]),
    "pqi_sample_round": (_int, [_vp, _vp, _i64, _dbl, ctypes.c_uint64, _i64, _i64, _int, _vp, _vp, _vp]),
    "pqi_batch_create": (_int, [_vp, _vp, _i64, _i64, _vp]),
    "pqi_batch_free": (_int, [_vp]),
    "pqi_batch_set_candidates": (_int, [_vp, _vp]),
    "pqi_batch_eval": (_int, [_vp, _int, _vp, _vp, _vp]),
    "pqi_batch_autotune": (_int, [_vp, _int, _vp, _vp, _int]),
    "pqi_batch_set_variant": (_int, [_vp, _int]),
    "pqi_batch_info": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pqi_batch_local_search": (_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "pqi_mpnn_create": (_int, [_vp, ctypes.c_uint64, _vp]),
    "pqi_mpnn_free": (_int, [_vp]),
    "pqi_mpnn_num_params": (_int, []),
    "pqi_mpnn_set_precision": (_int, [_vp, _int]),
    "pqi_mpnn_get_params": (_int, [_vp, _vp]),
    "pqi_mpnn_set_params": (_int, [_vp, _vp]),
    "pqi_mpnn_bind": (_int, [_vp, _vp, _i64, _vp]),
    "pqi_mpnn_forward": (_int, [_vp, _vp, _vp]),
    "pqi_mpnn_loss_grad": (_int, [_vp, _vp, _vp, _vp, _vp]),
    "pqi_mpnn_train_step": (_int, [_vp, _vp, _vp, ctypes.c_float, _vp]),
    "pqi_learned_search": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, ctypes.c_float, _vp]),
    "pqi_bp_marginals": (_int, [_vp, _i64, _dbl, _vp, _i64, _vp, _vp, _vp]),
    "pqi_bnb_eval": (_int, [_vp, _vp, _i64, _int, _vp, _vp, _vp, _vp]),
    "pqi_bnb": (_int, [_vp, _vp, _vp, _int, _int, _i64, _dbl, _vp]),
}

_lib = None
_lock = threading.Lock()

# Every live device object, so that they can be released in order (objects
# before their context) while the HIP runtime is still up.  After release at
# interpreter exit, close() calls are no-ops.
_live = weakref.WeakSet()
_shutdown = False


def _release_all():
    global _shutdown, _default_ctx
    if _shutdown:
        return
    for rank in (0, 1, 2):
        for obj in list(_live):
            if getattr(obj, "_release_rank", 1) == rank:
                try:
                    obj.close()
                except Exception:
                    pass
    _default_ctx = None
    _shutdown = True


atexit.register(_release_all)


def load():
    """Load libpqi.so and bind every entry point (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PqiError(PQI_ERR_UNSUPPORTED,
                           f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != PQI_OK:
        msg = load().pqi_last_error()
        raise PqiError(code, msg.decode() if msg else "")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count():
    c = ctypes.c_int(0)
    r = load().pqi_device_count(ctypes.byref(c))
    return c.value if r == PQI_OK else 0


class Context:
    """One device and one HIP stream (the reference's global OpenCL queue,
    pyqrackising/maxcut_tfim_util.py:103-104).  Calls are serialised by a lock."""

    _release_rank = 2

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(load().pqi_create(int(device), ctypes.byref(self._h)))
        self.lock = threading.RLock()
        _live.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and not _shutdown:
            load().pqi_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = Context(int(os.environ.get("PQI_DEVICE", "0")))
        return _default_ctx


def _dtype_code(a):
    if a.dtype == np.int8:
        return a, DTYPE_I8
    if a.dtype == np.int32:
        return a, DTYPE_I32
    if a.dtype == np.float32:
        return a, DTYPE_F32
    if a.dtype == np.float64:
        return a, DTYPE_F64
    if a.dtype == np.bool_ or np.issubdtype(a.dtype, np.integer):
        if a.size and (a.max() > np.iinfo(np.int32).max or a.min() < np.iinfo(np.int32).min):
            return a.astype(np.float64), DTYPE_F64
        return a.astype(np.int32), DTYPE_I32
    return a.astype(np.float64), DTYPE_F64


class DeviceGraph:
    """A problem instance resident in HBM (replaces make_G_m_buf /
    make_G_m_csr_buf, pyqrackising/maxcut_tfim_util.py:224-259)."""

    _release_rank = 1

    def __init__(self, ctx, handle, n):
        self.ctx = ctx
        self._h = handle
        self.n = n
        _live.add(self)

    @classmethod
    def dense(cls, W, ctx=None):
        ctx = ctx or default_context()
        W = np.ascontiguousarray(W)
        if W.ndim != 2 or W.shape[0] != W.shape[1]:
            raise ValueError("W must be square")
        W, code = _dtype_code(W)
        W = np.ascontiguousarray(W)
        h = ctypes.c_void_p()
        with ctx.lock:
            check(load().pqi_graph_dense(ctx.handle, _p(W), code, W.shape[0], ctypes.byref(h)))
        return cls(ctx, h, W.shape[0])

    @classmethod
    def csr(cls, G, ctx=None):
        """G: scipy CSR, strictly upper triangular (util:533-542)."""
        ctx = ctx or default_context()
        n = G.shape[0]
        indptr = np.ascontiguousarray(G.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(G.indices, dtype=np.int64)
        data, code = _dtype_code(np.ascontiguousarray(G.data))
        data = np.ascontiguousarray(data)
        h = ctypes.c_void_p()
        with ctx.lock:
            check(load().pqi_graph_csr(ctx.handle, _p(indptr), _p(indices), _p(data), code, n,
                                       int(indices.size), ctypes.byref(h)))
        return cls(ctx, h, n)

    def info(self):
        i = GraphInfo()
        check(load().pqi_graph_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def close(self):
        if self._h and not _shutdown:
            load().pqi_graph_free(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- entry points -------------------------------------------------------

    def _theta(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.uint8)
        if t.shape != (self.n,):
            raise ValueError(f"theta must have shape ({self.n},)")
        return t

    def eval_batch(self, bits, is_spin_glass, want_values=True, engine=ENGINE_AUTO, return_error=False):
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        B = bits.shape[0]
        if bits.ndim != 2 or bits.shape[1] != (self.n + 31) >> 5:
            raise ValueError("bits must be B x ceil(n/32) uint32")
        vals = np.empty(B, dtype=np.float64) if want_values else None
        bi = ctypes.c_int64()
        bv = ctypes.c_double()
        err = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_eval_batch_ex(self._h, _p(bits), B, int(bool(is_spin_glass)), int(engine),
                                           _p(vals) if want_values else None, ctypes.byref(bi), ctypes.byref(bv),
                                           ctypes.byref(err)))
        if return_error:
            return bi.value, bv.value, vals, err.value
        return bi.value, bv.value, vals

    def eval_batch_noalloc(self, bits, is_spin_glass, want_values=True):
        """pqi_eval_batch_noalloc: the exact engine, with no device allocation
        (a static 32 MiB scratch array per GPU)."""
        bits = np.ascontiguousarray(bits, dtype=np.uint32)
        B = bits.shape[0]
        if bits.ndim != 2 or bits.shape[1] != (self.n + 31) >> 5:
            raise ValueError("bits must be B x ceil(n/32) uint32")
        vals = np.empty(B, dtype=np.float64) if want_values else None
        bi = ctypes.c_int64()
        bv = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_eval_batch_noalloc(self._h, _p(bits), B, int(bool(is_spin_glass)),
                                                _p(vals) if want_values else None, ctypes.byref(bi),
                                                ctypes.byref(bv)))
        return bi.value, bv.value, vals

    def single_flip_best(self, theta, is_spin_glass):
        t = self._theta(theta)
        k = ctypes.c_int64()
        g = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_single_flip_best(self._h, _p(t), int(bool(is_spin_glass)), ctypes.byref(k),
                                              ctypes.byref(g)))
        return g.value, k.value

    def single_flip_descent(self, theta, is_spin_glass, max_steps=2 ** 62):
        """Repeated first-argmax single flips while the gain is positive.
        Returns (new theta, flips applied, summed gain)."""
        t = self._theta(theta).copy()
        a = ctypes.c_int64()
        g = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_single_flip_descent(self._h, _p(t), int(bool(is_spin_glass)), int(max_steps),
                                                 ctypes.byref(a), ctypes.byref(g)))
        return t, a.value, g.value

    def double_flip_best(self, theta, is_spin_glass, thread_count, mode=DOUBLE_REFERENCE):
        t = self._theta(theta)
        k, l = ctypes.c_int64(), ctypes.c_int64()
        g = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_double_flip_best(self._h, _p(t), int(bool(is_spin_glass)), int(thread_count), int(mode),
                                              ctypes.byref(k), ctypes.byref(l), ctypes.byref(g)))
        return g.value, k.value, l.value

    def gray_search(self, theta, params):
        t = self._theta(theta)
        out = np.empty(self.n, dtype=np.uint8)
        sg, wg = ctypes.c_double(), ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_gray_search(self._h, _p(t), ctypes.byref(params), _p(out), ctypes.byref(sg),
                                         ctypes.byref(wg)))
        return sg.value, wg.value, out

    def local_search(self, theta, params):
        t = self._theta(theta).copy()
        st = SearchStats()
        with self.ctx.lock:
            check(load().pqi_local_search(self._h, _p(t), ctypes.byref(params), ctypes.byref(st)))
        return t, st.as_dict()

    def score(self, theta):
        t = self._theta(theta)
        c, e = ctypes.c_double(), ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_score(self._h, _p(t), ctypes.byref(c), ctypes.byref(e)))
        return c.value, e.value

    def sample_candidates(self, cum_prob, repulsion_base, seed, shot_offset, shots):
        """shots x n bool candidates from the device sampler (shot indices
        shot_offset .. shot_offset + shots - 1)."""
        cum = np.ascontiguousarray(cum_prob, dtype=np.float64)
        out = np.empty((int(shots), self.n), dtype=np.uint8)
        with self.ctx.lock:
            check(load().pqi_sample_candidates(self._h, _p(cum), cum.shape[0], float(repulsion_base),
                                               int(seed) & (2 ** 64 - 1), int(shot_offset), int(shots), _p(out)))
        return out.astype(bool)

    def sample_round(self, cum_prob, repulsion_base, seed, shot_offset, shots, is_spin_glass):
        """Sample, score and reduce one round on the device: (index, value, bits)."""
        cum = np.ascontiguousarray(cum_prob, dtype=np.float64)
        bi, bv = ctypes.c_int64(), ctypes.c_double()
        bits = np.empty(self.n, dtype=np.uint8)
        with self.ctx.lock:
            check(load().pqi_sample_round(self._h, _p(cum), cum.shape[0], float(repulsion_base),
                                          int(seed) & (2 ** 64 - 1), int(shot_offset), int(shots),
                                          int(bool(is_spin_glass)), ctypes.byref(bi), ctypes.byref(bv), _p(bits)))
        return bi.value, bv.value, bits.astype(bool)

    def bp_marginals(self, max_iterations, damping, init=None):
        """Belief-propagation marginals (pqi_bp_marginals): (marginals,
        sweeps run, number of arcs)."""
        marg = np.empty(self.n, dtype=np.float64)
        it, arcs = ctypes.c_int64(), ctypes.c_int64()
        ip, cnt = None, 0
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.float64)
            ip, cnt = _p(init), init.size
        with self.ctx.lock:
            check(load().pqi_bp_marginals(self._h, int(max_iterations), float(damping), ip, cnt, _p(marg),
                                          ctypes.byref(it), ctypes.byref(arcs)))
        return marg, it.value, arcs.value

    def bnb_eval(self, nodes, is_spin_glass, return_time=False):
        """Bound, leaf value and branch variable of count x n nodes (pqi_bnb_eval)."""
        F = np.ascontiguousarray(nodes, dtype=np.int8)
        if F.ndim != 2 or F.shape[1] != self.n:
            raise ValueError(f"nodes must have shape (count, {self.n})")
        cnt = F.shape[0]
        ub = np.empty(cnt, dtype=np.float64)
        leaf = np.empty(cnt, dtype=np.float64)
        br = np.empty(cnt, dtype=np.int32)
        ks = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_bnb_eval(self._h, _p(F), cnt, int(bool(is_spin_glass)), _p(ub), _p(leaf), _p(br),
                                      ctypes.byref(ks)))
        return (ub, leaf, br, ks.value) if return_time else (ub, leaf, br)

    def bnb(self, warm_bits, warm_value, is_spin_glass, batch_cap, time_limit=None, energy_bound=False):
        """Warm-started branch and bound (pqi_bnb): (bits, value, stats)."""
        bits = self._theta(warm_bits).copy()
        val = ctypes.c_double(float(warm_value))
        st = BnbStats()
        tl = -1.0 if time_limit is None else float(time_limit)
        with self.ctx.lock:
            check(load().pqi_bnb(self._h, _p(bits), ctypes.byref(val), int(bool(is_spin_glass)),
                                 int(bool(energy_bound)), int(batch_cap), tl, ctypes.byref(st)))
        return bits.astype(bool), val.value, st.as_dict()


class Batch:
    """Candidate scoring for many dense int8 instances in one persistent,
    graph-captured launch (pqi_batch_*)."""

    N_VARIANTS = 10

    def __init__(self, graphs, cands):
        self.graphs = list(graphs)
        self.ctx = self.graphs[0].ctx
        self.cands = int(cands)
        arr = (ctypes.c_void_p * len(self.graphs))(*[g._h.value for g in self.graphs])
        self._h = ctypes.c_void_p()
        with self.ctx.lock:
            check(load().pqi_batch_create(self.ctx.handle, arr, len(self.graphs), self.cands, ctypes.byref(self._h)))
        _live.add(self)

    _release_rank = 0

    def set_candidates(self, bits_list):
        """bits_list[i]: cands x ceil(n_i/32) uint32 for instance i."""
        flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(b, dtype=np.uint32).ravel()
                                                    for b in bits_list]))
        with self.ctx.lock:
            check(load().pqi_batch_set_candidates(self._h, _p(flat)))

    def eval(self, is_spin_glass, want_values=False):
        k = len(self.graphs)
        bi = np.empty(k, dtype=np.int64)
        bv = np.empty(k, dtype=np.float64)
        vals = np.empty(k * self.cands, dtype=np.float64) if want_values else None
        with self.ctx.lock:
            check(load().pqi_batch_eval(self._h, int(bool(is_spin_glass)), _p(bi), _p(bv),
                                        _p(vals) if want_values else None))
        return bi, bv, (vals.reshape(k, self.cands) if want_values else None)

    def autotune(self, reps=3):
        ms = np.zeros(self.N_VARIANTS, dtype=np.float64)
        ch = ctypes.c_int()
        with self.ctx.lock:
            check(load().pqi_batch_autotune(self._h, int(reps), ctypes.byref(ch), _p(ms), self.N_VARIANTS))
        return ch.value, ms

    def set_variant(self, v):
        with self.ctx.lock:
            check(load().pqi_batch_set_variant(self._h, int(v)))

    def info(self):
        nt, ops, exe, wb = (ctypes.c_int64() for _ in range(4))
        v = ctypes.c_int()
        check(load().pqi_batch_info(self._h, ctypes.byref(nt), ctypes.byref(ops), ctypes.byref(exe),
                                    ctypes.byref(wb), ctypes.byref(v)))
        return {"tiles": nt.value, "algorithmic_ops": ops.value, "executed_ops": exe.value,
                "weight_bytes": wb.value, "variant": v.value}

    def close(self):
        if self._h and not _shutdown:
            load().pqi_batch_free(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_local_search(graphs, thetas, params):
    """pqi_batch_local_search: returns (states, list of stats dicts)."""
    graphs = list(graphs)
    ctx = graphs[0].ctx
    k = len(graphs)
    states, tarr = _batch_states(graphs, thetas)
    garr = (ctypes.c_void_p * k)(*[g._h.value for g in graphs])
    st = (SearchStats * k)()
    with ctx.lock:
        check(load().pqi_batch_local_search(ctx.handle, garr, k, tarr, ctypes.byref(params), st))
    return states, _stats_records(st)


class Mpnn:
    """The flip-proposal network (pqi_mpnn_*)."""

    def __init__(self, seed=0, ctx=None, precision=1):
        self.ctx = ctx or default_context()
        self._h = ctypes.c_void_p()
        check(load().pqi_mpnn_create(self.ctx.handle, ctypes.c_uint64(seed), ctypes.byref(self._h)))
        self.set_precision(precision)
        self.n_params = load().pqi_mpnn_num_params()
        self.N = 0
        _live.add(self)

    _release_rank = 0

    def set_precision(self, precision):
        check(load().pqi_mpnn_set_precision(self._h, int(precision)))

    def params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(load().pqi_mpnn_get_params(self._h, _p(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(load().pqi_mpnn_set_params(self._h, _p(p)))

    def bind(self, graphs):
        graphs = list(graphs)
        arr = (ctypes.c_void_p * len(graphs))(*[g._h.value for g in graphs])
        N = ctypes.c_int64()
        with self.ctx.lock:
            check(load().pqi_mpnn_bind(self._h, arr, len(graphs), ctypes.byref(N)))
        self.N = N.value
        return self.N

    def forward(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        assert f.shape == (self.N, 4)
        out = np.empty(self.N, dtype=np.float32)
        with self.ctx.lock:
            check(load().pqi_mpnn_forward(self._h, _p(f), _p(out)))
        return out

    def loss_grad(self, feats, labels):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_mpnn_loss_grad(self._h, _p(f), _p(y), ctypes.byref(loss), _p(g)))
        return loss.value, g

    def train_step(self, feats, labels, lr=1e-3):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        loss = ctypes.c_double()
        with self.ctx.lock:
            check(load().pqi_mpnn_train_step(self._h, _p(f), _p(y), ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def close(self):
        if self._h and not _shutdown:
            load().pqi_mpnn_free(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def learned_search(net, graphs, thetas, params, targets=None, passes=8, min_score=-np.inf, fallback=False):
    """pqi_learned_search: returns (states, list of stats mappings).  With
    fallback (targets and passes > 0 required), PQI_FLAG_FALLBACK: instances
    that end below their target are searched again by the plain batched loop
    from their starts within the same call (their record's "fallback" is 1)."""
    graphs = list(graphs)
    ctx = graphs[0].ctx
    k = len(graphs)
    states, tarr = _batch_states(graphs, thetas)
    garr = (ctypes.c_void_p * k)(*[g._h.value for g in graphs])
    tg = None if targets is None else np.ascontiguousarray(targets, dtype=np.float64)
    if fallback:
        params = SearchParams.from_buffer_copy(params)
        params.flags |= FLAG_FALLBACK
    st = (SearchStats * k)()
    with ctx.lock:
        check(load().pqi_learned_search(ctx.handle, net._h, garr, k, tarr, ctypes.byref(params),
                                        None if tg is None else _p(tg), int(passes), ctypes.c_float(min_score),
                                        st))
    return states, _stats_records(st)


def make_params(n, is_spin_glass=True, gray_iterations=None, walkers=None, seed_multiple=None,
                double_mode=DOUBLE_REFERENCE, max_rounds=0, seed_engine=SEEDS_AUTO, flags=0):
    """Defaults follow pyqrackising/spin_glass_solver.py:556-560, 603."""
    if seed_multiple is None:
        seed_multiple = max(2, os.cpu_count() or 2)
    if walkers is None:
        walkers = seed_multiple
    p = SearchParams()
    p.is_spin_glass = int(bool(is_spin_glass))
    p.double_mode = int(double_mode)
    p.gray_iterations = n * n if gray_iterations is None else int(gray_iterations)
    p.walkers = int(walkers)
    p.seed_multiple = int(seed_multiple)
    p.max_rounds = int(max_rounds)
    p.seed_engine = int(seed_engine)
    p.flags = int(flags)
    return p


def pack_bits(X):
    """B x n bool -> B x ceil(n/32) uint32, LSB-first
    (pyqrackising/maxcut_tfim_util.py:304-316)."""
    X = np.asarray(X, dtype=bool)
    B, n = X.shape
    nw = (n + 31) >> 5
    pad = np.zeros((B, nw * 32), dtype=bool)
    pad[:, :n] = X
    by = np.packbits(pad.reshape(B, nw * 4, 8), axis=2, bitorder="little").reshape(B, nw * 4)
    return np.ascontiguousarray(by).view(np.uint32).reshape(B, nw)
