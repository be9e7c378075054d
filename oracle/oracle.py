"""ctypes wrapper for the CPU oracle (oracle/pqi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the pyqrackising_amd package.

All arrays are converted to the oracle's types: float64 dense matrices,
uint8 bit vectors, int64 CSR index arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpqi_oracle.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [
        ("is_spin_glass", ctypes.c_int32),
        ("double_mode", ctypes.c_int32),
        ("gray_iterations", ctypes.c_int64),
        ("walkers", ctypes.c_int64),
        ("seed_multiple", ctypes.c_int64),
        ("max_rounds", ctypes.c_int64),
    ]


class Result(ctypes.Structure):
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
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_compute_cut": (_dbl, [_vp, _i64, _vp]),
            "orc_compute_energy": (_dbl, [_vp, _i64, _vp]),
            "orc_cut_diff": (_dbl, [_i64, _vp, _vp, _i64]),
            "orc_cut_diff_2": (_dbl, [_i64, _i64, _vp, _vp, _i64]),
            "orc_cut_diff_between": (_dbl, [_vp, _vp, _vp, _i64]),
            "orc_compute_energy_sparse": (_dbl, [_vp, _vp, _vp, _vp, _i64]),
            "orc_compute_cut_sparse": (_dbl, [_vp, _vp, _vp, _vp, _i64]),
            "orc_eval_batch": (_i64, [_vp, _i64, _vp, _i64, _int, _vp]),
            "orc_eval_batch_sparse": (_i64, [_vp, _vp, _vp, _i64, _vp, _i64, _int, _vp]),
            "orc_unrank_pair": (None, [_i64, _i64, _vp, _vp]),
            "orc_gray_flip_pos": (_i64, [_i64]),
            "orc_gray_mutation": (None, [_i64, _vp, _i64, _i64, _vp]),
            "orc_single_bit_flips": (_dbl, [_vp, _int, _vp, _i64, _vp]),
            "orc_double_last_index": (_i64, [_i64, _i64, _i64]),
            "orc_double_bit_flips": (_dbl, [_vp, _int, _vp, _i64, _i64, _int, _vp, _vp]),
            "orc_pick_gray_seeds": (_i64, [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp]),
            "orc_gray_optimization": (_i64, [_vp, _vp, _i64, _i64, _int, _vp, _i64, _vp, _vp]),
            "orc_dense_local_search": (_int, [_vp, _i64, _vp, _vp, _vp]),
            "orc_sparse_single": (_dbl, [_vp, _int, _vp, _vp, _vp, _i64, _vp]),
            "orc_sparse_double": (_dbl, [_vp, _int, _vp, _vp, _vp, _i64, _i64, _int, _vp, _vp]),
            "orc_sparse_pick_gray_seeds": (_i64, [_vp, _i64, _i64, _vp, _vp, _vp, _i64, _int, _vp, _vp]),
            "orc_sparse_gray_optimization": (_i64, [_vp, _vp, _i64, _i64, _int, _vp, _vp, _vp, _i64, _vp, _vp]),
            "orc_sparse_local_search": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp]),
            "orc_exact": (_dbl, [_vp, _i64, _int, _vp]),
            "orc_first_improvement": (_i64, [_vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _W(W):
    return np.ascontiguousarray(W, dtype=np.float64)


def _x(x):
    return np.ascontiguousarray(x, dtype=np.uint8)


def _csr(G):
    """(data float64, indptr int64, indices int64, n) from a scipy CSR matrix."""
    return (
        np.ascontiguousarray(G.data, dtype=np.float64),
        np.ascontiguousarray(G.indptr, dtype=np.int64),
        np.ascontiguousarray(G.indices, dtype=np.int64),
        G.shape[0],
    )


def compute_cut(W, x):
    W, x = _W(W), _x(x)
    return lib().orc_compute_cut(_p(W), W.shape[0], _p(x))


def compute_energy(W, x):
    W, x = _W(W), _x(x)
    return lib().orc_compute_energy(_p(W), W.shape[0], _p(x))


def cut_diff(u, x, W):
    W, x = _W(W), _x(x)
    return lib().orc_cut_diff(u, _p(x), _p(W), W.shape[0])


def cut_diff_2(k, l, x, W):
    W, x = _W(W), _x(x)
    return lib().orc_cut_diff_2(k, l, _p(x), _p(W), W.shape[0])


def cut_diff_between(o, nw, W):
    W, o, nw = _W(W), _x(o), _x(nw)
    return lib().orc_cut_diff_between(_p(o), _p(nw), _p(W), W.shape[0])


def compute_cut_sparse(x, G):
    d, r, c, n = _csr(G)
    x = _x(x)
    return lib().orc_compute_cut_sparse(_p(x), _p(d), _p(r), _p(c), n)


def compute_energy_sparse(x, G):
    d, r, c, n = _csr(G)
    x = _x(x)
    return lib().orc_compute_energy_sparse(_p(x), _p(d), _p(r), _p(c), n)


def pack_bits(X):
    """B x n bool -> B x ceil(n/32) uint32, LSB-first (util:304-316)."""
    X = np.asarray(X, dtype=bool)
    B, n = X.shape
    nw = (n + 31) >> 5
    pad = np.zeros((B, nw * 32), dtype=bool)
    pad[:, :n] = X
    by = np.packbits(pad.reshape(B, nw * 4, 8), axis=2, bitorder="little").reshape(B, nw * 4)
    return np.ascontiguousarray(by).view(np.uint32).reshape(B, nw)


def eval_batch(W, X, is_spin_glass):
    W = _W(W)
    bits = pack_bits(X)
    vals = np.empty(bits.shape[0], dtype=np.float64)
    best = lib().orc_eval_batch(_p(W), W.shape[0], _p(bits), bits.shape[0], int(is_spin_glass), _p(vals))
    return int(best), vals


def eval_batch_sparse(G, X, is_spin_glass):
    d, r, c, n = _csr(G)
    bits = pack_bits(X)
    vals = np.empty(bits.shape[0], dtype=np.float64)
    best = lib().orc_eval_batch_sparse(_p(d), _p(r), _p(c), n, _p(bits), bits.shape[0], int(is_spin_glass), _p(vals))
    return int(best), vals


def unrank_pair(s, n):
    i, j = ctypes.c_int64(), ctypes.c_int64()
    lib().orc_unrank_pair(s, n, ctypes.byref(i), ctypes.byref(j))
    return i.value, j.value


def gray_flip_pos(i):
    return lib().orc_gray_flip_pos(i)


def gray_mutation(index, seed_bits, offset):
    s = _x(seed_bits)
    out = np.empty_like(s)
    lib().orc_gray_mutation(index, _p(s), offset, s.shape[0], _p(out))
    return out


def single_bit_flips(theta, is_spin_glass, W):
    W, t = _W(W), _x(theta)
    k = ctypes.c_int64()
    e = lib().orc_single_bit_flips(_p(t), int(is_spin_glass), _p(W), W.shape[0], ctypes.byref(k))
    return e, k.value


def double_last_index(t, combo_count, thread_count):
    return lib().orc_double_last_index(t, combo_count, thread_count)


def double_bit_flips(theta, is_spin_glass, W, thread_count, mode=0):
    W, t = _W(W), _x(theta)
    k, l = ctypes.c_int64(), ctypes.c_int64()
    e = lib().orc_double_bit_flips(_p(t), int(is_spin_glass), _p(W), W.shape[0], thread_count, mode,
                                   ctypes.byref(k), ctypes.byref(l))
    return e, k.value, l.value


def pick_gray_seeds(theta, walkers, seed_multiple, W, want_all=False):
    W, t = _W(W), _x(theta)
    n = W.shape[0]
    seeds = np.empty((walkers, n), dtype=np.uint8)
    en = np.empty(walkers, dtype=np.float64)
    idx = np.empty(walkers, dtype=np.int64)
    count = walkers * seed_multiple * ((n + 63) >> 6)
    alle = np.empty(count, dtype=np.float64)
    lib().orc_pick_gray_seeds(_p(t), walkers, seed_multiple, _p(W), n, _p(seeds), _p(en), _p(idx), _p(alle))
    if want_all:
        return seeds, en, idx, alle
    return seeds, en, idx


def gray_optimization(iterators, energies, gray_iterations, walkers, is_spin_glass, W):
    W = _W(W)
    it = np.ascontiguousarray(iterators, dtype=np.uint8).copy()
    en = np.ascontiguousarray(energies, dtype=np.float64).copy()
    g = ctypes.c_double()
    ev = ctypes.c_int64()
    bw = lib().orc_gray_optimization(_p(it), _p(en), gray_iterations, walkers, int(is_spin_glass), _p(W),
                                     W.shape[0], ctypes.byref(g), ctypes.byref(ev))
    return int(bw), g.value, it, en, ev.value


def make_params(n, is_spin_glass=True, gray_iterations=None, walkers=8, seed_multiple=None,
                double_mode=0, max_rounds=0):
    p = Params()
    p.is_spin_glass = int(bool(is_spin_glass))
    p.double_mode = int(double_mode)
    p.gray_iterations = n * n if gray_iterations is None else int(gray_iterations)
    p.walkers = int(walkers)
    p.seed_multiple = int(walkers if seed_multiple is None else seed_multiple)
    p.max_rounds = int(max_rounds)
    return p


def dense_local_search(W, theta, **kw):
    W = _W(W)
    n = W.shape[0]
    best = _x(theta).copy()
    p = make_params(n, **kw)
    r = Result()
    lib().orc_dense_local_search(_p(W), n, _p(best), ctypes.byref(p), ctypes.byref(r))
    return best, r.as_dict()


def sparse_single(theta, is_spin_glass, G):
    d, r, c, n = _csr(G)
    t = _x(theta)
    k = ctypes.c_int64()
    e = lib().orc_sparse_single(_p(t), int(is_spin_glass), _p(d), _p(r), _p(c), n, ctypes.byref(k))
    return e, k.value


def sparse_double(theta, is_spin_glass, G, thread_count, mode=0):
    d, r, c, n = _csr(G)
    t = _x(theta)
    k, l = ctypes.c_int64(), ctypes.c_int64()
    e = lib().orc_sparse_double(_p(t), int(is_spin_glass), _p(d), _p(r), _p(c), n, thread_count, mode,
                                ctypes.byref(k), ctypes.byref(l))
    return e, k.value, l.value


def sparse_local_search(G, theta, **kw):
    d, r, c, n = _csr(G)
    best = _x(theta).copy()
    p = make_params(n, **kw)
    res = Result()
    lib().orc_sparse_local_search(_p(d), _p(r), _p(c), n, _p(best), ctypes.byref(p), ctypes.byref(res))
    return best, res.as_dict()


def exact(W, is_spin_glass):
    W = _W(W)
    n = W.shape[0]
    bits = np.empty(n, dtype=np.uint8)
    v = lib().orc_exact(_p(W), n, int(is_spin_glass), _p(bits))
    return bits, v


def first_improvement(W, x, order, max_passes):
    W = _W(W)
    xs = _x(x).copy()
    o = np.ascontiguousarray(order, dtype=np.int64)
    p = ctypes.c_int64()
    g = ctypes.c_double()
    f = lib().orc_first_improvement(_p(W), W.shape[0], _p(xs), _p(o), int(o.size), int(max_passes), ctypes.byref(p),
                                    ctypes.byref(g))
    return xs, int(f), p.value, g.value
