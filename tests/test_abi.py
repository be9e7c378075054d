"""CPU-side checks of the C ABI boundary: the library loads, exports every
entry point include/pqi.h declares, and reports the documented errors without
a GPU.  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from pyqrackising_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "pqi.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"PQI_API\s+(?:const\s+char\s*\*|int)\s+(pqi_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) >= 15
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes bindings and header differ"


def test_structs_match_header_sizes():
    # field order/size of the three ABI structs (all 8-byte aligned members)
    assert ctypes.sizeof(_lib.GraphInfo) == 40
    assert ctypes.sizeof(_lib.SearchParams) == 48
    assert ctypes.sizeof(_lib.SearchStats) == 8 * 35


def test_stats_records_match_the_structs():
    # the column conversion of a batched call's records gives each field's
    # value, as the per-struct conversion does
    st = (_lib.SearchStats * 5)()
    for i in range(5):
        st[i].max_energy = 1.5 * i - 2.0
        st[i].rounds = 7 + i
        st[i].learned_min_gain = float("inf") if i == 2 else -0.25 * i
        st[i].fallback = i % 2
        st[i].replaced_single_accepts = 100 * i
    recs = _lib._stats_records(st)
    assert len(recs) == 5
    for i in range(5):
        d = st[i].as_dict()
        assert dict(recs[i]) == d
        assert list(recs[i]) == list(d)
        assert type(recs[i]["rounds"]) is int and type(recs[i]["max_energy"]) is float


def test_batch_states_are_copies_with_pointers():
    class G:
        def __init__(self, n):
            self.n = n

    rng = np.random.default_rng(0)
    for ns in ([16, 16, 16], [5, 9, 7]):
        th = [rng.integers(0, 2, n).astype(np.uint8) for n in ns]
        states, tarr = _lib._batch_states([G(n) for n in ns], th)
        for t, s_, p in zip(th, states, tarr):
            assert np.array_equal(t, s_) and s_.ctypes.data == p
            assert not np.shares_memory(t, s_)
    with pytest.raises(ValueError):
        _lib._batch_states([G(4), G(4)], [np.zeros(4, np.uint8), np.zeros(5, np.uint8)])


def test_version_and_error_plumbing():
    L = _lib.load()
    assert L.pqi_version() == 1
    r = L.pqi_graph_dense(None, None, 1, 4, None)
    assert r == _lib.PQI_ERR_INVALID
    assert b"null" in L.pqi_last_error()
    with pytest.raises(_lib.PqiError):
        _lib.check(r)


@pytest.mark.skipif(_lib.device_count() > 0, reason="a GPU is present")
def test_no_device_is_reported():
    L = _lib.load()
    c = ctypes.c_int(-1)
    assert L.pqi_device_count(ctypes.byref(c)) == _lib.PQI_ERR_NO_DEVICE
    assert c.value == 0
    h = ctypes.c_void_p()
    assert L.pqi_create(0, ctypes.byref(h)) == _lib.PQI_ERR_NO_DEVICE
    with pytest.raises(_lib.PqiError) as e:
        _lib.Context(0)
    assert e.value.code == _lib.PQI_ERR_NO_DEVICE


def test_pack_bits_matches_reference_layout():
    X = np.zeros((1, 70), bool)
    X[0, [0, 31, 32, 69]] = True
    P = _lib.pack_bits(X)
    assert P.tolist() == [[1 | (1 << 31), 1, 1 << 5]]


def test_product_has_no_cpu_fallback():
    """The package must not import the oracle anywhere."""
    pkg = os.path.join(ROOT, "pyqrackising_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                with open(os.path.join(dirpath, f)) as fh:
                    assert "oracle" not in fh.read(), f
