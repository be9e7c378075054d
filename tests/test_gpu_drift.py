"""Drift control of the incrementally maintained fp64 field (real weights).

Every round of the local search starts by recomputing the field h = W s
exactly from W and the state and replacing the maintained one.  These tests
force a large drift from the host (PQI_FLAG_TEST_DRIFT adds 1e-3 ((v mod 7) - 3)
to every entry before the first round) and check that the first recompute
detects it (counted entries, largest deviation) and repairs it before any move
uses it: the whole search then ends in the same state, with the same counters
and the same max_energy, as the search without the injected drift.  Integer
weights keep exact int32 fields and are never recomputed."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import _lib

pytestmark = pytest.mark.gpu

KEYS = ("rounds", "single_accepts", "double_accepts", "seed_accepts", "walk_accepts", "reheat_accepts",
        "single_evals", "double_evals", "seed_evals", "gray_evals", "max_energy", "cut_value", "min_energy")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def real_graph(n, seed, fmt):
    """dense: 30 % of pairs (the flips update from dense rows); csr: 5 %."""
    rng = np.random.default_rng(seed)
    p = 0.05 if fmt == "csr" else 0.3
    U = np.triu(rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < p), 1).astype(np.float32)
    if fmt == "csr":
        return _lib.DeviceGraph.csr(sp.csr_matrix(U))
    return _lib.DeviceGraph.dense(U + U.T)


@pytest.mark.parametrize("fmt", ["dense", "csr"])
@pytest.mark.parametrize("host_loop", [False, True])
def test_injected_drift_is_detected_and_repaired(fmt, host_loop):
    n = 400
    log(f"drift fmt={fmt} host_loop={host_loop}")
    dg = real_graph(n, 5, fmt)
    x0 = np.random.default_rng(6).integers(0, 2, n).astype(np.uint8)
    base_flags = _lib.FLAG_HOST_LOOP if host_loop else 0
    P0 = _lib.make_params(n, walkers=4, seed_multiple=4, gray_iterations=2000, flags=base_flags)
    P1 = _lib.make_params(n, walkers=4, seed_multiple=4, gray_iterations=2000,
                          flags=base_flags | _lib.FLAG_TEST_DRIFT)
    xa, sa = dg.local_search(x0, P0)
    xb, sb = dg.local_search(x0, P1)
    dg.close()
    np.testing.assert_array_equal(xa, xb)
    for k in KEYS:
        assert sa[k] == sb[k], k
    # the clean run recomputes every round and finds only rounding-level drift
    assert sa["field_checks"] == sa["rounds"]
    assert sa["field_repairs"] == 0 and sa["max_field_drift"] < 1e-9
    # the injected drift is counted once (first round) and repaired there
    assert sb["field_checks"] == sb["rounds"]
    expected = int(np.count_nonzero(np.arange(n) % 7 != 3))
    assert sb["field_repairs"] == expected
    assert sb["max_field_drift"] == pytest.approx(3e-3, rel=1e-6)


def test_integer_fields_are_not_recomputed():
    rng = np.random.default_rng(3)
    n = 256
    U = np.triu(rng.choice([-1, 1], (n, n)) * (rng.random((n, n)) < 0.05), 1).astype(np.int8)
    dg = _lib.DeviceGraph.dense(U + U.T)
    x0 = rng.integers(0, 2, n).astype(np.uint8)
    _, st = dg.local_search(x0, _lib.make_params(n, walkers=4, seed_multiple=4, gray_iterations=1000,
                                                 flags=_lib.FLAG_TEST_DRIFT))
    dg.close()
    assert st["field_checks"] == 0 and st["field_repairs"] == 0
