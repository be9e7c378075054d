"""pqi_eval_batch_noalloc (candidate scoring with no device allocation, on a
static scratch array) against pqi_eval_batch_ex with the exact engine: the
same values, best index and best value, bit for bit, on two calls in a row
whose second input is smaller; and the size limit is refused up front."""
import sys

import numpy as np
import pytest

from pyqrackising_amd import _lib, graphs

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert _lib.device_count() > 0, "no GPU visible"
    return _lib.default_context()


def int8_graph(n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.random_regular_edges(n, 3, rng)
    return graphs.dense_from_edges(n, lo, hi, graphs.pm1(lo.size, rng), dtype=np.int8)


def real_graph(n, seed, dtype):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges(n, 0.2, rng)
    return graphs.dense_from_edges(n, lo, hi, rng.normal(size=lo.size), dtype=dtype)


def csr_graph(n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = graphs.erdos_renyi_edges_sparse(n, 8.0 / (n - 1), rng)
    return graphs.upper_csr(n, lo, hi, rng.uniform(-1.0, 1.0, lo.size), dtype=np.float32)


def same_as_exact(dg, X, sg):
    bits = _lib.pack_bits(X)
    bi, bv, vals = dg.eval_batch_noalloc(bits, sg)
    ebi, ebv, evals = dg.eval_batch(bits, sg, engine=_lib.ENGINE_EXACT)
    assert np.array_equal(vals, evals)
    assert bi == ebi and bv == ebv
    # without the values output
    bi2, bv2, _ = dg.eval_batch_noalloc(bits, sg, want_values=False)
    assert bi2 == ebi and bv2 == ebv


# (n, B) of the first call, then of the second, smaller call
SHAPES = [((2048, 4096), (300, 700)), ((1000, 600), (130, 17)), ((258, 512), (96, 3))]


@pytest.mark.parametrize("first,second", SHAPES)
@pytest.mark.parametrize("sg", [0, 1])
def test_noalloc_int8_two_calls(ctx, first, second, sg):
    for call, (n, B) in enumerate((first, second)):
        log(f"noalloc int8 call {call}: n={n} B={B} sg={sg}")
        dg = _lib.DeviceGraph.dense(int8_graph(n, seed=n + call))
        assert dg.info()["weights"] == _lib.WEIGHTS_I8
        X = np.random.default_rng(B + call).integers(0, 2, (B, n)).astype(bool)
        X[B // 2] = X[0]  # a tie: the first index wins
        same_as_exact(dg, X, sg)
        dg.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_real_two_calls(ctx, dtype):
    # B > 16 takes the FP64 tiles, B <= 16 the FP64 MFMA re-score kernel
    for call, (n, B) in enumerate(((600, 300), (200, 12))):
        log(f"noalloc real {dtype.__name__} call {call}: n={n} B={B}")
        dg = _lib.DeviceGraph.dense(real_graph(n, seed=7 + call, dtype=dtype))
        X = np.random.default_rng(11 + call).integers(0, 2, (B, n)).astype(bool)
        same_as_exact(dg, X, 1)
        dg.close()


def test_noalloc_csr_two_calls(ctx):
    for call, (n, B) in enumerate(((4096, 256), (500, 40))):
        log(f"noalloc csr call {call}: n={n} B={B}")
        dg = _lib.DeviceGraph.csr(csr_graph(n, seed=21 + call))
        X = np.random.default_rng(31 + call).integers(0, 2, (B, n)).astype(bool)
        same_as_exact(dg, X, 1)
        dg.close()


def test_noalloc_refuses_above_its_scratch(ctx):
    """B * n too large for the 32 MiB array: refused before any device work,
    and the next call still works."""
    n = 4096
    dg = _lib.DeviceGraph.dense(int8_graph(n, seed=3))
    B = 70000  # the rows alone: 70000 * 4096 / 8 bytes = 35.8 MB > 32 MiB
    bits = np.zeros((B, n // 32), np.uint32)
    with pytest.raises(_lib.PqiError) as e:
        dg.eval_batch_noalloc(bits, 1)
    assert e.value.code == _lib.PQI_ERR_INVALID and "scratch" in str(e.value)
    X = np.random.default_rng(5).integers(0, 2, (64, n)).astype(bool)
    same_as_exact(dg, X, 0)
    dg.close()
