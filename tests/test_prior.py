"""Host-side TFIM prior inputs: init_J_and_z_sparse (vectorised over the CSR)
against the reference's per-row loop over all n columns of the
upper-triangular matrix (pyqrackising/maxcut_tfim_sparse.py:204-224), run
literally on the dense form, including all-negative, all-equal negative and
all-positive weights and empty rows."""
import numpy as np
import pytest
import scipy.sparse as sp

from pyqrackising_amd import sampler as smp


def literal(U):
    n = U.shape[0]
    G_min = U.min()
    J_eff = np.empty(n)
    deg = np.empty(n, dtype=np.uint32)
    for r in range(n):
        d, J = 0, 0.0
        for c in range(n):
            val = U[r, c] - G_min
            if val <= smp.EPSILON:
                continue
            d += 1
            J += val
        if d > 0:
            J = -J / d
        deg[r] = d
        J_eff[r] = J
    return J_eff, deg


@pytest.mark.parametrize("case", ["mixed", "negative", "uniform_negative", "positive", "empty_rows"])
def test_sparse_prior_matches_reference_loop(case):
    rng = np.random.default_rng(3)
    n = 30
    mask = rng.random((n, n)) < 0.2
    if case == "mixed":
        vals = rng.integers(-3, 4, (n, n))
    elif case == "negative":
        vals = -rng.integers(1, 4, (n, n))
    elif case == "uniform_negative":  # every stored entry is the minimum
        vals = -np.ones((n, n), dtype=np.int64)
    else:
        vals = rng.integers(1, 4, (n, n))
    U = np.triu(vals * mask, 1).astype(np.float64)
    if case == "empty_rows":
        U[5:9] = 0.0
    J, z = smp.init_J_and_z_sparse(sp.csr_matrix(U))
    Jr, zr = literal(U)
    np.testing.assert_array_equal(z, zr)
    np.testing.assert_allclose(J, Jr, rtol=1e-12, atol=0)
