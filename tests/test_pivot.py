"""Partial-pivoted LU: the CPU restatement (oracle) pinned against LAPACK's
getrf (scipy.linalg.lu_factor: same pivot rule, largest magnitude with the
first row on ties) and against what the factorization must satisfy.

  * the permutation equals LAPACK's (its row interchanges replayed on the
    identity) on random blocks, square and ragged;
  * L U reproduces the permuted block to roundoff and |L| <= 1 (partial
    pivoting's defining bound);
  * a column that is zero at and below the diagonal gives info = its index
    + 1 and is left as it is;
  * FP32 uses the same rule in single precision (L U = P A to FP32 roundoff);
  * a planted mistake (multipliers taken from the unpermuted row) breaks the
    reconstruction check.
"""
import numpy as np
import pytest
import scipy.linalg as sl


def lapack_perm(a):
    _, piv = sl.lu_factor(a)
    perm = np.arange(a.shape[0])
    for i, q in enumerate(piv):
        perm[i], perm[q] = perm[q], perm[i]
    return perm


def factors(lu):
    ny, nx = lu.shape
    k = min(ny, nx)
    L = np.tril(lu[:, :k], -1) + np.eye(ny, k)
    U = np.triu(lu[:k, :])
    return L, U


@pytest.mark.parametrize("n", [1, 2, 7, 64, 200, 256])
def test_permutation_matches_lapack(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    lu, p, info = orc.getrf_pivoted(a)
    assert info == 0
    assert np.array_equal(p, lapack_perm(a))
    lf, _ = sl.lu_factor(a)
    assert np.abs(lf - lu).max() < 1e-10


@pytest.mark.parametrize("ny,nx", [(100, 60), (60, 100), (256, 32), (33, 257)])
def test_reconstruction_ragged(orc, ny, nx):
    rng = np.random.default_rng(ny * 1000 + nx)
    a = rng.standard_normal((ny, nx))
    lu, p, info = orc.getrf_pivoted(a)
    assert info == 0
    assert sorted(p.tolist()) == list(range(ny))
    L, U = factors(lu)
    assert np.abs(L @ U - a[p]).max() < 1e-12 * max(ny, nx)
    assert np.abs(L).max() <= 1.0


def test_zero_column_info(orc):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((40, 40))
    a[:, 3] = 0.0  # zero at and below the diagonal after any interchange
    lu, p, info = orc.getrf_pivoted(a)
    assert info == 4
    L, U = factors(lu)
    assert np.abs(L @ U - a[p]).max() < 1e-12
    z = np.zeros((8, 8))
    lu, p, info = orc.getrf_pivoted(z)
    assert info == 1 and np.array_equal(p, np.arange(8)) and not lu.any()


def test_ties_take_the_first_row(orc):
    a = np.array([[1.0, 2.0], [-1.0, 3.0]])
    _, p, _ = orc.getrf_pivoted(a)
    assert p.tolist() == [0, 1]
    # column 0 ties between rows 1 and 2 (row 1 wins); after that step
    # column 1 holds 3.5 (row 1) and 4 (row 2), so row 2 comes second
    a = np.array([[0.5, 2.0], [-1.0, 3.0], [1.0, 1.0]])
    _, p, _ = orc.getrf_pivoted(a)
    assert p.tolist() == [1, 2, 0]


def test_fp32_reconstruction(orc):
    rng = np.random.default_rng(9)
    a = rng.standard_normal((128, 128)).astype(np.float32)
    lu, p, info = orc.getrf_pivoted(a)
    assert lu.dtype == np.float32 and info == 0
    L, U = factors(lu.astype(np.float64))
    assert np.abs(L @ U - a[p].astype(np.float64)).max() < 1e-4
    assert np.array_equal(p, lapack_perm(a.astype(np.float64)))


def test_planted_mistake_is_caught(orc):
    rng = np.random.default_rng(11)
    a = rng.standard_normal((32, 32))
    lu, p, _ = orc.getrf_pivoted(a)
    L, U = factors(lu)
    assert np.abs(L @ U - a[p]).max() < 1e-12
    # forgetting the interchange when forming the product must fail the check
    assert np.abs(L @ U - a).max() > 1e-3
