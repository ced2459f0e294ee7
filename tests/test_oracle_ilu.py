"""The oracle's gmm::ilu_precond (src/gmm/gmm_precond_ilu.h:104-231) of
GetFEM's column-major matrix: the rows of A^T factored, applied as the
lower solve with U^T and the unit upper solve with L^T.

Pins: where ILU(0) drops no fill (tridiagonal, dense) M^-1 r is the exact
solve of A (numpy); on any symmetric pattern (L U)_ij = (A^T)_ij for every
(i, j) of the pattern, the defining property of a level-0 incomplete LU;
the apply equals the dense triangular solves with those factors; a zero
first pivot is replaced by 1 as gmm does."""
import numpy as np
import scipy.sparse as sp

import oracle


def csr(A):
    m = sp.csr_matrix(A)
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.int32), m.data.astype(np.float64)


def factors(ip, ix, F, n):
    L, U = np.eye(n), np.zeros((n, n))
    for i in range(n):
        for p in range(ip[i], ip[i + 1]):
            j = ix[p]
            (L if j < i else U)[i, j] = F[p]
    return L, U


def test_no_fill_is_exact():
    rng = np.random.default_rng(1)
    n = 12
    T = np.diag(4 + rng.random(n)) + np.diag(rng.standard_normal(n - 1), 1) + np.diag(rng.standard_normal(n - 1), -1)
    D = rng.standard_normal((7, 7)) + 7 * np.eye(7)
    for A in (T, D):
        ip, ix, av = csr(A)
        F, rep = oracle.ilu(ip, ix, av)
        assert rep == 0
        r = rng.standard_normal(A.shape[0])
        z = oracle.ilu_apply(ip, ix, F, r)
        assert np.abs(z - np.linalg.solve(A, r)).max() <= 1e-12 * np.abs(z).max()


def test_incomplete_lu_property_and_apply():
    rng = np.random.default_rng(2)
    n = 40
    P = (rng.random((n, n)) < 0.15)
    P = P | P.T | np.eye(n, dtype=bool)
    A = np.where(P, rng.standard_normal((n, n)), 0.0) + n * np.eye(n)
    ip, ix, av = csr(A)
    F, rep = oracle.ilu(ip, ix, av)
    assert rep == 0
    L, U = factors(ip, ix, F, n)
    LU = L @ U
    B = A.T
    assert np.abs((LU - B)[P]).max() <= 1e-12 * np.abs(B).max()
    assert np.abs(LU[~P] - B[~P]).max() > 1e-8  # fill was dropped: an incomplete factorization
    r = rng.standard_normal(n)
    z = oracle.ilu_apply(ip, ix, F, r)
    y = np.linalg.solve(U.T, r)
    assert np.abs(z - np.linalg.solve(L.T, y)).max() <= 1e-12 * np.abs(z).max()


def test_zero_first_pivot_is_replaced():
    A = np.array([[1.0, 1.0, 0.0], [1.0, 2.0, 1.0], [0.0, 1.0, 3.0]])
    ip, ix, av = csr(A)
    av[0] = 0.0  # an explicit zero diagonal in the pattern
    F, rep = oracle.ilu(ip, ix, av)
    assert rep == 1 and F[0] == 1.0
