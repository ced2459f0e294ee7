"""The oracle's gmm::ildlt_precond restatement (src/gmm/gmm_precond_ildlt.h)
pinned by what the mathematics fixes: on patterns without fill the level-0
factor is the exact LDL^T, on any pattern L D L^T reproduces A on A's
pattern, the apply is the solve with L D L^T, a zero pivot is replaced by 1,
and PCG with it converges in fewer iterations than with the diagonal."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from tests import problems


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64)


def _spd(n, density, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng)
    A = (M + M.T) + sp.identity(n) * (2.0 * n * density + 1.0)
    return sp.csr_matrix(A)


def test_tridiagonal_is_exact():
    n = 50
    rng = np.random.default_rng(1)
    off = rng.uniform(-1, 1, n - 1)
    A = sp.diags([off, 4.0 + rng.uniform(0, 1, n), off], [-1, 0, 1]).tocsr()
    P = oracle.Ildlt(*_csr(A))
    L, D = P.factors()
    assert P.replaced == 0
    np.testing.assert_allclose((L @ sp.diags(D) @ L.T).toarray(), A.toarray(), rtol=0, atol=1e-14)
    # equals the exact Cholesky: L D^1/2 = chol(A)
    C = np.linalg.cholesky(A.toarray())
    np.testing.assert_allclose(L.toarray() * np.sqrt(D), C, rtol=1e-13, atol=1e-15)


def test_dense_pattern_is_exact_cholesky():
    n = 12
    rng = np.random.default_rng(2)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    P = oracle.Ildlt(*_csr(A))
    L, D = P.factors()
    C = np.linalg.cholesky(A)
    np.testing.assert_allclose(L.toarray() * np.sqrt(D), C, rtol=1e-12, atol=1e-13)
    r = rng.standard_normal(n)
    np.testing.assert_allclose(P.apply(r), np.linalg.solve(A, r), rtol=1e-11)


@pytest.mark.parametrize("seed", [3, 4])
def test_product_matches_a_on_its_pattern(seed):
    """Level-0 incomplete LDL^T: (L D L^T)_ij = A_ij for (i, j) in A's
    pattern (fill outside it is dropped, so elsewhere it differs)."""
    A = _spd(200, 0.03, seed)
    P = oracle.Ildlt(*_csr(A))
    L, D = P.factors()
    LDLt = (L @ sp.diags(D) @ L.T).tocsr()
    Ad = A.toarray()
    rows, cols = A.nonzero()
    diff = np.abs(LDLt.toarray()[rows, cols] - Ad[rows, cols])
    assert diff.max() <= 1e-12 * np.abs(Ad).max()
    # and some dropped fill makes it differ off the pattern
    assert np.abs(LDLt.toarray() - Ad).max() > 1e-6
    # apply is the solve with L D L^T
    r = np.random.default_rng(seed).standard_normal(200)
    np.testing.assert_allclose(P.apply(r), np.linalg.solve(LDLt.toarray(), r), rtol=1e-10, atol=1e-12)


def test_zero_pivot_is_replaced_by_one():
    # A_00 = 0: GetFEM sets U_00 = 1 before factoring (gmm_precond_ildlt.h:168-171)
    A = np.array([[0.0, 1.0, 0.0], [1.0, 3.0, 1.0], [0.0, 1.0, 2.0]])
    P = oracle.Ildlt(*_csr(A))
    D = P.diag()
    assert D[0] == 1.0
    assert D[1] == pytest.approx(3.0 - 1.0)
    assert D[2] == pytest.approx(2.0 - 1.0 / 2.0)
    # a later zero pivot: A = [[1, 1], [1, 1]] gives U_11 = 0 <= maxp, replaced
    P2 = oracle.Ildlt(*_csr(np.array([[1.0, 1.0], [1.0, 1.0]])))
    assert P2.replaced == 1 and P2.diag()[1] == 1.0


def test_pcg_on_p2_elasticity_needs_fewer_iterations_than_jacobi():
    case = problems.c2(3)
    ip, ix, v = case.oracle_csr()
    n = ip.size - 1
    A = sp.csr_matrix((v, ix, ip), shape=(n, n))
    # clamp the dofs of the first 12 nodes (no rigid modes), unit diagonal
    fixed = np.zeros(n, bool)
    fixed[:36] = True
    A = A.tolil()
    for i in np.flatnonzero(fixed):
        A[i, :] = 0.0
        A[:, i] = 0.0
        A[i, i] = 1.0
    A = A.tocsr()
    A.sort_indices()
    ipA, ixA, vA = A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data
    b = np.random.default_rng(5).standard_normal(n)
    b[fixed] = 0.0
    P = oracle.Ildlt(ipA, ixA, vA)
    assert P.replaced == 0
    x1, it1, r1 = oracle.pcg(ipA, ixA, vA, b, 1e-10, 5000, P.apply)
    dinv = 1.0 / A.diagonal()
    x2, it2, r2 = oracle.pcg(ipA, ixA, vA, b, 1e-10, 5000, lambda r: dinv * r)
    assert r1 <= 1e-10 and r2 <= 1e-10
    assert it1 < it2
    np.testing.assert_allclose(A @ x1, b, atol=1e-8 * np.linalg.norm(b))
