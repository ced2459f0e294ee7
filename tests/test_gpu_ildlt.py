"""The device ILDLT preconditioner (gmm::ildlt_precond) against the oracle's
restatement on the same device-assembled matrix: the factor and the apply
bit for bit, and CG with it (gmm::cg) against the oracle's PCG on the host."""
import numpy as np
import pytest
import scipy.sparse as sp

import getfem_amd as ga
import oracle
from getfem_amd import _lib
from getfem_amd._lib import GfxError
from tests import problems

pytestmark = pytest.mark.gpu


def constrained(ctx, case, fem, qdim):
    """Plan assembled on the device, dofs of the nodes on the face x = 0
    clamped (Dirichlet by simplification on the device CSR)."""
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, fem, qdim)
    mim = ga.MeshIm(mesh, case.im)
    term = ga.TERM_LAPLACE if case.term == oracle.TERM_LAPLACE else ga.TERM_ISO_ELAST
    plan = ga.Plan(mf, mim, term, case.coefs())
    plan.build_pattern()
    plan.assemble()
    nodes = mf.basic_dof_nodes()  # one row per node; dof = qdim node + component
    on_face = np.flatnonzero(np.abs(nodes[:, 0]) < 1e-12)
    fixed = (qdim * on_face[:, None] + np.arange(qdim)[None, :]).ravel().astype(np.int32)
    plan.dof_constraints(fixed, np.zeros(fixed.size),
                         flags=_lib.BUILD_MATRIX | _lib.MODEL_LINEAR | _lib.MODEL_SYMMETRIC)
    ip, ix, v = plan.csr()
    return plan, (ip, ix, v), fixed, (mesh, mf, mim)


def factor_positions(ip, ix, P):
    """The oracle's U slots mapped onto the CSR positions: F_ref aligned with
    the CSR (D on the diagonal, U_ij / D_i at (i, j) and at (j, i))."""
    n = ip.size - 1
    F = np.full(ix.size, np.nan)
    A = sp.csr_matrix((np.arange(ix.size, dtype=np.float64), ix, ip), shape=(n, n))
    for i in range(n):
        s0, s1 = P.uptr[i], P.uptr[i + 1]
        cols = P.uind[s0:s1]
        vals = P.uval[s0:s1]
        for c, val in zip(cols, vals):
            F[int(A[i, c])] = val
            if c != i:
                F[int(A[c, i])] = val
    return F


CASES = {
    "p2_elast_n3": (lambda: problems.c2(3), "FEM_PK(3,2)", 3),
    "p1_poisson_n6": (lambda: problems.c1(6), "FEM_PK(3,1)", 1),
    "p3_elast_n2": (lambda: problems.c5(2), "FEM_PK(3,3)", 3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_factor_and_apply_bitwise(ctx, name):
    import torch
    mk, fem, q = CASES[name]
    case = mk()
    plan, (ip, ix, v), fixed, keep = constrained(ctx, case, fem, q)
    print("case", name, "ndof", ip.size - 1, "nnz", ix.size, flush=True)
    P = oracle.Ildlt(ip, ix, v)
    assert P.replaced == 0
    nf, nb = plan.ildlt_factor()
    print("levels forward", nf, "backward", nb)
    F = plan.ildlt_values()
    Fr = factor_positions(ip, ix, P)
    assert not np.isnan(Fr).any()
    assert np.array_equal(F.view(np.int64), Fr.view(np.int64))
    r = np.random.default_rng(7).standard_normal(ip.size - 1)
    dr = torch.tensor(r, device="cuda")
    dz = torch.empty_like(dr)
    plan.ildlt_apply(dr.data_ptr(), dz.data_ptr())
    plan.ctx.synchronize()
    z = dz.cpu().numpy()
    assert np.array_equal(z.view(np.int64), P.apply(r).view(np.int64))


@pytest.mark.parametrize("name", ["p2_elast_n3", "p1_poisson_n6"])
def test_pcg_matches_oracle_pcg(ctx, name):
    """gmm::cg with ILDLT on both sides: the same iteration count up to one
    (the dot products sum in different orders), both converged, and fewer
    iterations than block Jacobi on the device."""
    import torch
    mk, fem, q = CASES[name]
    case = mk()
    plan, (ip, ix, v), fixed, keep = constrained(ctx, case, fem, q)
    n = ip.size - 1
    b = np.random.default_rng(8).standard_normal(n)
    b[fixed] = 0.0
    P = oracle.Ildlt(ip, ix, v)
    x_ref, it_ref, res_ref = oracle.pcg(ip, ix, v, b, 1e-10, 5000, P.apply)
    db = torch.tensor(b, device="cuda")
    dx = torch.zeros(n, dtype=torch.float64, device="cuda")
    it, res = plan.solve_pcg(db.data_ptr(), dx.data_ptr(), ga.PRECOND_ILDLT, rtol=1e-10, maxit=5000)
    x = dx.cpu().numpy()
    K = sp.csr_matrix((v, ix, ip), shape=(n, n))
    host_res = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
    dx.zero_()
    it_bj, res_bj = plan.solve_pcg(db.data_ptr(), dx.data_ptr(), ga.PRECOND_BLOCK_JACOBI, rtol=1e-10, maxit=5000)
    print(name, "ILDLT device", it, res, "oracle", it_ref, res_ref, "block Jacobi", it_bj, res_bj)
    assert res <= 1e-10 and host_res <= 2e-10 and res_ref <= 1e-10
    assert abs(it - it_ref) <= 1
    assert it < it_bj
    np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-8 * np.abs(x_ref).max())


def test_singular_matrix_follows_the_reference(ctx):
    """Pure Neumann elasticity (rigid modes in the kernel): where the
    reference replaces a pivot the device reports it, otherwise the factors
    agree bit for bit."""
    case = problems.c2(2)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, case.im)
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, case.coefs())
    plan.build_pattern()
    plan.assemble()
    ip, ix, v = plan.csr()
    P = oracle.Ildlt(ip, ix, v)
    if P.replaced:
        with pytest.raises(GfxError, match="pivot"):
            plan.ildlt_factor()
    else:
        plan.ildlt_factor()
        F = plan.ildlt_values()
        assert np.array_equal(F.view(np.int64), factor_positions(ip, ix, P).view(np.int64))


def test_repeated_apply_bitwise(ctx):
    """Three applications to new right-hand sides written into the same
    buffers (torch work and synchronisations in between) each equal the
    oracle's apply bit for bit."""
    import torch
    mk, fem, q = CASES["p2_elast_n3"]
    case = mk()
    plan, (ip, ix, v), fixed, keep = constrained(ctx, case, fem, q)
    P = oracle.Ildlt(ip, ix, v)
    plan.ildlt_factor()
    n = ip.size - 1
    dr = torch.empty(n, dtype=torch.float64, device="cuda")
    dz = torch.empty_like(dr)
    for k in range(3):
        r = np.random.default_rng(20 + k).standard_normal(n)
        dr.copy_(torch.tensor(r))
        dz.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.ildlt_apply(dr.data_ptr(), dz.data_ptr())
        plan.ctx.synchronize()
        z = dz.cpu().numpy()
        assert np.array_equal(z.view(np.int64), P.apply(r).view(np.int64)), "replay %d" % k
