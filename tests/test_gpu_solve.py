"""Sparse operations on the assembled device CSR: the product K x against the
oracle's CSC, and the block-Jacobi-preconditioned CG solve (gmm::cg's stopping test
||b - K x|| <= rtol ||b||, checked on the host with the oracle's matrix)."""
import numpy as np
import pytest
import scipy.sparse as sp

import getfem_amd as ga
import oracle
from tests import problems

pytestmark = pytest.mark.gpu


def elasticity(ctx, n):
    case = problems.c2(n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [case.lam, case.mu])
    plan.build_pattern()
    plan.assemble()
    ip, ix, vr = case.oracle_csr()
    K = sp.csr_matrix((vr, ix, ip), shape=(mf.nbdof(), mf.nbdof()))
    return plan, K, (mesh, mf, mim)


def test_spmv_matches_oracle(ctx):
    import torch
    plan, K, keep = elasticity(ctx, 3)
    n = K.shape[0]
    x = np.random.default_rng(3).standard_normal(n)
    dx = torch.tensor(x, device="cuda")
    dy = torch.empty(n, dtype=torch.float64, device="cuda")
    plan.spmv(dx.data_ptr(), dy.data_ptr())
    plan.ctx.synchronize()
    assert problems.rel_frob(dy.cpu().numpy(), K @ x) <= 1e-13


@pytest.mark.parametrize("n", [2, 4])
def test_cg_consistent_system(ctx, n):
    """b = K x_true (K has the six rigid modes as kernel): CG from x = 0 reaches
    ||b - K x|| <= 1e-10 ||b||, and x differs from x_true by a rigid motion
    only (K (x - x_true) = b - K x)."""
    import torch
    plan, K, keep = elasticity(ctx, n)
    nd = K.shape[0]
    x_true = np.random.default_rng(11).standard_normal(nd)
    b = K @ x_true
    db = torch.tensor(b, device="cuda")
    dx = torch.zeros(nd, dtype=torch.float64, device="cuda")
    it, res = plan.solve_cg(db.data_ptr(), dx.data_ptr(), rtol=1e-10, maxit=5000)
    x = dx.cpu().numpy()
    host_res = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
    print("n", n, "ndof", nd, "iterations", it, "device rel res", res, "host rel res", host_res)
    assert res <= 1e-10 and host_res <= 2e-10
    assert 0 < it < 5000


def test_cg_zero_rhs_and_maxit(ctx):
    import torch
    plan, K, keep = elasticity(ctx, 2)
    nd = K.shape[0]
    db = torch.zeros(nd, dtype=torch.float64, device="cuda")
    dx = torch.ones(nd, dtype=torch.float64, device="cuda")
    it, res = plan.solve_cg(db.data_ptr(), dx.data_ptr())
    assert it == 0 and float(dx.abs().max()) == 0.0
    # maxit reached: not an error, residual reported above the tolerance
    b = K @ np.random.default_rng(5).standard_normal(nd)
    db = torch.tensor(b, device="cuda")
    dx = torch.zeros(nd, dtype=torch.float64, device="cuda")
    it, res = plan.solve_cg(db.data_ptr(), dx.data_ptr(), rtol=1e-14, maxit=3)
    assert it == 3 and res > 1e-14
    x = dx.cpu().numpy()
    assert abs(np.linalg.norm(b - K @ x) / np.linalg.norm(b) - res) <= 1e-8
