"""The ga_workspace drop-in on the device: expressions compiled by the GWFL
stage assemble the same matrices as the oracle, every spelling of the
reference's differential tests gives the same matrix (mat_norminf < 1e-10 as
tests/test_assembly.cc:514-548 requires), and gf.asm('generic')-style calls
return GetFEM's CSC orientation."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd.workspace import GaWorkspace, asm_generic
from tests import problems
from tests.test_gwfl_compile import ELASTICITY_SPELLINGS, LAPLACE_SPELLINGS

pytestmark = pytest.mark.gpu


def objects(ctx, n, fem, q, im):
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    return mesh, ga.MeshFem(mesh, fem, q), ga.MeshIm(mesh, im)


def test_workspace_p0_coefficients_match_oracle(ctx):
    case = problems.c2(4)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mf0 = ga.MeshFem(mesh, "FEM_PK_DISCONTINUOUS(3,0)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf)
    d0 = mf0.basic_dof_table()[:, 0]
    L = np.zeros(mf0.nbdof()); L[d0] = case.lam
    M = np.zeros(mf0.nbdof()); M[d0] = case.mu
    ws.add_fem_constant("lambda", mf0, L)
    ws.add_fem_constant("mu", mf0, M)
    ws.add_expression("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", mim)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    ipr, ixr, vr = case.oracle_csr()
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12
    # new coefficient values keep the pattern and change the values linearly
    ws.set_values("lambda", 2 * L)
    ws.set_values("mu", 2 * M)
    ws.assembly(2)
    v2 = ws.assembled_matrix_csr()[2]
    assert problems.rel_frob(v2, 2 * vr) <= 1e-12


def test_elasticity_spellings_same_matrix(ctx):
    mesh, mf, mim = objects(ctx, 3, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)")
    ref = None
    for e in ELASTICITY_SPELLINGS:
        K = asm_generic(mim, 2, e, -1, "u", 1, mf, np.zeros(mf.nbdof()), "lambda", 0, 3.0, "mu", 0, 2.0)
        if ref is None:
            ref = K
        assert abs(K - ref).max() < 1e-10, e


def test_laplacian_spellings_same_matrix(ctx):
    mesh, mf, mim = objects(ctx, 3, "FEM_PK(3,2)", 1, "IM_TETRAHEDRON(5)")
    ref = None
    for e in LAPLACE_SPELLINGS:
        K = asm_generic(mim, 2, e, -1, "p", 1, mf, np.zeros(mf.nbdof()))
        if ref is None:
            ref = K
        assert abs(K - ref).max() < 1e-10, e


def test_asm_generic_is_getfem_csc(ctx):
    case = problems.c1(6)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    K = asm_generic(mim, 2, "Grad_u.Grad_Test_u", -1, "u", 1, mf, np.zeros(mf.nbdof()))
    cp, r, v = case.oracle_problem().assemble_csc()
    assert np.array_equal(K.indptr, cp) and np.array_equal(K.indices, r)
    assert problems.rel_frob(K.data, v) <= 1e-12
