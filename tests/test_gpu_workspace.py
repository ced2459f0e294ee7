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


def test_order1_linear_is_K_u(ctx):
    """assembly(1) of a bilinear expression: the vector of its first
    derivative at the current u, i.e. K u (oracle CSC times u)."""
    case = problems.c2(3)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    rng = np.random.default_rng(7)
    u = rng.standard_normal(mf.nbdof())
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf, u)
    ws.add_fixed_size_constant("lambda", 3.0)
    ws.add_fixed_size_constant("mu", 2.0)
    ws.add_expression("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", mim)
    ws.assembly(1)
    V = ws.assembled_vector()
    P = oracle.Problem(oracle.GT_PK31, case.pts, case.cvx, oracle.FEM_PK, 2, 3, mf.basic_dof_table(),
                       mf.nbdof(), "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, scal_a=3.0, scal_b=2.0)
    import scipy.sparse as sp
    cp, r, v = P.assemble_csc()
    K = sp.csc_matrix((v, r, cp), shape=(mf.nbdof(), mf.nbdof()))
    assert problems.rel_frob(V, K @ u) <= 1e-12
    # new values of the unknown: same workspace, new vector
    ws.set_values("u", 2 * u)
    ws.assembly(1)
    assert problems.rel_frob(ws.assembled_vector(), 2 * (K @ u)) <= 1e-12
    # asm_generic order 1 gives the same vector
    V1 = asm_generic(mim, 1, "lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", -1,
                     "u", 1, mf, u, "lambda", 0, 3.0, "mu", 0, 2.0)
    assert problems.rel_frob(V1, K @ u) <= 1e-12


NH_EXPR = "((Id(meshdim)+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(Grad_u,[lambda;mu]))):Grad_Test_u"


def test_workspace_neohookean_tangent_and_residual(ctx):
    """The finite strain brick's expression through the workspace: order 2 is
    the tangent (pattern rebuilt for each new state, as GetFEM's stored
    pattern depends on the values) and order 1 the residual."""
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)")
    mesh = ga.Mesh(ctx, "GT_QK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_QK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_GAUSS_PARALLELEPIPED(3,5)")
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf, np.zeros(mf.nbdof()))
    ws.add_fixed_size_constant("lambda", 10.0)
    ws.add_fixed_size_constant("mu", 1.0)
    ws.add_expression(NH_EXPR, mim)
    assert "neo-hookean" in ws.compile()
    for scale in (0.0, 0.05, 0.02):
        U = meshes.sine_modes()(mf.basic_dof_nodes(), 0.05).ravel() * (scale / 0.05)
        ws.set_values("u", U)
        ws.assembly(1)
        R = ws.assembled_vector()
        ws.assembly(2)
        ip, ix, v = ws.assembled_matrix_csr()
        P = oracle.Problem(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, 2, 3, mf.basic_dof_table(), mf.nbdof(),
                           "IM_GAUSS_PARALLELEPIPED(3,5)", oracle.TERM_NEOHOOKEAN, scal_a=10.0, scal_b=1.0, U=U)
        ipr, ixr, vr = oracle.csc_to_csr(*P.assemble_csc(), mf.nbdof())
        assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), scale
        assert problems.rel_frob(v, vr) <= 1e-12
        Rr = P.assemble_vector()
        if scale == 0.0:
            assert np.abs(R).max() == 0.0 and np.abs(Rr).max() == 0.0
        else:
            assert problems.rel_frob(R, Rr) <= 1e-12
