"""The model layer on the device (getfem_amd.model.Model, after getfem::model +
standard_solve): linear elasticity with Dirichlet conditions by simplification
solved by device CG, and a Neo-Hookean Newton solve, both against the same
algorithm on the host with the oracle's matrices and a direct solver
(scipy.sparse.linalg.spsolve)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd.model import Model
from tests import problems

pytestmark = pytest.mark.gpu


def simplify(K, rhs, dofs, g, linear, U=None):
    """model::assembly's dof-constraint simplification (symmetric model) on
    host scipy matrices."""
    K = K.tolil(copy=True)
    rhs = rhs.copy()
    if linear:
        rhs -= K[:, dofs] @ g
        rhs[dofs] = g
    else:
        rhs[dofs] = g - U[dofs]
    K[dofs, :] = 0
    K[:, dofs] = 0
    K[dofs, dofs] = 1.0
    return K.tocsr(), rhs


def face_dofs(mf, axis, value, comps=(0, 1, 2)):
    nodes = mf.basic_dof_nodes()
    on = np.nonzero(np.abs(nodes[:, axis] - value) < 1e-10)[0]
    q = mf.qdim
    return np.sort(np.concatenate([on * q + c for c in comps]))


def test_linear_elasticity_dirichlet_cg(ctx):
    case = problems.c2(3)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    md = Model(ctx)
    md.add_fem_variable("u", mf)
    md.add_initialized_data("lambda", 3.0)
    md.add_initialized_data("mu", 2.0)
    md.add_initialized_data("pull", [0.01, 0.0, 0.0])
    md.add_linear_term(mim, "lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
                       is_symmetric=True, is_coercive=True)
    left, right = face_dofs(mf, 0, 0.0), face_dofs(mf, 0, 1.0)
    md.add_Dirichlet_condition_with_simplification("u", left)
    md.add_Dirichlet_condition_with_simplification("u", right, "pull")
    its, ok, log = md.solve(max_res=1e-12)
    U = md.variable("u")
    # host: same simplification, direct solve
    n = mf.nbdof()
    P = oracle.Problem(oracle.GT_PK31, case.pts, case.cvx, oracle.FEM_PK, 2, 3, mf.basic_dof_table(), n,
                       "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, scal_a=3.0, scal_b=2.0)
    cp, r, v = P.assemble_csc()
    K = sp.csc_matrix((v, r, cp), shape=(n, n))
    dofs = np.concatenate([left, right])
    g = np.concatenate([np.zeros(left.size), np.array([0.01, 0, 0])[right % 3]])
    Ks, rhs = simplify(K, np.zeros(n), dofs, g, True)
    Uref = spla.spsolve(Ks.tocsc(), rhs)
    err = np.linalg.norm(U - Uref) / np.linalg.norm(Uref)
    print("cg iterations", log, "rel err vs direct", err)
    assert ok and err <= 1e-9
    assert np.abs(U[left]).max() < 1e-10 and np.abs(U[right] - g[left.size:]).max() < 1e-10


def test_neohookean_newton(ctx):
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)")
    mesh = ga.Mesh(ctx, "GT_QK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_QK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_GAUSS_PARALLELEPIPED(3,5)")
    md = Model(ctx)
    md.add_fem_variable("u", mf)
    md.add_initialized_data("lambda", 10.0)
    md.add_initialized_data("mu", 1.0)
    md.add_initialized_data("lift", [0.0, 0.0, 0.08])
    md.add_nonlinear_term(
        mim, "((Id(meshdim)+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(Grad_u,[lambda;mu]))):Grad_Test_u",
        is_symmetric=True, is_coercive=True)
    bottom, top = face_dofs(mf, 2, 0.0), face_dofs(mf, 2, 1.0)
    md.add_Dirichlet_condition_with_simplification("u", bottom)
    md.add_Dirichlet_condition_with_simplification("u", top, "lift")
    its, ok, log = md.solve(max_iter=30, max_res=1e-10, noisy=True)
    U = md.variable("u")
    assert ok, (its, log)
    # host Newton: same constraints, oracle tangent/residual, direct solves
    nd = mf.nbdof()
    dofs = np.concatenate([bottom, top])
    g = np.concatenate([np.zeros(bottom.size), np.array([0.0, 0.0, 0.08])[top % 3]])
    Uh = np.zeros(nd)
    for k in range(30):
        P = oracle.Problem(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, 2, 3, mf.basic_dof_table(), nd,
                           "IM_GAUSS_PARALLELEPIPED(3,5)", oracle.TERM_NEOHOOKEAN, scal_a=10.0, scal_b=1.0,
                           U=Uh)
        cp, r, v = P.assemble_csc()
        K = sp.csc_matrix((v, r, cp), shape=(nd, nd))
        Ks, rhs = simplify(K, -P.assemble_vector(), dofs, g, False, Uh)
        dU = spla.spsolve(Ks.tocsc(), rhs)
        Uh = Uh + dU
        if np.abs(dU).sum() / max(1e-25, np.abs(Uh).sum()) < 1e-13:
            break
    err = np.linalg.norm(U - Uh) / np.linalg.norm(Uh)
    print("newton", its, "cg", log, "rel err vs host Newton", err)
    assert err <= 1e-8
    assert np.abs(U[top] - g[bottom.size:]).max() < 1e-10



IM5 = "IM_TETRAHEDRON(5)"


def _host_pieces(pts, cvx, du, dv, kappa):
    """The linear tangent of  Grad_u.Grad_Test_u + Grad_v.Grad_Test_v
    + kappa (u - v)(Test_u - Test_v)  and u's mass matrix, from the oracle's
    block element matrices, over the global vector (u's interval first)."""
    nu = int(du.max()) + 1
    n = nu + int(dv.max()) + 1
    fu, fv = (oracle.FEM_PK, 2, 1, du), (oracle.FEM_PK, 1, 1, dv)

    def diag(k):
        def fn(vals):
            D = np.zeros((4, 4))
            D[0, 0] = k
            D[1:, 1:] = np.eye(3)
            return D
        return fn

    def mass(k):
        def fn(vals):
            D = np.zeros((4, 4))
            D[0, 0] = k
            return D
        return fn

    rv = nu + dv
    Euu = oracle.block_element_matrices(pts, cvx, IM5, fu, fu, {}, diag(kappa))
    Evv = oracle.block_element_matrices(pts, cvx, IM5, fv, fv, {}, diag(kappa))
    Euv = oracle.block_element_matrices(pts, cvx, IM5, fu, fv, {}, mass(-kappa))
    Mu = oracle.block_element_matrices(pts, cvx, IM5, fu, fu, {}, mass(1.0))
    K = oracle.assemble_blocks(cvx.shape[0], n, [(du, du, Euu), (du, rv, Euv),
                                                 (rv, du, np.transpose(Euv, (0, 2, 1))), (rv, rv, Evv)])
    M = oracle.assemble_blocks(cvx.shape[0], n, [(du, du, Mu)])
    return (sp.csc_matrix((K[2], K[1], K[0]), shape=(n, n)),
            sp.csc_matrix((M[2], M[1], M[0]), shape=(n, n)))


def _u2_mass(pts, cvx, du, n, Uu):
    """int u^2 phi_i phi_j over u's block (the cubic term's weight)"""
    def fn(vals):
        D = np.zeros((4, 4))
        D[0, 0] = vals["u"][0][0] ** 2
        return D
    f = (oracle.FEM_PK, 2, 1, du)
    E = oracle.block_element_matrices(pts, cvx, IM5, f, f, {"u": (oracle.FEM_PK, 2, 1, du, Uu)}, fn)
    cp, r, v = oracle.assemble_blocks(cvx.shape[0], n, [(du, du, E)])
    return sp.csc_matrix((v, r, cp), shape=(n, n))


@pytest.mark.parametrize("c", [0.0, 2.0])
def test_two_variable_model_matches_host_newton(ctx, c):
    """Two scalar fields (u P2, v P1) coupled by kappa (u - v), a source on u,
    u = v = 0 on the boundary, and for c > 0 the nonlinear term c u^3: the
    device model (polynomial route, Newton with device CG) against Newton on
    the host with the oracle's matrices and a direct solver.  v is added
    before u: the intervals follow the names (u first)."""
    print("case: two-variable model c=%g" % c, flush=True)
    n, kappa, f = 3, 5.0, 10.0
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mfu = ga.MeshFem(mesh, "FEM_PK(3,2)", 1)
    mfv = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, IM5)
    md = Model(ctx)
    md.add_fem_variable("v", mfv)
    md.add_fem_variable("u", mfu)
    md.add_initialized_data("kappa", kappa)
    md.add_initialized_data("f", f)
    md.add_initialized_data("c", c)
    md.add_linear_term(mim, "Grad_u.Grad_Test_u + Grad_v.Grad_Test_v + kappa*(u-v)*(Test_u-Test_v)",
                       is_symmetric=True, is_coercive=True)
    md.add_linear_term(mim, "-f*Test_u", is_symmetric=True, is_coercive=True)
    if c:
        md.add_nonlinear_term(mim, "c*u*u*u*Test_u", is_symmetric=True, is_coercive=True)

    def boundary(mf):
        x = mf.basic_dof_nodes()
        return np.nonzero(np.any((np.abs(x) < 1e-10) | (np.abs(x - 1) < 1e-10), axis=1))[0]

    bu, bv = boundary(mfu), boundary(mfv)
    md.add_Dirichlet_condition_with_simplification("u", bu)
    md.add_Dirichlet_condition_with_simplification("v", bv)
    its, ok, log = md.solve(max_res=1e-11)
    nu = mfu.nbdof()
    U = np.concatenate([md.variable("u"), md.variable("v")])
    # host Newton on the same system
    du, dv = mfu.basic_dof_table(), mfv.basic_dof_table()
    dofs = np.concatenate([bu, nu + bv])
    Kl, Mu = _host_pieces(pts, cvx, du, dv, kappa)
    F = f * np.asarray(Mu.sum(axis=1)).ravel()  # f int phi_i on u's interval
    X = np.zeros_like(F)
    for _ in range(30):
        Kt, R = Kl, Kl @ X - F
        if c:
            M2 = _u2_mass(pts, cvx, du, Kl.shape[0], X[:nu])
            Kt, R = Kl + 3 * c * M2, R + c * (M2 @ X)
        Ks, rhs = simplify(Kt, -R, dofs, np.zeros(dofs.size), False, X)
        dX = spla.spsolve(Ks.tocsc(), rhs)
        X += dX
        if np.abs(dX).max() <= 1e-14 * np.abs(X).max():
            break
    err = np.linalg.norm(U - X) / np.linalg.norm(X)
    print("newton", its, "cg", log, "rel err vs host", err)
    assert ok and err <= 1e-8
    assert np.abs(U[dofs]).max() == 0.0


def test_mixed_pressure_model_gmres(ctx):
    """Nearly incompressible elasticity in mixed form, u P2 vector and p P1:
    2 mu Sym(Grad_u):Grad_Test_u - p Div_Test_u - Test_p Div_u - p Test_p / lambda,
    symmetric but indefinite (not coercive), so the model solves it with
    GMRES; u clamped at x = 0 and pulled at x = 1.  Against the host direct
    solve of the oracle's block matrices.  The pressure rows come first (name
    order), where ILU(0)'s pivots are the tiny -M/lambda entries and the
    factorization is useless (as GetFEM's would be): the diagonal
    preconditioner; at lambda = 100 (it bounds the error less tightly as
    lambda grows: 3e-6 at lambda = 1e4)."""
    print("case: mixed u-p model, GMRES", flush=True)
    n, mu, lam = 3, 1.0, 100.0
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mfu = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mfp = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, IM5)
    md = Model(ctx)
    md.add_fem_variable("u", mfu)
    md.add_fem_variable("p", mfp)
    md.add_initialized_data("mu", mu)
    md.add_initialized_data("ilam", 1.0 / lam)
    md.add_initialized_data("pull", [0.02, 0.0, 0.0])
    md.add_linear_term(mim, "2*mu*Sym(Grad_u):Grad_Test_u - p*Div_Test_u - Test_p*Div_u - ilam*p*Test_p",
                       is_symmetric=True, is_coercive=False)
    left, right = face_dofs(mfu, 0, 0.0), face_dofs(mfu, 0, 1.0)
    md.add_Dirichlet_condition_with_simplification("u", left)
    md.add_Dirichlet_condition_with_simplification("u", right, "pull")
    its, ok, log = md.solve(max_res=1e-10, precond="diagonal")
    npp, nu = mfp.nbdof(), mfu.nbdof()
    X = np.concatenate([md.variable("p"), md.variable("u")])  # name order: p, then u
    dp, du = mfp.basic_dof_table(), mfu.basic_dof_table()
    rows_u = npp + (du[:, :, None] + np.arange(3)).reshape(cvx.shape[0], -1)
    rows_p = dp
    fu, fp = (oracle.FEM_PK, 2, 3, du), (oracle.FEM_PK, 1, 1, dp)

    def uu(vals):
        d = np.eye(3)
        D = np.zeros((12, 12))
        D[3:, 3:] = (mu * (np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))).reshape(9, 9)
        return D

    def up(vals):
        D = np.zeros((12, 4))
        for a in range(3):
            D[3 + 4 * a, 0] = -1.0
        return D

    def pp(vals):
        D = np.zeros((4, 4))
        D[0, 0] = -1.0 / lam
        return D

    Euu = oracle.block_element_matrices(pts, cvx, IM5, fu, fu, {}, uu)
    Eup = oracle.block_element_matrices(pts, cvx, IM5, fu, fp, {}, up)
    Epp = oracle.block_element_matrices(pts, cvx, IM5, fp, fp, {}, pp)
    N = npp + nu
    cp, r, v = oracle.assemble_blocks(cvx.shape[0], N, [(rows_u, rows_u, Euu), (rows_u, rows_p, Eup),
                                                        (rows_p, rows_u, np.transpose(Eup, (0, 2, 1))),
                                                        (rows_p, rows_p, Epp)])
    K = sp.csc_matrix((v, r, cp), shape=(N, N))
    dofs = npp + np.concatenate([left, right])
    g = np.concatenate([np.zeros(left.size), np.array([0.02, 0, 0])[right % 3]])
    Ks, rhs = simplify(K, np.zeros(N), dofs, g, True)
    Xref = spla.spsolve(Ks.tocsc(), rhs)
    err = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
    print("newton", its, "gmres", log, "rel err vs direct", err)
    assert ok and err <= 1e-8


def test_convection_diffusion_model_gmres(ctx):
    """One scalar P2 variable with a convection term (not symmetric, outside
    the matrix kernels: the polynomial route) and a source, u = 0 on the
    boundary: GMRES through Newton's residual form, against the host direct
    solve with the oracle's block matrix."""
    print("case: convection-diffusion model, GMRES", flush=True)
    n, beta, f = 3, np.array([3.0, -2.0, 1.0]), 5.0
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 1)
    mim = ga.MeshIm(mesh, IM5)
    md = Model(ctx)
    md.add_fem_variable("u", mf)
    for name, val in zip(("bx", "by", "bz"), beta):
        md.add_initialized_data(name, val)
    md.add_initialized_data("f", f)
    md.add_linear_term(mim, "Grad_u.Grad_Test_u + (bx*Grad_u(1) + by*Grad_u(2) + bz*Grad_u(3))*Test_u"
                       " - f*Test_u")
    x = mf.basic_dof_nodes()
    bd = np.nonzero(np.any((np.abs(x) < 1e-10) | (np.abs(x - 1) < 1e-10), axis=1))[0]
    md.add_Dirichlet_condition_with_simplification("u", bd)
    its, ok, log = md.solve(max_res=1e-10)
    U = md.variable("u")
    dofs = mf.basic_dof_table()
    fu = (oracle.FEM_PK, 2, 1, dofs)

    def conv(vals):
        D = np.zeros((4, 4))
        D[1:, 1:] = np.eye(3)
        D[0, 1:] = beta
        return D

    def mass(vals):
        D = np.zeros((4, 4))
        D[0, 0] = 1.0
        return D

    nd = mf.nbdof()
    E = oracle.block_element_matrices(pts, cvx, IM5, fu, fu, {}, conv)
    M = oracle.block_element_matrices(pts, cvx, IM5, fu, fu, {}, mass)
    cp, r, v = oracle.assemble_blocks(cvx.shape[0], nd, [(dofs, dofs, E)])
    cm, rm, vm = oracle.assemble_blocks(cvx.shape[0], nd, [(dofs, dofs, M)])
    K = sp.csc_matrix((v, r, cp), shape=(nd, nd)).tolil()
    F = f * np.asarray(sp.csc_matrix((vm, rm, cm), shape=(nd, nd)).sum(axis=1)).ravel()
    K[bd, :] = 0
    K[bd, bd] = 1.0
    F[bd] = 0.0
    Uref = spla.spsolve(K.tocsc(), F)
    err = np.linalg.norm(U - Uref) / np.linalg.norm(Uref)
    print("newton", its, "gmres", log, "rel err vs direct", err)
    assert ok and err <= 1e-8
