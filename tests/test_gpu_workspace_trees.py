"""Several expressions and trees in one workspace.  GetFEM merges expressions
with the same (mesh_im, region, test functions) into one tree
(src/getfem_generic_assembly_workspace.cc:474-495: one integrand, one element
matrix, one threshold) and adds every tree into the assembled matrix
(compile_and_exec.cc:7910-8643); the matrix of a workspace that does not own
it is added into, not cleared (workspace.cc:804-807).  Each case is compared
with the oracle: pattern bit-exact, values within 1e-12."""
import numpy as np
import pytest
import torch

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd.workspace import GaWorkspace
from tests import problems

pytestmark = pytest.mark.gpu

ELAST = "lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u"


def setup_case(ctx, n=3):
    case = problems.c2(n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mf0 = ga.MeshFem(mesh, "FEM_PK_DISCONTINUOUS(3,0)", 1)
    d0 = mf0.basic_dof_table()[:, 0]
    L = np.zeros(mf0.nbdof()); L[d0] = case.lam
    M = np.zeros(mf0.nbdof()); M[d0] = case.mu
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf)
    ws.add_fem_constant("lambda", mf0, L)
    ws.add_fem_constant("mu", mf0, M)
    return case, mesh, mf, ws


def oracle_csr(case, im, term, **coefs):
    P = oracle.Problem(oracle.GT_PK31, case.pts, case.cvx, oracle.FEM_PK, 2, 3, case.dofs, case.ndof, im, term,
                       **coefs)
    return oracle.csc_to_csr(*P.assemble_csc(), case.ndof)


def csr_sum(n, *mats):
    """Structural union of CSR matrices with summed values (entries that sum
    to zero stay stored, as in GetFEM's rsvector)."""
    acc = {}
    for ip, ix, v in mats:
        for r in range(n):
            for k in range(ip[r], ip[r + 1]):
                acc[(r, ix[k])] = acc.get((r, ix[k]), 0.0) + v[k]
    keys = sorted(acc)
    ip = np.zeros(n + 1, dtype=np.int64)
    for r, _ in keys:
        ip[r + 1] += 1
    return np.cumsum(ip), np.array([c for _, c in keys], dtype=np.int32), np.array([acc[k] for k in keys])


def test_laplace_plus_elasticity_one_tree(ctx):
    """Same mesh_im: one merged tree, thresholded once on the sum."""
    print("case: workspace elasticity + 3*Laplace, one mesh_im", flush=True)
    case, mesh, mf, ws = setup_case(ctx)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws.add_expression(ELAST, mim)
    ws.add_expression("3*Grad_u:Grad_Test_u", mim)
    ws.assembly(2)
    assert ws.nb_trees() == 1
    assert ws.describe().startswith("isotropic")
    ip, ix, v = ws.assembled_matrix_csr()
    ipr, ixr, vr = oracle_csr(case, "IM_TETRAHEDRON(5)", oracle.TERM_ISOTROPIC, coef_a=case.lam, coef_b=case.mu,
                              coef_c=case.mu + 3.0)
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12


def test_navier_form(ctx):
    print("case: workspace mu*Grad:Grad + (lambda+mu)*Div*Div", flush=True)
    case, mesh, mf, ws = setup_case(ctx)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws.add_expression("mu*Grad_u:Grad_Test_u + (lambda+mu)*Div_u*Div_Test_u", mim)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    ipr, ixr, vr = oracle_csr(case, "IM_TETRAHEDRON(5)", oracle.TERM_ISOTROPIC, coef_a=case.lam + case.mu,
                              scal_b=0.0, coef_c=case.mu)
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12


def test_two_trees_on_two_mesh_ims(ctx):
    """Different mesh_ims: two trees, each thresholded on its own element
    matrices, both added into one matrix whose pattern is the union."""
    print("case: workspace elasticity on IM_TETRAHEDRON(5) + Laplace on IM_TETRAHEDRON(2)", flush=True)
    case, mesh, mf, ws = setup_case(ctx)
    mim5 = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    mim2 = ga.MeshIm(mesh, "IM_TETRAHEDRON(2)")
    ws.add_expression(ELAST, mim5)
    ws.add_expression("2*Grad_u:Grad_Test_u", mim2)
    for rep in range(2):  # a second assembly reuses the union
        ws.assembly(2)
        assert ws.nb_trees() == 2
        ip, ix, v = ws.assembled_matrix_csr()
        A = oracle_csr(case, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, coef_a=case.lam, coef_b=case.mu)
        B = oracle_csr(case, "IM_TETRAHEDRON(2)", oracle.TERM_LAPLACE, scal_a=2.0)
        ipr, ixr, vr = csr_sum(case.ndof, A, B)
        assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), rep
        assert problems.rel_frob(v, vr) <= 1e-12
    # GetFEM's CSC orientation of the same sum
    K = ws.assembled_matrix()
    assert np.array_equal(K.tocsr().indptr, ipr)
    # order 1: the sum of the trees' K u
    rng = np.random.default_rng(3)
    u = rng.standard_normal(case.ndof)
    ws.set_values("u", u)
    ws.assembly(1)
    import scipy.sparse as sp
    Kr = sp.csr_matrix((vr, ixr, ipr), shape=(case.ndof, case.ndof))
    assert problems.rel_frob(ws.assembled_vector(), Kr @ u) <= 1e-12


def test_caller_owned_matrix_is_added_into(ctx):
    """set_assembled_matrix: the assembly adds into the caller's matrix and
    never clears it (two assemblies give twice the values)."""
    print("case: workspace into a caller-owned CSR, twice", flush=True)
    case, mesh, mf, ws = setup_case(ctx)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws.add_expression(ELAST, mim)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    dev = torch.device("cuda:0")
    tip, tix = torch.tensor(ip, device=dev), torch.tensor(ix, device=dev)
    tv = torch.zeros(len(v), dtype=torch.float64, device=dev)
    ws.set_assembled_matrix(tip, tix, tv)
    ws.assembly(2)
    ws.assembly(2)
    torch.cuda.synchronize()
    assert problems.rel_frob(tv.cpu().numpy(), 2 * v) <= 1e-14
    ws.set_assembled_matrix()  # back to the workspace's own matrix
    ws.assembly(2)
    assert problems.rel_frob(ws.assembled_matrix_csr()[2], v) <= 1e-14


def test_plan_add_to_csc_target(ctx):
    """A plan's values added into a caller CSC in GetFEM's orientation with a
    larger pattern (the full topological pattern of the mesh)."""
    print("case: plan values into a GetFEM-orientation CSC superset", flush=True)
    case = problems.c2(2)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [case.lam, case.mu])
    plan.build_pattern()
    plan.assemble()
    ip, ix, v = plan.csr()
    n = case.ndof
    cols = [set() for _ in range(n)]
    for e in range(case.cvx.shape[0]):
        d = (case.dofs[e][:, None] + np.arange(3)[None, :]).ravel()
        for r in d:
            cols[r].update(int(c) for c in d)
    # CSC over columns c with sorted rows r: the topological pattern is symmetric
    cp = np.zeros(n + 1, dtype=np.int64)
    for c in range(n):
        cp[c + 1] = cp[c] + len(cols[c])
    ri = np.concatenate([np.array(sorted(cols[c]), dtype=np.int32) for c in range(n)])
    dev = torch.device("cuda:0")
    tv = torch.zeros(len(ri), dtype=torch.float64, device=dev)
    plan.add_to_matrix(torch.tensor(cp, device=dev), torch.tensor(ri, device=dev), tv, ga.CSC)
    got = tv.cpu().numpy()
    import scipy.sparse as sp
    K = sp.csr_matrix((v, ix, ip), shape=(n, n)).tocsc()
    dense = np.zeros((n, n))
    for c in range(n):
        for k in range(cp[c], cp[c + 1]):
            dense[ri[k], c] = got[k]
    assert np.abs(dense - K.toarray()).max() <= 1e-15 * np.abs(v).max()
    # a target that lacks an entry (the first of row 0) is refused
    tip = ip.copy()
    tip[1:] -= 1
    with pytest.raises(ga.GfxError) as e:
        plan.add_to_matrix(torch.tensor(tip, device=dev), torch.tensor(ix[1:], device=dev),
                           torch.zeros(len(ix) - 1, dtype=torch.float64, device=dev), ga.CSR)
    assert e.value.code == 7


def test_order_one_follows_expression_orders(ctx):
    """assembly(1) runs only expressions of test-function order <= 1: an
    order-2 expression ('Grad_Test2_u:Grad_Test_u') has no order-1 tree."""
    print("case: workspace order-2-only expression + elasticity, assembly(1)", flush=True)
    case, mesh, mf, ws = setup_case(ctx)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    rng = np.random.default_rng(9)
    u = rng.standard_normal(case.ndof)
    ws.set_values("u", u)
    ws.add_expression("3*Grad_Test2_u:Grad_Test_u", mim)
    ws.assembly(1)
    assert np.abs(ws.assembled_vector()).max() == 0.0
    ws.add_expression(ELAST, mim)
    ws.assembly(2)  # both: the merged tree
    ip, ix, v = ws.assembled_matrix_csr()
    ipr, ixr, vr = oracle_csr(case, "IM_TETRAHEDRON(5)", oracle.TERM_ISOTROPIC, coef_a=case.lam, coef_b=case.mu,
                              coef_c=case.mu + 3.0)
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12
    ws.assembly(1)  # only the elasticity expression: K_elast u
    import scipy.sparse as sp
    ipe, ixe, ve = oracle_csr(case, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, coef_a=case.lam, coef_b=case.mu)
    Ke = sp.csr_matrix((ve, ixe, ipe), shape=(case.ndof, case.ndof))
    assert problems.rel_frob(ws.assembled_vector(), Ke @ u) <= 1e-12


def test_device_state_order_one(ctx):
    """The unknown's values from a device vector give the same order-1 vector."""
    print("case: workspace set_values_device", flush=True)
    case, mesh, mf, ws = setup_case(ctx, 2)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws.add_expression(ELAST, mim)
    u = np.random.default_rng(2).standard_normal(case.ndof)
    ws.set_values("u", u)
    ws.assembly(1)
    ref = ws.assembled_vector()
    du = torch.tensor(2 * u, device="cuda:0")
    ws.set_values_device("u", du.data_ptr())
    ws.assembly(1)
    assert problems.rel_frob(ws.assembled_vector(), 2 * ref) <= 1e-14
