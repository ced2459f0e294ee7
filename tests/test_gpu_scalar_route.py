"""The workspace's order-0 / order-1 route on the device: GetFEM's symbolic
derivative route to the SIMP compliance gradients (u registered as a fem
constant, rho as a fem variable on a P0 mesh_fem; the potential
E(rho) (lambda Div_u^2 + 2 mu Sym(Grad_u):Sym(Grad_u)); assembly(0) gives J,
assembly(1) dJ/drho, src/getfem_generic_assembly_workspace.cc:520-541,
:832-847), and the pulled-back coordinates route to dJ/dX (a P1 vector
variable w at zero, Grad_u*Inv(Id+Grad_w) and Det(Id+Grad_w)).  Checked
against the oracle's analytic compliance gradients at n = 3 and against the
hand-written adjoint kernel at 32^3 x 6, to 1e-12."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd.workspace import GaWorkspace

pytestmark = pytest.mark.gpu

NU = 0.3
LAM = NU / ((1 + NU) * (1 - 2 * NU))
MU = 1 / (2 * (1 + NU))
EMIN = 1e-3
SIMP = "(0.001 + pow(rho,3)*(1-0.001))"
ELAST = "(lambda*sqr(Div_u) + 2*mu*Sym(Grad_u):Sym(Grad_u))"
FINV = "Inv(Id(meshdim)+Grad_w)"
PULLED = ("%s*(lambda*sqr(Trace(Grad_u*%s)) + 2*mu*Sym(Grad_u*%s):Sym(Grad_u*%s))*Det(Id(meshdim)+Grad_w)"
          % (SIMP, FINV, FINV, FINV))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def setup(ctx, n):
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf_u = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    mf_rho = ga.MeshFem(mesh, "FEM_PK_DISCONTINUOUS(3,0)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    rho = meshes.simp_density(cvx.shape[0])
    u = meshes.random_state(mf_u.nbdof())
    rdofs = mf_rho.basic_dof_table().reshape(-1)
    rho_dofs = np.zeros(mf_rho.nbdof())
    rho_dofs[rdofs] = rho
    return pts, cvx, mesh, mf_u, mf_rho, mim, rho, rho_dofs, rdofs, u


def workspace(ctx, mf_u, mf_rho, mim, u, rho_dofs, expr, mf_w=None):
    ws = GaWorkspace(ctx)
    ws.add_fem_constant("u", mf_u, u)
    ws.add_fem_variable("rho", mf_rho, rho_dofs, first=0)
    if mf_w is not None:
        ws.add_fem_variable("w", mf_w, np.zeros(mf_w.nbdof()), first=mf_rho.nbdof())
    ws.add_fixed_size_constant("lambda", LAM)
    ws.add_fixed_size_constant("mu", MU)
    ws.add_expression(expr, mim, 1)
    return ws


def point_gradient(mf_w, cvx, V, n_pts):
    """dJ/dw on the P1 vector mesh_fem's dofs -> per mesh point (x, y, z)."""
    dofs = mf_w.basic_dof_table().reshape(cvx.shape)
    out = np.zeros((n_pts, 3))
    out[cvx.reshape(-1)] = V[dofs.reshape(-1)[:, None] + np.arange(3)]
    return out


@pytest.mark.parametrize("n", [2, 3])
def test_potential_and_rho_gradient_match_oracle(ctx, n):
    print("case: scalar route SIMP n=%d" % n, flush=True)
    pts, cvx, mesh, mf_u, mf_rho, mim, rho, rho_dofs, rdofs, u = setup(ctx, n)
    ws = workspace(ctx, mf_u, mf_rho, mim, u, rho_dofs, SIMP + "*" + ELAST)
    prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 1, 3, mf_u.basic_dof_table(),
                          mf_u.nbdof(), "IM_TETRAHEDRON(1)", oracle.TERM_ISO_ELAST, scal_a=LAM, scal_b=MU, U=u)
    Jr, drr, dxr = prob.compliance(rho)
    ws.assembly(0)
    J = ws.assembled_potential()
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    ws.assembly(1)
    V = ws.assembled_vector()
    assert V.shape == (mf_rho.nbdof(),)
    assert rel(V[rdofs], drr) <= 1e-12


def test_pulled_back_coordinates_match_oracle(ctx):
    print("case: scalar route pulled-back coordinates n=3", flush=True)
    pts, cvx, mesh, mf_u, mf_rho, mim, rho, rho_dofs, rdofs, u = setup(ctx, 3)
    mf_w = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    ws = workspace(ctx, mf_u, mf_rho, mim, u, rho_dofs, PULLED, mf_w)
    prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 1, 3, mf_u.basic_dof_table(),
                          mf_u.nbdof(), "IM_TETRAHEDRON(1)", oracle.TERM_ISO_ELAST, scal_a=LAM, scal_b=MU, U=u)
    Jr, drr, dxr = prob.compliance(rho)
    ws.assembly(0)
    assert abs(ws.assembled_potential() - Jr) <= 1e-12 * abs(Jr)
    ws.assembly(1)
    V = ws.assembled_vector()
    nr = mf_rho.nbdof()
    assert V.shape == (nr + mf_w.nbdof(),)
    assert rel(V[:nr][rdofs], drr) <= 1e-12
    assert rel(point_gradient(mf_w, cvx, V[nr:], pts.shape[0]), dxr) <= 1e-12
    # the expansion of Inv around w = 0 is refused at another value of w
    ws.set_values("w", np.full(mf_w.nbdof(), 1e-3))
    with pytest.raises(ga.GfxError) as e:
        ws.assembly(1)
    assert "expanded around w = 0" in str(e.value)


def test_full_size_matches_adjoint_kernel(ctx):
    """32^3 x 6 = 196,608 tets: the GWFL route (symbolic derivatives, generic
    kernel) against the hand-written adjoint kernel behind the autograd
    Function."""
    import torch

    from getfem_amd.torch_ops import Compliance, compliance
    print("case: scalar route vs adjoint kernel n=32", flush=True)
    pts, cvx, mesh, mf_u, mf_rho, mim, rho, rho_dofs, rdofs, u = setup(ctx, 32)
    mf_w = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    ws = workspace(ctx, mf_u, mf_rho, mim, u, rho_dofs, PULLED, mf_w)
    ws.assembly(0)
    J = ws.assembled_potential()
    ws.assembly(1)
    V = ws.assembled_vector()
    op = Compliance(mf_u, mim, LAM, MU, EMIN, 3.0)
    dev = torch.device("cuda:0")
    tr = torch.tensor(rho, device=dev, requires_grad=True)
    tx = torch.tensor(pts, device=dev, requires_grad=True)
    Jk = compliance(tr, tx, torch.tensor(u, device=dev), op)
    gr, gx = torch.autograd.grad(Jk, [tr, tx])
    nr = mf_rho.nbdof()
    print("J route %.17g kernel %.17g" % (J, float(Jk)), flush=True)
    assert abs(J - float(Jk)) <= 1e-12 * abs(float(Jk))
    assert rel(V[:nr][rdofs], gr.cpu().numpy()) <= 1e-12
    assert rel(point_gradient(mf_w, cvx, V[nr:], pts.shape[0]), gx.cpu().numpy()) <= 1e-12
