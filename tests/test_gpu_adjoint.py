"""Adjoint of the SIMP compliance J = u^T K(rho) u (P1 tets): libgfx through
the torch.autograd.Function against the oracle's analytic gradients and
against central finite differences (h = 1e-6) on 64 random directions."""
import numpy as np
import pytest

import oracle
from getfem_amd import meshes

pytestmark = pytest.mark.gpu

NU = 0.3
LAM = NU / ((1 + NU) * (1 - 2 * NU))
MU = 1 / (2 * (1 + NU))


def setup(n, ctx):
    import torch

    import getfem_amd as ga
    from getfem_amd.torch_ops import Compliance
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    op = Compliance(mf, mim, LAM, MU, 1e-3, 3.0)
    dev = torch.device("cuda:0")
    rho = torch.tensor(meshes.simp_density(cvx.shape[0]), device=dev)
    u = torch.tensor(meshes.random_state(mf.nbdof()), device=dev)
    X = torch.tensor(pts, device=dev)
    dofs = mf.basic_dof_table()
    prob = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 1, 3, dofs, mf.nbdof(),
                          "IM_TETRAHEDRON(1)", oracle.TERM_ISO_ELAST, scal_a=LAM, scal_b=MU,
                          U=u.cpu().numpy())
    return op, rho, X, u, prob, (mesh, mf, mim)


@pytest.mark.parametrize("n", [3, 32])
def test_adjoint_matches_oracle(ctx, n):
    import torch
    from getfem_amd.torch_ops import compliance
    op, rho, X, u, prob, keep = setup(n, ctx)
    rho.requires_grad_(True)
    X.requires_grad_(True)
    u.requires_grad_(True)
    J = compliance(rho, X, u, op)
    J.backward()
    Jr, drr, dxr = prob.compliance(rho.detach().cpu().numpy())
    print("n", n, "J", float(J.detach()), "oracle", Jr, flush=True)
    assert abs(float(J.detach()) - Jr) <= 1e-12 * abs(Jr)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    assert rel(rho.grad.cpu().numpy(), drr) <= 1e-12
    assert rel(X.grad.cpu().numpy(), dxr) <= 1e-12
    # dJ/du = 2 K(rho) u with K assembled by the oracle
    E = 1e-3 + rho.detach().cpu().numpy() ** 3 * (1 - 1e-3)
    Q = oracle.Problem(oracle.GT_PK31, prob.pts, prob.cvx, oracle.FEM_PK, 1, 3, prob.dofs, prob.ndof,
                       "IM_TETRAHEDRON(1)", oracle.TERM_ISO_ELAST, coef_a=E * LAM, coef_b=E * MU)
    import scipy.sparse as sp
    cp, r, v = Q.assemble_csc()
    K = sp.csc_matrix((v, r, cp), shape=(prob.ndof, prob.ndof))
    assert rel(u.grad.cpu().numpy(), 2 * (K @ u.detach().cpu().numpy())) <= 1e-12


def test_adjoint_finite_differences_64_directions(ctx):
    """Full size: 32^3 x 6 = 196,608 tets; directions with N(0,1) entries in
    rho and in every nodal coordinate."""
    import torch
    from getfem_amd.torch_ops import compliance
    op, rho, X, u, prob, keep = setup(32, ctx)
    rho.requires_grad_(True)
    X.requires_grad_(True)
    J = compliance(rho, X, u, op)
    gr, gx = torch.autograd.grad(J, [rho, X])
    gen = torch.Generator(device="cpu").manual_seed(1234)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for k in range(64):
            dr = torch.randn(rho.shape, generator=gen, dtype=torch.float64).to(rho.device)
            dx = torch.randn(X.shape, generator=gen, dtype=torch.float64).to(X.device)
            Jp = op.eval(rho + h * dr, X + h * dx, u, want_grads=False)[0]
            Jm = op.eval(rho - h * dr, X - h * dx, u, want_grads=False)[0]
            fd = float((Jp - Jm) / (2 * h))
            an = float(torch.dot(gr, dr) + torch.sum(gx * dx))
            err = abs(fd - an) / max(abs(an), 1e-30)
            worst = max(worst, err)
    print("worst relative FD error over 64 directions", worst)
    assert worst <= 1e-6


def test_compliance_graph_replays_match_eager(ctx):
    """gfx_compliance_eval is kernels only: captured once into a graph, every
    replay (with eager evaluations and synchronisations in between, on
    changing inputs) gives the eager result."""
    import torch
    op, rho, X, u, prob, keep = setup(3, ctx)
    outs = (torch.empty((), dtype=torch.float64, device="cuda"), torch.empty_like(rho),
            torch.empty_like(X), torch.empty_like(u))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        op.eval(rho, X, u, out=outs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        op.eval(rho, X, u, out=outs)
    torch.cuda.synchronize()
    for k in range(4):
        with torch.no_grad():
            rho.mul_(0.9).add_(0.05)
            u.mul_(1.1)
        J, dr, dx, du = op.eval(rho.clone(), X, u.clone())  # eager, fresh outputs
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((J, dr, dx, du), outs):
            assert torch.equal(a, b), k
