"""Re-thresholding after the values change.  GetFEM thresholds every element
matrix at every assembly (src/getfem_generic_assembly_compile_and_exec.cc:
5441-5453), so its stored pattern and the entries each element contributes
follow the current coefficients, node positions and state.  A plan whose
values changed re-derives its thresholds and keep bits at the next assembly
and rebuilds its pattern when GetFEM's would differ; every case here is
compared with the oracle on the new values: pattern bit-exact, values within
1e-12 relative Frobenius."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from tests import problems

pytestmark = pytest.mark.gpu


def check_against_oracle(plan, case, lam, mu, pts=None, tol=1e-12):
    if pts is not None:
        case.pts = pts
    case.lam, case.mu = lam, mu
    ip, ix, v = plan.csr()
    ipr, ixr, vr = case.oracle_csr()
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), "pattern differs from the oracle"
    assert problems.rel_frob(v, vr) <= tol


def build(ctx, case, form=ga.FORM_FACTORIZED, prec=ga.PREC_FP64):
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, case.fem, case.qdim)
    mim = ga.MeshIm(mesh, case.im)
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [case.lam, case.mu])
    plan.set_form(form, prec)
    plan.build_pattern()
    plan.assemble()
    return mesh, plan


@pytest.mark.parametrize("form", [ga.FORM_FACTORIZED, ga.FORM_PERQP])
def test_nonuniform_coefficients_zeroed_and_tiny_elements(ctx, form):
    print("case: P2 elasticity n=3, zeroed / 1e-16-scaled elements, form", form, flush=True)
    case = problems.c2(3)
    lam0, mu0 = case.lam.copy(), case.mu.copy()
    mesh, plan = build(ctx, case, form)
    check_against_oracle(plan, case, lam0, mu0)
    n_cv = case.cvx.shape[0]
    rng = np.random.default_rng(11)
    lam, mu = case.lam * rng.uniform(0.5, 2.0, n_cv), case.mu * rng.uniform(0.5, 2.0, n_cv)
    zero = rng.choice(n_cv, 20, replace=False)
    tiny = rng.choice(np.setdiff1d(np.arange(n_cv), zero), 20, replace=False)
    lam[zero] = mu[zero] = 0.0
    lam[tiny] *= 1e-16
    mu[tiny] *= 1e-16
    id0 = plan.pattern_id()
    plan.set_coef(0, lam)
    plan.set_coef(1, mu)
    plan.assemble()
    check_against_oracle(plan, case, lam, mu)
    # zeroed elements insert nothing: entries only they carried leave the pattern
    assert plan.pattern_id() != id0
    # entries carried only by the 1e-16-scaled elements keep their (tiny)
    # values: the thresholds follow the element's own new scale
    ip, ix, v = plan.csr()
    ipr, ixr, vr = case.oracle_csr()
    node_elems = {}
    for e in range(n_cv):
        for d in case.dofs[e]:
            node_elems.setdefault(int(d), set()).add(e)
    tiny_set, checked = set(int(t) for t in tiny), 0
    for e in tiny:
        for a in case.dofs[e]:
            for b in case.dofs[e]:
                if not (node_elems[int(a)] & node_elems[int(b)]) <= tiny_set:
                    continue
                for r in range(int(a), int(a) + 3):
                    for k in range(ipr[r], ipr[r + 1]):
                        if int(b) <= ixr[k] < int(b) + 3 and vr[k] != 0.0:
                            assert abs(v[k] - vr[k]) <= 1e-9 * abs(vr[k])
                            checked += 1
    assert checked > 0
    # and back: elements zero at the last pattern build contribute again
    plan.set_coef(0, lam0)
    plan.set_coef(1, mu0)
    plan.assemble()
    check_against_oracle(plan, case, lam0, mu0)


def test_same_pattern_after_new_values_is_kept(ctx):
    print("case: P2 elasticity n=3, coefficients doubled on half the elements", flush=True)
    case = problems.c2(3)
    mesh, plan = build(ctx, case)
    id0 = plan.pattern_id()
    lam, mu = case.lam.copy(), case.mu.copy()
    lam[::2] *= 2.0
    mu[::2] *= 3.0
    plan.set_coef(0, lam)
    plan.set_coef(1, mu)
    plan.assemble()
    check_against_oracle(plan, case, lam, mu)
    assert plan.pattern_id() == id0  # the thresholds moved, the pattern did not


@pytest.mark.parametrize("prec", [ga.PREC_FP64, ga.PREC_FP32])
def test_moved_nodes(ctx, prec):
    """Unjittered mesh first: the flat boundary makes some element-matrix
    entries vanish, which GetFEM does not store.  Jittering the nodes brings
    entries in; flattening them again takes them out."""
    print("case: P2 elasticity n=3 unjittered -> jittered -> unjittered, prec", prec, flush=True)
    case = problems.Case(3, 2, 3, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, jitter=False)
    case.lam, case.mu = meshes.lognormal_lame(case.cvx.shape[0])
    flat = case.pts.copy()
    mesh, plan = build(ctx, case, ga.FORM_FACTORIZED, prec)
    nnz_flat = plan.dims()[1]
    rng = np.random.default_rng(5)
    moved = flat + 0.02 * rng.uniform(-1, 1, flat.shape)  # every node, boundary ones too
    # the pattern is GetFEM's (FP64 thresholds) in both precisions; the FP32
    # form's values are within its stated 1e-6
    tol = 1e-12 if prec == ga.PREC_FP64 else 1e-6
    mesh.set_points(moved)
    plan.assemble()
    check_against_oracle(plan, case, case.lam, case.mu, pts=moved, tol=tol)
    assert plan.dims()[1] > nnz_flat
    mesh.set_points(flat)
    plan.assemble()
    check_against_oracle(plan, case, case.lam, case.mu, pts=flat, tol=tol)
    assert plan.dims()[1] == nnz_flat


def test_degenerate_element_after_a_move(ctx):
    """A node moved onto a neighbour flattens the elements around both
    (det K exactly 0 on the dyadic coordinates of the 2^3 mesh): the next
    assembly raises GetFEM's 'Non invertible matrix'."""
    print("case: P1 Poisson n=2, node moved onto its neighbour", flush=True)
    pts, cvx = meshes.regular_unit_mesh(2)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    plan = ga.Plan(mf, mim, ga.TERM_LAPLACE, [1.0])
    plan.build_pattern()
    plan.assemble()
    bad = pts.copy()
    a, b = cvx[0, 0], cvx[0, 1]
    bad[a] = pts[b]
    mesh.set_points(bad)
    with pytest.raises(ga.GfxError) as e:
        plan.assemble()
    assert e.value.code == 3
    mesh.set_points(pts)
    plan.assemble()  # the mesh is valid again


def test_pinned_pattern_is_not_rebuilt(ctx):
    """While zero-copy torch views of the CSR are alive the pattern is pinned:
    a change of values that would change GetFEM's pattern raises instead of
    freeing the arrays the views point to."""
    print("case: P2 elasticity n=2 pinned by torch views", flush=True)
    case = problems.c2(2)
    mesh, plan = build(ctx, case)
    views = plan.torch_csr_view()
    lam = case.lam.copy()
    mu = case.mu.copy()
    lam[:10] = mu[:10] = 0.0
    plan.set_coef(0, lam)
    plan.set_coef(1, mu)
    with pytest.raises(ga.GfxError) as e:
        plan.assemble()
    assert e.value.code == 6
    del views
    plan.assemble()
    check_against_oracle(plan, case, lam, mu)


def test_neohookean_state_change_keeps_or_rebuilds_pattern(ctx):
    """The nonlinear plan's pattern follows the state: each new state is
    re-thresholded at assembly; the tangent and residual match the oracle
    and the pattern is rebuilt only when its presence bits change."""
    print("case: Q2 Neo-Hookean n=2 over three states", flush=True)
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)")
    mesh = ga.Mesh(ctx, "GT_QK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_QK(3,2)", 3)
    mim = ga.MeshIm(mesh, "IM_GAUSS_PARALLELEPIPED(3,5)")
    plan = ga.Plan(mf, mim, ga.TERM_NEOHOOKEAN, [10.0, 1.0])
    plan.set_state(np.zeros(mf.nbdof()))
    plan.build_pattern()
    ids = []
    for amp in (0.0, 0.05, 0.03):
        U = meshes.sine_modes()(mf.basic_dof_nodes(), 0.05).ravel() * (amp / 0.05)
        plan.set_state(U)
        plan.assemble()
        ids.append(plan.pattern_id())
        ip, ix, v = plan.csr()
        P = oracle.Problem(oracle.GT_QK31, pts, cvx, oracle.FEM_QK, 2, 3, mf.basic_dof_table(), mf.nbdof(),
                           "IM_GAUSS_PARALLELEPIPED(3,5)", oracle.TERM_NEOHOOKEAN, scal_a=10.0, scal_b=1.0, U=U)
        ipr, ixr, vr = oracle.csc_to_csr(*P.assemble_csc(), mf.nbdof())
        assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), amp
        assert problems.rel_frob(v, vr) <= 1e-12
        if amp:
            assert problems.rel_frob(plan.residual(), P.assemble_vector()) <= 1e-12
    # two deformed states with the same pattern: no rebuild between them
    assert ids[1] == ids[2]


def test_workspace_nonuniform_constant_values(ctx):
    print("case: workspace P2 elasticity n=3, P0 constants zeroed / scaled", flush=True)
    from getfem_amd.workspace import GaWorkspace
    case = problems.c2(3)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mf0 = ga.MeshFem(mesh, "FEM_PK_DISCONTINUOUS(3,0)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    d0 = mf0.basic_dof_table()[:, 0]
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf)
    L = np.zeros(mf0.nbdof()); L[d0] = case.lam
    M = np.zeros(mf0.nbdof()); M[d0] = case.mu
    ws.add_fem_constant("lambda", mf0, L)
    ws.add_fem_constant("mu", mf0, M)
    ws.add_expression("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", mim)
    ws.assembly(2)
    lam, mu = case.lam.copy(), case.mu.copy()
    lam[:15] = mu[:15] = 0.0
    lam[15:30] *= 1e-16
    mu[15:30] *= 1e-16
    L[d0] = lam
    M[d0] = mu
    ws.set_values("lambda", L)
    ws.set_values("mu", M)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    case.lam, case.mu = lam, mu
    ipr, ixr, vr = case.oracle_csr()
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12


def test_compliance_degenerate_coordinates_reported(ctx):
    import torch
    from getfem_amd.torch_ops import Compliance, compliance
    print("case: compliance P1 n=2 with a flattened element in X", flush=True)
    pts, cvx = meshes.regular_unit_mesh(2)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    op = Compliance(mf, mim, 1.0, 0.5)
    dev = torch.device("cuda:0")
    rho = torch.full((cvx.shape[0],), 0.5, dtype=torch.float64, device=dev, requires_grad=True)
    u = torch.tensor(meshes.random_state(mf.nbdof()), device=dev)
    bad = pts.copy()
    bad[cvx[0, 0]] = pts[cvx[0, 1]]
    with pytest.raises(ga.GfxError) as e:
        compliance(rho, torch.tensor(bad, device=dev), u, op)
    assert e.value.code == 3
    J = compliance(rho, None, u, op)  # the flag was cleared; the mesh's own points are fine
    J.backward()
    assert torch.isfinite(rho.grad).all()
