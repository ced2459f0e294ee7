"""The workspace's order-2 polynomial route on the device (several fem
variables, blocks between two mesh_fems, data fields, polynomial
nonlinearities): ga_workspace::assembly(2) with
ga_instruction_matrix_assembly_mf_mf and add_elem_matrix
(src/getfem_generic_assembly_compile_and_exec.cc:5074-5189, :4853-4936).

Checked against the oracle (numpy element matrices of hand-derived second
derivatives, assembled by the oracle's add_elem_matrix): the CSR pattern
bit-exact and the values within 1e-12 (relative, Frobenius and per entry).
The elasticity block is also checked against the matrix kernels' own plan."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd.workspace import GaWorkspace

pytestmark = pytest.mark.gpu

LAM, MU, K3 = 1.3, 0.7, 0.4


def elasticity(vals):
    d = np.eye(3)
    C = LAM * np.einsum("ab,cd->abcd", d, d) + MU * (np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))
    D = np.zeros((12, 12))
    D[3:, 3:] = C.reshape(9, 9)
    return D


def coupling_up(vals):  # -p Div(Test_u) along Test2_p, scaled by rho
    D = np.zeros((12, 4))
    for a in range(3):
        D[3 + 4 * a, 0] = -vals["rho"][0][0]
    return D


def stiff_pp(vals):  # rho Grad_p . Grad_Test_p
    D = np.zeros((4, 4))
    D[1:, 1:] = vals["rho"][0][0] * np.eye(3)
    return D


def gpu_csr(ws):
    ip, ix, v = ws.assembled_matrix_csr()
    return ip, ix, v


def check(ip, ix, v, ref, n):
    rip, rix, rv = oracle.csc_to_csr(*ref, n)
    assert np.array_equal(ip, rip), "row pointers"
    assert np.array_equal(ix, rix), "columns"
    assert np.linalg.norm(v - rv) <= 1e-12 * np.linalg.norm(rv)
    assert np.abs(v - rv).max() <= 1e-12 * np.abs(rv).max()


def setup(ctx, n):
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf_u = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
    mf_p = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mf_r = ga.MeshFem(mesh, "FEM_PK_DISCONTINUOUS(3,0)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    return pts, cvx, mesh, mf_u, mf_p, mf_r, mim


@pytest.mark.parametrize("n", [2, 3])
def test_two_variable_blocks_match_oracle(ctx, n):
    print("case: polynomial route, u P2 vector + p P1, n=%d" % n, flush=True)
    pts, cvx, mesh, mf_u, mf_p, mf_r, mim = setup(ctx, n)
    nu, npp = mf_u.nbdof(), mf_p.nbdof()
    rng = np.random.default_rng(7 + n)
    U, P = rng.standard_normal(nu), rng.uniform(0.5, 1.5, npp)
    rho = rng.uniform(0.5, 2.0, cvx.shape[0])
    rho[::5] = 0.0  # zero element matrices in the (u,p), (p,u), (p,p) blocks: nothing added
    rdofs = mf_r.basic_dof_table().reshape(-1)
    R = np.zeros(mf_r.nbdof())
    R[rdofs] = rho
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf_u, U, first=0)
    ws.add_fem_variable("p", mf_p, P, first=nu)
    ws.add_fem_constant("rho", mf_r, R)
    ws.add_fixed_size_constant("lambda", LAM)
    ws.add_fixed_size_constant("mu", MU)
    ws.add_expression("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", mim)
    ws.add_expression("-rho*p*Div_Test_u - rho*Test_p*Div_u", mim)
    ws.add_expression("rho*Grad_p.Grad_Test_p", mim)
    ws.assembly(2)
    ip, ix, v = gpu_csr(ws)
    assert ip.shape == (nu + npp + 1,)
    du, dp = mf_u.basic_dof_table(), mf_p.basic_dof_table()
    fu = (oracle.FEM_PK, 2, 3, du)
    fp = (oracle.FEM_PK, 1, 1, dp)
    fields = {"rho": (oracle.FEM_PK_DISC0, 0, 1, mf_r.basic_dof_table(), R)}
    im = "IM_TETRAHEDRON(5)"
    Euu = oracle.block_element_matrices(pts, cvx, im, fu, fu, fields, elasticity)
    Eup = oracle.block_element_matrices(pts, cvx, im, fu, fp, fields, coupling_up)
    Epp = oracle.block_element_matrices(pts, cvx, im, fp, fp, fields, stiff_pp)
    rows_u = (du[:, :, None] + np.arange(3)).reshape(cvx.shape[0], -1)
    rows_p = nu + dp
    ref = oracle.assemble_blocks(cvx.shape[0], nu + npp, [
        (rows_u, rows_u, Euu), (rows_u, rows_p, Eup), (rows_p, rows_u, np.transpose(Eup, (0, 2, 1))),
        (rows_p, rows_p, Epp)])
    check(ip, ix, v, ref, nu + npp)
    # the elasticity block equals the matrix kernels' plan of the same term
    ws1 = GaWorkspace(ctx)
    ws1.add_fem_variable("u", mf_u, U)
    ws1.add_fixed_size_constant("lambda", LAM)
    ws1.add_fixed_size_constant("mu", MU)
    ws1.add_expression("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", mim)
    ws1.assembly(2)
    kip, kix, kv = ws1.assembled_matrix_csr()
    lo, hi = ip[:nu], ip[1:nu + 1]
    cols = [ix[a:b] for a, b in zip(lo, hi)]
    assert all(np.array_equal(c[c < nu], kix[kip[r]:kip[r + 1]]) for r, c in enumerate(cols))
    vals = np.concatenate([v[a:b][ix[a:b] < nu] for a, b in zip(lo, hi)])
    assert np.abs(vals - kv).max() <= 1e-12 * np.abs(kv).max()


def test_nonlinear_scalar_tangent_and_device_values(ctx):
    """Grad_u.Grad_Test_u + k u^3 Test_u (u P2 scalar): the tangent
    Grad_Test2_u.Grad_Test_u + 3 k u^2 Test2_u Test_u, at a state given on
    the host and then on the device; repeated assemblies are bitwise equal."""
    import torch
    print("case: polynomial route, nonlinear P2 scalar n=3", flush=True)
    pts, cvx, mesh, _, _, _, mim = setup(ctx, 3)
    mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 1)
    nd = mf.nbdof()
    rng = np.random.default_rng(3)
    U = rng.uniform(-1.0, 1.0, nd)
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mf, np.zeros(nd))
    ws.add_fixed_size_constant("k", K3)
    ws.add_expression("Grad_u.Grad_Test_u + k*u*u*u*Test_u", mim)
    dofs = mf.basic_dof_table()
    f = (oracle.FEM_PK, 2, 1, dofs)

    def tangent(vals):
        D = np.zeros((4, 4))
        D[0, 0] = 3 * K3 * vals["u"][0][0] ** 2
        D[1:, 1:] = np.eye(3)
        return D

    E = oracle.block_element_matrices(pts, cvx, "IM_TETRAHEDRON(5)", f, f,
                                      {"u": (oracle.FEM_PK, 2, 1, dofs, U)}, tangent)
    ref = oracle.assemble_blocks(cvx.shape[0], nd, [(dofs, dofs, E)])
    ws.set_values("u", U)
    ws.assembly(2)
    a = ws.assembled_matrix_csr()
    check(*a, ref, nd)
    ws.assembly(2)
    b = ws.assembled_matrix_csr()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    d_U = torch.tensor(U, device="cuda")
    torch.cuda.synchronize()
    ws.set_values("u", np.zeros(nd))
    ws.set_values_device("u", d_U.data_ptr())
    ws.assembly(2)
    c = ws.assembled_matrix_csr()
    assert all(np.array_equal(x, y) for x, y in zip(a, c))


def test_p3_vector_block_and_size_limit(ctx):
    """The largest element matrices of the route: P3 vector x P3 vector
    (3,600 entries, IM_TETRAHEDRON(8)) beside a P1 variable, against the
    matrix kernels' plan of the same elasticity term; a P4 vector block
    (11,025 entries) is refused with a message, before any launch."""
    print("case: polynomial route P3 vector block, P4 refusal", flush=True)
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mfu = ga.MeshFem(mesh, "FEM_PK(3,3)", 3)
    mfp = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(8)")
    nu = mfu.nbdof()
    elast = "lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u"
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mfu, np.zeros(nu), first=0)
    ws.add_fem_variable("p", mfp, np.zeros(mfp.nbdof()), first=nu)
    ws.add_fixed_size_constant("lambda", LAM)
    ws.add_fixed_size_constant("mu", MU)
    ws.add_expression(elast, mim)
    ws.add_expression("p*Test_p", mim)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    ws1 = GaWorkspace(ctx)
    ws1.add_fem_variable("u", mfu, np.zeros(nu))
    ws1.add_fixed_size_constant("lambda", LAM)
    ws1.add_fixed_size_constant("mu", MU)
    ws1.add_expression(elast, mim)
    ws1.assembly(2)
    kip, kix, kv = ws1.assembled_matrix_csr()
    assert np.array_equal(ip[:nu + 1], kip) and np.array_equal(ix[:kip[-1]], kix)
    assert np.abs(v[:kip[-1]] - kv).max() <= 1e-12 * np.abs(kv).max()
    mf4 = ga.MeshFem(mesh, "FEM_PK(3,4)", 3)
    ws2 = GaWorkspace(ctx)
    ws2.add_fem_variable("u", mf4, np.zeros(mf4.nbdof()), first=0)
    ws2.add_fem_variable("p", mfp, np.zeros(mfp.nbdof()), first=mf4.nbdof())
    ws2.add_expression("Grad_u:Grad_Test_u + p*Test_p", mim)
    with pytest.raises(ga.GfxError) as e:
        ws2.assembly(2)
    assert "larger than the polynomial route's element matrices" in str(e.value)


def test_hexahedra_two_variables(ctx):
    """Q1 hexahedra (the geometry at every point): u Q2 and v Q1 coupled by
    kappa (u - v) with k u^3 on u, tangent against the oracle's blocks."""
    print("case: polynomial route on hexahedra, u Q2 + v Q1", flush=True)
    n, kappa = 2, 3.0
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)")
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_QK(3,1)", pts, cvx)
    mfu = ga.MeshFem(mesh, "FEM_QK(3,2)", 1)
    mfv = ga.MeshFem(mesh, "FEM_QK(3,1)", 1)
    im = "IM_GAUSS_PARALLELEPIPED(3,5)"
    mim = ga.MeshIm(mesh, im)
    nu, nv = mfu.nbdof(), mfv.nbdof()
    rng = np.random.default_rng(8)
    U = rng.uniform(-1.0, 1.0, nu)
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mfu, U, first=0)
    ws.add_fem_variable("v", mfv, np.zeros(nv), first=nu)
    ws.add_fixed_size_constant("kappa", kappa)
    ws.add_fixed_size_constant("k", K3)
    ws.add_expression("Grad_u.Grad_Test_u + Grad_v.Grad_Test_v + kappa*(u-v)*(Test_u-Test_v)"
                      " + k*u*u*u*Test_u", mim)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    du, dv = mfu.basic_dof_table(), mfv.basic_dof_table()
    fu, fv = (oracle.FEM_QK, 2, 1, du), (oracle.FEM_QK, 1, 1, dv)
    fields = {"u": (oracle.FEM_QK, 2, 1, du, U)}

    def uu(vals):
        D = np.zeros((4, 4))
        D[0, 0] = kappa + 3 * K3 * vals["u"][0][0] ** 2
        D[1:, 1:] = np.eye(3)
        return D

    def vv(vals):
        D = np.zeros((4, 4))
        D[0, 0] = kappa
        D[1:, 1:] = np.eye(3)
        return D

    def uv(vals):
        D = np.zeros((4, 4))
        D[0, 0] = -kappa
        return D

    Euu = oracle.block_element_matrices(pts, cvx, im, fu, fu, fields, uu)
    Evv = oracle.block_element_matrices(pts, cvx, im, fv, fv, fields, vv)
    Euv = oracle.block_element_matrices(pts, cvx, im, fu, fv, fields, uv)
    rv = nu + dv
    ref = oracle.assemble_blocks(cvx.shape[0], nu + nv, [(du, du, Euu), (du, rv, Euv),
                                                          (rv, du, np.transpose(Euv, (0, 2, 1))), (rv, rv, Evv)])
    check(ip, ix, v, ref, nu + nv)


def test_hexahedra_residual_is_matrix_times_state(ctx):
    """Order 1 on Q1 hexahedra: for an expression linear in the variables
    the residual is the tangent times the state (up to the entries the
    threshold drops, none here)."""
    print("case: polynomial route on hexahedra, order 1 = K U", flush=True)
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n, "GT_QK(3,1)")
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_QK(3,1)", pts, cvx)
    mfu = ga.MeshFem(mesh, "FEM_QK(3,2)", 1)
    mfv = ga.MeshFem(mesh, "FEM_QK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_GAUSS_PARALLELEPIPED(3,5)")
    nu, nv = mfu.nbdof(), mfv.nbdof()
    rng = np.random.default_rng(12)
    X = rng.standard_normal(nu + nv)
    ws = GaWorkspace(ctx)
    ws.add_fem_variable("u", mfu, X[:nu], first=0)
    ws.add_fem_variable("v", mfv, X[nu:], first=nu)
    ws.add_fixed_size_constant("kappa", 3.0)
    ws.add_expression("Grad_u.Grad_Test_u + Grad_v.Grad_Test_v + kappa*(u-v)*(Test_u-Test_v)", mim)
    ws.assembly(2)
    K = ws.assembled_matrix()
    ws.assembly(1)
    V = ws.assembled_vector()
    KX = K @ X  # rows: test dofs, columns: trial dofs
    assert np.linalg.norm(V - KX) <= 1e-12 * np.linalg.norm(KX)


def test_asm_generic_two_variables_select_output(ctx):
    """gf.asm('generic') with two variables: consecutive intervals in argument
    order, select_output the (u, p) block and u's part of the vector."""
    from getfem_amd.workspace import asm_generic
    print("case: asm_generic, two variables, select_output", flush=True)
    pts, cvx, mesh, mf_u, mf_p, mf_r, mim = setup(ctx, 2)
    nu, npp = mf_u.nbdof(), mf_p.nbdof()
    rng = np.random.default_rng(6)
    U, P = rng.standard_normal(nu), rng.standard_normal(npp)
    expr = "2*Sym(Grad_u):Grad_Test_u - p*Div_Test_u - Test_p*Div_u"
    args = ("u", 1, mf_u, U, "p", 1, mf_p, P)
    K = asm_generic(mim, 2, expr, -1, *args)
    Kup = asm_generic(mim, 2, expr, -1, *args, select_output=("u", "p"))
    assert K.shape == (nu + npp, nu + npp) and Kup.shape == (nu, npp)
    assert (abs(K[:nu, nu:] - Kup)).max() == 0.0
    assert abs(K[:nu, nu:] - K[nu:, :nu].T).max() <= 1e-14 * abs(K[:nu, nu:]).max()
    V = asm_generic(mim, 1, expr, -1, *args)
    Vu = asm_generic(mim, 1, expr, -1, *args, select_output=("u",))
    assert np.array_equal(V[:nu], Vu)
    X = np.concatenate([U, P])
    assert np.linalg.norm(V - K @ X) <= 1e-12 * np.linalg.norm(K @ X)
