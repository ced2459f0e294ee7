"""GMRES's ILU on the device against the oracle's restatement of
gmm::ilu_precond (src/gmm/gmm_precond_ilu.h:104-231) on matrices the
polynomial route assembles: a mixed u-p elasticity matrix (symmetric,
indefinite) and a convection-diffusion matrix (not symmetric).  The factor
and its application follow the reference's operations in order, so the
device's M^-1 r must equal the oracle's bit for bit."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from getfem_amd._lib import check, lib
from getfem_amd.workspace import GaWorkspace

pytestmark = pytest.mark.gpu


def workspace(ctx, case):
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(5)")
    ws = GaWorkspace(ctx)
    rng = np.random.default_rng(5)
    if case == "mixed":
        mfu = ga.MeshFem(mesh, "FEM_PK(3,2)", 3)
        mfp = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
        ws.add_fem_variable("p", mfp, np.zeros(mfp.nbdof()), first=0)
        ws.add_fem_variable("u", mfu, np.zeros(mfu.nbdof()), first=mfp.nbdof())
        ws.add_fixed_size_constant("ilam", 1e-4)
        ws.add_expression("2*Sym(Grad_u):Grad_Test_u - p*Div_Test_u - Test_p*Div_u - ilam*p*Test_p", mim)
    else:
        mf = ga.MeshFem(mesh, "FEM_PK(3,2)", 1)
        ws.add_fem_variable("u", mf, np.zeros(mf.nbdof()))
        for k, v in zip(("bx", "by", "bz"), rng.standard_normal(3) * 5):
            ws.add_fixed_size_constant(k, v)
        ws.add_expression("Grad_u.Grad_Test_u + (bx*Grad_u(1) + by*Grad_u(2) + bz*Grad_u(3))*Test_u", mim)
    return ws


@pytest.mark.parametrize("case", ["mixed", "convection"])
def test_ilu_apply_bitwise(ctx, case):
    import ctypes
    import torch
    print("case: ILU apply, %s" % case, flush=True)
    ws = workspace(ctx, case)
    ws.assembly(2)
    ip, ix, v = ws.assembled_matrix_csr()
    n = ip.size - 1
    rep = ctypes.c_int64()
    check(lib().gfx_ws_ilu_factor(ws.h, ctypes.byref(rep)), ctx.h)
    F, rep_ref = oracle.ilu(ip, ix, v)
    assert rep.value == rep_ref
    rng = np.random.default_rng(9)
    for _ in range(2):
        r = rng.standard_normal(n)
        d_r = torch.tensor(r, device="cuda")
        d_z = torch.empty_like(d_r)
        torch.cuda.synchronize()
        check(lib().gfx_ws_ilu_apply(ws.h, ctypes.c_void_p(d_r.data_ptr()), ctypes.c_void_p(d_z.data_ptr())), ctx.h)
        check(lib().gfx_ctx_synchronize(ctx.h), ctx.h)
        z_ref = oracle.ilu_apply(ip, ix, F, r)
        assert np.array_equal(d_z.cpu().numpy(), z_ref)
