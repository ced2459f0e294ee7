"""gfx_compliance_eval_noalloc (static scratch, no allocation) against
gfx_compliance_eval (the compliance object's scratch) through the C ABI:
the same kernels in the same order, so J and the three gradients must be
bit-identical.  Two calls in a row, the second on a smaller mesh, then the
first mesh again (the shared scratch keeps the previous call's contents), and
a mesh beyond the scratch's limit is rejected with a clear error."""
import ctypes

import numpy as np
import pytest

import getfem_amd as ga
from getfem_amd import meshes
from getfem_amd._lib import GfxError, check, lib

pytestmark = pytest.mark.gpu

NU = 0.3
LAM = NU / ((1 + NU) * (1 - 2 * NU))
MU = 1 / (2 * (1 + NU))


def _problem(ctx, n, seed):
    import torch
    from getfem_amd.torch_ops import Compliance
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    op = Compliance(mf, mim, LAM, MU, 1e-3, 3.0)
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    rho = torch.tensor(rng.uniform(0.1, 1.0, cvx.shape[0]), device=dev)
    u = torch.tensor(rng.standard_normal(mf.nbdof()), device=dev)
    X = torch.tensor(pts, device=dev)
    return (mesh, mf, mim), op, rho, X, u


def _call(fn, op, rho, X, u, with_x=True):
    import torch
    J = torch.full((), float("nan"), dtype=torch.float64, device=rho.device)
    dr = torch.full_like(rho, float("nan"))
    dx = torch.full_like(X, float("nan")) if with_x else None
    du = torch.full_like(u, float("nan"))
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    check(fn(op.h, p(X), p(rho), p(u), p(J), p(dr), p(dx), p(du)), op.ctx.h)
    check(lib().gfx_ctx_synchronize(op.ctx.h), op.ctx.h)
    op.check()
    return [t.cpu().numpy() for t in (J, dr, dx, du) if t is not None]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.all(np.isfinite(x))
        assert np.array_equal(np.asarray(x).view(np.int64), np.asarray(y).view(np.int64))


def test_noalloc_matches_eval_twice_second_smaller():
    import torch
    torch.cuda.init()
    ctx = ga.Context(0)
    big = _problem(ctx, 8, 1)
    small = _problem(ctx, 5, 2)
    for keep, op, rho, X, u in (big, small, big):
        print("compliance noalloc n_cv", op.mf.mesh.cvx.shape[0], flush=True)
        ref = _call(lib().gfx_compliance_eval, op, rho, X, u)
        got = _call(lib().gfx_compliance_eval_noalloc, op, rho, X, u)
        _same(got, ref)
    # only some outputs requested (dJ/dX omitted): the scratch layout shifts
    keep, op, rho, X, u = small
    ref = _call(lib().gfx_compliance_eval, op, rho, X, u, with_x=False)
    got = _call(lib().gfx_compliance_eval_noalloc, op, rho, X, u, with_x=False)
    _same(got, ref)


def test_noalloc_rejects_meshes_beyond_the_scratch():
    import torch
    torch.cuda.init()
    ctx = ga.Context(0)
    keep, op, rho, X, u = _problem(ctx, 40, 3)  # 384,000 tets: 73.7 MB of scratch with both gradients
    with pytest.raises(GfxError, match="static scratch holds"):
        _call(lib().gfx_compliance_eval_noalloc, op, rho, X, u)
