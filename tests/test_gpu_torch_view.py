"""Zero-copy torch views of a plan's device CSR (__cuda_array_interface__):
equal to the copied arrays, sharing the plan's device memory, following later
assemblies, and guarding the pattern rebuild that would reallocate them."""
import gc

import numpy as np
import pytest

import getfem_amd as ga
from tests import problems

pytestmark = pytest.mark.gpu


def make_plan(ctx):
    case = problems.c2(4)
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, case.fem, case.qdim)
    mim = ga.MeshIm(mesh, case.im)
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, case.coefs())
    plan.build_pattern()
    plan.assemble()
    return plan, case, (mesh, mf, mim)


def test_torch_csr_view_is_zero_copy_and_follows_assembly(ctx):
    import torch
    plan, case, keep = make_plan(ctx)
    ip, ix, v = plan.torch_csr_view()
    ipc, ixc, vc = plan.torch_csr()
    assert ip.dtype == torch.int64 and ix.dtype == torch.int32 and v.dtype == torch.float64
    assert torch.equal(ip, ipc) and torch.equal(ix, ixc) and torch.equal(v, vc)
    assert (ip.data_ptr(), ix.data_ptr(), v.data_ptr()) == plan.device_csr()
    # new coefficients, same pattern: the view shows the new values
    lam, mu = case.coefs()
    plan.set_coef(0, 3.0 * np.asarray(lam))
    plan.set_coef(1, 3.0 * np.asarray(mu))
    plan.assemble()
    ctx.synchronize()
    assert problems.rel_frob(v.cpu().numpy(), 3.0 * vc.cpu().numpy()) <= 1e-13
    with pytest.raises(RuntimeError):
        plan.build_pattern()
    del ip, ix, v
    gc.collect()
    assert plan.build_pattern() == ipc[-1].item()


def test_torch_view_keeps_the_plan_alive(ctx):
    import torch
    plan, case, keep = make_plan(ctx)
    ref = plan.torch_csr()[2].clone()
    v = plan.torch_csr_view()[2]
    del plan
    gc.collect()
    torch.cuda.synchronize()
    assert torch.equal(v, ref)
