"""torch surface of libgfx: the SIMP compliance adjoint as a
torch.autograd.Function.

J(rho, X, u) = u^T K(rho, X) u for isotropic elasticity on a P1 vector
mesh_fem with E_e = emin + rho_e^penal (1 - emin).  One libgfx call computes
J and the three gradients on the device (stream = torch's current stream);
backward scales them by the incoming gradient.  GetFEM's route to the same
derivatives is the order-1 assembly of the order-0 potential with rho and the
pulled-back coordinates as variables (src/getfem_generic_assembly_workspace.cc:
520-541, :832-847).
"""
import ctypes
import sys

import torch

from ._lib import check, lib


class Compliance:
    """gfx_compliance bound to a MeshFem (FEM_PK(3,1), qdim 3) and MeshIm."""

    def __init__(self, mf, mim, lam, mu, emin=1e-3, penal=3.0):
        self.mf, self.mim = mf, mim
        self.ctx = mf.mesh.ctx
        h = ctypes.c_void_p()
        check(lib().gfx_compliance_create(mf.h, mim.h, float(lam), float(mu), float(emin),
                                          float(penal), ctypes.byref(h)), self.ctx.h)
        self.h = h

    def eval(self, rho, X, u, want_grads=True, out=None):
        """All tensors float64 on the device; X may be None (mesh points).
        out: optional (J, drho, dX, du) device tensors to write into."""
        dev = rho.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        if stream != (self.ctx.stream or 0):  # the context may be shared: compare with its stream
            self.ctx.set_stream(stream)
        rho = rho.contiguous()
        u = u.contiguous()
        Xc = X.contiguous() if X is not None else None
        if out is not None:
            J, dr, dx, du = out
        else:
            J = torch.empty((), dtype=torch.float64, device=dev)
            dr = torch.empty_like(rho) if want_grads else None
            n_pts = self.mf.mesh.pts.shape[0]
            dx = torch.empty((n_pts, 3), dtype=torch.float64, device=dev) if want_grads else None
            du = torch.empty_like(u) if want_grads else None
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        check(lib().gfx_compliance_eval(self.h, p(Xc), p(rho), p(u), p(J), p(dr), p(dx), p(du)),
              self.ctx.h)
        return J, dr, dx, du

    def check(self):
        """Raise GfxError (GFX_E_DEGENERATE_ELEMENT) if an evaluation since the
        last check met an element with det K = 0; synchronises the stream."""
        check(lib().gfx_compliance_status(self.h), self.ctx.h)

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_compliance_destroy(self.h)
        self.h = None


class ComplianceFunction(torch.autograd.Function):
    """J = ComplianceFunction.apply(rho, X, u, op)."""

    @staticmethod
    def forward(ctx, rho, X, u, op):
        J, dr, dx, du = op.eval(rho, X, u, want_grads=True)
        op.check()  # a degenerate element raises, as GetFEM's assembly does
        ctx.save_for_backward(dr, dx, du)
        return J

    @staticmethod
    def backward(ctx, gJ):
        dr, dx, du = ctx.saved_tensors
        need = ctx.needs_input_grad  # X may be None (the mesh points): no gradient for it
        return (gJ * dr if need[0] else None, gJ * dx if need[1] else None,
                gJ * du if need[2] else None, None)


def compliance(rho, X, u, op):
    return ComplianceFunction.apply(rho, X, u, op)
