"""The autograd surface of the compliance adjoint, on the CPU with a stand-in
for the device operator (the gradients are the operator's; what is tested is
how backward hands them to autograd): X may be None (the mesh's own points),
and then no gradient is returned for it."""
import torch

from getfem_amd.torch_ops import compliance


class FakeOp:
    """J = sum(rho) + sum(X) + sum(u^2) with its exact gradients."""

    def __init__(self, n_pts):
        self.n_pts = n_pts
        self.checked = 0

    def eval(self, rho, X, u, want_grads=True, out=None):
        Xv = X if X is not None else torch.zeros(self.n_pts, 3, dtype=torch.float64)
        J = rho.sum() + Xv.sum() + (u * u).sum()
        return J.detach(), torch.ones_like(rho), torch.ones_like(Xv), 2 * u.detach()

    def check(self):
        self.checked += 1


def test_backward_without_coordinates():
    rho = torch.rand(5, dtype=torch.float64, requires_grad=True)
    u = torch.rand(7, dtype=torch.float64, requires_grad=True)
    op = FakeOp(4)
    J = compliance(rho, None, u, op)
    J.backward()
    assert torch.equal(rho.grad, torch.ones(5, dtype=torch.float64))
    assert torch.allclose(u.grad, 2 * u.detach())
    assert op.checked == 1


def test_backward_only_requested_inputs():
    rho = torch.rand(5, dtype=torch.float64, requires_grad=True)
    X = torch.rand(4, 3, dtype=torch.float64, requires_grad=True)
    u = torch.rand(7, dtype=torch.float64)  # no gradient wanted
    J = compliance(rho, X, u, FakeOp(4))
    J.backward()
    assert torch.equal(X.grad, torch.ones(4, 3, dtype=torch.float64))
    assert u.grad is None
