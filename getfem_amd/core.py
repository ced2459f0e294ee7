"""Python objects over the libgfx C ABI, named after their GetFEM counterparts
(getfem.Mesh, getfem.MeshFem, getfem.MeshIm) plus the assembly Plan."""
import ctypes
import sys
import weakref

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Context:
    """One device + one HIP stream (gfx_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        code = lib().gfx_ctx_create(device, ctypes.byref(h))
        if code:
            raise _lib.GfxError(code, "cannot open HIP device %d" % device)
        self.h = h

    def set_stream(self, stream_handle):
        check(lib().gfx_ctx_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)), self.h)

    @property
    def stream(self):
        return lib().gfx_ctx_stream(self.h)

    def synchronize(self):
        check(lib().gfx_ctx_synchronize(self.h), self.h)

    def close(self):
        """Release the context; objects created from it stay valid (the
        library frees the context after the last of them)."""
        if self.h:
            lib().gfx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """getfem::mesh restricted to one geometric transformation."""

    def __init__(self, ctx, geotrans, pts, cvx):
        self.ctx = ctx
        self.pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.cvx = np.ascontiguousarray(cvx, dtype=np.int32)
        self.geotrans = geotrans
        h = ctypes.c_void_p()
        check(lib().gfx_mesh_create(ctx.h, geotrans.encode(), self.pts.shape[0], _ptr(self.pts),
                                    self.cvx.shape[0], _ptr(self.cvx), ctypes.byref(h)), ctx.h)
        self.h = h

    @property
    def nbcvx(self):
        return self.cvx.shape[0]

    def set_points(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        check(lib().gfx_mesh_set_points(self.h, _ptr(pts)), self.ctx.h)
        self.pts = pts

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_mesh_destroy(self.h)
        self.h = None


class MeshFem:
    """getfem::mesh_fem with one Lagrange element and Qdim."""

    def __init__(self, mesh, fem_name, qdim=1, basic_dofs=None, ndof=None):
        self.mesh = mesh
        self.fem_name = fem_name
        self.qdim = qdim
        h = ctypes.c_void_p()
        if basic_dofs is None:
            check(lib().gfx_mesh_fem_create(mesh.h, fem_name.encode(), qdim, ctypes.byref(h)),
                  mesh.ctx.h)
        else:
            bd = np.ascontiguousarray(basic_dofs, dtype=np.int32)
            check(lib().gfx_mesh_fem_create_from_dofs(mesh.h, fem_name.encode(), qdim, _ptr(bd),
                                                      int(ndof), ctypes.byref(h)), mesh.ctx.h)
        self.h = h

    def nbdof(self):
        return int(lib().gfx_mesh_fem_nb_dof(self.h))

    def nb_dof_per_element(self):
        return int(lib().gfx_mesh_fem_nb_dof_per_element(self.h))

    def basic_dof_table(self):
        out = np.zeros((self.mesh.nbcvx, self.nb_dof_per_element()), dtype=np.int32)
        check(lib().gfx_mesh_fem_basic_dofs(self.h, _ptr(out)), self.mesh.ctx.h)
        return out

    def basic_dof_nodes(self):
        out = np.zeros((self.nbdof() // self.qdim, 3))
        check(lib().gfx_mesh_fem_dof_nodes(self.h, _ptr(out)), self.mesh.ctx.h)
        return out

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_mesh_fem_destroy(self.h)
        self.h = None


class MeshIm:
    """getfem::mesh_im with one integration method."""

    def __init__(self, mesh, im_name):
        self.mesh = mesh
        self.im_name = im_name
        h = ctypes.c_void_p()
        check(lib().gfx_mesh_im_create(mesh.h, im_name.encode(), ctypes.byref(h)), mesh.ctx.h)
        self.h = h

    def points(self):
        n = lib().gfx_mesh_im_nb_points(self.h)
        p = np.zeros((n, 3))
        w = np.zeros(n)
        check(lib().gfx_mesh_im_points(self.h, _ptr(p), _ptr(w)), self.mesh.ctx.h)
        return p, w

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_mesh_im_destroy(self.h)
        self.h = None


def _coef(v, n_cv):
    if np.isscalar(v):
        return _lib.GfxCoef(_lib.COEF_CONST, float(v), None), None
    a = np.ascontiguousarray(v, dtype=np.float64)
    if a.shape != (n_cv,):
        raise ValueError("per-element coefficient must have one value per element")
    return _lib.GfxCoef(_lib.COEF_PER_ELEMENT, 0.0, ctypes.c_void_p(a.ctypes.data)), a


class _DeviceBuffer:
    """A plan-owned device array exported through __cuda_array_interface__
    (version 2: the consumer orders its work after the context's stream).
    Holding the plan keeps the buffer allocated for as long as a tensor made
    from it lives."""

    def __init__(self, plan, ptr, n, typestr):
        self.plan = plan
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "strides": None, "version": 2}


class Plan:
    """One compiled order-2 term (gfx_plan): pattern + device CSR values."""

    def __init__(self, mf, mim, term, coefs):
        self.mf, self.mim, self.term = mf, mim, term
        self.ctx = mf.mesh.ctx
        n_cv = mf.mesh.nbcvx
        arr = (_lib.GfxCoef * len(coefs))()
        keep = []
        for i, c in enumerate(coefs):
            arr[i], a = _coef(c, n_cv)
            keep.append(a)
        h = ctypes.c_void_p()
        check(lib().gfx_plan_create(mf.h, mim.h, term, arr, len(coefs), ctypes.byref(h)), self.ctx.h)
        self.h = h
        self.nnz = None
        self._views = weakref.WeakSet()

    def set_coef(self, slot, value):
        c, a = _coef(value, self.mf.mesh.nbcvx)
        check(lib().gfx_plan_set_coef(self.h, slot, ctypes.byref(c)), self.ctx.h)

    def set_form(self, form=_lib.FORM_FACTORIZED, precision=_lib.PREC_FP64):
        """Contraction form and precision of the next assemblies: FORM_FACTORIZED
        (reference tensor, affine elements) or FORM_PERQP (per integration
        point); PREC_FP64, PREC_FP32 (FP32 element contraction, FP64
        accumulation) or PREC_BF16X2 (isotropic elasticity on P2/P3 vector
        fields only: element matrices B^T D B on the matrix cores with a bf16
        hi + lo split, gathered into the same CSR; the form is ignored).  An
        unsupported combination raises GfxError."""
        check(lib().gfx_plan_set_form(self.h, form, precision), self.ctx.h)

    def build_pattern(self):
        if len(self._views):
            raise RuntimeError("torch views of this plan's CSR are alive: a new pattern "
                               "reallocates the arrays they point to")
        self._sync_pin()
        nnz = ctypes.c_int64()
        check(lib().gfx_plan_build_pattern(self.h, ctypes.byref(nnz)), self.ctx.h)
        self.nnz = nnz.value
        return self.nnz

    def assemble(self):
        """Order-2 assembly.  After new coefficients, moved nodes or a new
        state the thresholds are re-derived and the pattern is rebuilt if
        GetFEM's would differ; while torch views of the CSR are alive the
        pattern is pinned and such a change raises GfxError instead."""
        self._sync_pin()
        check(lib().gfx_plan_assemble(self.h), self.ctx.h)

    def _sync_pin(self):
        check(lib().gfx_plan_pin_pattern(self.h, 1 if len(self._views) else 0), self.ctx.h)

    def pattern_id(self):
        """Counter of pattern (re)builds: device pointers taken from
        device_csr() are valid while it does not change."""
        i = ctypes.c_int64()
        check(lib().gfx_plan_pattern_id(self.h, ctypes.byref(i)), self.ctx.h)
        return i.value

    def set_state_device(self, d_U):
        """State of a nonlinear plan from a device vector (raw address)."""
        check(lib().gfx_plan_set_state_device(self.h, ctypes.c_void_p(d_U)), self.ctx.h)

    def add_to_matrix(self, indptr, indices, values, orientation=_lib.CSR):
        """Add the assembled values into a caller-owned device matrix given
        by torch tensors (CSR rows = test dofs, or CSC in GetFEM's orientation)
        whose pattern holds this plan's."""
        n = indptr.numel() - 1
        check(lib().gfx_plan_add_to_matrix(self.h, orientation, n, ctypes.c_void_p(indptr.data_ptr()),
                                           ctypes.c_void_p(indices.data_ptr()),
                                           ctypes.c_void_p(values.data_ptr())), self.ctx.h)

    def set_state(self, U):
        """Current displacement of a nonlinear (Neo-Hookean) plan."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().gfx_plan_set_state(self.h, _ptr(U)), self.ctx.h)

    def assemble_residual(self):
        """Order-1 assembly alone (nonlinear plans, after build_pattern)."""
        check(lib().gfx_plan_assemble_residual(self.h), self.ctx.h)

    def spmv(self, d_x, d_y):
        """d_y = K d_x on the device (raw device addresses, e.g. data_ptr())."""
        check(lib().gfx_plan_spmv(self.h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y)), self.ctx.h)

    def copy_residual_device(self, d_V):
        """The assembled residual into a device buffer (raw address)."""
        check(lib().gfx_plan_copy_residual_device(self.h, ctypes.c_void_p(d_V)), self.ctx.h)

    def dof_constraints(self, dofs, values, d_pr=None, d_rhs=None, flags=0):
        """Dirichlet condition by simplification on the device CSR (and on the
        device right-hand side d_rhs); flags from BUILD_MATRIX, BUILD_RHS,
        MODEL_LINEAR, MODEL_SYMMETRIC."""
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        values = np.ascontiguousarray(np.broadcast_to(values, dofs.shape), dtype=np.float64)
        v = lambda x: ctypes.c_void_p(x) if x else None  # noqa: E731
        check(lib().gfx_plan_dof_constraints(self.h, dofs.size, _ptr(dofs), _ptr(values), v(d_pr),
                                             v(d_rhs), flags), self.ctx.h)

    def solve_cg(self, d_b, d_x, rtol=1e-10, maxit=10000):
        """Block-Jacobi-preconditioned CG on the device (3x3 node blocks for a
        vector field, the diagonal otherwise): d_x (initial guess in,
        solution out) and d_b are raw device addresses.  Returns (iterations,
        relative residual)."""
        it, res = ctypes.c_int(), ctypes.c_double()
        check(lib().gfx_plan_solve_cg(self.h, ctypes.c_void_p(d_b), ctypes.c_void_p(d_x), rtol, maxit,
                                      ctypes.byref(it), ctypes.byref(res)), self.ctx.h)
        return it.value, res.value

    def solve_pcg(self, d_b, d_x, precond=_lib.PRECOND_ILDLT, rtol=1e-10, maxit=10000):
        """CG on the device with PRECOND_ILDLT (gmm::ildlt_precond, rebuilt
        from the current values) or PRECOND_BLOCK_JACOBI.  Returns
        (iterations, relative residual)."""
        it, res = ctypes.c_int(), ctypes.c_double()
        check(lib().gfx_plan_solve_pcg(self.h, precond, ctypes.c_void_p(d_b), ctypes.c_void_p(d_x), rtol, maxit,
                                       ctypes.byref(it), ctypes.byref(res)), self.ctx.h)
        return it.value, res.value

    def ildlt_factor(self):
        """Factor the current values; returns the (forward, backward) level counts."""
        f, b = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_plan_ildlt_factor(self.h, ctypes.byref(f), ctypes.byref(b)), self.ctx.h)
        return f.value, b.value

    def ildlt_values(self):
        """The factor aligned with the CSR: D on the diagonal, U_ij / D_i at
        (i, j) and (j, i)."""
        F = np.zeros(self.dims()[1])
        check(lib().gfx_plan_ildlt_values(self.h, _ptr(F)), self.ctx.h)
        return F

    def ildlt_apply(self, d_r, d_z):
        check(lib().gfx_plan_ildlt_apply(self.h, ctypes.c_void_p(d_r), ctypes.c_void_p(d_z)), self.ctx.h)

    def residual(self):
        n, _ = self.dims()
        V = np.zeros(n)
        check(lib().gfx_plan_residual(self.h, _ptr(V)), self.ctx.h)
        return V

    def dims(self):
        n, z = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_plan_dims(self.h, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        return n.value, z.value

    def csr(self, values=True):
        n, nnz = self.dims()
        indptr = np.zeros(n + 1, dtype=np.int64)
        indices = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz) if values else None
        check(lib().gfx_plan_export_csr(self.h, _ptr(indptr), _ptr(indices), _ptr(vals)), self.ctx.h)
        return indptr, indices, vals

    def csc(self):
        n, nnz = self.dims()
        colptr = np.zeros(n + 1, dtype=np.int64)
        rows = np.zeros(nnz, dtype=np.uint32)
        vals = np.zeros(nnz)
        check(lib().gfx_plan_export_csc(self.h, _ptr(colptr), _ptr(rows), _ptr(vals)), self.ctx.h)
        return colptr, rows, vals

    def device_csr(self):
        ip, ix, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().gfx_plan_device_csr(self.h, ctypes.byref(ip), ctypes.byref(ix), ctypes.byref(v)),
              self.ctx.h)
        return ip.value, ix.value, v.value

    def copy_csr_device(self, d_indptr=None, d_indices=None, d_vals=None):
        """Device-to-device copy into caller buffers (raw device addresses,
        e.g. torch tensors' data_ptr())."""
        v = lambda x: ctypes.c_void_p(x) if x else None  # noqa: E731
        check(lib().gfx_plan_copy_csr_device(self.h, v(d_indptr), v(d_indices), v(d_vals)),
              self.ctx.h)

    def torch_csr(self, device="cuda"):
        """The CSR arrays as torch tensors (copied on the device)."""
        import torch
        n, nnz = self.dims()
        ip = torch.empty(n + 1, dtype=torch.int64, device=device)
        ix = torch.empty(nnz, dtype=torch.int32, device=device)
        vv = torch.empty(nnz, dtype=torch.float64, device=device)
        self.copy_csr_device(ip.data_ptr(), ix.data_ptr(), vv.data_ptr())
        self.ctx.synchronize()
        return ip, ix, vv

    def torch_csr_view(self):
        """The plan's own device CSR arrays as torch tensors, without a copy
        (indptr int64, indices int32, values float64).  The values follow every
        later assemble(); the tensors keep the plan alive, and build_pattern()
        refuses to run while any of them does."""
        import torch
        n, nnz = self.dims()
        ip, ix, vv = self.device_csr()
        out = []
        for ptr, cnt, ts in ((ip, n + 1, "<i8"), (ix, nnz, "<i4"), (vv, nnz, "<f8")):
            buf = _DeviceBuffer(self, ptr, cnt, ts)
            self._views.add(buf)
            out.append(torch.as_tensor(buf, device="cuda"))
        self._sync_pin()
        return tuple(out)

    def set_timing(self, enable=True):
        check(lib().gfx_plan_set_timing(self.h, 1 if enable else 0), self.ctx.h)

    def timing_summary(self):
        """(calls, summed geometry ms, summed contraction+scatter ms)."""
        n, g, a = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().gfx_plan_timing_summary(self.h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(a)),
              self.ctx.h)
        return n.value, g.value, a.value

    def timing_calls(self, max_calls=100000):
        """Per-call (geometry ms, contraction + scatter ms) arrays of the
        recorded calls; resets the record."""
        g = np.zeros(max_calls, dtype=np.float32)
        a = np.zeros(max_calls, dtype=np.float32)
        n = ctypes.c_int()
        check(lib().gfx_plan_timing_calls(self.h, max_calls, ctypes.byref(n), _ptr(g), _ptr(a)), self.ctx.h)
        return g[:n.value].astype(np.float64), a[:n.value].astype(np.float64)

    def last_timings(self):
        g, a = ctypes.c_float(), ctypes.c_float()
        check(lib().gfx_plan_last_timings(self.h, ctypes.byref(g), ctypes.byref(a)), self.ctx.h)
        return g.value, a.value

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_plan_destroy(self.h)
        self.h = None


# ---------------------------------------------------------------- host-only
def integration_rule(im_name):
    """Points (n,3) and weights (n,) of an integration method."""
    n = ctypes.c_int()
    check(lib().gfx_integration_rule(im_name.encode(), ctypes.byref(n), None, None))
    p = np.zeros((n.value, 3))
    w = np.zeros(n.value)
    check(lib().gfx_integration_rule(im_name.encode(), ctypes.byref(n), _ptr(p), _ptr(w)))
    return p, w


def fem_eval(fem_name, x):
    """Base values (nd,) and reference gradients (nd,3) at x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nd = ctypes.c_int()
    check(lib().gfx_fem_eval(fem_name.encode(), _ptr(x), ctypes.byref(nd), None, None))
    v = np.zeros(nd.value)
    g = np.zeros((nd.value, 3))
    check(lib().gfx_fem_eval(fem_name.encode(), _ptr(x), ctypes.byref(nd), _ptr(v), _ptr(g)))
    return v, g


def enumerate_dofs(geotrans, fem_name, qdim, pts, cvx):
    """First-encounter dof numbering: (basic dof table (n_cv, nd), ndof)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cvx = np.ascontiguousarray(cvx, dtype=np.int32)
    _, g = fem_eval(fem_name, np.array([0.1, 0.1, 0.1]))
    out = np.zeros((cvx.shape[0], g.shape[0]), dtype=np.int32)
    ndof = ctypes.c_int64()
    check(lib().gfx_enumerate_dofs(geotrans.encode(), fem_name.encode(), qdim, pts.shape[0],
                                   _ptr(pts), cvx.shape[0], _ptr(cvx), _ptr(out),
                                   ctypes.byref(ndof)))
    return out, ndof.value
