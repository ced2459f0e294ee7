"""Python mirror of GetFEM's generic assembly entry points over libgfx.

``GaWorkspace`` follows getfem::ga_workspace (add_fem_variable,
add_fem_constant, add_fixed_size_constant, add_expression, assembly) and
``asm_generic`` follows the Python interface's ``gf.asm('generic', mim, order,
expression, region, name, is_variable, mf, value, ...)``
(interface/src/gf_asm.cc:421-548, :762-821): order-2 results are returned as
scipy CSC matrices, like the interface's gsparse export
(interface/src/getfemint.cc:712-737), order-1 results as numpy vectors.
"""
import ctypes
import sys

import numpy as np

from ._lib import CSR, FORM_FACTORIZED, PREC_FP64, check, lib
from .core import MeshFem


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def gwfl_compile(exprs, decls):
    """Run the compile stage alone (no device): returns the normal form.
    decls: {name: ("var"|"const"|"p0"|"p0var"|"data", n)}."""
    d = ";".join("%s:%s:%d" % (k, v[0], v[1]) for k, v in decls.items())
    e = "\n".join([exprs] if isinstance(exprs, str) else exprs)
    buf = ctypes.create_string_buffer(1024)
    code = lib().gfx_gwfl_compile(e.encode(), d.encode(), buf, len(buf))
    if code:
        from ._lib import GfxError
        raise GfxError(code, buf.value.decode())
    return buf.value.decode()


def gwfl_order(expr, decls):
    """Test-function order of one expression (0, 1 or 2), no device."""
    d = ";".join("%s:%s:%d" % (k, v[0], v[1]) for k, v in decls.items())
    o = ctypes.c_int()
    check(lib().gfx_gwfl_order(expr.encode(), d.encode(), ctypes.byref(o)))
    return o.value


class GaWorkspace:
    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().gfx_ws_create(ctx.h, ctypes.byref(h)), ctx.h)
        self.h = h
        self._keep = []
        self._ext = None

    def add_fem_variable(self, name, mf, V=None, first=0):
        V = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        check(lib().gfx_ws_add_fem_variable(self.h, name.encode(), mf.h, first, _ptr(V)), self.ctx.h)
        self._keep.append(mf)

    def add_fem_constant(self, name, mf, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        check(lib().gfx_ws_add_fem_constant(self.h, name.encode(), mf.h, _ptr(V)), self.ctx.h)
        self._keep.append(mf)

    def add_fixed_size_constant(self, name, V):
        V = np.atleast_1d(np.ascontiguousarray(V, dtype=np.float64))
        check(lib().gfx_ws_add_fixed_size_constant(self.h, name.encode(), _ptr(V), V.size), self.ctx.h)

    def set_values(self, name, V):
        V = np.ascontiguousarray(np.atleast_1d(V), dtype=np.float64)
        check(lib().gfx_ws_set_values(self.h, name.encode(), _ptr(V)), self.ctx.h)

    def add_expression(self, expr, mim, der_order=2):
        check(lib().gfx_ws_add_expression(self.h, expr.encode(), mim.h, der_order), self.ctx.h)
        self._keep.append(mim)

    def set_form(self, form=FORM_FACTORIZED, precision=PREC_FP64):
        check(lib().gfx_ws_set_form(self.h, form, precision), self.ctx.h)

    def compile(self):
        check(lib().gfx_ws_compile(self.h), self.ctx.h)
        return self.describe()

    def describe(self):
        buf = ctypes.create_string_buffer(1024)
        check(lib().gfx_ws_describe(self.h, buf, len(buf)), self.ctx.h)
        return buf.value.decode()

    def assembly(self, order=2):
        check(lib().gfx_ws_assembly(self.h, order), self.ctx.h)

    def assembled_vector(self):
        """The order-1 result: the unknown's nb_dof values on the matrix
        kernels' route, every fem variable's block at its interval on the
        order-0/1 route."""
        n = ctypes.c_int64()
        check(lib().gfx_ws_vector_size(self.h, ctypes.byref(n)), self.ctx.h)
        V = np.zeros(n.value)
        check(lib().gfx_ws_assembled_vector(self.h, _ptr(V)), self.ctx.h)
        return V

    def assembled_potential(self):
        """The order-0 result (ga_workspace::assembled_potential)."""
        J = ctypes.c_double()
        check(lib().gfx_ws_assembled_potential(self.h, ctypes.byref(J)), self.ctx.h)
        return J.value

    def set_values_device(self, name, d_V):
        """New values of the fem variable from a device vector (raw address,
        e.g. a torch tensor's data_ptr()); stream-ordered, no host copy."""
        check(lib().gfx_ws_set_values_device(self.h, name.encode(), ctypes.c_void_p(d_V)), self.ctx.h)

    def set_assembled_matrix(self, indptr=None, indices=None, values=None, orientation=CSR):
        """ga_workspace::set_assembled_matrix: assembly(2) then adds into this
        caller-owned device matrix (torch tensors; CSR rows = test dofs, or CSC
        in GetFEM's orientation) without clearing it.  No arguments: back to
        the workspace's own matrix."""
        if indptr is None:
            check(lib().gfx_ws_set_assembled_matrix(self.h, orientation, 0, None, None, None), self.ctx.h)
            self._ext = None
            return
        self._ext = (indptr, indices, values)  # keep the caller's storage alive
        check(lib().gfx_ws_set_assembled_matrix(self.h, orientation, indptr.numel() - 1,
                                                ctypes.c_void_p(indptr.data_ptr()),
                                                ctypes.c_void_p(indices.data_ptr()),
                                                ctypes.c_void_p(values.data_ptr())), self.ctx.h)

    def nb_trees(self):
        return int(lib().gfx_ws_nb_trees(self.h))

    def _plan_handle(self):
        h = ctypes.c_void_p()
        check(lib().gfx_ws_plan(self.h, ctypes.byref(h)), self.ctx.h)
        return h

    def _nbdof(self):
        n, z = ctypes.c_int64(), ctypes.c_int64()
        code = lib().gfx_ws_matrix_dims(self.h, ctypes.byref(n), ctypes.byref(z))
        if code == 0:
            return n.value
        p = self._plan_handle()
        check(lib().gfx_plan_dims(p, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        return n.value

    def _dims(self):
        n, z = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_ws_matrix_dims(self.h, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        return n.value, z.value

    def assembled_matrix_csr(self):
        """(indptr int64, indices int32, values float64), rows = Test_u: the
        sum of the workspace's trees."""
        n, z = self._dims()
        ip = np.zeros(n + 1, dtype=np.int64)
        ix = np.zeros(z, dtype=np.int32)
        v = np.zeros(z)
        check(lib().gfx_ws_export_csr(self.h, _ptr(ip), _ptr(ix), _ptr(v)), self.ctx.h)
        return ip, ix, v

    def assembled_matrix(self):
        """scipy.sparse.csc_matrix in GetFEM's orientation (columns = trial dofs)."""
        import scipy.sparse as sp
        n, z = self._dims()
        cp = np.zeros(n + 1, dtype=np.int64)
        r = np.zeros(z, dtype=np.uint32)
        v = np.zeros(z)
        check(lib().gfx_ws_export_csc(self.h, _ptr(cp), _ptr(r), _ptr(v)), self.ctx.h)
        return sp.csc_matrix((v, r, cp), shape=(n, n))

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_ws_destroy(self.h)
        self.h = None


def asm_generic(mim, order, expression, region, *args, ctx=None, select_output=None):
    """gf.asm('generic', mim, order, expression, region, [name, is_variable,
    [mf,] value]..., ['select_output', var1[, var2]]) (interface/src/gf_asm.cc:
    421-548): the variables take consecutive intervals in argument order
    (:486-488); select_output=(var1,) or (var1, var2) returns the part of
    the vector or matrix at those variables' intervals (:519-541).  region
    must be -1 (all elements)."""
    if region not in (-1, None):
        raise ValueError("only region -1 (all elements) is supported")
    ctx = ctx or mim.mesh.ctx
    ws = GaWorkspace(ctx)
    intervals = {}
    nbdof = 0
    i = 0
    while i < len(args):
        name, is_var = args[i], args[i + 1]
        if i + 2 < len(args) and isinstance(args[i + 2], MeshFem):
            mf, val = args[i + 2], args[i + 3]
            if is_var:
                ws.add_fem_variable(name, mf, val, first=nbdof)
                intervals[name] = (nbdof, nbdof + mf.nbdof())
                nbdof += mf.nbdof()
            else:
                ws.add_fem_constant(name, mf, val)
            i += 4
        else:
            if is_var:
                raise ValueError("a variable needs a mesh_fem")
            ws.add_fixed_size_constant(name, args[i + 2])
            i += 3
    ws.add_expression(expression, mim, 2)
    ws.assembly(order)
    sel = [intervals[v] for v in (select_output or ())]
    if order == 2:
        K = ws.assembled_matrix()
        if sel:
            (a, b), (c, d) = sel[0], sel[-1]
            K = K[a:b, c:d]
        return K
    if order == 1:
        V = ws.assembled_vector()
        return V[sel[0][0]:sel[0][1]] if sel else V
    if order == 0:
        return ws.assembled_potential()
    raise ValueError("order must be 0, 1 or 2")
