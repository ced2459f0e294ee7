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

from ._lib import FORM_FACTORIZED, PREC_FP64, check, lib
from .core import MeshFem


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def gwfl_compile(exprs, decls):
    """Run the compile stage alone (no device): returns the normal form.
    decls: {name: ("var"|"const"|"p0", n)}."""
    d = ";".join("%s:%s:%d" % (k, v[0], v[1]) for k, v in decls.items())
    e = "\n".join([exprs] if isinstance(exprs, str) else exprs)
    buf = ctypes.create_string_buffer(1024)
    code = lib().gfx_gwfl_compile(e.encode(), d.encode(), buf, len(buf))
    if code:
        from ._lib import GfxError
        raise GfxError(code, buf.value.decode())
    return buf.value.decode()


class GaWorkspace:
    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().gfx_ws_create(ctx.h, ctypes.byref(h)), ctx.h)
        self.h = h
        self._keep = []

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
        """The order-1 result (nb_dof values of the unknown's mesh_fem)."""
        p = self._plan_handle()
        n, z = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_plan_dims(p, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        V = np.zeros(n.value)
        check(lib().gfx_ws_assembled_vector(self.h, _ptr(V)), self.ctx.h)
        return V

    def _plan_handle(self):
        h = ctypes.c_void_p()
        check(lib().gfx_ws_plan(self.h, ctypes.byref(h)), self.ctx.h)
        return h

    def assembled_matrix_csr(self):
        """(indptr int64, indices int32, values float64), rows = Test_u."""
        p = self._plan_handle()
        n, z = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_plan_dims(p, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        ip = np.zeros(n.value + 1, dtype=np.int64)
        ix = np.zeros(z.value, dtype=np.int32)
        v = np.zeros(z.value)
        check(lib().gfx_plan_export_csr(p, _ptr(ip), _ptr(ix), _ptr(v)), self.ctx.h)
        return ip, ix, v

    def assembled_matrix(self):
        """scipy.sparse.csc_matrix in GetFEM's orientation (columns = trial dofs)."""
        import scipy.sparse as sp
        p = self._plan_handle()
        n, z = ctypes.c_int64(), ctypes.c_int64()
        check(lib().gfx_plan_dims(p, ctypes.byref(n), ctypes.byref(z)), self.ctx.h)
        cp = np.zeros(n.value + 1, dtype=np.int64)
        r = np.zeros(z.value, dtype=np.uint32)
        v = np.zeros(z.value)
        check(lib().gfx_plan_export_csc(p, _ptr(cp), _ptr(r), _ptr(v)), self.ctx.h)
        return sp.csc_matrix((v, r, cp), shape=(n.value, n.value))

    def __del__(self):
        if getattr(self, "h", None) and not sys.is_finalizing():
            lib().gfx_ws_destroy(self.h)
        self.h = None


def asm_generic(mim, order, expression, region, *args, ctx=None):
    """gf.asm('generic', mim, order, expression, region, [name, is_variable,
    [mf,] value]...).  region must be -1 (all elements)."""
    if region not in (-1, None):
        raise ValueError("only region -1 (all elements) is supported")
    ctx = ctx or mim.mesh.ctx
    ws = GaWorkspace(ctx)
    i = 0
    while i < len(args):
        name, is_var = args[i], args[i + 1]
        if i + 2 < len(args) and isinstance(args[i + 2], MeshFem):
            mf, val = args[i + 2], args[i + 3]
            if is_var:
                ws.add_fem_variable(name, mf, val)
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
    if order == 2:
        return ws.assembled_matrix()
    if order == 1:
        return ws.assembled_vector()
    raise ValueError("only orders 1 and 2 are supported")
