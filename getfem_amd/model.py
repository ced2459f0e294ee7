"""A model over the device assembly, after getfem::model (src/getfem_models.cc)
and the Python interface's gf.Model('real'): fem variables, generic
assembly terms and Dirichlet conditions by simplification.

Variables are laid out one after the other in the global vector in the
order of their names, as model::resize_global_system gives each its
interval from the name-ordered variable map (src/getfem_models.cc:305-316).
One variable whose terms the matrix kernels cover assembles and solves
through that tree's plan (3x3 block Jacobi CG for a vector field);
otherwise (several variables, coupling blocks, polynomial nonlinearities)
the workspace takes the polynomial route, the tangent is the matrix over
all intervals and every solve is a Newton iteration on the residual form
(a linear model converges in one).

What runs where: the tangent matrix and the residual are assembled on the
device by the workspace (ga_workspace inside model::assembly,
src/getfem_models.cc:2646-2736, one device call instead of one per OpenMP
partition); the dof constraints are applied to the device CSR
(src/getfem_models.cc:2806-2870); the linear systems are solved on the device
by preconditioned CG (coercive models) or restarted GMRES (the others); the
Newton loop follows classical_Newton
(src/getfem/getfem_model_solvers.h:548-605) with basic_newton_line_search
(:336-360).  The state vector stays on the device between iterations; the
host sees scalars only.  A Newton iteration's tangent and residual come from
one element pass (the workspace reuses the residual of the tangent assembly),
and the pattern is rebuilt only when the new state changes GetFEM's
threshold-filtered pattern.

Sign convention as model::assembly: rhs = -(order-1 assembly of the
nonlinear terms); a linear model's rhs holds only the constraint data.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .workspace import GaWorkspace


class Model:
    def __init__(self, ctx):
        self.ctx = ctx
        self.ws = GaWorkspace(ctx)
        self.var = None     # the first variable (single-variable models)
        self.mf = None
        self.vars = {}      # name -> (mesh_fem, first)
        self.nvec = 0
        self.data = {}
        self.terms = []  # (expr, mim, linear, symmetric, coercive)
        self.dirichlet = []  # (variable, dofs, dataname or None)
        self.U = None  # device state (torch)
        self._rhs = None
        self._built = False
        self._route = None  # the workspace's route, known after its first assembly

    # -------------------------------------------------------------- variables
    def add_fem_variable(self, name, mf):
        """model::add_fem_variable.  The intervals are re-laid out in name
        order; the variables' values move with them."""
        import torch
        if name in self.vars or name in self.data:
            raise ValueError("name already used: " + name)
        old = {k: self._slice(k).clone() for k in self.vars}
        self.vars[name] = (mf, 0)
        first = 0
        for k in sorted(self.vars):
            self.vars[k] = (self.vars[k][0], first)
            first += self.vars[k][0].nbdof()
        self.nvec = first
        if self.var is None:
            self.var, self.mf = name, mf
        self.U = torch.zeros(self.nvec, dtype=torch.float64, device="cuda")
        for k, v in old.items():
            self._slice(k).copy_(v)
        for k in sorted(self.vars):
            m, f = self.vars[k]
            self.ws.add_fem_variable(k, m, np.zeros(m.nbdof()), first=f)
        self._built = False
        self._route = None

    def _ws_route(self):
        """True when the workspace assembles through the polynomial route
        (several variables, or terms outside the matrix kernels): constraints
        and solves then work on the workspace's matrix, and the right-hand
        side is always the residual form.  Decided by the compile stage."""
        if len(self.vars) > 1:
            return True
        if self._route is None:
            h = ctypes.c_void_p()
            self._route = lib().gfx_ws_plan(self.ws.h, ctypes.byref(h)) != 0
        return self._route

    def _slice(self, name):
        mf, first = self.vars[name]
        return self.U[first:first + mf.nbdof()]

    def add_initialized_data(self, name, value):
        """model::add_initialized_fixed_size_data."""
        v = np.atleast_1d(np.asarray(value, dtype=np.float64))
        self.data[name] = v
        self.ws.add_fixed_size_constant(name, v)
        self._built = False

    def add_initialized_fem_data(self, name, mf, V):
        """model::add_initialized_fem_data (P0 data on the variable's mesh)."""
        V = np.asarray(V, dtype=np.float64)
        self.data[name] = V
        self.ws.add_fem_constant(name, mf, V)

    def set_variable(self, name, V):
        import torch
        if name in self.vars:
            self._slice(name).copy_(torch.as_tensor(np.asarray(V, dtype=np.float64)))
        else:
            V = np.atleast_1d(np.asarray(V, dtype=np.float64))
            self.data[name] = V
            self.ws.set_values(name, V)

    def variable(self, name):
        if name in self.vars:
            return self._slice(name).cpu().numpy()
        return self.data[name]

    # ------------------------------------------------------------------ bricks
    def add_linear_term(self, mim, expr, region=-1, is_symmetric=False, is_coercive=False):
        """getfem::add_linear_term (src/getfem_models.cc, generic linear
        assembly brick)."""
        self._add_term(mim, expr, region, True, is_symmetric, is_coercive)

    def add_nonlinear_term(self, mim, expr, region=-1, is_symmetric=False, is_coercive=False):
        """getfem::add_nonlinear_term (generic nonlinear assembly brick)."""
        self._add_term(mim, expr, region, False, is_symmetric, is_coercive)

    def _add_term(self, mim, expr, region, linear, sym, coercive):
        if region not in (-1, None):
            raise ValueError("only region -1 (all elements) is supported")
        self.terms.append((expr, mim, linear, sym, coercive))
        self.ws.add_expression(expr, mim, 2)
        self._route = None

    def add_Dirichlet_condition_with_simplification(self, varname, dofs, dataname=None):
        """getfem::add_Dirichlet_condition_with_simplification
        (src/getfem_models.cc:4864-4940, :5019-5030).  `dofs` are the
        constrained dofs (GetFEM takes a mesh region and uses
        mf.dof_on_region(rg)).  dataname: fixed-size data of qdim values
        (value of dof i = data[i % qdim]) or fem data on the same mesh_fem;
        None = homogeneous."""
        if varname not in self.vars:
            raise ValueError("unknown variable " + varname)
        self.dirichlet.append((varname, np.asarray(dofs, dtype=np.int64), dataname))

    def is_linear(self):
        return all(t[2] for t in self.terms)

    def is_symmetric(self):
        return all(t[3] for t in self.terms)

    def is_coercive(self):
        return all(t[4] for t in self.terms)

    # ---------------------------------------------------------------- assembly
    def _constraints(self):
        """(global dofs, values): each variable's dofs at its interval."""
        dofs, vals = [], []
        for var, d, name in self.dirichlet:
            mf, first = self.vars[var]
            q = mf.qdim
            if name is None:
                v = np.zeros(d.size)
            else:
                A = self.data[name]
                v = A[d % q] if A.size == q else A[d]
            dofs.append(d + first)
            vals.append(v)
        if not dofs:
            return np.zeros(0, np.int32), np.zeros(0)
        d = np.concatenate(dofs)
        v = np.concatenate(vals)
        order = np.argsort(d, kind="stable")  # a dof constrained twice keeps its last value
        d, v = d[order], v[order]
        last = np.r_[d[1:] != d[:-1], True]
        return d[last].astype(np.int32), v[last]

    def _plan(self):
        h = ctypes.c_void_p()
        check(lib().gfx_ws_plan(self.ws.h, ctypes.byref(h)), self.ctx.h)
        return h

    def approx_external_load(self):
        """model::approx_external_load: here the 1-norm of the Dirichlet data."""
        return float(np.abs(self._constraints()[1]).sum())

    def assembly(self, option="build_all"):
        """model::assembly(BUILD_ALL | BUILD_MATRIX | BUILD_RHS): tangent
        matrix into the workspace plan's device CSR, rhs into a device vector."""
        import torch
        build_matrix = option in ("build_all", "build_matrix")
        build_rhs = option in ("build_all", "build_rhs")
        # the state goes to the workspace device to device (no host copy);
        # torch's writes to U are complete before the context's stream reads it
        torch.cuda.current_stream().synchronize()
        for name in self.vars:
            self.ws.set_values_device(name, self._slice(name).data_ptr())
        n = self.nvec
        rhs = torch.zeros(n, dtype=torch.float64, device="cuda")
        if build_matrix or not self._built:
            self.ws.assembly(2)
            self._built = True
        if build_rhs and (self._ws_route() or not self.is_linear()):
            self.ws.assembly(1)
            check(lib().gfx_ws_copy_assembled_vector_device(self.ws.h, ctypes.c_void_p(rhs.data_ptr())),
                  self.ctx.h)
            check(lib().gfx_ctx_synchronize(self.ctx.h), self.ctx.h)
            rhs.neg_()
        dofs, vals = self._constraints()
        flags = (_lib.BUILD_MATRIX if build_matrix else 0) | (_lib.BUILD_RHS if build_rhs else 0)
        flags |= _lib.MODEL_LINEAR if (self.is_linear() and not self._ws_route()) else 0
        flags |= _lib.MODEL_SYMMETRIC if self.is_symmetric() else 0
        if dofs.size:
            args = (dofs.size, ctypes.c_void_p(dofs.ctypes.data), ctypes.c_void_p(vals.ctypes.data),
                    ctypes.c_void_p(self.U.data_ptr()), ctypes.c_void_p(rhs.data_ptr()), flags)
            if self._ws_route():
                check(lib().gfx_ws_dof_constraints(self.ws.h, *args), self.ctx.h)
            else:
                check(lib().gfx_plan_dof_constraints(self._plan(), *args), self.ctx.h)
        check(lib().gfx_ctx_synchronize(self.ctx.h), self.ctx.h)
        self._rhs = rhs
        return rhs

    def tangent_matrix(self):
        return self.ws.assembled_matrix()

    def rhs(self):
        return self._rhs.cpu().numpy()

    def _linear_solve(self, rhs, dx, rtol, maxit, precond="ilu"):
        """default_linear_solver's iterative branch
        (src/getfem/getfem_model_solvers.h:652-660): CG for a coercive model
        (block Jacobi on a single plan, the diagonal on the workspace's
        matrix), restarted GMRES(500) otherwise, preconditioned by GetFEM's
        ILU ("ilu") or the inverse diagonal ("diagonal").  ILU(0) breaks down
        on saddle points whose multiplier or pressure rows come first (their
        tiny pivots), as GetFEM's does; there "diagonal" converges."""
        it, res = ctypes.c_int(), ctypes.c_double()
        b, x = ctypes.c_void_p(rhs.data_ptr()), ctypes.c_void_p(dx.data_ptr())
        if not (self.is_symmetric() and self.is_coercive()):
            pc = _lib.PRECOND_ILU if precond == "ilu" else _lib.PRECOND_DIAGONAL
            check(lib().gfx_ws_solve_gmres(self.ws.h, pc, b, x, rtol, maxit, 500, ctypes.byref(it), ctypes.byref(res)),
                  self.ctx.h)
        elif self._ws_route():
            check(lib().gfx_ws_solve_cg(self.ws.h, b, x, rtol, maxit, ctypes.byref(it), ctypes.byref(res)), self.ctx.h)
        else:
            check(lib().gfx_plan_solve_cg(self._plan(), b, x, rtol, maxit, ctypes.byref(it), ctypes.byref(res)),
                  self.ctx.h)
        return it.value, res.value

    # ------------------------------------------------------------------- solve
    def solve(self, max_iter=100, max_res=1e-6, cg_maxit=20000, noisy=False, precond="ilu"):
        """getfem::standard_solve for this model: a linear model of one
        variable is one linear solve of K U = rhs; otherwise classical_Newton
        with basic_newton_line_search.  Linear solves: CG for symmetric
        coercive models, GMRES otherwise (precond: "ilu" or "diagonal", see
        _linear_solve).  Returns (newton iterations,
        converged, list of (linear solver iterations, residual))."""
        import torch
        cg_rtol = max_res / 20.0  # iter_linsolv0.set_resmax(iter.get_resmax()/20)
        log = []
        if self.is_linear() and len(self.vars) == 1:
            rhs = self.assembly("build_all")
            if not self._ws_route():
                it, res = self._linear_solve(rhs, self.U, cg_rtol, cg_maxit, precond)
                log.append((it, res))
                return 1, res <= cg_rtol, log
            # the polynomial route's rhs is the residual form: Newton below
            # (one iteration for a linear model)
        eln = self.approx_external_load() or 1.0
        rhs = self.assembly("build_rhs")
        res = float(rhs.abs().sum())
        crit = res / eln
        k = 0
        while crit > max_res and k < max_iter:
            rhs = self.assembly("build_all")
            dU = torch.zeros_like(self.U)
            it, cres = self._linear_solve(rhs, dU, cg_rtol, cg_maxit, precond)
            log.append((it, cres))
            if cres > cg_rtol:
                break
            # basic_newton_line_search: alpha = 1, 0.6, 0.36, ...
            U0 = self.U.clone()
            first = prev = res
            alpha, tries = 1.0, 0
            while True:
                conv_alpha = alpha
                alpha *= 3.0 / 5.0
                tries += 1
                self.U.copy_(U0 + conv_alpha * dU)
                r = float(self.assembly("build_rhs").abs().sum())
                if (k == 0 or r < first / 2 or (conv_alpha <= 1e-3 and r < 2 * first)
                        or (tries > 1 and r > prev and prev < 5.0 / 3.0 * first)):
                    break
                prev = r
            res = r
            k += 1
            crit = min(res / eln, float(dU.abs().sum()) / max(1e-25, float(self.U.abs().sum())))
            if noisy:
                print("newton %d: alpha %g residual %.3e crit %.3e (cg %d its)" % (k, conv_alpha, res, crit, it))
        return k, crit <= max_res, log
