"""CPU oracle for GetFEM's generic-assembly path (test infrastructure only).

ctypes front-end of ``oracle/gfx_oracle.c``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package; the getfem_amd product never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

GT_PK31, GT_QK31 = 0, 1
FEM_PK, FEM_QK, FEM_PK_DISC0 = 0, 1, 2
TERM_LAPLACE, TERM_ISO_ELAST, TERM_NEOHOOKEAN, TERM_ISOTROPIC = 0, 1, 2, 3

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64, i32, d = ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.orc_im_npts.argtypes = [ctypes.c_char_p]
        L.orc_im_rule.argtypes = [ctypes.c_char_p, P, P]
        L.orc_fem_nb_dof.argtypes = [i32, i32]
        L.orc_fem_nodes.argtypes = [i32, i32, P]
        L.orc_fem_eval.argtypes = [i32, i32, P, P, P]
        L.orc_regular_mesh.argtypes = [i32, i32, P, P]
        L.orc_regular_mesh.restype = i64
        L.orc_enumerate_dof.argtypes = [i32, i64, P, i64, P, i32, i32, i32, P]
        L.orc_enumerate_dof.restype = i64
        L.orc_element_matrix.argtypes = [P, i64, P]
        L.orc_element_vector.argtypes = [P, i64, P]
        L.orc_assemble_matrix.argtypes = [P, i64, i64, i32]
        L.orc_assemble_matrix.restype = P
        L.orc_mat_nnz.argtypes = [P]
        L.orc_mat_nnz.restype = i64
        L.orc_mat_ncols.argtypes = [P]
        L.orc_mat_ncols.restype = i64
        L.orc_mat_export_csc.argtypes = [P, P, P, P]
        L.orc_mat_free.argtypes = [P]
        L.orc_assemble_vector.argtypes = [P, P]
        L.orc_compliance.argtypes = [P, P, d, d, P, P]
        L.orc_compliance.restype = d
        L.orc_ildlt_pattern.argtypes = [i64, P, P, P, P]
        L.orc_ildlt_pattern.restype = i64
        L.orc_ildlt.argtypes = [i64, P, P, P, P, P, P]
        L.orc_ildlt.restype = i64
        L.orc_ildlt_apply.argtypes = [i64, P, P, P, P, P]
        L.orc_assemble_blocks.argtypes = [i64, i64, i32, P]
        L.orc_assemble_blocks.restype = P
        L.orc_ilu.argtypes = [i64, P, P, P, P]
        L.orc_ilu.restype = i64
        L.orc_ilu_apply.argtypes = [i64, P, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class _Problem(ctypes.Structure):
    _fields_ = [
        ("gt", ctypes.c_int), ("n_pts", ctypes.c_int64), ("pts", ctypes.c_void_p),
        ("n_cv", ctypes.c_int64), ("cvx", ctypes.c_void_p),
        ("fem_kind", ctypes.c_int), ("fem_k", ctypes.c_int), ("qdim", ctypes.c_int),
        ("dofs", ctypes.c_void_p), ("ndof", ctypes.c_int64),
        ("nq", ctypes.c_int), ("qp", ctypes.c_void_p), ("qw", ctypes.c_void_p),
        ("term", ctypes.c_int), ("coef_a", ctypes.c_void_p), ("coef_b", ctypes.c_void_p),
        ("scal_a", ctypes.c_double), ("scal_b", ctypes.c_double),
        ("U", ctypes.c_void_p), ("coef_c", ctypes.c_void_p), ("scal_c", ctypes.c_double),
    ]


def im_rule(name):
    n = lib().orc_im_npts(name.encode())
    if n <= 0:
        raise ValueError("unknown integration method " + name)
    pts = np.zeros((n, 3))
    w = np.zeros(n)
    lib().orc_im_rule(name.encode(), _ptr(pts), _ptr(w))
    return pts, w


def fem_nodes(kind, k):
    nd = lib().orc_fem_nb_dof(kind, k)
    out = np.zeros((nd, 3))
    lib().orc_fem_nodes(kind, k, _ptr(out))
    return out


def fem_eval(kind, k, x):
    nd = lib().orc_fem_nb_dof(kind, k)
    x = np.ascontiguousarray(x, dtype=np.float64)
    val = np.zeros(nd)
    grad = np.zeros((nd, 3))
    lib().orc_fem_eval(kind, k, _ptr(x), _ptr(val), _ptr(grad))
    return val, grad


def regular_mesh(gt, n):
    npts = (n + 1) ** 3
    nv = 4 if gt == GT_PK31 else 8
    ncv = 6 * n ** 3 if gt == GT_PK31 else n ** 3
    pts = np.zeros((npts, 3))
    cvx = np.zeros((ncv, nv), dtype=np.int32)
    got = lib().orc_regular_mesh(gt, n, _ptr(pts), _ptr(cvx))
    return pts[:got].copy(), cvx


def enumerate_dof(gt, pts, cvx, fem_kind, fem_k, qdim):
    nd = lib().orc_fem_nb_dof(fem_kind, fem_k)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cvx = np.ascontiguousarray(cvx, dtype=np.int32)
    dofs = np.zeros((cvx.shape[0], nd), dtype=np.int32)
    ndof = lib().orc_enumerate_dof(gt, pts.shape[0], _ptr(pts), cvx.shape[0], _ptr(cvx),
                                   fem_kind, fem_k, qdim, _ptr(dofs))
    return dofs, int(ndof)


class Problem:
    """Holds the arrays of one assembly problem alive for the C side."""

    def __init__(self, gt, pts, cvx, fem_kind, fem_k, qdim, dofs, ndof, im_name,
                 term, coef_a=None, coef_b=None, scal_a=1.0, scal_b=0.0, U=None,
                 coef_c=None, scal_c=0.0):
        self.pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.cvx = np.ascontiguousarray(cvx, dtype=np.int32)
        self.dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        self.qp, self.qw = im_rule(im_name)
        self.coef_a = None if coef_a is None else np.ascontiguousarray(coef_a, dtype=np.float64)
        self.coef_b = None if coef_b is None else np.ascontiguousarray(coef_b, dtype=np.float64)
        self.U = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        self.coef_c = None if coef_c is None else np.ascontiguousarray(coef_c, dtype=np.float64)
        self.nd = lib().orc_fem_nb_dof(fem_kind, fem_k)
        self.nloc = self.nd * qdim
        self.ndof = ndof
        self.s = _Problem(gt, self.pts.shape[0], _ptr(self.pts), self.cvx.shape[0], _ptr(self.cvx),
                          fem_kind, fem_k, qdim, _ptr(self.dofs), ndof,
                          self.qw.shape[0], _ptr(self.qp), _ptr(self.qw), term,
                          _ptr(self.coef_a), _ptr(self.coef_b), scal_a, scal_b, _ptr(self.U),
                          _ptr(self.coef_c), scal_c)

    @property
    def n_cv(self):
        return self.cvx.shape[0]

    def element_matrix(self, cv):
        out = np.zeros(self.nloc * self.nloc)
        r = lib().orc_element_matrix(ctypes.byref(self.s), cv, _ptr(out))
        if r < 0:
            raise ValueError("unsupported term")
        return out.reshape(self.nloc, self.nloc).T.copy()  # rows = test

    def element_vector(self, cv):
        out = np.zeros(self.nloc)
        if lib().orc_element_vector(ctypes.byref(self.s), cv, _ptr(out)) < 0:
            raise ValueError("unsupported term")
        return out

    def assemble_csc(self, cv_begin=0, cv_end=None, nthreads=1):
        """GetFEM's col_matrix<rsvector>: returns (colptr, rows, vals)."""
        if cv_end is None:
            cv_end = self.n_cv
        m = lib().orc_assemble_matrix(ctypes.byref(self.s), cv_begin, cv_end, nthreads)
        try:
            nnz = lib().orc_mat_nnz(m)
            colptr = np.zeros(self.ndof + 1, dtype=np.int64)
            rows = np.zeros(nnz, dtype=np.int32)
            vals = np.zeros(nnz)
            lib().orc_mat_export_csc(m, _ptr(colptr), _ptr(rows), _ptr(vals))
        finally:
            lib().orc_mat_free(m)
        return colptr, rows, vals

    def assemble_timed(self, cv_begin, cv_end, nthreads=1):
        """Assemble and drop the result; returns the stored nnz."""
        m = lib().orc_assemble_matrix(ctypes.byref(self.s), cv_begin, cv_end, nthreads)
        nnz = lib().orc_mat_nnz(m)
        lib().orc_mat_free(m)
        return nnz

    def assemble_vector(self):
        V = np.zeros(self.ndof)
        if lib().orc_assemble_vector(ctypes.byref(self.s), _ptr(V)) < 0:
            raise ValueError("unsupported term")
        return V

    def compliance(self, rho, emin=1e-3, penal=3.0, grads=True):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        dr = np.zeros(self.n_cv) if grads else None
        dx = np.zeros_like(self.pts) if grads else None
        J = lib().orc_compliance(ctypes.byref(self.s), _ptr(rho), emin, penal, _ptr(dr), _ptr(dx))
        return J, dr, dx


class _Block(ctypes.Structure):
    _fields_ = [("s1", ctypes.c_int), ("s2", ctypes.c_int), ("rows", ctypes.c_void_p),
                ("cols", ctypes.c_void_p), ("elem", ctypes.c_void_p)]


def assemble_blocks(n_cv, ndof, blocks):
    """A matrix over several variables from its blocks' element matrices, as
    ga_instruction_matrix_assembly_mf_mf::exec adds them
    (src/getfem_generic_assembly_compile_and_exec.cc:5074-5189): element by
    element, block by block, zero element matrices skipped, the others added
    by add_elem_matrix (:4853-4936) with the threshold 1e-14 max|elem|.
    blocks: (rows (n_cv, s1) global row dofs, cols (n_cv, s2) global column
    dofs, elem (n_cv, s1, s2) with rows = test dofs).  Returns GetFEM's CSC
    (colptr, rows, vals) of the ndof x ndof matrix."""
    keep = []
    arr = (_Block * max(len(blocks), 1))()
    for b, (r, c, e) in enumerate(blocks):
        r = np.ascontiguousarray(r, dtype=np.int32)
        c = np.ascontiguousarray(c, dtype=np.int32)
        ecm = np.ascontiguousarray(np.transpose(np.asarray(e, dtype=np.float64), (0, 2, 1)))  # column-major
        keep += [r, c, ecm]
        arr[b] = _Block(r.shape[1], c.shape[1], r.ctypes.data, c.ctypes.data, ecm.ctypes.data)
    m = lib().orc_assemble_blocks(n_cv, ndof, len(blocks), arr)
    try:
        nnz = lib().orc_mat_nnz(m)
        colptr = np.zeros(ndof + 1, dtype=np.int64)
        rows = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz)
        lib().orc_mat_export_csc(m, _ptr(colptr), _ptr(rows), _ptr(vals))
    finally:
        lib().orc_mat_free(m)
    return colptr, rows, vals


def tet_geometry(pts, cv):
    """Affine tetrahedron: K = [x1 - x0, x2 - x0, x3 - x0], J = det K and
    B = K^{-T} (src/bgeot_geometric_trans.cc:357-413)."""
    X = pts[cv]
    K = (X[1:] - X[0]).T
    return np.linalg.det(K), np.linalg.inv(K).T


def block_element_matrices(pts, cvx, im_name, test, trial, fields, integrand2):
    """Element matrices of one block (test variable x trial variable) of a
    polynomial integrand on affine tetrahedra or Q1 hexahedra (cvx with 8
    vertices: the geometry at every point), accumulated over the
    integration points as ga_exec runs the element loop
    (src/getfem_generic_assembly_compile_and_exec.cc:8750-8869) and the
    matrix instruction adds coeff * t to elem (:5074-5100).
    test, trial: (fem_kind, k, qdim, dofs (n_cv, nd)).  fields: name ->
    (fem_kind, k, qdim, dofs, U), the fields the integrand reads.
    integrand2(vals) -> D (4 qt, 4 qs): the integrand's second derivative with
    respect to a test slot and a trial slot (slot a < q the value of component
    a, q + 3a + b its gradient along b), with vals[name] = (value (q,),
    gradient (q, 3)) at the point.  Returns (n_cv, nd_t qt, nd_s qs), rows and
    columns ordered node-major (node i, component a) -> i q + a."""
    qp, qw = im_rule(im_name)
    tabs = {}

    def table(kind, k):
        if (kind, k) not in tabs:
            ev = [fem_eval(kind, k, x) for x in qp]
            tabs[(kind, k)] = (np.array([v for v, _ in ev]), np.array([g for _, g in ev]))
        return tabs[(kind, k)]

    def slots(kind, k, q, cv_dofs, B, iq):
        phi, gref = table(kind, k)
        g = gref[iq] @ B.T  # real gradients (nd, 3)
        nd = phi.shape[1]
        S = np.zeros((nd * q, 4 * q))
        for i in range(nd):
            for a in range(q):
                S[i * q + a, a] = phi[iq, i]
                S[i * q + a, q + 3 * a:q + 3 * a + 3] = g[i]
        return S

    n_cv = cvx.shape[0]
    hexa = cvx.shape[1] == 8  # Q1 hexahedra: the geometry at every point
    gq = [fem_eval(FEM_QK, 1, x)[1] for x in qp] if hexa else None
    tk, tkk, tq, tdofs = test
    sk, skk, sq, sdofs = trial
    out = np.zeros((n_cv, tdofs.shape[1] * tq, sdofs.shape[1] * sq))
    for e in range(n_cv):
        if not hexa:
            J, B = tet_geometry(pts, cvx[e])
        for iq in range(len(qw)):
            if hexa:  # K = sum_v X_v grad(N_v) (bgeot_geometric_trans.cc:270-287)
                K = pts[cvx[e]].T @ gq[iq]
                J, B = np.linalg.det(K), np.linalg.inv(K).T
            vals = {}
            for name, (kind, k, q, dofs, U) in fields.items():
                phi, gref = table(kind, k)
                g = gref[iq] @ B.T
                loc = np.array([[U[d + c] for c in range(q)] for d in dofs[e]])  # (nd, q)
                vals[name] = (phi[iq] @ loc, loc.T @ g)
            D = integrand2(vals)
            St = slots(tk, tkk, tq, tdofs[e], B, iq)
            Ss = slots(sk, skk, sq, sdofs[e], B, iq)
            out[e] += qw[iq] * abs(J) * (St @ D @ Ss.T)
    return out


def ilu(indptr, indices, vals):
    """gmm::ilu_precond of GetFEM's column-major matrix (the rows of A^T
    factored): F aligned with A's CSR and the count of replaced pivots."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    av = np.ascontiguousarray(vals, dtype=np.float64)
    F = np.zeros(av.size)
    rep = lib().orc_ilu(ip.size - 1, _ptr(ip), _ptr(ix), _ptr(av), _ptr(F))
    return F, int(rep)


def ilu_apply(indptr, indices, F, r):
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    F = np.ascontiguousarray(F, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    lib().orc_ilu_apply(ip.size - 1, _ptr(ip), _ptr(ix), _ptr(F), _ptr(r), _ptr(z))
    return z


def csc_to_csr(colptr, rows, vals, n):
    """Transpose a CSC (n x n) into CSR arrays of the same matrix."""
    import scipy.sparse as sp
    m = sp.csc_matrix((vals, rows, colptr), shape=(n, n))
    r = m.tocsr()
    r.sort_indices()
    return r.indptr.astype(np.int64), r.indices.astype(np.int32), r.data


class Ildlt:
    """gmm::ildlt_precond of a CSR matrix (sorted columns): U's slots per row
    (diagonal first, then the columns above it), D on the diagonal slots,
    U_ij / D_i elsewhere; ``replaced`` counts pivots replaced by 1."""

    def __init__(self, indptr, indices, vals):
        ip = np.ascontiguousarray(indptr, dtype=np.int64)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        av = np.ascontiguousarray(vals, dtype=np.float64)
        self.n = n = ip.size - 1
        self.uptr = np.zeros(n + 1, np.int64)
        m = lib().orc_ildlt_pattern(n, _ptr(ip), _ptr(ix), _ptr(self.uptr), None)
        self.uind = np.zeros(m, np.int32)
        lib().orc_ildlt_pattern(n, _ptr(ip), _ptr(ix), _ptr(self.uptr), _ptr(self.uind))
        self.uval = np.zeros(m)
        self.replaced = lib().orc_ildlt(n, _ptr(ip), _ptr(ix), _ptr(av), _ptr(self.uptr), _ptr(self.uind),
                                        _ptr(self.uval))

    def diag(self):
        return self.uval[self.uptr[:-1]].copy()

    def factors(self):
        """(L unit lower as scipy CSR, D) with L = U^T."""
        import scipy.sparse as sp
        n = self.n
        rows = np.repeat(np.arange(n), np.diff(self.uptr))
        off = self.uind != rows
        U = sp.csr_matrix((self.uval[off], (rows[off], self.uind[off])), shape=(n, n)) + sp.identity(n)
        return U.T.tocsr(), self.diag()

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        lib().orc_ildlt_apply(self.n, _ptr(self.uptr), _ptr(self.uind), _ptr(self.uval), _ptr(r), _ptr(z))
        return z


def pcg(indptr, indices, vals, b, rtol, maxit, precond):
    """gmm::cg (src/gmm/gmm_solver_cg.h:84-118) from x = 0 with the stopping
    test of gmm::iteration (|r| <= rtol |b|, checked before each step).
    precond(r) -> z.  Returns (x, iterations, |r| / |b|)."""
    import scipy.sparse as sp
    n = len(indptr) - 1
    A = sp.csr_matrix((vals, indices, indptr), shape=(n, n))
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, 0.0
    r = b - A @ x
    z = precond(r)
    rho = z @ r
    p = z.copy()
    it = 0
    while np.linalg.norm(r) > rtol * bn and it < maxit:
        if it > 0:
            z = precond(r)
            rho_new = z @ r
            p = z + (rho_new / rho) * p
            rho = rho_new
        q = A @ p
        a = rho / (q @ p)
        x += a * p
        r -= a * q
        it += 1
    return x, it, np.linalg.norm(r) / bn
