"""Edge cases of the device path, as the reference's own code treats them:
an empty mesh (empty matrix), a single element, elements whose matrix is
entirely zero (skipped: GetFEM inserts nothing for them,
compile_and_exec.cc:5441-5453), and a degenerate element (GetFEM throws
"Non invertible matrix", bgeot_geometric_trans.cc:380)."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import _lib, meshes
from tests import problems

pytestmark = pytest.mark.gpu


def assemble(ctx, pts, cvx, fem, q, im, term, coefs):
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, fem, q)
    mim = ga.MeshIm(mesh, im)
    plan = ga.Plan(mf, mim, term, coefs)
    nnz = plan.build_pattern()
    plan.assemble()
    return plan, mf, nnz, (mesh, mim)


def oracle_csr(pts, cvx, k, q, im, term, lam, mu, mf):
    dofs = mf.basic_dof_table()
    kw = dict(coef_a=lam, coef_b=mu) if term == oracle.TERM_ISO_ELAST else dict(scal_a=1.0)
    P = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, q, dofs, mf.nbdof(), im, term, **kw)
    return oracle.csc_to_csr(*P.assemble_csc(), mf.nbdof())


def test_empty_mesh(ctx):
    pts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    cvx = np.zeros((0, 4), dtype=np.int32)
    plan, mf, nnz, keep = assemble(ctx, pts, cvx, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)",
                                   ga.TERM_ISO_ELAST, [1.0, 1.0])
    assert nnz == 0 and mf.nbdof() == 0
    ip, ix, v = plan.csr()
    assert ip.tolist() == [0] and ix.size == 0


def test_single_tetrahedron(ctx):
    pts = np.array([[0.0, 0, 0], [1, 0.1, 0], [0.2, 1, 0], [0.1, 0.3, 0.9]])
    cvx = np.array([[0, 1, 2, 3]], dtype=np.int32)
    lam, mu = np.array([2.0]), np.array([0.5])
    plan, mf, nnz, keep = assemble(ctx, pts, cvx, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)",
                                   ga.TERM_ISO_ELAST, [lam, mu])
    ip, ix, v = plan.csr()
    ipr, ixr, vr = oracle_csr(pts, cvx, 2, 3, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, lam, mu, mf)
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12
    assert nnz <= 30 * 30


def test_zero_coefficient_elements_are_skipped(ctx):
    """Elements with lambda = mu = 0 have an all-zero matrix: nothing is
    inserted for them, so rows touched only by such elements are empty."""
    n = 3
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    zero = np.arange(cvx.shape[0]) % 3 == 0
    lam[zero] = 0.0
    mu[zero] = 0.0
    plan, mf, nnz, keep = assemble(ctx, pts, cvx, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)",
                                   ga.TERM_ISO_ELAST, [lam, mu])
    ip, ix, v = plan.csr()
    ipr, ixr, vr = oracle_csr(pts, cvx, 2, 3, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, lam, mu, mf)
    assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr)
    assert problems.rel_frob(v, vr) <= 1e-12
    # all elements zero: an empty pattern
    plan0, mf0, nnz0, keep0 = assemble(ctx, pts, cvx, "FEM_PK(3,2)", 3, "IM_TETRAHEDRON(5)",
                                       ga.TERM_ISO_ELAST, [0.0, 0.0])
    assert nnz0 == 0


def test_degenerate_element_is_rejected(ctx):
    pts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1]])
    cvx = np.array([[0, 1, 2, 4], [0, 1, 2, 3]], dtype=np.int32)  # the second is flat
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    plan = ga.Plan(mf, mim, ga.TERM_LAPLACE, [1.0])
    with pytest.raises(_lib.GfxError) as exc:
        plan.build_pattern()
    assert exc.value.code == 3 and "Non invertible" in str(exc.value)


def test_any_destruction_order():
    """Destroying a context, mesh or mesh_fem while objects made from them are
    alive only marks it; those objects keep working and the last one releases
    it (so garbage-collection order cannot free a parent under its child)."""
    c = ga.Context(0)
    n = 2
    pts, cvx = meshes.regular_unit_mesh(n)
    mesh = ga.Mesh(c, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,1)", 1)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(1)")
    plan = ga.Plan(mf, mim, ga.TERM_LAPLACE, [1.0])
    lib = _lib.lib()
    # parents first
    assert lib.gfx_mesh_destroy(mesh.h) == 0
    mesh.h = None
    assert lib.gfx_mesh_fem_destroy(mf.h) == 0
    mf.h = None
    assert lib.gfx_ctx_destroy(c.h) == 0
    c.h = None
    nnz = plan.build_pattern()
    plan.assemble()
    ip, ix, v = plan.csr()
    assert nnz > 0 and ip[-1] == nnz and np.all(np.isfinite(v))
    del plan  # releases mf, then the mesh once mim goes, then the context
    del mim
