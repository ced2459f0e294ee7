"""libgfx (HIP, gfx950) against the oracle on identical inputs: the integer
pattern bit-exact, FP64 values within 1e-12 relative Frobenius, FP32 element
contractions within 1e-6, bf16 split element matrices within 1e-4."""
import numpy as np
import pytest

import getfem_amd as ga
import oracle
from getfem_amd import meshes
from tests import problems

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_FP32 = 1e-6


def gpu_plan(ctx, case):
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", case.pts, case.cvx)
    mf = ga.MeshFem(mesh, case.fem, case.qdim)
    mim = ga.MeshIm(mesh, case.im)
    term = ga.TERM_LAPLACE if case.term == 0 else ga.TERM_ISO_ELAST
    plan = ga.Plan(mf, mim, term, case.coefs())
    plan.build_pattern()
    return plan, (mesh, mf, mim)


CASES = {
    "c1_p1_poisson_n20": lambda: problems.c1(20),
    "p1_poisson_n6_coef": lambda: problems.Case(6, 1, 1, "IM_TETRAHEDRON(1)", 0,
                                               a=np.linspace(0.5, 2.0, 6 * 6 ** 3)),
    "p2_poisson_n5": lambda: problems.Case(5, 2, 1, "IM_TETRAHEDRON(5)", 0),
    "p2_elast_n4": lambda: problems.c2(4),
    "p2_elast_n8": lambda: problems.c2(8),
    "p2_elast_n3_unjittered": lambda: problems.Case(3, 2, 3, "IM_TETRAHEDRON(5)", 1,
                                                   lam=np.full(162, 2.0), mu=np.full(162, 0.7),
                                                   jitter=False),
    "p1_elast_n5": lambda: problems.Case(5, 1, 3, "IM_TETRAHEDRON(1)", 1,
                                        lam=np.full(750, 1.3), mu=np.full(750, 0.4)),
    "p3_elast_n3": lambda: problems.c5(3),
    # no integer form of the reference tensor (irrational points, inexact
    # rule): the floating-point table path of the factorized form
    "p3_elast_n2_im2": lambda: problems.Case(2, 3, 3, "IM_TETRAHEDRON(2)", 1,
                                             lam=np.linspace(0.5, 2.0, 48), mu=np.linspace(1.5, 0.3, 48)),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("form", [ga.FORM_FACTORIZED, ga.FORM_PERQP])
def test_fp64_parity(ctx, name, form):
    case = CASES[name]()
    print("case", name, "form", form, "ndof", case.ndof, flush=True)
    ip_ref, ix_ref, v_ref = case.oracle_csr()
    plan, keep = gpu_plan(ctx, case)
    plan.set_form(form, ga.PREC_FP64)
    plan.assemble()
    ip, ix, v = plan.csr()
    assert np.array_equal(ip, ip_ref), "row pointers differ"
    assert np.array_equal(ix, ix_ref), "column indices differ"
    err = problems.rel_frob(v, v_ref)
    print("rel frobenius error", err)
    assert err <= TOL_FP64


@pytest.mark.parametrize("name", ["p2_elast_n8", "c1_p1_poisson_n20", "p3_elast_n3", "p3_elast_n2_im2"])
@pytest.mark.parametrize("form", [ga.FORM_FACTORIZED, ga.FORM_PERQP])
def test_fp32_element_error(ctx, name, form):
    case = CASES[name]()
    ip_ref, ix_ref, v_ref = case.oracle_csr()
    plan, keep = gpu_plan(ctx, case)
    plan.set_form(form, ga.PREC_FP32)
    plan.assemble()
    ip, ix, v = plan.csr()
    assert np.array_equal(ix, ix_ref)
    err = problems.rel_frob(v, v_ref)
    print("fp32 element form: rel frobenius error vs FP64 oracle", err)
    assert err <= TOL_FP32


TOL_BF16X2 = 1e-4  # relative Frobenius: bf16 hi + lo products, FP32 element accumulation


BF16X2_CASES = dict(CASES)
# a rule with a negative weight (IM_TETRAHEDRON(3): -2/15 at the centroid)
BF16X2_CASES["p2_elast_n4_im3"] = lambda: problems.Case(4, 2, 3, "IM_TETRAHEDRON(3)", 1,
                                                        lam=np.linspace(0.5, 2.0, 384),
                                                        mu=np.linspace(1.5, 0.3, 384))


# the element-batched B^T D B forms on the matrix cores and their tolerances
# against the FP64 oracle (relative Frobenius)
MFMA_FORMS = {"bf16x2": (ga.PREC_BF16X2, TOL_BF16X2), "fp64_mfma": (ga.PREC_FP64_MFMA, TOL_FP64),
              "fp32_mfma": (ga.PREC_FP32_MFMA, TOL_FP32)}


@pytest.mark.parametrize("mf_form", sorted(MFMA_FORMS))
@pytest.mark.parametrize("name", ["p2_elast_n8", "p2_elast_n3_unjittered", "p3_elast_n3", "p2_elast_n4_im3"])
def test_bf16x2_element_batched(ctx, name, mf_form):
    """Element matrices B^T D B on the matrix cores (bf16 hi + lo split, FP64,
    FP32), gathered over the packed waves and the large-patch launches into
    GetFEM's pattern."""
    prec, tol = MFMA_FORMS[mf_form]
    case = BF16X2_CASES[name]()
    print("case", name, mf_form, "ndof", case.ndof, flush=True)
    ip_ref, ix_ref, v_ref = case.oracle_csr()
    plan, keep = gpu_plan(ctx, case)
    plan.set_form(ga.FORM_FACTORIZED, prec)
    plan.assemble()
    ip, ix, v = plan.csr()
    assert np.array_equal(ip, ip_ref) and np.array_equal(ix, ix_ref)
    err = problems.rel_frob(v, v_ref)
    print(mf_form, "element-batched form: rel frobenius error vs FP64 oracle", err)
    assert err <= tol
    # the same plan returns to the FP64 form
    plan.set_form(ga.FORM_FACTORIZED, ga.PREC_FP64)
    plan.assemble()
    assert problems.rel_frob(plan.csr()[2], v_ref) <= TOL_FP64


def test_bf16x2_rejects_other_terms(ctx):
    plan, keep = gpu_plan(ctx, CASES["p2_poisson_n5"]())
    with pytest.raises(ga.GfxError):
        plan.set_form(ga.FORM_FACTORIZED, ga.PREC_BF16X2)


def test_bf16x2_rejects_indefinite_material(ctx):
    """3 lambda + 2 mu <= 0 on one element: D has no real square root."""
    lam = np.full(162, 1.0)
    lam[7] = -0.5  # with mu = 0.7: 3 lambda + 2 mu = -0.1
    case = problems.Case(3, 2, 3, "IM_TETRAHEDRON(5)", 1, lam=lam, mu=np.full(162, 0.7))
    plan, keep = gpu_plan(ctx, case)
    plan.set_form(ga.FORM_FACTORIZED, ga.PREC_BF16X2)
    with pytest.raises(ga.GfxError):
        plan.assemble()


def test_csc_export_is_getfem_orientation(ctx):
    case = CASES["p2_elast_n4"]()
    plan, keep = gpu_plan(ctx, case)
    plan.assemble()
    colptr, rows, vals = plan.csc()
    cp, r, v = case.oracle_problem().assemble_csc()
    assert np.array_equal(colptr, cp)
    assert np.array_equal(rows.astype(np.int64), r.astype(np.int64))
    assert problems.rel_frob(vals, v) <= TOL_FP64


@pytest.mark.parametrize("name", ["p2_elast_n8", "p3_elast_n3"])
@pytest.mark.parametrize("prec", [ga.PREC_FP64, ga.PREC_FP32])
def test_repeated_assembly_is_bitwise_identical(ctx, name, prec):
    """The summation order of every CSR value is fixed (incidence lists
    sorted at pattern build, passes in list order, each wave's rows private
    to it): repeated assemblies, and a second plan built on the same inputs,
    give bit-identical values."""
    case = CASES[name]()
    plan, keep = gpu_plan(ctx, case)
    plan.set_form(ga.FORM_FACTORIZED, prec)
    plan.assemble()
    a = plan.csr()[2]
    for _ in range(4):
        plan.assemble()
        assert np.array_equal(plan.csr()[2].view(np.int64), a.view(np.int64))
    plan2, keep2 = gpu_plan(ctx, case)
    plan2.set_form(ga.FORM_FACTORIZED, prec)
    plan2.assemble()
    assert np.array_equal(plan2.csr()[2].view(np.int64), a.view(np.int64))


@pytest.mark.parametrize("name,n,fem,im,nd", [
    ("p2_elasticity_64cube_tet5", 64, "FEM_PK(3,2)", "IM_TETRAHEDRON(5)", 10),
    ("p3_elasticity_32cube_tet8", 32, "FEM_PK(3,3)", "IM_TETRAHEDRON(8)", 20),
])
def test_full_size_stored_nnz_equals_reference(ctx, name, n, fem, im, nd):
    """At full size the stored pattern has exactly the entry count GetFEM
    itself produced on the same inputs, the columns of every row are strictly
    increasing, and 48 sampled rows equal the rows rebuilt from the oracle's
    element matrices (columns bit-exact, values within the FP64 tolerance)."""
    import torch
    kat = problems.golden(name)
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, fem, 3)
    mim = ga.MeshIm(mesh, im)
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [lam, mu])
    nnz = plan.build_pattern()
    plan.assemble()
    print(name, "nnz", nnz, "ndof", mf.nbdof(), flush=True)
    assert nnz == kat["stored_nnz"] and mf.nbdof() == kat["ndof"]
    ip, ix, v = plan.torch_csr_view()
    d = ix[1:].to(torch.int64) - ix[:-1].to(torch.int64)
    row_start = torch.zeros(ix.numel(), dtype=torch.bool, device=ix.device)
    row_start[ip[:-1][ip[:-1] < ix.numel()]] = True
    assert bool(((d > 0) | row_start[1:]).all())  # strictly increasing inside each row
    assert bool(torch.isfinite(v).all())
    # Values at full size: sampled rows rebuilt from the oracle's element
    # matrices of the elements around the row's node, each element's entries
    # kept above its own 1e-14 max|K_e| (compile_and_exec.cc:5441-5453).
    k = {"FEM_PK(3,2)": 2, "FEM_PK(3,3)": 3}[fem]
    dofs = mf.basic_dof_table()
    D = np.asarray(dofs).reshape(cvx.shape[0], nd)
    P = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, k, 3, dofs, mf.nbdof(), im,
                       oracle.TERM_ISO_ELAST, coef_a=lam, coef_b=mu)
    rows = np.random.default_rng(7).choice(mf.nbdof(), 48, replace=False)
    ip_h = ip.cpu().numpy()
    worst = 0.0
    for row in rows:
        node_dof, comp = row - row % 3, row % 3
        acc = {}
        for e in np.nonzero((D == node_dof).any(axis=1))[0]:
            Ke = P.element_matrix(int(e))
            thr = 1e-14 * np.abs(Ke).max()
            li = int(np.nonzero(D[e] == node_dof)[0][0])
            krow = Ke[3 * li + comp]
            for j in range(nd):
                for b in range(3):
                    val = krow[3 * j + b]
                    if abs(val) > thr:
                        col = int(D[e, j]) + b
                        acc[col] = acc.get(col, 0.0) + val
        cols_ref = np.array(sorted(acc))
        vals_ref = np.array([acc[c] for c in cols_ref])
        s0, s1 = int(ip_h[row]), int(ip_h[row + 1])
        cols = ix[s0:s1].cpu().numpy().astype(np.int64)
        vals = v[s0:s1].cpu().numpy()
        assert np.array_equal(cols, cols_ref), "row %d: columns differ" % row
        worst = max(worst, np.linalg.norm(vals - vals_ref) / np.linalg.norm(vals_ref))
    print(name, "sampled rows", len(rows), "worst rel error", worst, flush=True)
    assert worst <= TOL_FP64


@pytest.mark.parametrize("mf_form", sorted(MFMA_FORMS))
def test_full_size_bf16x2_against_fp64(ctx, mf_form):
    """P3 elasticity 32^3x6 (196,608 tets, IM_TETRAHEDRON(8)): the matrix-core
    element-batched forms on the full mesh against the FP64 factorized form of
    the same plan (itself pinned to the oracle above): relative Frobenius
    within the form's tolerance, no entry off by more than that of the largest."""
    prec, tol = MFMA_FORMS[mf_form]
    import torch
    pts, cvx = meshes.regular_unit_mesh(32)
    pts = meshes.jitter(pts, 32)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    mesh = ga.Mesh(ctx, "GT_PK(3,1)", pts, cvx)
    mf = ga.MeshFem(mesh, "FEM_PK(3,3)", 3)
    mim = ga.MeshIm(mesh, "IM_TETRAHEDRON(8)")
    plan = ga.Plan(mf, mim, ga.TERM_ISO_ELAST, [lam, mu])
    plan.build_pattern()
    plan.set_form(ga.FORM_FACTORIZED, ga.PREC_FP64)
    plan.assemble()
    v64 = plan.torch_csr_view()[2].clone()
    plan.set_form(ga.FORM_FACTORIZED, prec)
    plan.assemble()
    v = plan.torch_csr_view()[2]
    err = float(torch.linalg.vector_norm(v - v64) / torch.linalg.vector_norm(v64))
    worst = float((v - v64).abs().max() / v64.abs().max())
    print(mf_form, "vs fp64 at full size: rel frobenius", err, "max entry", worst, flush=True)
    assert err <= tol and worst <= max(tol, 1e-12)
