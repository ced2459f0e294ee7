"""The C++ facade (include/gfx_workspace.hpp) used the way GetFEM's own test
uses ga_workspace (tests/test_assembly.cc:812-818): same matrix as the oracle,
unsupported terms surface as a C++ exception."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from getfem_amd import meshes
from tests import problems

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "_build", "ga_workspace_demo")


def test_cpp_workspace_matches_oracle(tmp_path):
    assert os.path.exists(DEMO), "build first (__graft_entry__.build())"
    n = 3
    pts, cvx = meshes.regular_unit_mesh(n)
    pts = meshes.jitter(pts, n)
    pts.astype(np.float64).tofile(tmp_path / "points.f64")
    cvx.astype(np.int32).tofile(tmp_path / "convexes.i32")
    out = subprocess.run([DEMO, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "rejected:" in out.stdout
    cp = np.fromfile(tmp_path / "colptr.i64", dtype=np.int64)
    rows = np.fromfile(tmp_path / "rows.u32", dtype=np.uint32)
    vals = np.fromfile(tmp_path / "values.f64")
    dofs, ndof = oracle.enumerate_dof(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 2, 3)
    P = oracle.Problem(oracle.GT_PK31, pts, cvx, oracle.FEM_PK, 2, 3, dofs, ndof, "IM_TETRAHEDRON(5)",
                       oracle.TERM_ISO_ELAST, coef_a=np.full(cvx.shape[0], 3.0),
                       coef_b=np.full(cvx.shape[0], 2.0))
    cpr, rr, vr = P.assemble_csc()
    assert np.array_equal(cp, cpr) and np.array_equal(rows.astype(np.int64), rr.astype(np.int64))
    assert problems.rel_frob(vals, vr) <= 1e-12
    import scipy.sparse as sp
    K = sp.csc_matrix((vr, rr, cpr), shape=(ndof, ndof))
    V = np.fromfile(tmp_path / "V.f64")
    assert problems.rel_frob(V, K @ np.sin(np.arange(ndof, dtype=np.float64))) <= 1e-12
    # the order-2 spelling has no order-1 tree (GetFEM's assembly(1) adds nothing for it)
    assert np.abs(np.fromfile(tmp_path / "V2.f64")).max() == 0.0
