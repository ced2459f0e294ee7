"""regular_unit_mesh and first-encounter dof numbering.

Checks: the oracle's two-stage construction (structured mesh with point
dedup, then add_convex_by_points renumbering, getfem_regular_meshes.cc:237-285)
equals the vectorised generator of getfem_amd.meshes; the first points and
convexes follow by hand from the Kuhn table
(bgeot_convex_ref_simplexified.cc:34-39); conformity and total volume; dof
counts; the oracle's coordinate-matched numbering equals libgfx's lattice-key
numbering."""
import numpy as np
import pytest

import oracle
from getfem_amd import core, meshes


@pytest.mark.parametrize("n", [1, 2, 3, 6])
@pytest.mark.parametrize("gt,name", [(oracle.GT_PK31, "GT_PK(3,1)"), (oracle.GT_QK31, "GT_QK(3,1)")])
def test_generators_agree(n, gt, name):
    po, co = oracle.regular_mesh(gt, n)
    pp, cp = meshes.regular_unit_mesh(n, name)
    assert po.shape == ((n + 1) ** 3, 3)
    assert np.array_equal(co, cp)
    assert np.array_equal(po, pp)


def test_first_points_by_hand():
    n = 4
    h = 1.0 / n
    pts, cvx = meshes.regular_unit_mesh(n)
    expect = np.array([[h, h, 0], [h, h, h], [0, 0, 0], [h, 0, 0], [h, 0, h], [0, 0, h],
                       [0, h, 0], [0, h, h]])
    assert np.allclose(pts[:8], expect, atol=0)
    assert np.array_equal(cvx[:6], [[0, 1, 2, 3], [1, 2, 4, 5], [1, 2, 3, 4], [0, 1, 2, 6],
                                    [7, 1, 2, 5], [7, 1, 2, 6]])


@pytest.mark.parametrize("jit", [False, True])
def test_conformity_and_volume(jit):
    n = 5
    pts, cvx = meshes.regular_unit_mesh(n)
    if jit:
        pts = meshes.jitter(pts, n)
    faces = np.sort(np.concatenate([cvx[:, [1, 2, 3]], cvx[:, [0, 2, 3]], cvx[:, [0, 1, 3]],
                                    cvx[:, [0, 1, 2]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert set(counts.tolist()) <= {1, 2}
    assert (counts == 1).sum() == 12 * n * n
    K = np.stack([pts[cvx[:, c]] - pts[cvx[:, 0]] for c in (1, 2, 3)], axis=2)
    vol = np.abs(np.linalg.det(K)).sum() / 6
    assert abs(vol - 1.0) < 1e-12


def test_jitter_keeps_border():
    n = 6
    pts, _ = meshes.regular_unit_mesh(n)
    j = meshes.jitter(pts, n)
    border = np.any((np.abs(pts) < 1e-10) | (np.abs(pts - 1) < 1e-10), axis=1)
    assert np.array_equal(j[border], pts[border])
    d = np.abs(j[~border] - pts[~border])
    assert d.max() <= 0.1 / n and d.min() > 0


DOFCASES = [("GT_PK(3,1)", "FEM_PK(3,1)", oracle.GT_PK31, oracle.FEM_PK, 1, 1, lambda n: (n + 1) ** 3),
            ("GT_PK(3,1)", "FEM_PK(3,2)", oracle.GT_PK31, oracle.FEM_PK, 2, 3, lambda n: 3 * (2 * n + 1) ** 3),
            ("GT_PK(3,1)", "FEM_PK(3,3)", oracle.GT_PK31, oracle.FEM_PK, 3, 3, lambda n: 3 * (3 * n + 1) ** 3),
            ("GT_QK(3,1)", "FEM_QK(3,2)", oracle.GT_QK31, oracle.FEM_QK, 2, 3, lambda n: 3 * (2 * n + 1) ** 3),
            ("GT_PK(3,1)", "FEM_PK_DISCONTINUOUS(3,0)", oracle.GT_PK31, oracle.FEM_PK_DISC0, 0, 1,
             lambda n: 6 * n ** 3)]


@pytest.mark.parametrize("gtn,femn,gt,kind,k,q,count", DOFCASES)
def test_dof_numbering(gtn, femn, gt, kind, k, q, count):
    n = 3
    pts, cvx = meshes.regular_unit_mesh(n, gtn)
    pts = meshes.jitter(pts, n)
    d1, n1 = core.enumerate_dofs(gtn, femn, q, pts, cvx)
    d2, n2 = oracle.enumerate_dof(gt, pts, cvx, kind, k, q)
    assert n1 == n2 == count(n)
    assert np.array_equal(d1, d2)
    # first-encounter: new nodes appear in increasing order, Qdim apart
    flat = d1.ravel()
    _, first = np.unique(flat, return_index=True)
    order = flat[np.sort(first)]
    assert np.array_equal(order, np.arange(0, n1, q))


def test_dof_table_rejects_bad_mesh():
    pts, cvx = meshes.regular_unit_mesh(2)
    cvx = cvx.copy()
    cvx[0, 0] = 10 ** 6
    with pytest.raises(Exception):
        core.enumerate_dofs("GT_PK(3,1)", "FEM_PK(3,1)", 1, pts, cvx)
