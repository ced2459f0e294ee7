"""Shared construction of the parity problems (same inputs for libgfx and the
oracle)."""
import numpy as np

import oracle
from getfem_amd import meshes

FEMS = {1: "FEM_PK(3,1)", 2: "FEM_PK(3,2)", 3: "FEM_PK(3,3)"}


class Case:
    def __init__(self, n, k, qdim, im, term, lam=None, mu=None, a=1.0, jitter=True):
        self.n, self.k, self.qdim, self.im, self.term = n, k, qdim, im, term
        pts, cvx = meshes.regular_unit_mesh(n)
        self.pts = meshes.jitter(pts, n) if jitter else pts
        self.cvx = cvx
        self.lam, self.mu, self.a = lam, mu, a
        self.dofs, self.ndof = oracle.enumerate_dof(oracle.GT_PK31, self.pts, cvx, oracle.FEM_PK, k, qdim)

    @property
    def fem(self):
        return FEMS[self.k]

    def coefs(self):
        if self.term == oracle.TERM_LAPLACE:
            return [self.a]
        return [self.lam, self.mu]

    def oracle_problem(self):
        if self.term == oracle.TERM_LAPLACE:
            return oracle.Problem(oracle.GT_PK31, self.pts, self.cvx, oracle.FEM_PK, self.k, self.qdim,
                                  self.dofs, self.ndof, self.im, self.term,
                                  coef_a=None if np.isscalar(self.a) else self.a,
                                  scal_a=float(self.a) if np.isscalar(self.a) else 0.0)
        return oracle.Problem(oracle.GT_PK31, self.pts, self.cvx, oracle.FEM_PK, self.k, self.qdim,
                              self.dofs, self.ndof, self.im, self.term,
                              coef_a=self.lam, coef_b=self.mu)

    def oracle_csr(self):
        cp, r, v = self.oracle_problem().assemble_csc()
        return oracle.csc_to_csr(cp, r, v, self.ndof)


def c1(n=20):
    return Case(n, 1, 1, "IM_TETRAHEDRON(1)", oracle.TERM_LAPLACE)


def c2(n=64):
    pts, cvx = meshes.regular_unit_mesh(n)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    return Case(n, 2, 3, "IM_TETRAHEDRON(5)", oracle.TERM_ISO_ELAST, lam=lam, mu=mu)


def c5(n=32):
    pts, cvx = meshes.regular_unit_mesh(n)
    lam, mu = meshes.lognormal_lame(cvx.shape[0])
    return Case(n, 3, 3, "IM_TETRAHEDRON(8)", oracle.TERM_ISO_ELAST, lam=lam, mu=mu)


def rel_frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def golden(name):
    """Known answers recorded from the reference (tests/golden/reference_counts.json)."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_counts.json")
    with open(path) as f:
        return json.load(f)[name]
