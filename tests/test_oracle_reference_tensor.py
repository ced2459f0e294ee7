"""The patch kernel stores the factorized form's reference tensor
R_ij = sum_q w_q gradref_i(xi_q) gradref_j(xi_q)^T as small integers N with
R = N / L (int8 for P1/P2, int16 for P3).  That rests on the rule integrating
the products of basis gradients exactly, which makes R the exact (rational)
integral over the reference tetrahedron.  Pinned here on the oracle's own
bases and rules (src/getfem_fem.cc:726-767, getfem_im_list.h):

* P1 and P2: R equals the exact integral computed independently -- the
  gradients are affine, so each is fitted from its values at the four
  vertices and the products are integrated with the moment formula
  int x^a y^b z^c = a! b! c! / (a+b+c+3)!;
* R * L is integral for L = 6 (P1), 30 (P2), and some L <= 362880 with
  |N| < 2^15 for P3."""
import math

import numpy as np
import pytest

import oracle


def reference_tensor(k, im):
    r = oracle.im_rule(im)
    pts, w = np.asarray(r[0]).reshape(-1, 3), np.asarray(r[1])
    nd = (k + 1) * (k + 2) * (k + 3) // 6
    R = np.zeros((nd, nd, 3, 3))
    for q in range(len(w)):
        g = np.asarray(oracle.fem_eval(oracle.FEM_PK, k, pts[q])[1]).reshape(nd, 3)
        R += w[q] * np.einsum("ia,jb->ijab", g, g)
    return R, nd


def moment(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def exact_affine_gradient_tensor(k, nd):
    """grad phi_i(x) = c_i + M_i x for k <= 2; int over the tetrahedron of
    (c_i + M_i x)(c_j + M_j x)^T from the moments of order 0, 1, 2."""
    verts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    G = np.array([np.asarray(oracle.fem_eval(oracle.FEM_PK, k, v)[1]).reshape(nd, 3) for v in verts])
    c = G[0]                                 # (nd, 3)
    M = np.stack([G[1] - G[0], G[2] - G[0], G[3] - G[0]], axis=2)  # (nd, 3 comps, 3 coords)
    m0 = moment(0, 0, 0)
    m1 = np.array([moment(*e) for e in np.eye(3, dtype=int)])
    m2 = np.array([[moment(*(np.eye(3, dtype=int)[p] + np.eye(3, dtype=int)[q])) for q in range(3)]
                   for p in range(3)])
    R = (m0 * np.einsum("ia,jb->ijab", c, c)
         + np.einsum("ia,jbp,p->ijab", c, M, m1)
         + np.einsum("iap,jb,p->ijab", M, c, m1)
         + np.einsum("iap,jbq,pq->ijab", M, M, m2))
    return R


@pytest.mark.parametrize("k,im,L,nmax", [(1, "IM_TETRAHEDRON(1)", 6, 1), (2, "IM_TETRAHEDRON(5)", 30, 8)])
def test_reference_tensor_is_the_exact_integral_and_integral_times_L(k, im, L, nmax):
    R, nd = reference_tensor(k, im)
    Rx = exact_affine_gradient_tensor(k, nd)
    assert np.abs(R - Rx).max() <= 1e-14
    N = R * L
    assert np.abs(N - np.rint(N)).max() <= 1e-12
    assert np.abs(np.rint(N)).max() == nmax


def test_p3_reference_tensor_has_an_int16_form():
    R, nd = reference_tensor(3, "IM_TETRAHEDRON(8)")
    for L in (2520, 5040, 10080, 20160, 40320, 80640, 362880):
        N = R * L
        if np.abs(N - np.rint(N)).max() <= 1e-9 * max(1.0, np.abs(N).max()):
            assert np.abs(np.rint(N)).max() < 2 ** 15
            return
    pytest.fail("no common denominator found for the P3 reference tensor")



def test_an_inexact_rule_with_irrational_points_has_no_integer_form():
    """P3 under IM_TETRAHEDRON(2) (degree 2 < 4, points (5 -+ sqrt 5)/20):
    R is not N/L for any candidate L, so the plan keeps the floating-point
    table -- the case the GPU parity suite runs to cover that path."""
    R, nd = reference_tensor(3, "IM_TETRAHEDRON(2)")
    for L in (1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 30, 36, 40, 45, 48, 60, 72, 80, 90,
              96, 120, 144, 180, 240, 280, 315, 360, 420, 480, 504, 560, 630, 720, 840, 1008, 1120,
              1260, 1440, 1680, 2016, 2240, 2520, 2880, 3360, 4032, 5040, 6720, 7560, 10080, 15120,
              20160, 30240, 40320, 60480, 80640, 120960, 181440, 362880):
        N = R * L
        assert not np.all(np.abs(N - np.rint(N)) <= 1e-9 * np.maximum(1.0, np.abs(N)))
