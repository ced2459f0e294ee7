"""Integration rules: the oracle and libgfx against the mathematics.

Point counts are the reference's own (IM_TETRAHEDRON(1)=1, (5)=15, (8)=43,
IM_GAUSS_PARALLELEPIPED(3,5)=27, as GetFEM's tables in getfem_im_list.h give them); every
monomial up to the rule degree must integrate exactly on the reference
tetrahedron (a! b! c! / (a+b+c+3)!) or cube (1/((a+1)(b+1)(c+1)))."""
from math import factorial

import numpy as np
import pytest

import oracle
from getfem_amd import core

TET = {1: 1, 2: 4, 3: 5, 5: 15, 6: 24, 8: 43}


def tet_moment(a, b, c):
    return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)


@pytest.mark.parametrize("k", sorted(TET))
def test_tet_rule_exactness(k):
    p, w = oracle.im_rule("IM_TETRAHEDRON(%d)" % k)
    assert len(w) == TET[k]
    assert abs(w.sum() - 1.0 / 6.0) < 1e-15
    for a in range(k + 1):
        for b in range(k + 1 - a):
            for c in range(k + 1 - a - b):
                got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - tet_moment(a, b, c)) < 2e-15, (a, b, c)
    # points inside the reference simplex
    assert np.all(p >= -1e-15) and np.all(p.sum(1) <= 1 + 1e-15)


@pytest.mark.parametrize("k", [1, 3, 5, 7, 9])
def test_gauss_parallelepiped_exactness(k):
    p, w = oracle.im_rule("IM_GAUSS_PARALLELEPIPED(3,%d)" % k)
    m = k // 2 + 1
    assert len(w) == m ** 3
    x, wx = np.polynomial.legendre.leggauss(m)
    assert np.allclose(np.unique(p[:, 0]), np.sort(0.5 + 0.5 * x), atol=1e-15)
    for a in range(k + 1):
        for b in range(k + 1):
            for c in range(k + 1):
                got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - 1.0 / ((a + 1) * (b + 1) * (c + 1))) < 1e-14
    # product order: first coordinate fastest
    if m > 1:
        assert p[1, 0] > p[0, 0] and p[1, 1] == p[0, 1]


def test_tet5_not_exact_at_degree_six():
    """The degree claim is tight (a wrong table entry could only raise it by luck)."""
    p, w = oracle.im_rule("IM_TETRAHEDRON(5)")
    errs = [abs(np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** (6 - a - b)) - tet_moment(a, b, 6 - a - b))
            for a in range(7) for b in range(7 - a)]
    assert max(errs) > 1e-8


@pytest.mark.parametrize("name", ["IM_TETRAHEDRON(1)", "IM_TETRAHEDRON(5)", "IM_TETRAHEDRON(8)",
                                  "IM_GAUSS_PARALLELEPIPED(3,5)"])
def test_library_rules_equal_oracle(name):
    p1, w1 = core.integration_rule(name)
    p2, w2 = oracle.im_rule(name)
    assert np.array_equal(p1, p2) and np.abs(w1 - w2).max() < 1e-17


def test_unknown_rule_rejected():
    with pytest.raises(Exception):
        core.integration_rule("IM_TETRAHEDRON(4)")
