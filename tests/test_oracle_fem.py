"""Lagrange elements: Kronecker property at the nodes, partition of unity,
gradients against finite differences, closed-form P2 bases, node order of
the reference (first reference coordinate fastest)."""
import numpy as np
import pytest

import oracle
from getfem_amd import core

ELEMS = [(oracle.FEM_PK, 1, "FEM_PK(3,1)"), (oracle.FEM_PK, 2, "FEM_PK(3,2)"),
         (oracle.FEM_PK, 3, "FEM_PK(3,3)"), (oracle.FEM_QK, 1, "FEM_QK(3,1)"),
         (oracle.FEM_QK, 2, "FEM_QK(3,2)")]


@pytest.mark.parametrize("kind,k,name", ELEMS)
def test_kronecker_and_unity(kind, k, name):
    nodes = oracle.fem_nodes(kind, k)
    for j, x in enumerate(nodes):
        v, _ = oracle.fem_eval(kind, k, x)
        e = np.zeros(len(nodes))
        e[j] = 1.0
        assert np.allclose(v, e, atol=1e-14)
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.uniform(0, 0.3, 3)
        v, g = oracle.fem_eval(kind, k, x)
        assert abs(v.sum() - 1) < 1e-13 and np.abs(g.sum(0)).max() < 1e-12


@pytest.mark.parametrize("kind,k,name", ELEMS)
def test_gradients_finite_differences(kind, k, name):
    x = np.array([0.21, 0.17, 0.29])
    _, g = oracle.fem_eval(kind, k, x)
    h = 1e-6
    for c in range(3):
        d = np.zeros(3)
        d[c] = h
        fd = (oracle.fem_eval(kind, k, x + d)[0] - oracle.fem_eval(kind, k, x - d)[0]) / (2 * h)
        assert np.allclose(fd, g[:, c], atol=1e-8)


@pytest.mark.parametrize("kind,k,name", ELEMS)
def test_library_bases_equal_oracle(kind, k, name):
    for x in ([0.1, 0.2, 0.3], [0.25, 0.25, 0.25], [0.6, 0.1, 0.05]):
        v1, g1 = core.fem_eval(name, x)
        v2, g2 = oracle.fem_eval(kind, k, x)
        assert np.allclose(v1, v2, atol=1e-15) and np.allclose(g1, g2, atol=1e-14)


def test_p2_closed_form_and_node_order():
    # node order v0, m01, v1, m02, m12, v2, m03, m13, m23, v3
    nodes = oracle.fem_nodes(oracle.FEM_PK, 2)
    expect = np.array([[0, 0, 0], [.5, 0, 0], [1, 0, 0], [0, .5, 0], [.5, .5, 0], [0, 1, 0],
                       [0, 0, .5], [.5, 0, .5], [0, .5, .5], [0, 0, 1]])
    assert np.array_equal(nodes, expect)
    x = np.array([0.13, 0.31, 0.22])
    l = np.array([1 - x.sum(), x[0], x[1], x[2]])
    vert = lambda a: l[a] * (2 * l[a] - 1)  # noqa: E731
    edge = lambda a, b: 4 * l[a] * l[b]  # noqa: E731
    closed = [vert(0), edge(0, 1), vert(1), edge(0, 2), edge(1, 2), vert(2), edge(0, 3), edge(1, 3),
              edge(2, 3), vert(3)]
    v, _ = oracle.fem_eval(oracle.FEM_PK, 2, x)
    assert np.allclose(v, closed, atol=1e-15)


def test_q2_product_order():
    nodes = oracle.fem_nodes(oracle.FEM_QK, 2)
    assert nodes.shape == (27, 3)
    assert np.array_equal(nodes[1], [0.5, 0, 0]) and np.array_equal(nodes[3], [0, 0.5, 0])
    assert np.array_equal(nodes[9], [0, 0, 0.5])
