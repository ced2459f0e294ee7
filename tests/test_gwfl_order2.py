"""The order-2 compile stage of the polynomial route (ga_workspace::add_tree
deriving expressions twice, src/getfem_generic_assembly_workspace.cc:
439-543) on the CPU: the programs of each (test variable, trial variable)
block, evaluated at random symbol values, against central differences of the
order-1 programs with respect to the trial variable's symbols, and against
closed forms (isotropic elasticity's tensor, the second derivatives of det
and of the Neumann series of Inv at zero)."""
import numpy as np
import pytest

from tests.test_gwfl_scalar import compile_scalar, evaluate


def sym_names(var, q):
    """symbol name of each slot of variable var (value a, gradient q + 3a + b)"""
    sa = (lambda a: "") if q == 1 else (lambda a: str(a))
    names = ["%s:V_%s" % (var, sa(a)) for a in range(q)]
    for a in range(q):
        names += ["%s:G_%s%d" % (var, sa(a), b) for b in range(3)]
    return names


def random_values(rng, decls):
    v = {}
    for d in decls.split(";"):
        name, kind, q = d.split(":")
        q = int(q)
        if kind == "const":
            v["c:" + name] = rng.uniform(0.5, 2.0)
        else:
            for s in sym_names(name, q):
                v[s] = rng.uniform(-1.0, 1.0)
    return v


def hessian_block_fd(exprs, decls, tv, qt, sv, qs, values, h=1e-6):
    """d(order-1 program of tv)/d(symbols of sv), central differences"""
    p1 = compile_scalar(exprs, decls, 1, tv)
    D = np.zeros((4 * qt, 4 * qs))
    for beta, s in enumerate(sym_names(sv, qs)):
        vp, vm = dict(values), dict(values)
        vp[s] = values.get(s, 0.0) + h
        vm[s] = values.get(s, 0.0) - h
        D[:, beta] = (evaluate(p1, vp) - evaluate(p1, vm)) / (2 * h)
    return D


def block(exprs, decls, tv, qt, sv, qs, values):
    prog = compile_scalar(exprs, decls, 2, "%s,%s" % (tv, sv))
    assert prog[0] == 16 * qt * qs
    return evaluate(prog, values).reshape(4 * qt, 4 * qs)


@pytest.mark.parametrize("seed", range(3))
def test_nonlinear_scalar_block_is_the_derivative_of_the_residual(seed):
    exprs = ["Grad_u.Grad_Test_u + c*u*u*u*Test_u - f*Test_u"]
    decls = "u:var:1;c:const:1;f:const:1"
    v = random_values(np.random.default_rng(seed), decls)
    D = block(exprs, decls, "u", 1, "u", 1, v)
    assert np.abs(D - hessian_block_fd(exprs, decls, "u", 1, "u", 1, v)).max() <= 1e-8
    assert D[0, 0] == pytest.approx(3 * v["c:c"] * v["u:V_"] ** 2, rel=1e-14)
    assert np.array_equal(D[1:, 1:], np.eye(3))


@pytest.mark.parametrize("seed", range(3))
def test_two_variable_potential_blocks(seed):
    """An order-0 potential in u (P2 vector) and p (scalar), derived twice:
    the four blocks are the Hessian, (u,p) the transpose of (p,u)."""
    exprs = ["0.5*sqr(Div_u - p) + k*p*p*p + Grad_p.Grad_p*Trace(Grad_u)"]
    decls = "u:var:3;p:var:1;k:const:1"
    v = random_values(np.random.default_rng(10 + seed), decls)
    for tv, qt in (("u", 3), ("p", 1)):
        for sv, qs in (("u", 3), ("p", 1)):
            D = block(exprs, decls, tv, qt, sv, qs, v)
            assert np.abs(D - hessian_block_fd(exprs, decls, tv, qt, sv, qs, v)).max() <= 1e-7
    assert np.array_equal(block(exprs, decls, "u", 3, "p", 1, v), block(exprs, decls, "p", 1, "u", 3, v).T)


def test_order1_expressions_join_their_test_variable_block():
    """-p Div(Test_u) - Test_p Div(u): block (u,p) from the first, (p,u) from
    the second, (u,u) and (p,p) empty."""
    exprs = ["-p*Div_Test_u", "-Test_p*Div_u"]
    decls = "u:var:3;p:var:1"
    v = random_values(np.random.default_rng(0), decls)
    Dup = block(exprs, decls, "u", 3, "p", 1, v)
    Dpu = block(exprs, decls, "p", 1, "u", 3, v)
    expect = np.zeros((12, 4))
    for a in range(3):
        expect[3 + 4 * a, 0] = -1.0
    assert np.array_equal(Dup, expect)
    assert np.array_equal(Dpu, expect.T)
    for tv, qt in (("u", 3), ("p", 1)):
        assert compile_scalar(exprs, decls, 2, "%s,%s" % (tv, tv))[3] == []


def test_isotropic_elasticity_tensor():
    exprs = ["lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u"]
    decls = "u:var:3;lambda:const:1;mu:const:1"
    v = random_values(np.random.default_rng(1), decls)
    D = block(exprs, decls, "u", 3, "u", 3, v)
    lam, mu = v["c:lambda"], v["c:mu"]
    d = np.eye(3)
    C = (lam * np.einsum("ab,cd->abcd", d, d) + mu * (np.einsum("ac,bd->abcd", d, d)
                                                       + np.einsum("ad,bc->abcd", d, d)))
    assert np.abs(D[3:, 3:] - C.reshape(9, 9)).max() <= 1e-14
    assert np.all(D[:3] == 0) and np.all(D[:, :3] == 0)


def test_det_and_inv_second_derivatives_at_zero():
    """d2 det(I + G) / dG_ab dG_cd = d_ab d_cd - d_ad d_cb and
    d2 tr Inv(I + G) / dG_ab dG_cd = 2 d_ad d_bc at G = 0: the Neumann series
    kept to degree 2 for an order-2 assembly."""
    d = np.eye(3)
    v = {s: 0.0 for s in sym_names("w", 3)}
    Dd = block(["Det(Id(meshdim)+Grad_w)"], "w:var:3", "w", 3, "w", 3, v)
    assert np.abs(Dd[3:, 3:] - (np.einsum("ab,cd->abcd", d, d) - np.einsum("ad,cb->abcd", d, d)).reshape(9, 9)
                  ).max() <= 1e-14
    Di = block(["Trace(Inv(Id(meshdim)+Grad_w))"], "w:var:3", "w", 3, "w", 3, v)
    assert np.abs(Di[3:, 3:] - 2 * np.einsum("ad,bc->abcd", d, d).reshape(9, 9)).max() <= 1e-14


def test_expression_with_two_test_variables_splits_by_block():
    """kappa (u - v)(Test_u - Test_v): GetFEM makes one tree per test
    variable; each block keeps the monomials of its own test functions."""
    exprs = ["kappa*(u-v)*(Test_u-Test_v)"]
    decls = "u:var:1;v:var:1;kappa:const:1"
    v = random_values(np.random.default_rng(4), decls)
    k, du = v["c:kappa"], v["u:V_"] - v["v:V_"]
    assert evaluate(compile_scalar(exprs, decls, 1, "u"), v)[0] == pytest.approx(k * du, rel=1e-14)
    assert evaluate(compile_scalar(exprs, decls, 1, "v"), v)[0] == pytest.approx(-k * du, rel=1e-14)
    for tv, sv, sign in (("u", "u", 1), ("u", "v", -1), ("v", "u", -1), ("v", "v", 1)):
        D = block(exprs, decls, tv, 1, sv, 1, v)
        expect = np.zeros((4, 4))
        expect[0, 0] = sign * k
        assert np.abs(D - expect).max() <= 1e-15


def test_explicit_vectors_and_matrices():
    """GWFL's explicit tensors (doc/sphinx/source/userdoc/gasm_high.rst:
    511-520): [a,b,c] and [a;b;c] are vectors, [a,b;c,d] a matrix by lines
    and [[a,c],[b,d]] the same matrix by columns."""
    decls = "u:var:1;bx:const:1;by:const:1;bz:const:1"
    v = random_values(np.random.default_rng(2), decls)
    expect = block(["(bx*Grad_u(1)+by*Grad_u(2)+bz*Grad_u(3))*Test_u"], decls, "u", 1, "u", 1, v)
    assert np.any(expect != 0)
    for e in ["([bx,by,bz].Grad_u)*Test_u", "([bx;by;bz].Grad_u)*Test_u"]:
        assert np.array_equal(block([e], decls, "u", 1, "u", 1, v), expect)
    cols = block(["([[1,2],[3,4]]:[1,3;2,4])*u*Test_u"], "u:var:1", "u", 1, "u", 1, {"u:V_": 0.0})
    rows = block(["([[1,2],[3,4]]:[1,2;3,4])*u*Test_u"], "u:var:1", "u", 1, "u", 1, {"u:V_": 0.0})
    assert cols[0, 0] == 30.0 and rows[0, 0] == 29.0
