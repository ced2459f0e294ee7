"""The GWFL compile stage: every spelling the reference's own tests use for
the Laplacian and for linear elasticity (tests/test_assembly.cc:773-884,
MAT_TEST_2) compiles to the same kernel parameters; order-0 and order-1
expressions are derived like ga_workspace::add_tree; terms outside the
subset are rejected with a message."""
import pytest

from getfem_amd import GfxError
from getfem_amd.workspace import gwfl_compile

ELAST = {"u": ("var", 3), "lambda": ("const", 1), "mu": ("const", 1)}
LAP = {"p": ("var", 1)}

ELASTICITY_SPELLINGS = [
    "lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u",
    "(lambda*Trace(Grad_Test_u)*Id(qdim(u)) + mu*(Grad_Test_u'+Grad_Test_u)):Grad_Test2_u",
    "lambda*Div_Test_u*Div_Test2_u + mu*(Grad_Test_u'+Grad_Test_u):Grad_Test2_u",
    "lambda*((Grad_Test2_u@Grad_Test_u):Id(meshdim)):Id(meshdim) + mu*(Grad_Test_u'+Grad_Test_u):Grad_Test2_u",
    "lambda*Id(meshdim)@Id(meshdim)*Grad_Test_u:Grad_Test2_u + mu*(Grad_Test_u'+Grad_Test_u):Grad_Test2_u",
    "lambda*(Id(meshdim)*Id(meshdim))@Id(meshdim)*Grad_Test_u:Grad_Test2_u+ mu*(Grad_Test_u'+Grad_Test_u):Grad_Test2_u",
    "lambda*Trace(Grad_Test_u)*Trace(Grad_Test2_u) +mu*(Grad_Test_u'(:,1)+Grad_Test_u(:,1)):Grad_Test2_u(:,1)"
    "+mu*(Grad_Test_u'(:,2)+Grad_Test_u(:,2)):Grad_Test2_u(:,2)+mu*(Grad_Test_u'(:,3)+Grad_Test_u(:,3)):Grad_Test2_u(:,3)",
    "lambda*Trace(Grad_Test_u)*Trace(Grad_Test2_u) + mu*(Grad_Test_u'(1,:)+Grad_Test_u(1,:)):Grad_Test2_u(1,:)"
    "+ mu*(Grad_Test_u'(2,:)+Grad_Test_u(2,:)):Grad_Test2_u(2,:)+mu*(Grad_Test_u'(3,:)+Grad_Test_u(3,:)):Grad_Test2_u(3,:)",
    "(lambda*Div_u*Div_u + 2*mu*Sym(Grad_u):Sym(Grad_u))/2",
]

LAPLACE_SPELLINGS = [
    "Grad_Test_p:Grad_Test2_p", "(Grad_p:Grad_p)/2", "sqr(Norm(Grad_p))/2", "Norm_sqr(Grad_p)/2",
    "(sqr(Grad_p(1)) + sqr(Grad_p(2)) + sqr(Grad_p(3)))/2",
    "(Grad_p(1)*Grad_p(1) + Grad_p(2)*Grad_p(2)+ Grad_p(3)*Grad_p(3))/2",
    "([Grad_p(1); Grad_p(3); Grad_p(2)].[Grad_p(1); Grad_p(3); Grad_p(2)])/2",
    "Grad_p.Grad_Test_p",
]


@pytest.mark.parametrize("expr", ELASTICITY_SPELLINGS)
def test_elasticity_spellings(expr):
    assert gwfl_compile(expr, ELAST) == "isotropic elasticity unknown=u a/lambda=1*lambda mu=1*mu"


@pytest.mark.parametrize("expr", LAPLACE_SPELLINGS)
def test_laplacian_spellings(expr):
    assert gwfl_compile(expr, LAP) == "laplace unknown=p a/lambda=1"


def test_summed_expressions_and_coefficients():
    d = dict(ELAST, E=("p0", 1))
    got = gwfl_compile(["E*lambda*Div_u*Div_Test_u", "2*E*mu*Sym(Grad_u):Grad_Test_u"], d)
    assert got == "isotropic elasticity unknown=u a/lambda=1*E*lambda mu=1*E*mu"
    got = gwfl_compile("(lambda+2*mu)*Div_u*Div_Test_u + mu*(Grad_u+Grad_u'):Grad_Test_u", ELAST)
    assert got.startswith("isotropic elasticity") and "a/lambda=1*lambda+2*mu" in got


def test_vector_laplacian_and_scaled():
    assert gwfl_compile("3*Grad_u:Grad_Test_u", {"u": ("var", 3)}) == "laplace unknown=u a/lambda=3"


def test_neo_hookean_recognised():
    e = "((Id(meshdim)+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(Grad_u,[lambda;mu]))):Grad_Test_u"
    assert gwfl_compile(e, ELAST).startswith("neo-hookean(bonet)")


@pytest.mark.parametrize("expr,msg", [
    ("Test_u.Test2_u", "value terms"),
    ("Div_u*Div_u*Div_Test_u", "depends on the unknown"),
    ("lambda*Grad_u(1,2)*Grad_Test_u(2,1)", "not one of the supported"),
    ("Grad_v:Grad_Test_u", "unknown name"),
    ("lambda*Div_u*Div_Test_u +", "unexpected end"),
])
def test_rejected(expr, msg):
    with pytest.raises(GfxError) as e:
        gwfl_compile(expr, ELAST)
    assert msg in str(e.value)


# Sums GetFEM merges into one tree (same mesh_im and test functions,
# src/getfem_generic_assembly_workspace.cc:474-495) that are neither a plain
# Laplacian nor isotropic elasticity: the three-coefficient isotropic term
# alpha*Div.Div + beta*(Grad_u)':Grad_Test_u + gamma*Grad_u:Grad_Test_u.
def _coefs(desc):
    out = {}
    for part in desc.split()[2:]:
        k, v = part.split("=")
        out[k] = v
    return out


def test_navier_form_is_isotropic():
    got = gwfl_compile("mu*Grad_u:Grad_Test_u + (lambda+mu)*Div_u*Div_Test_u", ELAST)
    assert got.startswith("isotropic unknown=u")
    c = _coefs(got)
    assert sorted(c["alpha"].split("+")) == ["1*lambda", "1*mu"] and c["gamma"] == "1*mu" and "beta" not in c


def test_laplace_plus_elasticity_merges():
    got = gwfl_compile(["lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", "3*Grad_u:Grad_Test_u"], ELAST)
    c = _coefs(got)
    assert got.startswith("isotropic unknown=u")
    assert c["alpha"] == "1*lambda" and c["beta"] == "1*mu" and sorted(c["gamma"].split("+")) == ["1*mu", "3"]


def test_transposed_gradient_alone():
    assert gwfl_compile("mu*(Grad_u'):Grad_Test_u", ELAST) == "isotropic unknown=u beta=1*mu"


@pytest.mark.parametrize("expr,order", [
    ("Grad_Test2_u:Grad_Test_u", 2),
    ("lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u", 1),
    ("(lambda*Div_u*Div_u + 2*mu*Sym(Grad_u):Sym(Grad_u))/2", 0),
    ("((Id(meshdim)+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(Grad_u,[lambda;mu]))):Grad_Test_u", 1),
    ("lambda*Div_Test_u*Div_Test2_u + mu*Grad_u:Grad_Test_u", 2),
])
def test_expression_order(expr, order):
    """The test-function order add_expression computes before deriving
    (src/getfem_generic_assembly_workspace.cc:545-600): it decides which
    assemblies an expression takes part in."""
    from getfem_amd.workspace import gwfl_order
    assert gwfl_order(expr, ELAST) == order
