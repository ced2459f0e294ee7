"""The order-0 / order-1 compile stage (ga_workspace::assembly(0 | 1) of a
potential and of its derivatives, src/getfem_generic_assembly_workspace.cc:
520-541, :832-847) on the CPU: the polynomial programs it emits, evaluated
at random symbol values, against the integrands computed independently in
numpy, and the derivatives against central finite differences of the
numpy integrand.  The pulled-back coordinates route (F = Id + Grad_w at
w = 0, Inv and Det, SURVEY of the adjoint route) is checked the same way."""
import ctypes

import numpy as np
import pytest

from getfem_amd import _lib

EMIN = 1e-3
SIMP = "(%g + pow(rho,3)*(1-%g))" % (EMIN, EMIN)
ELAST = "(lambda*sqr(Div_u) + 2*mu*Sym(Grad_u):Sym(Grad_u))"
DECLS = "u:data:3;rho:p0var:1;lambda:const:1;mu:const:1"
FINV = "Inv(Id(meshdim)+Grad_w)"
PULLED = ("%s*(lambda*sqr(Trace(Grad_u*%s)) + 2*mu*Sym(Grad_u*%s):Sym(Grad_u*%s))*Det(Id(meshdim)+Grad_w)"
          % (SIMP, FINV, FINV, FINV))


def compile_scalar(exprs, decls, order, var=""):
    buf = ctypes.create_string_buffer(1 << 22)
    code = _lib.lib().gfx_gwfl_compile_scalar("\n".join(exprs).encode(), decls.encode(), order,
                                              var.encode(), buf, len(buf))
    text = buf.value.decode()
    if code:
        raise _lib.GfxError(code, text)
    lines = text.split("\n")
    head = dict(kv.split("=", 1) for kv in lines[0].split())
    syms = lines[1][len("syms="):].split()
    monos = []
    for ln in lines[2:]:
        f = ln.split()
        monos.append((int(f[0]), float(f[1]), [int(x) for x in f[2:]]))
    return int(head["nout"]), [z for z in head["zero"].split(",") if z], syms, monos


def evaluate(prog, values):
    nout, _, syms, monos = prog
    out = np.zeros(nout)
    for o, c, s in monos:
        v = c
        for k in s:
            v *= values[syms[k]]
        out[o] += v
    return out


def symbol_values(H, rho, lam, mu, Gw=None):
    v = {"c:lambda": lam, "c:mu": mu, "rho:V_": rho}
    for a in range(3):
        for b in range(3):
            v["u:G_%d%d" % (a, b)] = H[a, b]
            if Gw is not None:
                v["w:G_%d%d" % (a, b)] = Gw[a, b]
    return v


def energy(H, rho, lam, mu):
    E = EMIN + rho ** 3 * (1 - EMIN)
    S = 0.5 * (H + H.T)
    return E * (lam * np.trace(H) ** 2 + 2 * mu * np.sum(S * S))


def pulled_energy(H, Gw, rho, lam, mu):
    F = np.eye(3) + Gw
    return energy(H @ np.linalg.inv(F), rho, lam, mu) * np.linalg.det(F)


@pytest.fixture
def sample():
    rng = np.random.default_rng(7)
    return rng.standard_normal((3, 3)), rng.uniform(0.1, 1.0), 0.5769, 0.3846


def test_potential(sample):
    H, rho, lam, mu = sample
    prog = compile_scalar([SIMP + "*" + ELAST], DECLS, 0)
    assert prog[0] == 1 and prog[1] == []
    got = evaluate(prog, symbol_values(H, rho, lam, mu))[0]
    assert abs(got - energy(H, rho, lam, mu)) <= 1e-13 * abs(got)


def test_derivative_wrt_p0_variable(sample):
    """d/drho in the direction Test_rho: E'(rho) (lambda div^2 + 2 mu eps:eps)."""
    H, rho, lam, mu = sample
    prog = compile_scalar([SIMP + "*" + ELAST], DECLS, 1, "rho")
    out = evaluate(prog, symbol_values(H, rho, lam, mu))
    assert prog[0] == 4 and np.all(out[1:] == 0.0)
    h = 1e-6
    fd = (energy(H, rho + h, lam, mu) - energy(H, rho - h, lam, mu)) / (2 * h)
    closed = 3 * rho ** 2 * (1 - EMIN) * energy(H, 1.0, lam, mu) / 1.0
    assert abs(out[0] - closed) <= 1e-12 * abs(closed)
    assert abs(out[0] - fd) <= 1e-7 * abs(fd)


def test_derivative_wrt_vector_variable(sample):
    """u a fem variable: the coefficients of Grad_Test_u are
    E (2 lambda tr(H) I + 4 mu Sym(H)), none on Test_u."""
    H, rho, lam, mu = sample
    prog = compile_scalar([SIMP + "*" + ELAST], DECLS.replace("u:data", "u:var"), 1, "u")
    out = evaluate(prog, symbol_values(H, rho, lam, mu))
    E = EMIN + rho ** 3 * (1 - EMIN)
    want = E * (2 * lam * np.trace(H) * np.eye(3) + 2 * mu * (H + H.T))
    assert prog[0] == 12 and np.all(out[:3] == 0.0)
    assert np.abs(out[3:].reshape(3, 3) - want).max() <= 1e-13 * np.abs(want).max()


def test_order_one_expression_joins_its_block():
    """An expression already of order 1 in Test_u joins u's block as it is;
    it is absent from another variable's block and from the potential."""
    d = "u:var:3;rho:p0var:1"
    prog = compile_scalar(["Grad_u:Grad_Test_u", "rho*rho"], d, 1, "u")
    vals = {"u:G_%d%d" % (a, b): 10 * a + b + 1.0 for a in range(3) for b in range(3)}
    vals["rho:V_"] = 2.0
    out = evaluate(prog, vals)
    assert np.all(out[3:] == [vals["u:G_%d%d" % (a, b)] for a in range(3) for b in range(3)])
    prog_rho = compile_scalar(["Grad_u:Grad_Test_u", "rho*rho"], d, 1, "rho")
    assert evaluate(prog_rho, vals)[0] == 4.0
    prog0 = compile_scalar(["Grad_u:Grad_Test_u", "rho*rho"], d, 0)
    assert evaluate(prog0, vals)[0] == 4.0


def test_pulled_back_route_at_zero(sample):
    """F = Id + Grad_w at w = 0: the potential is the plain one, and the
    coefficients of Grad_Test_w are the shape derivative, equal to central
    differences of det(F) W(H F^-1) in numpy (exact inverse)."""
    H, rho, lam, mu = sample
    decls = DECLS + ";w:var:3"
    prog0 = compile_scalar([PULLED], decls, 0)
    assert prog0[1] == ["w"]
    assert abs(evaluate(prog0, symbol_values(H, rho, lam, mu))[0] - energy(H, rho, lam, mu)) <= \
        1e-13 * energy(H, rho, lam, mu)
    prog = compile_scalar([PULLED], decls, 1, "w")
    out = evaluate(prog, symbol_values(H, rho, lam, mu, Gw=np.zeros((3, 3))))
    h = 1e-6
    fd = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            G = np.zeros((3, 3))
            G[a, b] = h
            fd[a, b] = (pulled_energy(H, G, rho, lam, mu) - pulled_energy(H, -G, rho, lam, mu)) / (2 * h)
    assert np.all(out[:3] == 0.0)
    assert np.abs(out[3:].reshape(3, 3) - fd).max() <= 1e-7 * np.abs(fd).max()
    # closed form of the shape derivative: W I - H^T P, P = dW/dH
    E = EMIN + rho ** 3 * (1 - EMIN)
    P = E * (2 * lam * np.trace(H) * np.eye(3) + 2 * mu * (H + H.T))
    closed = energy(H, rho, lam, mu) * np.eye(3) - H.T @ P
    assert np.abs(out[3:].reshape(3, 3) - closed).max() <= 1e-12 * np.abs(closed).max()
    # the same potential differentiated with respect to rho keeps w at zero
    prog_rho = compile_scalar([PULLED], decls, 1, "rho")
    ref = compile_scalar([SIMP + "*" + ELAST], DECLS, 1, "rho")
    v = symbol_values(H, rho, lam, mu)
    assert abs(evaluate(prog_rho, v)[0] - evaluate(ref, v)[0]) <= 1e-13 * abs(evaluate(ref, v)[0])


def test_det_and_pow_closed_forms():
    d = "u:var:3;k:const:1"
    prog = compile_scalar(["Det(Grad_u)", "pow(k,4)"], d, 0)
    rng = np.random.default_rng(3)
    H = rng.standard_normal((3, 3))
    v = {"u:G_%d%d" % (a, b): H[a, b] for a in range(3) for b in range(3)}
    v["c:k"] = 1.7
    assert abs(evaluate(prog, v)[0] - (np.linalg.det(H) + 1.7 ** 4)) <= 1e-13


@pytest.mark.parametrize("expr,decls,order,var,msg", [
    ("Inv(Grad_u):Grad_u", "u:data:3", 0, "", "Inv is supported"),
    ("pow(rho,0.5)", "rho:p0var:1", 0, "", "integer exponents"),
    ("Grad_Test_u:Grad_Test2_u", "u:var:3", 1, "u", None),
    ("rho*Test_rho", "rho:p0var:1;u:data:3", 1, "u", "not a fem variable"),
])
def test_rejected_or_empty(expr, decls, order, var, msg):
    if msg is None:  # an order-2 expression has no order-1 tree: empty program
        prog = compile_scalar([expr], decls, order, var)
        assert prog[3] == []
        return
    with pytest.raises(_lib.GfxError) as e:
        compile_scalar([expr], decls, order, var)
    assert msg in str(e.value)
