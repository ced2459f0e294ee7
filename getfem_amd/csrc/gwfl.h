// GWFL compile stage for the supported subset: parse an expression, evaluate
// it symbolically (every tensor entry a polynomial in the components of
// Grad_u, Grad_Test_u, Grad_Test2_u and the coefficient names), differentiate
// it with respect to the unknown up to order 2 (the role of ga_derivative,
// src/getfem_generic_assembly_semantic.cc:3965, called from
// ga_workspace::add_tree, src/getfem_generic_assembly_workspace.cc:520-541),
// and classify the order-2 integrand into the parameters of an assembly
// kernel.  This replaces the instruction lists that ga_compile emits
// (src/getfem_generic_assembly_compile_and_exec.cc:7910-8643).
#pragma once
#include <map>
#include <string>
#include <vector>

namespace gfx {

struct GwflVar {
  std::string name;
  bool is_fem = false;       // fem variable / constant (otherwise fixed size)
  bool is_constant = false;  // not differentiated
  int qdim = 1;              // components of the field (fem) / size (fixed)
  bool per_element = false;  // fem constant on a P0 element (one value per element)
};

// coefficient of the kernel: value = sum_k scale_k * prod(names_k)
struct GwflCoef {
  std::vector<std::pair<double, std::vector<std::string>>> terms;
};

struct GwflKernelParams {
  int term = -1;           // GFX_TERM_*
  std::string unknown;     // name of the variable
  GwflCoef coef[3];        // LAPLACE: a ; ISO_ELAST: lambda, mu ; ISOTROPIC: alpha, beta, gamma
  std::string describe;    // human-readable normal form
};

// The test-function order of one expression (0: potential, 1: residual form
// with Test_u, 2: bilinear form with Test_u and Test2_u), the number of
// test-function kinds it contains, as ga_workspace::add_expression computes
// it before deriving the higher orders (src/getfem_generic_assembly_
// workspace.cc:545-600).  Returns -1 with a message on a parse error.
int gwfl_expression_order(const std::string &expr, const std::map<std::string, GwflVar> &vars,
                          std::string &err);

// Compile the sum of expressions (each with its derivation order request) for
// an order-2 assembly.  Returns false with a message on unsupported input.
bool gwfl_compile_order2(const std::vector<std::string> &exprs,
                         const std::map<std::string, GwflVar> &vars, GwflKernelParams &out,
                         std::string &err);

// The order-0 / order-1 route (ga_workspace::assembly(0 | 1) of potentials
// and their derivatives, src/getfem_generic_assembly_workspace.cc:520-541,
// :832-847): the integrand as a polynomial program over data symbols (values
// "f:V_a" and gradients "f:G_ab" of fem fields at the point, constants
// "c:name").  order 0: one polynomial, the potential's integrand.  order 1,
// variable `var` (qdim q): the derivative's coefficients of Test_var
// (outputs 0..q-1) and of Grad_Test_var (output q + 3a + b); order-0
// expressions are differentiated, order-1 expressions with Test_var join as
// they are.  order 2, the block of test variable `var` (qdim q) and trial
// variable `var2` (qdim q2): the coefficients of the products of a Test_var
// slot (0..q-1 the value, q + 3a + b the gradient) and a Test2_var2 slot
// (the same numbering with q2), output slot1 * 4 q2 + slot2; order-0
// expressions are differentiated twice, order-1 expressions in Test_var once
// along Test2_var2, and order-2 expressions in Test_var and Test2_var2 join
// as they are (ga_workspace::add_tree, src/getfem_generic_assembly_
// workspace.cc:439-543).  Variables inside Inv(...) are expanded around zero
// (zero_vars: their values must be zero at assembly).
struct GwflScalarProgram {
  int order = 0;
  std::string var, var2;
  int nout = 1;
  std::vector<std::string> syms;
  struct Mono {
    double coef = 0.0;
    std::vector<int> s;  // indices into syms
    int out = 0;
  };
  std::vector<Mono> monos;
  std::vector<std::string> zero_vars;
};
bool gwfl_compile_scalar(const std::vector<std::string> &exprs, const std::map<std::string, GwflVar> &vars,
                         int order, const std::string &var, GwflScalarProgram &out, std::string &err,
                         const std::string &var2 = "");

}  // namespace gfx
