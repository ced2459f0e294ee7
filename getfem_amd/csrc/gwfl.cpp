// GWFL compile stage (subset): symbolic evaluation, derivation and
// classification of order-2 integrands.  Grammar and priorities follow
// GetFEM's parser (src/getfem_generic_assembly_tree.cc:36-66): '+' '-' (1),
// '*' '/' ':' '.' '@' (2, left associative), postfix quote and unary minus
// (3), Sym/Skew/Trace and other functions (4).
#include "gwfl.h"

#include <algorithm>
#include <cctype>
#include <memory>
#include <cmath>
#include <regex>
#include <set>
#include <sstream>

#include "../../include/gfx.h"

namespace gfx {

namespace {

// ----------------------------------------------------------- polynomials
using Mono = std::vector<int>;  // sorted symbol ids
using Poly = std::map<Mono, double>;

Poly pconst(double c) {
  Poly p;
  if (c != 0.0) p[Mono()] = c;
  return p;
}
Poly psym(int s) {
  Poly p;
  p[Mono{s}] = 1.0;
  return p;
}
void padd_to(Poly &a, const Poly &b, double s = 1.0) {
  for (auto &kv : b) {
    double &v = a[kv.first];
    v += s * kv.second;
    if (v == 0.0) a.erase(kv.first);
  }
}
// Truncation of the polynomials at a variable held at zero: the values and
// gradients of a variable that appears inside Inv(...) are expanded around 0
// (Inv is a Neumann series there), and an order-k assembly (k = 0, 1, 2) at
// that variable's zero value needs only the monomials of degree <= max(k, 1)
// in its symbols.  `zsym[s]` marks those symbols; monomials of higher degree in them
// are dropped by every product.
struct Trunc {
  std::vector<char> zsym;
  int max_deg = -1;  // -1: no truncation
  bool drops(const Mono &m) const {
    if (max_deg < 0) return false;
    int d = 0;
    for (int s : m)
      if (s < (int)zsym.size() && zsym[s] && ++d > max_deg) return true;
    return false;
  }
};
thread_local const Trunc *g_trunc = nullptr;

Poly pmul(const Poly &a, const Poly &b) {
  Poly r;
  for (auto &x : a)
    for (auto &y : b) {
      Mono m = x.first;
      m.insert(m.end(), y.first.begin(), y.first.end());
      std::sort(m.begin(), m.end());
      if (g_trunc && g_trunc->drops(m)) continue;
      double &v = r[m];
      v += x.second * y.second;
      if (v == 0.0) r.erase(m);
    }
  return r;
}
// derivative along  sum_s  dP/ds * dir(s)
Poly pderiv(const Poly &p, const std::map<int, int> &dir) {
  Poly r;
  for (auto &kv : p) {
    const Mono &m = kv.first;
    for (size_t i = 0; i < m.size(); ++i) {
      auto it = dir.find(m[i]);
      if (it == dir.end()) continue;
      if (i > 0 && m[i] == m[i - 1]) continue;  // each distinct symbol once
      int mult = 0;
      for (int s : m) mult += (s == m[i]);
      Mono n = m;
      n.erase(n.begin() + i);
      n.push_back(it->second);
      std::sort(n.begin(), n.end());
      double &v = r[n];
      v += mult * kv.second;
      if (v == 0.0) r.erase(n);
    }
  }
  return r;
}
bool pequal(const Poly &a, const Poly &b) {
  Poly d = a;
  padd_to(d, b, -1.0);
  double mx = 0.0, sc = 0.0;
  for (auto &kv : d) mx = std::max(mx, std::fabs(kv.second));
  for (auto &kv : a) sc = std::max(sc, std::fabs(kv.second));
  for (auto &kv : b) sc = std::max(sc, std::fabs(kv.second));
  return mx <= 1e-13 * std::max(sc, 1e-300);
}

// --------------------------------------------------------------- tensors
struct Tensor {
  std::vector<int> shape;  // {}, {n}, {n,m}, {n,m,p,q}
  std::vector<Poly> v;
  std::shared_ptr<Tensor> norm_arg;  // set by Norm(): only sqr(Norm(x)) is polynomial
  size_t size() const { return v.size(); }
  int order() const { return (int)shape.size(); }
};

Tensor scalar(const Poly &p) {
  Tensor t;
  t.v.push_back(p);
  return t;
}
Tensor zeros(std::vector<int> shape) {
  Tensor t;
  t.shape = shape;
  size_t n = 1;
  for (int s : shape) n *= (size_t)s;
  t.v.assign(n, Poly());
  return t;
}

struct Error {
  std::string msg;
};
[[noreturn]] void fail(const std::string &m) { throw Error{m}; }

Tensor tadd(const Tensor &a, const Tensor &b, double s) {
  if (a.shape != b.shape) fail("addition of tensors of different sizes");
  Tensor r = a;
  r.norm_arg.reset();
  for (size_t i = 0; i < r.size(); ++i) padd_to(r.v[i], b.v[i], s);
  return r;
}
Tensor tscale(const Tensor &a, const Poly &s) {
  Tensor r = zeros(a.shape);
  for (size_t i = 0; i < a.size(); ++i) r.v[i] = pmul(a.v[i], s);
  return r;
}
// contraction of the last k indices of a with the first k of b
Tensor tcontract(const Tensor &a, const Tensor &b, int k) {
  if (a.order() < k || b.order() < k) fail("contraction on too many indices");
  for (int i = 0; i < k; ++i)
    if (a.shape[a.order() - k + i] != b.shape[i]) fail("contraction of incompatible sizes");
  std::vector<int> sh(a.shape.begin(), a.shape.end() - k);
  sh.insert(sh.end(), b.shape.begin() + k, b.shape.end());
  Tensor r = zeros(sh);
  size_t inner = 1, outer_b = 1;
  for (int i = 0; i < k; ++i) inner *= (size_t)b.shape[i];
  for (int i = k; i < b.order(); ++i) outer_b *= (size_t)b.shape[i];
  const size_t outer_a = a.size() / inner;
  for (size_t i = 0; i < outer_a; ++i)
    for (size_t j = 0; j < outer_b; ++j) {
      Poly s;
      for (size_t l = 0; l < inner; ++l) padd_to(s, pmul(a.v[i * inner + l], b.v[l * outer_b + j]));
      r.v[i * outer_b + j] = s;
    }
  return r;
}
Tensor touter(const Tensor &a, const Tensor &b) {
  std::vector<int> sh = a.shape;
  sh.insert(sh.end(), b.shape.begin(), b.shape.end());
  Tensor r = zeros(sh);
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < b.size(); ++j) r.v[i * b.size() + j] = pmul(a.v[i], b.v[j]);
  return r;
}
Tensor ttranspose(const Tensor &a) {
  if (a.order() != 2) fail("transpose of a non-matrix");
  Tensor r = zeros({a.shape[1], a.shape[0]});
  for (int i = 0; i < a.shape[0]; ++i)
    for (int j = 0; j < a.shape[1]; ++j) r.v[j * a.shape[0] + i] = a.v[i * a.shape[1] + j];
  return r;
}

// --------------------------------------------------------------- parser
// the variable a symbol belongs to and its tag: "u:G_01" -> ("u", "G")
void split_symbol(const std::string &nm, std::string &var, std::string &tag) {
  const size_t c = nm.find(':');
  var = nm.substr(0, c);
  tag = c == std::string::npos ? "" : nm.substr(c + 1, nm.find('_', c) - c - 1);
}

struct Ctx {
  const std::map<std::string, GwflVar> *vars;
  std::vector<std::string> symnames;
  std::map<std::string, int> symid;
  // variables held at zero (they appear inside Inv) and the truncation of
  // products in their value/gradient symbols
  std::set<std::string> zero_vars;
  Trunc trunc;
  bool collect_inv = false;  // first pass: Inv only records its variables
  int sym(const std::string &n) {
    auto it = symid.find(n);
    if (it != symid.end()) return it->second;
    symid[n] = (int)symnames.size();
    symnames.push_back(n);
    std::string var, tag;
    split_symbol(n, var, tag);
    trunc.zsym.push_back(zero_vars.count(var) && (tag == "V" || tag == "G"));
    return symid[n];
  }
};

struct Parser {
  const std::string &s;
  size_t pos = 0;
  Ctx &C;
  Parser(const std::string &str, Ctx &c) : s(str), C(c) {}

  void skip() {
    while (pos < s.size() && std::isspace((unsigned char)s[pos])) ++pos;
  }
  bool eat(char c) {
    skip();
    if (pos < s.size() && s[pos] == c) { ++pos; return true; }
    return false;
  }
  char peek() {
    skip();
    return pos < s.size() ? s[pos] : '\0';
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "' at position " + std::to_string(pos));
  }

  Tensor expr() {
    Tensor a = term();
    for (;;) {
      if (eat('+')) a = tadd(a, term(), 1.0);
      else if (peek() == '-') { ++pos; a = tadd(a, term(), -1.0); }
      else return a;
    }
  }
  Tensor term() {
    Tensor a = unary();
    for (;;) {
      char c = peek();
      if (c == '*') { ++pos; a = mul(a, unary()); }
      else if (c == '/') {
        ++pos;
        Tensor b = unary();
        if (b.order() != 0 || b.v[0].size() > 1 || (b.v[0].size() == 1 && !b.v[0].begin()->first.empty()))
          fail("division by a non-constant");
        if (b.v[0].empty()) fail("division by zero");
        a = tscale(a, pconst(1.0 / b.v[0].begin()->second));
      } else if (c == ':') { ++pos; a = colon(a, unary()); }
      else if (c == '.' && !(pos + 1 < s.size() && (s[pos + 1] == '*' || s[pos + 1] == '/'))) {
        ++pos;
        a = tcontract(a, unary(), 1);
      } else if (c == '@') { ++pos; a = touter(a, unary()); }
      else return a;
    }
  }
  static Tensor mul(const Tensor &a, const Tensor &b) {
    if (a.order() == 0) return tscale(b, a.v[0]);
    if (b.order() == 0) return tscale(a, b.v[0]);
    if (a.order() == 2 && (b.order() == 2 || b.order() == 1)) return tcontract(a, b, 1);
    if (a.order() == 4 && b.order() == 2) return tcontract(a, b, 2);
    fail("unsupported '*' between tensors of orders " + std::to_string(a.order()) + " and " +
         std::to_string(b.order()));
  }
  static Tensor colon(const Tensor &a, const Tensor &b) {
    if (a.order() >= 2 && b.order() >= 2) return tcontract(a, b, 2);
    if (a.order() == 1 && b.order() == 1) return tcontract(a, b, 1);
    fail("unsupported ':' between tensors of orders " + std::to_string(a.order()) + " and " +
         std::to_string(b.order()));
  }
  Tensor unary() {
    if (eat('-')) {
      Tensor a = unary();
      return tscale(a, pconst(-1.0));
    }
    if (eat('+')) return unary();
    return postfix();
  }
  Tensor postfix() {
    Tensor a = primary();
    for (;;) {
      if (peek() == '\'') { ++pos; a = ttranspose(a); }
      else if (peek() == '(') a = index(a);
      else return a;
    }
  }
  Tensor index(const Tensor &a) {
    expect('(');
    std::vector<int> idx;  // -1 = ':'
    for (;;) {
      skip();
      if (peek() == ':') { ++pos; idx.push_back(-1); }
      else {
        Tensor i = expr();
        if (i.order() != 0 || i.v[0].size() > 1 || (i.v[0].size() == 1 && !i.v[0].begin()->first.empty()))
          fail("index must be a number");
        const double v = i.v[0].empty() ? 0.0 : i.v[0].begin()->second;
        idx.push_back((int)std::lround(v) - 1);
      }
      if (eat(')')) break;
      expect(',');
    }
    if ((int)idx.size() != a.order()) fail("wrong number of indices");
    std::vector<int> sh;
    for (size_t k = 0; k < idx.size(); ++k) {
      if (idx[k] == -1) sh.push_back(a.shape[k]);
      else if (idx[k] < 0 || idx[k] >= a.shape[k]) fail("index out of range");
    }
    Tensor r = zeros(sh);
    // enumerate the kept indices
    const size_t n = r.size();
    for (size_t f = 0; f < n; ++f) {
      size_t rem = f, off = 0, stride = 1;
      std::vector<int> full(idx.size());
      std::vector<int> kept;
      for (size_t k = idx.size(); k-- > 0;)
        if (idx[k] == -1) {
          full[k] = (int)(rem % (size_t)a.shape[k]);
          rem /= (size_t)a.shape[k];
        } else full[k] = idx[k];
      for (size_t k = idx.size(); k-- > 0;) { off += full[k] * stride; stride *= (size_t)a.shape[k]; }
      r.v[f] = a.v[off];
    }
    return r;
  }
  std::string name() {
    skip();
    size_t b = pos;
    while (pos < s.size() && (std::isalnum((unsigned char)s[pos]) || s[pos] == '_')) ++pos;
    return s.substr(b, pos - b);
  }
  Tensor primary() {
    skip();
    if (pos >= s.size()) fail("unexpected end of expression");
    const char c = s[pos];
    if (std::isdigit((unsigned char)c) || (c == '.' && pos + 1 < s.size() && std::isdigit((unsigned char)s[pos + 1]))) {
      size_t used = 0;
      double v = std::stod(s.substr(pos), &used);
      pos += used;
      return scalar(pconst(v));
    }
    if (c == '(') {
      ++pos;
      Tensor a = expr();
      expect(')');
      return a;
    }
    if (c == '[') return bracket();
    if (std::isalpha((unsigned char)c)) {
      const std::string n = name();
      if (peek() == '(' && is_function(n)) return function(n);
      return identifier(n);
    }
    fail(std::string("unexpected character '") + c + "'");
  }
  Tensor bracket() {
    expect('[');
    std::vector<std::vector<Tensor>> rows(1);
    for (;;) {
      rows.back().push_back(expr());
      if (eat(',')) continue;
      if (eat(';')) { rows.emplace_back(); continue; }
      expect(']');
      break;
    }
    // GWFL's explicit tensors: [a,b,c] and [a;b;c] are vectors, [a,b;c,d] a
    // matrix by lines, [[a,c],[b,d]] the same matrix by columns
    if (rows.size() == 1 && !rows[0].empty() && rows[0][0].order() == 1) {
      const int m = rows[0][0].shape[0], n = (int)rows[0].size();
      Tensor r = zeros({m, n});
      for (int j = 0; j < n; ++j) {
        const Tensor &col = rows[0][j];
        if (col.order() != 1 || col.shape[0] != m) fail("ragged [[...],[...]]");
        for (int i = 0; i < m; ++i) r.v[i * n + j] = col.v[i];
      }
      return r;
    }
    for (auto &r : rows)
      for (auto &t : r)
        if (t.order() != 0) fail("only scalar entries or column vectors in [...]");
    if (rows.size() == 1 || rows[0].size() == 1) {  // a vector
      Tensor r = zeros({(int)(rows.size() == 1 ? rows[0].size() : rows.size())});
      for (size_t i = 0; i < rows.size(); ++i) {
        if (rows.size() > 1 && rows[i].size() != 1) fail("ragged [...]");
        for (size_t j = 0; j < rows[i].size(); ++j) r.v[i + j] = rows[i][j].v[0];
      }
      return r;
    }
    const int m = (int)rows.size(), n = (int)rows[0].size();
    Tensor r = zeros({m, n});
    for (int i = 0; i < m; ++i) {
      if ((int)rows[i].size() != n) fail("ragged [...]");
      for (int j = 0; j < n; ++j) r.v[i * n + j] = rows[i][j].v[0];
    }
    return r;
  }
  static bool is_function(const std::string &n) {
    static const std::set<std::string> f = {"Sym", "Skew", "Trace", "Id", "sqr", "Norm", "Norm_sqr",
                                            "qdim", "Deviator", "pow", "Det", "Inv"};
    return f.count(n) > 0;
  }
  static bool is_number(const Tensor &t, double &v) {
    if (t.order() != 0 || t.v[0].size() > 1 || (t.v[0].size() == 1 && !t.v[0].begin()->first.empty()))
      return false;
    v = t.v[0].empty() ? 0.0 : t.v[0].begin()->second;
    return true;
  }
  // determinant of a 2x2 or 3x3 matrix of polynomials (cofactor expansion)
  static Poly det(const Tensor &a) {
    const int d = a.shape[0];
    auto e = [&](int i, int j) -> const Poly & { return a.v[i * d + j]; };
    if (d == 1) return e(0, 0);
    if (d == 2) {
      Poly r = pmul(e(0, 0), e(1, 1));
      padd_to(r, pmul(e(0, 1), e(1, 0)), -1.0);
      return r;
    }
    Poly r;
    for (int j = 0; j < 3; ++j) {
      Poly minor = pmul(e(1, (j + 1) % 3), e(2, (j + 2) % 3));
      padd_to(minor, pmul(e(1, (j + 2) % 3), e(2, (j + 1) % 3)), -1.0);
      padd_to(r, pmul(e(0, j), minor));
    }
    return r;
  }
  // Inv(M) for M = M0 + N: M0 a numeric invertible matrix, N's monomials in
  // the value/gradient symbols of fem variables that are held at zero.  Around
  // N = 0, Inv(M) = A - A N A + (A N)^2 A - ... with A = M0^{-1}; the products
  // truncate the series at the degree the assembly needs.  This is what the
  // pulled-back coordinates route uses: F = Id + Grad_w at w = 0.
  Tensor inverse(const Tensor &a) {
    if (a.order() != 2 || a.shape[0] != a.shape[1] || a.shape[0] > 3) fail("Inv of a non-square matrix");
    const int d = a.shape[0];
    std::vector<double> m0((size_t)d * d, 0.0);
    Tensor nn = zeros({d, d});
    for (int k = 0; k < d * d; ++k)
      for (auto &kv : a.v[k]) {
        if (kv.first.empty()) { m0[k] = kv.second; continue; }
        for (int sy : kv.first) {
          std::string var, tag;
          split_symbol(C.symnames[sy], var, tag);
          auto it = C.vars->find(var);
          if (var == "c" || tag == "T" || tag == "GT" || tag == "S" || tag == "GS" || it == C.vars->end() ||
              it->second.is_constant)
            fail("Inv is supported for Id + (gradients of variables held at zero) only");
          if (C.collect_inv) C.zero_vars.insert(var);
        }
        padd_to(nn.v[k], Poly{{kv.first, kv.second}});
      }
    // A = M0^{-1} by Gauss-Jordan with partial pivoting
    std::vector<double> g(m0), inv((size_t)d * d, 0.0);
    for (int i = 0; i < d; ++i) inv[i * d + i] = 1.0;
    for (int c = 0; c < d; ++c) {
      int piv = c;
      for (int r = c + 1; r < d; ++r)
        if (std::fabs(g[r * d + c]) > std::fabs(g[piv * d + c])) piv = r;
      if (std::fabs(g[piv * d + c]) < 1e-300) fail("Inv of a singular matrix");
      for (int k = 0; k < d; ++k) { std::swap(g[c * d + k], g[piv * d + k]); std::swap(inv[c * d + k], inv[piv * d + k]); }
      const double pv = g[c * d + c];
      for (int k = 0; k < d; ++k) { g[c * d + k] /= pv; inv[c * d + k] /= pv; }
      for (int r = 0; r < d; ++r) {
        if (r == c) continue;
        const double f = g[r * d + c];
        for (int k = 0; k < d; ++k) { g[r * d + k] -= f * g[c * d + k]; inv[r * d + k] -= f * inv[c * d + k]; }
      }
    }
    Tensor A = zeros({d, d});
    for (int k = 0; k < d * d; ++k) A.v[k] = pconst(inv[k]);
    if (C.collect_inv) return A;
    const int deg = C.trunc.max_deg < 0 ? 0 : C.trunc.max_deg;
    if (C.trunc.max_deg < 0) {
      bool none = true;
      for (auto &p : nn.v) none = none && p.empty();
      if (!none) fail("Inv of a matrix of variables is supported by the polynomial route only");
    }
    // sum_{k <= deg} (-A N)^k A
    Tensor term = A, total = A;
    const Tensor mAN = tscale(tcontract(A, nn, 1), pconst(-1.0));
    for (int k = 1; k <= deg; ++k) {
      term = tcontract(mAN, term, 1);
      total = tadd(total, term, 1.0);
    }
    return total;
  }
  Tensor function(const std::string &n) {
    expect('(');
    if (n == "qdim") {
      const std::string v = name();
      expect(')');
      auto it = C.vars->find(v);
      if (it == C.vars->end()) fail("unknown variable " + v);
      return scalar(pconst(it->second.qdim));
    }
    Tensor a = expr();
    if (n == "pow") {
      expect(',');
      Tensor b = expr();
      expect(')');
      double k;
      if (!is_number(b, k) || k < 0 || k != std::floor(k) || k > 16)
        fail("pow(x, k) is supported for integer exponents 0 <= k <= 16");
      if (a.order() != 0) fail("pow of a non-scalar");
      Poly r = pconst(1.0);
      for (int i = 0; i < (int)k; ++i) r = pmul(r, a.v[0]);
      return scalar(r);
    }
    expect(')');
    if (n == "Det") {
      if (a.order() != 2 || a.shape[0] != a.shape[1] || a.shape[0] > 3) fail("Det of a non-square matrix");
      return scalar(det(a));
    }
    if (n == "Inv") return inverse(a);
    if (n == "Id") {
      if (a.order() != 0 || a.v[0].size() != 1 || !a.v[0].begin()->first.empty()) fail("Id(n) needs a number");
      const int d = (int)std::lround(a.v[0].begin()->second);
      Tensor r = zeros({d, d});
      for (int i = 0; i < d; ++i) r.v[i * d + i] = pconst(1.0);
      return r;
    }
    if (n == "Trace") {
      if (a.order() != 2 || a.shape[0] != a.shape[1]) fail("Trace of a non-square matrix");
      Poly s;
      for (int i = 0; i < a.shape[0]; ++i) padd_to(s, a.v[i * a.shape[0] + i]);
      return scalar(s);
    }
    if (n == "Sym" || n == "Skew") {
      Tensor t = ttranspose(a);
      Tensor r = tadd(a, t, n == "Sym" ? 1.0 : -1.0);
      return tscale(r, pconst(0.5));
    }
    if (n == "Deviator") {
      if (a.order() != 2) fail("Deviator of a non-matrix");
      Poly tr;
      const int d = a.shape[0];
      for (int i = 0; i < d; ++i) padd_to(tr, a.v[i * d + i]);
      Tensor r = a;
      for (int i = 0; i < d; ++i) padd_to(r.v[i * d + i], tr, -1.0 / d);
      return r;
    }
    if (n == "Norm_sqr") return tcontract(a, a, a.order());
    if (n == "Norm") {
      Tensor r = scalar(Poly());
      r.norm_arg = std::make_shared<Tensor>(a);
      return r;
    }
    if (n == "sqr") {
      if (a.norm_arg) return tcontract(*a.norm_arg, *a.norm_arg, a.norm_arg->order());
      if (a.order() != 0) fail("sqr of a non-scalar");
      return scalar(pmul(a.v[0], a.v[0]));
    }
    fail("unsupported function " + n);
  }

  // symbol tags: V value, G gradient, T / GT test value / gradient, S / GS
  // second test function value / gradient; components after '_'
  static std::string tag_of(const std::string &kind) {
    if (kind == "Val") return "V";
    if (kind == "Grad" || kind == "Div") return "G";
    if (kind == "Test") return "T";
    if (kind == "Grad_Test" || kind == "Div_Test") return "GT";
    if (kind == "Test2") return "S";
    return "GS";  // Grad_Test2, Div_Test2
  }

  Tensor field(const std::string &var, const std::string &kind) {
    auto it = C.vars->find(var);
    if (it == C.vars->end()) fail("unknown variable " + var);
    const GwflVar &v = it->second;
    if (!v.is_fem) fail(kind + " of a non-fem quantity " + var);
    const int q = v.qdim;
    const std::string base = var + ":" + tag_of(kind) + "_";
    if (kind == "Grad" || kind == "Grad_Test" || kind == "Grad_Test2") {
      if (q == 1) {
        Tensor t = zeros({3});
        for (int b = 0; b < 3; ++b) t.v[b] = psym(C.sym(base + std::to_string(b)));
        return t;
      }
      Tensor t = zeros({q, 3});
      for (int a = 0; a < q; ++a)
        for (int b = 0; b < 3; ++b) t.v[a * 3 + b] = psym(C.sym(base + std::to_string(a) + std::to_string(b)));
      return t;
    }
    if (kind == "Div" || kind == "Div_Test" || kind == "Div_Test2") {
      if (q != 3) fail("Div of a non-vector field " + var);
      Poly s;
      for (int a = 0; a < 3; ++a) padd_to(s, psym(C.sym(base + std::to_string(a) + std::to_string(a))));
      return scalar(s);
    }
    if (q == 1) return scalar(psym(C.sym(base)));
    Tensor t = zeros({q});
    for (int a = 0; a < q; ++a) t.v[a] = psym(C.sym(base + std::to_string(a)));
    return t;
  }

  Tensor identifier(const std::string &n) {
    if (n == "meshdim") return scalar(pconst(3.0));
    if (n == "pi") return scalar(pconst(M_PI));
    static const char *pre[] = {"Grad_Test2_", "Grad_Test_", "Div_Test2_", "Div_Test_", "Test2_",
                                "Test_", "Grad_", "Div_"};
    for (const char *p : pre) {
      const std::string ps(p);
      if (n.compare(0, ps.size(), ps) == 0 && n.size() > ps.size()) {
        const std::string var = n.substr(ps.size());
        const std::string kind = ps.substr(0, ps.size() - 1);
        if (C.vars->count(var)) return field(var, kind == "Test2" ? "Test2" : kind);
      }
    }
    auto it = C.vars->find(n);
    if (it == C.vars->end()) fail("unknown name " + n);
    const GwflVar &v = it->second;
    if (v.is_fem && !v.is_constant) return field(n, "Val");
    if (v.is_fem && !v.per_element) return field(n, "Val");  // fem data: values at the point
    if (v.qdim != 1) fail("constant " + n + " is not a scalar");
    return scalar(psym(C.sym("c:" + n)));
  }
};

}  // namespace

namespace {

// the finite strain brick's expression (getfem_nonlinear_elasticity.cc:2319-2320)
bool match_neohookean(const std::string &expr, std::smatch &m, std::string &squeezed) {
  static const std::regex nh(
      R"(^\(\(Id\(meshdim\)\+Grad_(\w+)\)\*\(Compressible_Neo_Hookean_Bonet_PK2\(Grad_\1,\[(\w+);(\w+)\]\)\)\):Grad_Test_\1$)");
  squeezed.clear();
  for (char c : expr) if (!std::isspace((unsigned char)c)) squeezed += c;
  return std::regex_match(squeezed, m, nh);
}

// highest test-function order among the monomials of a polynomial
int test_order(const Poly &p, const Ctx &C) {
  int order = 0;
  for (auto &kv : p)
    for (int s : kv.first) {
      const std::string &nm = C.symnames[s];
      const size_t c = nm.find(':');
      if (c == std::string::npos || nm.substr(0, c) == "c") continue;
      const std::string tag = nm.substr(c + 1, nm.find('_', c) - c - 1);
      if (tag == "S" || tag == "GS") order = std::max(order, 2);
      else if (tag == "T" || tag == "GT") order = std::max(order, 1);
    }
  return order;
}

}  // namespace

int gwfl_expression_order(const std::string &expr, const std::map<std::string, GwflVar> &vars,
                          std::string &err) {
  std::smatch m;
  std::string sq;
  if (match_neohookean(expr, m, sq)) return 1;
  Ctx C;
  C.vars = &vars;
  try {
    Parser P(expr, C);
    Tensor t = P.expr();
    P.skip();
    if (P.pos != expr.size()) fail("trailing characters in expression");
    if (t.order() != 0) fail("the expression is not a scalar");
    return test_order(t.v[0], C);
  } catch (const Error &e) {
    err = e.msg;
  } catch (const std::exception &e) {
    err = std::string("parse error: ") + e.what();
  }
  return -1;
}

bool gwfl_compile_order2(const std::vector<std::string> &exprs,
                         const std::map<std::string, GwflVar> &vars, GwflKernelParams &out,
                         std::string &err) {
  out = GwflKernelParams();
  if (exprs.size() == 1) {
    std::string e;
    std::smatch m;
    if (match_neohookean(exprs[0], m, e)) {
      out.term = GFX_TERM_NEOHOOKEAN;
      out.unknown = m[1];
      out.coef[0].terms.push_back({1.0, {m[2]}});
      out.coef[1].terms.push_back({1.0, {m[3]}});
      out.describe = "neo-hookean(bonet) lambda=" + std::string(m[2]) + " mu=" + std::string(m[3]);
      return true;
    }
  }
  Ctx C;
  C.vars = &vars;
  Poly total;
  std::string unknown;
  try {
    for (const std::string &ex : exprs) {
      Parser P(ex, C);
      Tensor t = P.expr();
      P.skip();
      if (P.pos != ex.size()) fail("trailing characters in expression");
      if (t.order() != 0) fail("the expression is not a scalar");
      Poly p = t.v[0];
      // the variables it depends on and its test-function order
      std::set<std::string> dvars;
      int order = 0;
      for (auto &kv : p)
        for (int s : kv.first) {
          const std::string &nm = C.symnames[s];
          const size_t c = nm.find(':');
          if (c == std::string::npos) continue;
          const std::string var = nm.substr(0, c);
          if (var == "c") continue;
          const std::string tag = nm.substr(c + 1, nm.find('_', c) - c - 1);
          if (tag == "S" || tag == "GS") order = std::max(order, 2);
          else if (tag == "T" || tag == "GT") order = std::max(order, 1);
          else if (!vars.at(var).is_constant) dvars.insert(var);
        }
      for (const std::string &v : dvars) {
        if (!unknown.empty() && unknown != v) fail("several unknowns in one term are not supported");
        unknown = v;
      }
      // differentiate up to order 2 (ga_workspace::add_tree)
      while (order < 2) {
        if (unknown.empty()) fail("no unknown to differentiate with respect to");
        const GwflVar &v = vars.at(unknown);
        std::map<int, int> dir;
        const std::string tv = order == 0 ? "T" : "S", tg = order == 0 ? "GT" : "GS";
        for (int a = 0; a < v.qdim; ++a) {
          const std::string sa = v.qdim == 1 ? "" : std::to_string(a);
          dir[C.sym(unknown + ":V_" + sa)] = C.sym(unknown + ":" + tv + "_" + sa);
          for (int b = 0; b < 3; ++b)
            dir[C.sym(unknown + ":G_" + sa + std::to_string(b))] =
                C.sym(unknown + ":" + tg + "_" + sa + std::to_string(b));
        }
        p = pderiv(p, dir);
        ++order;
      }
      padd_to(total, p);
    }
    // classify: every monomial = coefficients x one Grad_Test x one Grad_Test2
    std::string tvar;
    std::map<std::pair<std::string, std::string>, Poly> D;
    for (auto &kv : total) {
      std::string t1, t2;
      Mono coef;
      for (int s : kv.first) {
        const std::string &nm = C.symnames[s];
        const size_t c = nm.find(':');
        const std::string var = nm.substr(0, c);
        if (var == "c") { coef.push_back(s); continue; }
        const size_t us = nm.find('_', c);
        const std::string tag = nm.substr(c + 1, us - c - 1), comp = nm.substr(us + 1);
        if (tag == "GS") {
          if (!t2.empty()) fail("term not bilinear");
          t2 = comp;
        } else if (tag == "GT") {
          if (!t1.empty()) fail("term not bilinear");
          t1 = comp;
        } else if (tag == "T" || tag == "S") {
          fail("value terms (Test_u, mass matrices) are not supported by the kernels");
        } else if (vars.at(var).is_constant) {
          fail("the order-2 term depends on the fem data " + var +
               " (only P0 and fixed-size coefficients are supported by the matrix kernels)");
        } else {
          fail("the order-2 term depends on the unknown (nonlinear form outside the supported subset)");
        }
        if (tvar.empty()) tvar = var;
        else if (tvar != var) fail("coupling between different variables is not supported");
      }
      if (t1.empty() || t2.empty()) fail("term not bilinear in the test functions");
      padd_to(D[{t1, t2}], Poly{{coef, kv.second}});
    }
    if (tvar.empty()) fail("empty order-2 term");
    const int q = vars.at(tvar).qdim;
    auto get = [&](const std::string &a, const std::string &b) {
      auto it = D.find({a, b});
      return it == D.end() ? Poly() : it->second;
    };
    auto idx = [&](int a, int b) { return q == 1 ? std::to_string(b) : std::to_string(a) + std::to_string(b); };
    auto coef_of = [&](const Poly &p, GwflCoef &c) {
      for (auto &kv : p) {
        std::vector<std::string> nm;
        for (int s : kv.first) nm.push_back(C.symnames[s].substr(2));
        c.terms.push_back({kv.second, nm});
      }
    };
    out.unknown = tvar;
    // Laplace: D_{ab,cd} = alpha delta_ac delta_bd
    const Poly alpha = get(idx(0, 0), idx(0, 0));
    bool lap = true;
    for (int a = 0; a < q && lap; ++a)
      for (int b = 0; b < 3 && lap; ++b)
        for (int c = 0; c < q && lap; ++c)
          for (int d = 0; d < 3 && lap; ++d)
            lap = pequal(get(idx(a, b), idx(c, d)), (a == c && b == d) ? alpha : Poly());
    if (lap) {
      out.term = GFX_TERM_LAPLACE;
      coef_of(alpha, out.coef[0]);
      out.describe = "laplace";
      return true;
    }
    if (q == 3) {
      // D_{ab,cd} = alpha d_ab d_cd + beta d_ad d_bc + gamma d_ac d_bd: the
      // div-div, transposed-gradient and gradient:gradient contractions
      const Poly alpha3 = get("00", "11"), beta = get("01", "10"), gamma = get("01", "01");
      bool iso = true;
      for (int a = 0; a < 3 && iso; ++a)
        for (int b = 0; b < 3 && iso; ++b)
          for (int c = 0; c < 3 && iso; ++c)
            for (int d = 0; d < 3 && iso; ++d) {
              Poly e;
              if (a == b && c == d) padd_to(e, alpha3);
              if (a == d && b == c) padd_to(e, beta);
              if (a == c && b == d) padd_to(e, gamma);
              iso = pequal(get(idx(a, b), idx(c, d)), e);
            }
      if (iso && pequal(beta, gamma)) {
        out.term = GFX_TERM_ISO_ELAST;
        coef_of(alpha3, out.coef[0]);
        coef_of(beta, out.coef[1]);
        out.describe = "isotropic elasticity";
        return true;
      }
      if (iso) {
        out.term = GFX_TERM_ISOTROPIC;
        coef_of(alpha3, out.coef[0]);
        coef_of(beta, out.coef[1]);
        coef_of(gamma, out.coef[2]);
        out.describe = "isotropic";
        return true;
      }
    }
    fail("the order-2 term is not one of the supported forms (Laplacian, isotropic elasticity, "
         "isotropic div-div + transposed-gradient + gradient terms)");
  } catch (const Error &e) {
    err = e.msg;
    return false;
  } catch (const std::exception &e) {
    err = std::string("parse error: ") + e.what();
    return false;
  }
}

namespace {
struct TruncGuard {
  const Trunc *prev;
  explicit TruncGuard(const Trunc *t) : prev(g_trunc) { g_trunc = t; }
  ~TruncGuard() { g_trunc = prev; }
};
}  // namespace

bool gwfl_compile_scalar(const std::vector<std::string> &exprs, const std::map<std::string, GwflVar> &vars,
                         int order, const std::string &var, GwflScalarProgram &out, std::string &err,
                         const std::string &var2) {
  out = GwflScalarProgram();
  out.order = order;
  out.var = var;
  out.var2 = var2;
  try {
    if (order < 0 || order > 2) fail("the polynomial route assembles orders 0, 1 and 2");
    int q = 1, q2 = 1;
    auto fem_var = [&](const std::string &n) {
      auto it = vars.find(n);
      if (it == vars.end() || !it->second.is_fem || it->second.is_constant)
        fail("order-" + std::to_string(order) + " assembly with respect to " + n + ", which is not a fem variable");
      return it->second.qdim;
    };
    if (order >= 1) q = fem_var(var);
    if (order == 2) q2 = fem_var(var2);
    for (const std::string &ex : exprs) {
      std::smatch m;
      std::string sq;
      if (match_neohookean(ex, m, sq)) fail("the Neo-Hookean term is assembled by its own kernels");
    }
    // first pass: the variables Inv(...) expands around zero
    Ctx C0;
    C0.vars = &vars;
    C0.collect_inv = true;
    for (const std::string &ex : exprs) {
      Parser P(ex, C0);
      P.expr();
    }
    Ctx C;
    C.vars = &vars;
    C.zero_vars = C0.zero_vars;
    C.trunc.max_deg = C.zero_vars.empty() ? -1 : std::max(order, 1);
    TruncGuard guard(&C.trunc);
    // the direction of a derivative with respect to variable v: its value and
    // gradient symbols to the test (T, GT) or second test (S, GS) symbols
    auto direction = [&](const std::string &v, int qv, const char *tv, const char *tg) {
      std::map<int, int> dir;
      for (int a = 0; a < qv; ++a) {
        const std::string sa = qv == 1 ? "" : std::to_string(a);
        dir[C.sym(v + ":V_" + sa)] = C.sym(v + ":" + tv + "_" + sa);
        for (int b = 0; b < 3; ++b)
          dir[C.sym(v + ":G_" + sa + std::to_string(b))] = C.sym(v + ":" + tg + "_" + sa + std::to_string(b));
      }
      return dir;
    };
    Poly total;
    for (const std::string &ex : exprs) {
      Parser P(ex, C);
      Tensor t = P.expr();
      P.skip();
      if (P.pos != ex.size()) fail("trailing characters in expression");
      if (t.order() != 0) fail("the expression is not a scalar");
      Poly p = t.v[0];
      const int o = test_order(p, C);
      if (o > order) continue;  // a tree of a higher-order assembly only
      // the potential's derivative in the direction Test_var, then (order 2)
      // the derivative in the direction Test2_var2; the monomials of other
      // test variables belong to other blocks and are dropped below
      if (order >= 1 && o == 0) p = pderiv(p, direction(var, q, "T", "GT"));
      if (order == 2 && o <= 1) p = pderiv(p, direction(var2, q2, "S", "GS"));
      padd_to(total, p);
    }
    // at the zero value of the variables Inv expanded, their symbols vanish
    for (auto it = total.begin(); it != total.end();) {
      bool z = false;
      for (int sy : it->first) z = z || (sy < (int)C.trunc.zsym.size() && C.trunc.zsym[sy]);
      it = z ? total.erase(it) : std::next(it);
    }
    out.zero_vars.assign(C.zero_vars.begin(), C.zero_vars.end());
    out.nout = order == 0 ? 1 : (order == 1 ? 4 * q : 16 * q * q2);
    std::map<int, int> data;  // symbol id -> index in out.syms
    for (auto &kv : total) {
      GwflScalarProgram::Mono mo;
      mo.coef = kv.second;
      int s1 = order == 0 ? 0 : -1, s2 = order == 2 ? -1 : 0;  // test and second test slots
      bool other = false;  // a test function of another variable: another block
      for (int sy : kv.first) {
        const std::string &nm = C.symnames[sy];
        std::string v, tag;
        split_symbol(nm, v, tag);
        if ((tag == "S" || tag == "GS") && order < 2) fail("a second test function in an order-0/1 expression");
        if (tag == "T" || tag == "GT" || tag == "S" || tag == "GS") {
          const bool second = tag == "S" || tag == "GS";
          const int qv = second ? q2 : q;
          int &slot = second ? s2 : s1;
          if (order == 0) fail("a test function in a potential");
          if (v != (second ? var2 : var)) {
            other = true;
            continue;
          }
          if (slot >= 0) fail("term not linear in the test function of " + v);
          const std::string comp = nm.substr(nm.find('_', nm.find(':')) + 1);
          const int a = qv == 1 ? 0 : comp[0] - '0';
          slot = (tag == "T" || tag == "S") ? a : qv + 3 * a + (comp.back() - '0');
          continue;
        }
        auto it = data.find(sy);
        if (it == data.end()) {
          it = data.emplace(sy, (int)out.syms.size()).first;
          out.syms.push_back(nm);
        }
        mo.s.push_back(it->second);
      }
      if (other || s1 < 0 || s2 < 0) continue;  // another block's term, or one without a test function (zero)
      mo.out = order == 2 ? s1 * 4 * q2 + s2 : s1;
      out.monos.push_back(mo);
    }
    return true;
  } catch (const Error &e) {
    err = e.msg;
  } catch (const std::exception &e) {
    err = std::string("parse error: ") + e.what();
  }
  return false;
}

}  // namespace gfx
