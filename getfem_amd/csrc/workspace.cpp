// gfx_workspace: the ga_workspace drop-in.  Variables and constants are
// registered by name (ga_workspace::add_fem_variable / add_fem_constant /
// add_fixed_size_constant, src/getfem_generic_assembly_workspace.cc:72-121),
// expressions are added as GWFL text (add_expression, :545-600) and
// assembly(2) (:791-936) compiles them into one plan and runs it.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "gfx_internal.h"
#include "gwfl.h"

using namespace gfx;

namespace {

enum { VAR_FEM = 0, CONST_FEM = 1, CONST_FIXED = 2 };

struct WsVar {
  int kind = VAR_FEM;
  gfx_mesh_fem *mf = nullptr;
  int64_t first = 0;
  std::vector<double> V;
};

struct WsExpr {
  std::string expr;
  gfx_mesh_im *mim = nullptr;
  int der_order = 2;
};

}  // namespace

struct gfx_workspace {
  gfx_ctx *ctx = nullptr;
  std::map<std::string, WsVar> vars;
  std::vector<WsExpr> exprs;
  gfx_plan *plan = nullptr;
  bool dirty = true, coefs_dirty = true;
  bool state_dirty = true;  // the unknown's values changed since the last upload
  bool matrix_valid = false;  // plan values = K at the current coefficients
  bool pattern_stale = false;  // Neo-Hookean: state or coefficients changed since the pattern
  GwflKernelParams params;
  int form = GFX_FORM_FACTORIZED, precision = GFX_PREC_FP64;
  // order-1 result: the unknown's values and the assembled vector (device)
  double *d_u = nullptr, *d_V = nullptr;
  const double *V_out = nullptr;  // d_V or the plan's residual
  int64_t ndof = 0;
};

namespace {

int wfail(gfx_workspace *w, int code, const std::string &m) {
  if (w && w->ctx) w->ctx->err = m;
  return code;
}

bool valid_name(const char *n) {
  if (!n || !*n) return false;
  for (const char *p = n; *p; ++p)
    if (!(std::isalnum((unsigned char)*p) || *p == '_')) return false;
  return std::isalpha((unsigned char)n[0]);
}

// per-element values (P0 fem constant) of a registered constant
bool element_values(gfx_workspace *w, const std::string &name, int64_t n_cv, std::vector<double> &out,
                    double &scalar, bool &is_scalar, std::string &err) {
  auto it = w->vars.find(name);
  if (it == w->vars.end()) { err = "unknown constant " + name; return false; }
  const WsVar &v = it->second;
  if (v.kind == CONST_FIXED) {
    if (v.V.size() != 1) { err = name + " is not a scalar"; return false; }
    scalar = v.V[0];
    is_scalar = true;
    return true;
  }
  if (v.mf->fem.kind != FEM_PK_DISC0 || v.mf->qdim != 1 || v.mf->mesh->n_cv != n_cv) {
    err = name + " is not a P0 field on the assembly mesh";
    return false;
  }
  out.resize((size_t)n_cv);
  for (int64_t e = 0; e < n_cv; ++e) out[e] = v.V[(size_t)v.mf->h_dofs[e]];
  is_scalar = false;
  return true;
}

int bind_coefs(gfx_workspace *w, int64_t n_cv, std::vector<gfx_coef> &cf,
               std::vector<std::vector<double>> &store) {
  const int n = w->params.term == GFX_TERM_LAPLACE ? 1 : 2;
  cf.assign(n, gfx_coef());
  store.assign(n, std::vector<double>());
  for (int s = 0; s < n; ++s) {
    bool all_scalar = true;
    double sval = 0.0;
    std::vector<double> arr;
    for (auto &t : w->params.coef[s].terms) {
      double sc = t.first;
      std::vector<double> prod;
      for (auto &nm : t.second) {
        std::vector<double> vals;
        double x = 0.0;
        bool isS = false;
        std::string err;
        if (!element_values(w, nm, n_cv, vals, x, isS, err)) return wfail(w, GFX_E_UNSUPPORTED, err);
        if (isS) sc *= x;
        else if (prod.empty()) prod = vals;
        else for (int64_t e = 0; e < n_cv; ++e) prod[e] *= vals[e];
      }
      if (prod.empty()) {
        sval += sc;
        if (!arr.empty()) for (auto &a : arr) a += sc;
      } else {
        if (all_scalar) { arr.assign((size_t)n_cv, sval); all_scalar = false; }
        for (int64_t e = 0; e < n_cv; ++e) arr[e] += sc * prod[e];
      }
    }
    if (all_scalar) {
      cf[s].kind = GFX_COEF_CONST;
      cf[s].value = sval;
    } else {
      store[s] = arr;
      cf[s].kind = GFX_COEF_PER_ELEMENT;
      cf[s].h_values = store[s].data();
    }
  }
  return GFX_OK;
}

}  // namespace

extern "C" {

int gfx_ws_create(gfx_ctx *ctx, gfx_workspace **out) {
  if (!ctx || !out) return GFX_E_INVALID;
  gfx_workspace *w = new (std::nothrow) gfx_workspace();
  if (!w) return GFX_E_OOM;
  w->ctx = ctx;
  ref(ctx);
  *out = w;
  return GFX_OK;
}

int gfx_ws_destroy(gfx_workspace *w) {
  if (!w) return GFX_E_INVALID;
  if (w->plan) gfx_plan_destroy(w->plan);
  (void)hipSetDevice(w->ctx->device);
  if (w->d_u) (void)hipFree(w->d_u);
  if (w->d_V) (void)hipFree(w->d_V);
  for (auto &kv : w->vars)
    if (kv.second.mf) unref(kv.second.mf);
  gfx_ctx *c = w->ctx;
  delete w;
  unref(c);
  return GFX_OK;
}

static int add_var(gfx_workspace *w, const char *name, int kind, gfx_mesh_fem *mf, int64_t first,
                   const double *V, int64_t n) {
  if (!w || !valid_name(name)) return wfail(w, GFX_E_INVALID, "invalid variable name");
  if (kind != CONST_FIXED && !mf) return wfail(w, GFX_E_INVALID, "missing mesh_fem");
  if (kind != CONST_FIXED && mf->mesh->ctx != w->ctx) return wfail(w, GFX_E_INVALID, "mesh_fem of another context");
  WsVar v;
  v.kind = kind; v.mf = mf; v.first = first;
  const int64_t size = kind == CONST_FIXED ? n : mf->ndof;
  if (size < 0) return wfail(w, GFX_E_INVALID, "negative size");
  v.V.assign((size_t)size, 0.0);
  if (V) std::memcpy(v.V.data(), V, 8 * (size_t)size);
  if (mf) ref(mf);  // the workspace keeps the mesh_fem alive while it names it
  auto old = w->vars.find(name);
  if (old != w->vars.end() && old->second.mf) unref(old->second.mf);
  w->vars[name] = v;
  w->dirty = true;
  return GFX_OK;
}

int gfx_ws_add_fem_variable(gfx_workspace *w, const char *name, gfx_mesh_fem *mf, int64_t first,
                            const double *h_V) {
  return add_var(w, name, VAR_FEM, mf, first, h_V, 0);
}

int gfx_ws_add_fem_constant(gfx_workspace *w, const char *name, gfx_mesh_fem *mf, const double *h_V) {
  return add_var(w, name, CONST_FEM, mf, 0, h_V, 0);
}

int gfx_ws_add_fixed_size_constant(gfx_workspace *w, const char *name, const double *h_V, int64_t n) {
  if (!h_V && n > 0) return wfail(w, GFX_E_INVALID, "missing values");
  return add_var(w, name, CONST_FIXED, nullptr, 0, h_V, n);
}

int gfx_ws_set_values(gfx_workspace *w, const char *name, const double *h_V) {
  if (!w || !name || !h_V) return GFX_E_INVALID;
  auto it = w->vars.find(name);
  if (it == w->vars.end()) return wfail(w, GFX_E_INVALID, std::string("unknown name ") + name);
  std::memcpy(it->second.V.data(), h_V, 8 * it->second.V.size());
  if (it->second.kind == VAR_FEM) w->state_dirty = true;
  else w->coefs_dirty = true;
  return GFX_OK;
}

int gfx_ws_add_expression(gfx_workspace *w, const char *expr, gfx_mesh_im *mim, int der_order) {
  if (!w || !expr || !mim) return GFX_E_INVALID;
  if (der_order != 1 && der_order != 2)
    return wfail(w, GFX_E_UNSUPPORTED, "derivation order must be 1 or 2");
  if (!w->exprs.empty() && w->exprs[0].mim != mim)
    return wfail(w, GFX_E_UNSUPPORTED, "all expressions must use the same mesh_im");
  w->exprs.push_back({expr, mim, der_order});
  w->dirty = true;
  return GFX_OK;
}

int gfx_ws_set_form(gfx_workspace *w, int form, int precision) {
  if (!w) return GFX_E_INVALID;
  w->form = form;
  w->precision = precision;
  if (w->plan) return gfx_plan_set_form(w->plan, form, precision);
  return GFX_OK;
}

int gfx_ws_compile(gfx_workspace *w) {
  if (!w) return GFX_E_INVALID;
  if (w->exprs.empty()) return wfail(w, GFX_E_STATE, "no expression");
  std::map<std::string, GwflVar> gv;
  for (auto &kv : w->vars) {
    GwflVar g;
    g.name = kv.first;
    g.is_fem = kv.second.kind != CONST_FIXED;
    g.is_constant = kv.second.kind != VAR_FEM;
    g.qdim = g.is_fem ? kv.second.mf->qdim : (int)kv.second.V.size();
    g.per_element = g.is_fem && kv.second.mf->fem.kind == FEM_PK_DISC0;
    gv[kv.first] = g;
  }
  std::vector<std::string> ex;
  for (auto &e : w->exprs) ex.push_back(e.expr);
  std::string err;
  if (!gwfl_compile_order2(ex, gv, w->params, err)) return wfail(w, GFX_E_UNSUPPORTED, err);
  return GFX_OK;
}

// ga_workspace::assembly(order) (src/getfem_generic_assembly_workspace.cc:
// 791-936).  order 2: the tangent matrix into the plan's CSR.  order 1: the
// vector of the first derivative at the unknown's current value; for the
// Neo-Hookean term the element residual gathered per node, for the linear
// terms (whose first derivative is bilinear in u and the test function) the
// product K u of the assembled matrix.
int gfx_ws_assembly(gfx_workspace *w, int order) {
  if (!w) return GFX_E_INVALID;
  if (order != 1 && order != 2)
    return wfail(w, GFX_E_UNSUPPORTED, "only order-1 (vector) and order-2 (matrix) assembly are supported");
  int r;
  if (w->dirty) {
    if ((r = gfx_ws_compile(w))) return r;
    const int t = w->params.term;
    if (t != GFX_TERM_LAPLACE && t != GFX_TERM_ISO_ELAST && t != GFX_TERM_NEOHOOKEAN)
      return wfail(w, GFX_E_UNSUPPORTED, "term compiled (" + w->params.describe + ") has no plan");
    for (auto &e : w->exprs)
      if (e.der_order < order)
        return wfail(w, GFX_E_STATE, "an expression was added with a lower derivation order");
    const WsVar &u = w->vars.at(w->params.unknown);
    if (u.kind != VAR_FEM) return wfail(w, GFX_E_INVALID, "the unknown must be a fem variable");
    std::vector<gfx_coef> cf;
    std::vector<std::vector<double>> store;
    if ((r = bind_coefs(w, u.mf->mesh->n_cv, cf, store))) return r;
    if (w->plan) { gfx_plan_destroy(w->plan); w->plan = nullptr; }
    if ((r = gfx_plan_create(u.mf, w->exprs[0].mim, t, cf.data(), (int)cf.size(), &w->plan)))
      return r;
    if ((r = gfx_plan_set_form(w->plan, w->form, w->precision))) return r;
    if (t == GFX_TERM_NEOHOOKEAN && (r = gfx_plan_set_state(w->plan, u.V.data()))) return r;
    if ((r = gfx_plan_build_pattern(w->plan, nullptr))) return r;
    if (w->ndof != u.mf->ndof) {
      if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
      if (w->d_u) (void)hipFree(w->d_u);
      if (w->d_V) (void)hipFree(w->d_V);
      w->d_u = w->d_V = nullptr;
      const size_t bytes = 8 * (size_t)std::max<int64_t>(u.mf->ndof, 1);
      if (hipMalloc((void **)&w->d_u, bytes) != hipSuccess || hipMalloc((void **)&w->d_V, bytes) != hipSuccess)
        return wfail(w, GFX_E_OOM, "workspace vectors");
      w->ndof = u.mf->ndof;
    }
    w->dirty = false;
    w->coefs_dirty = false;
    w->state_dirty = t != GFX_TERM_NEOHOOKEAN;
    w->pattern_stale = false;
    w->matrix_valid = false;
  } else if (w->coefs_dirty) {
    const WsVar &u = w->vars.at(w->params.unknown);
    std::vector<gfx_coef> cf;
    std::vector<std::vector<double>> store;
    if ((r = bind_coefs(w, u.mf->mesh->n_cv, cf, store))) return r;
    for (size_t s = 0; s < cf.size(); ++s)
      if ((r = gfx_plan_set_coef(w->plan, (int)s, &cf[s]))) return r;
    w->coefs_dirty = false;
    w->matrix_valid = false;
    w->pattern_stale = true;
  }
  const WsVar &u = w->vars.at(w->params.unknown);
  if (w->params.term == GFX_TERM_NEOHOOKEAN) {
    if (w->state_dirty) {
      if ((r = gfx_plan_set_state(w->plan, u.V.data()))) return r;
      w->state_dirty = false;
      w->pattern_stale = true;
    }
    w->V_out = w->plan->d_rhs;
    if (order == 1) return gfx_plan_assemble_residual(w->plan);
    // the stored pattern depends on the values (threshold on each element
    // matrix), as GetFEM's matrix is rebuilt by every assembly
    if (w->pattern_stale) {
      if ((r = gfx_plan_build_pattern(w->plan, nullptr))) return r;
      w->pattern_stale = false;
    }
    return gfx_plan_assemble(w->plan);
  }
  if (order == 2 || !w->matrix_valid) {
    if ((r = gfx_plan_assemble(w->plan))) return r;
    w->matrix_valid = true;
  }
  if (order == 1) {
    if (w->state_dirty) {
      if (hipMemcpyAsync(w->d_u, u.V.data(), 8 * (size_t)w->ndof, hipMemcpyHostToDevice,
                         w->ctx->stream) != hipSuccess ||
          hipStreamSynchronize(w->ctx->stream) != hipSuccess)
        return wfail(w, GFX_E_HIP, "upload of the unknown");
      w->state_dirty = false;
    }
    if ((r = gfx_plan_spmv(w->plan, w->d_u, w->d_V))) return r;
    w->V_out = w->d_V;
  }
  return GFX_OK;
}

int gfx_ws_copy_assembled_vector_device(gfx_workspace *w, double *d_V) {
  if (!w || !d_V) return GFX_E_INVALID;
  if (!w->V_out) return wfail(w, GFX_E_STATE, "no order-1 assembly done yet");
  if (hipSetDevice(w->ctx->device) != hipSuccess ||
      hipMemcpyAsync(d_V, w->V_out, 8 * (size_t)w->ndof, hipMemcpyDeviceToDevice, w->ctx->stream) !=
          hipSuccess)
    return wfail(w, GFX_E_HIP, "copy of the assembled vector");
  return GFX_OK;
}

int gfx_ws_assembled_vector(gfx_workspace *w, double *h_V) {
  if (!w || !h_V) return GFX_E_INVALID;
  if (!w->V_out) return wfail(w, GFX_E_STATE, "no order-1 assembly done yet");
  if (hipSetDevice(w->ctx->device) != hipSuccess ||
      hipStreamSynchronize(w->ctx->stream) != hipSuccess ||
      hipMemcpy(h_V, w->V_out, 8 * (size_t)w->ndof, hipMemcpyDeviceToHost) != hipSuccess)
    return wfail(w, GFX_E_HIP, "copy of the assembled vector");
  return GFX_OK;
}

static void describe_params(const GwflKernelParams &p, std::string &d) {
  d = p.describe + " unknown=" + p.unknown;
  const char *names[2] = {"a/lambda", "mu"};
  for (int s = 0; s < 2; ++s) {
    if (p.coef[s].terms.empty()) continue;
    d += std::string(" ") + names[s] + "=";
    bool first = true;
    for (auto &t : p.coef[s].terms) {
      char num[64];
      std::snprintf(num, sizeof num, "%.17g", t.first);
      d += (first ? "" : "+") + std::string(num);
      for (auto &n : t.second) d += "*" + n;
      first = false;
    }
  }
}

int gfx_gwfl_compile(const char *exprs, const char *decls, char *out, int64_t len) {
  if (!exprs || !decls || !out || len <= 0) return GFX_E_INVALID;
  std::map<std::string, GwflVar> gv;
  std::string ds(decls);
  size_t b = 0;
  while (b < ds.size()) {
    size_t e = ds.find(';', b);
    if (e == std::string::npos) e = ds.size();
    const std::string item = ds.substr(b, e - b);
    b = e + 1;
    if (item.empty()) continue;
    const size_t c1 = item.find(':'), c2 = item.find(':', c1 + 1);
    if (c1 == std::string::npos || c2 == std::string::npos) return GFX_E_INVALID;
    GwflVar g;
    g.name = item.substr(0, c1);
    const std::string kind = item.substr(c1 + 1, c2 - c1 - 1);
    g.qdim = std::atoi(item.c_str() + c2 + 1);
    g.is_fem = kind == "var" || kind == "p0";
    g.is_constant = kind != "var";
    g.per_element = kind == "p0";
    gv[g.name] = g;
  }
  std::vector<std::string> ex;
  std::string es(exprs);
  b = 0;
  while (b <= es.size()) {
    size_t e = es.find('\n', b);
    if (e == std::string::npos) e = es.size();
    if (e > b) ex.push_back(es.substr(b, e - b));
    b = e + 1;
  }
  GwflKernelParams p;
  std::string err, d;
  const bool ok = gwfl_compile_order2(ex, gv, p, err);
  if (ok) describe_params(p, d);
  else d = err;
  std::strncpy(out, d.c_str(), (size_t)len - 1);
  out[len - 1] = 0;
  return ok ? GFX_OK : GFX_E_UNSUPPORTED;
}

int gfx_ws_plan(gfx_workspace *w, gfx_plan **out) {
  if (!w || !out) return GFX_E_INVALID;
  if (!w->plan) return wfail(w, GFX_E_STATE, "no assembly done yet");
  *out = w->plan;
  return GFX_OK;
}

int gfx_ws_describe(gfx_workspace *w, char *buf, int64_t len) {
  if (!w || !buf || len <= 0) return GFX_E_INVALID;
  std::string d;
  describe_params(w->params, d);
  std::strncpy(buf, d.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return GFX_OK;
}

}  // extern "C"
