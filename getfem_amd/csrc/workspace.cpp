// gfx_workspace: the ga_workspace drop-in.  Variables and constants are
// registered by name (ga_workspace::add_fem_variable / add_fem_constant /
// add_fixed_size_constant, src/getfem_generic_assembly_workspace.cc:72-121),
// expressions are added as GWFL text (add_expression, :545-600) and
// assembly(order) (:791-936) compiles them into plans and runs them.
//
// Trees.  GetFEM keeps one tree per (mesh_im, region, test functions) and
// merges the expressions that share them (add_tree, :474-495): their
// integrands are summed before the element matrix is formed and
// thresholded.  Here every mesh_im gets one tree, compiled into one plan (the
// supported order-2 terms of one field on one mesh_im always merge into one
// of the plan terms).  Several trees (several mesh_ims) are assembled by their
// own plans and added into the workspace's matrix, whose pattern is the union
// of theirs, as ga_exec adds every tree into K.
//
// Orders.  An expression's test-function order (0 potential, 1 residual
// form, 2 bilinear form) decides which assemblies it takes part in:
// add_expression derives it up to der_order, so it has a tree of every order
// from its own up to der_order (:520-541).  assembly(2) runs the expressions
// with der_order 2; assembly(1) those of order <= 1.
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
  int order = -1;  // test-function order, set by the compile stage
};

struct WsTree {
  gfx_mesh_im *mim = nullptr;
  std::vector<int> e2, e1;      // expressions of the order-2 and order-1 trees
  GwflKernelParams p2, p1;
  gfx_plan *plan = nullptr;     // order-2 tree (null when e2 is empty)
  gfx_plan *plan1 = nullptr;    // order-1 tree of linear terms, when it differs from the order-2 tree
  bool same1 = true;            // e1 == e2
  bool valid2 = false, valid1 = false;  // plan values at the current coefficients
  int64_t union_id = -1;        // plan pattern id the workspace matrix's map was built with
  int64_t *d_map = nullptr;     // positions of the plan's entries in the workspace matrix
};

}  // namespace

struct gfx_workspace {
  gfx_ctx *ctx = nullptr;
  std::map<std::string, WsVar> vars;
  std::vector<WsExpr> exprs;
  std::vector<WsTree> trees;
  std::string unknown;
  bool nonlinear = false;       // the (single) tree is the Neo-Hookean term
  bool dirty = true, coefs_dirty = true;
  bool state_dirty = true;      // the unknown's values changed since the last upload
  bool res_valid = false;       // Neo-Hookean: the plan's residual is at the current values
  int form = GFX_FORM_FACTORIZED, precision = GFX_PREC_FP64;
  // order-1 result: the unknown's values and the assembled vector (device)
  double *d_u = nullptr, *d_V = nullptr;
  const double *V_out = nullptr;  // d_V or the plan's residual
  int64_t ndof = 0;
  // several trees: the workspace's own matrix (union of the trees' patterns)
  int64_t *d_ip = nullptr;
  int32_t *d_ix = nullptr;
  double *d_v = nullptr;
  int64_t nnz = -1;
  // set_assembled_matrix: a caller-owned device matrix the assembly adds into
  bool ext = false;
  int ext_orient = GFX_CSR;
  int64_t ext_n = 0;
  const int64_t *ext_ptr = nullptr;
  const int32_t *ext_ind = nullptr;
  double *ext_vals = nullptr;
};

namespace {

int wfail(gfx_workspace *w, int code, const std::string &m) {
  if (w && w->ctx) w->ctx->err = m;
  return code;
}

int whip(gfx_workspace *w, hipError_t e, const char *what) {
  return wfail(w, e == hipErrorOutOfMemory ? GFX_E_OOM : GFX_E_HIP,
               std::string(what) + ": " + hipGetErrorString(e));
}

#define WHIP(w, call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return whip((w), e_, #call);  \
  } while (0)

template <typename T>
void wfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

bool valid_name(const char *n) {
  if (!n || !*n) return false;
  for (const char *p = n; *p; ++p)
    if (!(std::isalnum((unsigned char)*p) || *p == '_')) return false;
  return std::isalpha((unsigned char)n[0]);
}

int term_coefs(int term) { return term == GFX_TERM_LAPLACE ? 1 : (term == GFX_TERM_ISOTROPIC ? 3 : 2); }

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

int bind_coefs(gfx_workspace *w, const GwflKernelParams &params, int64_t n_cv, std::vector<gfx_coef> &cf,
               std::vector<std::vector<double>> &store) {
  const int n = term_coefs(params.term);
  cf.assign(n, gfx_coef());
  store.assign(n, std::vector<double>());
  for (int s = 0; s < n; ++s) {
    bool all_scalar = true;
    double sval = 0.0;
    std::vector<double> arr;
    for (auto &t : params.coef[s].terms) {
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

void drop_plans(gfx_workspace *w) {
  for (auto &t : w->trees) {
    if (t.plan) gfx_plan_destroy(t.plan);
    if (t.plan1) gfx_plan_destroy(t.plan1);
    wfree(t.d_map);
  }
  w->trees.clear();
  wfree(w->d_ip); wfree(w->d_ix); wfree(w->d_v);
  w->nnz = -1;
}

int plan_for(gfx_workspace *w, const GwflKernelParams &params, gfx_mesh_im *mim, gfx_plan **out) {
  const WsVar &u = w->vars.at(w->unknown);
  std::vector<gfx_coef> cf;
  std::vector<std::vector<double>> store;
  int r;
  if ((r = bind_coefs(w, params, u.mf->mesh->n_cv, cf, store))) return r;
  if ((r = gfx_plan_create(u.mf, mim, params.term, cf.data(), (int)cf.size(), out))) return r;
  if ((r = gfx_plan_set_form(*out, w->form, w->precision))) return r;
  if (params.term == GFX_TERM_NEOHOOKEAN && (r = gfx_plan_set_state(*out, u.V.data()))) return r;
  return gfx_plan_build_pattern(*out, nullptr);
}

int rebind(gfx_workspace *w, gfx_plan *p, const GwflKernelParams &params) {
  if (!p) return GFX_OK;
  const WsVar &u = w->vars.at(w->unknown);
  std::vector<gfx_coef> cf;
  std::vector<std::vector<double>> store;
  int r;
  if ((r = bind_coefs(w, params, u.mf->mesh->n_cv, cf, store))) return r;
  for (size_t s = 0; s < cf.size(); ++s)
    if ((r = gfx_plan_set_coef(p, (int)s, &cf[s]))) return r;
  return GFX_OK;
}

// the plans of the trees with an order-2 part
std::vector<WsTree *> matrix_trees(gfx_workspace *w) {
  std::vector<WsTree *> v;
  for (auto &t : w->trees)
    if (t.plan) v.push_back(&t);
  return v;
}

// The workspace's own matrix for several trees: the union of their patterns
// (device merges), with each plan's entry positions in it.  Rebuilt when a
// plan's pattern changed.
int build_union(gfx_workspace *w) {
  std::vector<WsTree *> mt = matrix_trees(w);
  bool stale = w->nnz < 0;
  for (WsTree *t : mt) stale |= t->union_id != t->plan->pattern_id;
  if (!stale) return GFX_OK;
  hipStream_t s = w->ctx->stream;
  const int64_t n = w->ndof;
  wfree(w->d_ip); wfree(w->d_ix); wfree(w->d_v);
  for (WsTree *t : mt) wfree(t->d_map);
  w->nnz = -1;
  // U = first plan's pattern, then U <- U u P_k with the earlier maps composed
  gfx_plan *p0 = mt[0]->plan;
  int64_t *ip = nullptr, *tmp = nullptr, *cnt = nullptr, *mapu = nullptr;
  int32_t *ix = nullptr;
  int64_t nnz = p0->nnz;
  WHIP(w, hipMalloc((void **)&ip, 8 * (size_t)(n + 1)));
  WHIP(w, hipMalloc((void **)&ix, 4 * (size_t)std::max<int64_t>(nnz, 1)));
  WHIP(w, hipMemcpyAsync(ip, p0->d_indptr, 8 * (size_t)(n + 1), hipMemcpyDeviceToDevice, s));
  WHIP(w, hipMemcpyAsync(ix, p0->d_indices, 4 * (size_t)nnz, hipMemcpyDeviceToDevice, s));
  // first map: identity (positions 0..nnz-1)
  {
    std::vector<int64_t> id((size_t)std::max<int64_t>(nnz, 1));
    for (int64_t k = 0; k < nnz; ++k) id[(size_t)k] = k;
    WHIP(w, hipMalloc((void **)&mt[0]->d_map, 8 * id.size()));
    WHIP(w, hipMemcpy(mt[0]->d_map, id.data(), 8 * id.size(), hipMemcpyHostToDevice));
  }
  for (size_t k = 1; k < mt.size(); ++k) {
    gfx_plan *pk = mt[k]->plan;
    int64_t *ipn = nullptr, *mapk = nullptr;
    int32_t *ixn = nullptr;
    WHIP(w, hipMalloc((void **)&cnt, 8 * (size_t)std::max<int64_t>(n, 1)));
    WHIP(w, hipMalloc((void **)&tmp, 8 * (size_t)scan_tmp_elems(n)));
    WHIP(w, hipMalloc((void **)&ipn, 8 * (size_t)(n + 1)));
    WHIP(w, launch_union_count(s, n, ip, ix, pk->d_indptr, pk->d_indices, cnt));
    WHIP(w, scan_exclusive_i64(s, cnt, ipn, n, tmp));
    int64_t nn = 0;
    WHIP(w, hipMemcpyAsync(&nn, ipn + n, 8, hipMemcpyDeviceToHost, s));
    WHIP(w, hipStreamSynchronize(s));
    wfree(cnt); wfree(tmp);
    WHIP(w, hipMalloc((void **)&ixn, 4 * (size_t)std::max<int64_t>(nn, 1)));
    WHIP(w, hipMalloc((void **)&mapu, 8 * (size_t)std::max<int64_t>(nnz, 1)));
    WHIP(w, hipMalloc((void **)&mapk, 8 * (size_t)std::max<int64_t>(pk->nnz, 1)));
    WHIP(w, launch_union_fill(s, n, ip, ix, pk->d_indptr, pk->d_indices, ipn, ixn, mapu, mapk));
    for (size_t j = 0; j < k; ++j) {  // earlier maps: old union position -> new one
      gfx_plan *pj = mt[j]->plan;
      int64_t *m2 = nullptr;
      WHIP(w, hipMalloc((void **)&m2, 8 * (size_t)std::max<int64_t>(pj->nnz, 1)));
      WHIP(w, launch_gather_i64(s, pj->nnz, mt[j]->d_map, mapu, m2));
      WHIP(w, hipStreamSynchronize(s));
      wfree(mt[j]->d_map);
      mt[j]->d_map = m2;
    }
    WHIP(w, hipStreamSynchronize(s));
    wfree(mapu); wfree(ip); wfree(ix);
    mt[k]->d_map = mapk;
    ip = ipn; ix = ixn; nnz = nn;
  }
  w->d_ip = ip; w->d_ix = ix; w->nnz = nnz;
  WHIP(w, hipMalloc((void **)&w->d_v, 8 * (size_t)std::max<int64_t>(nnz, 1)));
  for (WsTree *t : mt) t->union_id = t->plan->pattern_id;
  return GFX_OK;
}

int compile(gfx_workspace *w) {
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
  std::string err;
  for (auto &e : w->exprs) {
    e.order = gwfl_expression_order(e.expr, gv, err);
    if (e.order < 0) return wfail(w, GFX_E_UNSUPPORTED, err);
    if (e.order > e.der_order)
      return wfail(w, GFX_E_INVALID, "the expression's order is above its derivation order: " + e.expr);
  }
  drop_plans(w);
  std::vector<WsTree> trees;
  for (size_t i = 0; i < w->exprs.size(); ++i) {
    const WsExpr &e = w->exprs[i];
    WsTree *t = nullptr;
    for (auto &x : trees)
      if (x.mim == e.mim) t = &x;
    if (!t) {
      trees.push_back(WsTree());
      t = &trees.back();
      t->mim = e.mim;
    }
    if (e.der_order == 2) t->e2.push_back((int)i);
    if (e.order <= 1) t->e1.push_back((int)i);
  }
  std::string unknown;
  bool nonlinear = false;
  auto texts = [&](const std::vector<int> &ix) {
    std::vector<std::string> v;
    for (int i : ix) v.push_back(w->exprs[i].expr);
    return v;
  };
  for (auto &t : trees) {
    t.same1 = t.e1 == t.e2;
    GwflKernelParams *ps[2] = {&t.p2, &t.p1};
    const std::vector<int> *es[2] = {&t.e2, &t.e1};
    for (int k = 0; k < 2; ++k) {
      if (es[k]->empty() || (k == 1 && t.same1)) continue;
      if (!gwfl_compile_order2(texts(*es[k]), gv, *ps[k], err)) return wfail(w, GFX_E_UNSUPPORTED, err);
      if (!unknown.empty() && unknown != ps[k]->unknown)
        return wfail(w, GFX_E_UNSUPPORTED, "terms in several unknowns are not supported");
      unknown = ps[k]->unknown;
      nonlinear |= ps[k]->term == GFX_TERM_NEOHOOKEAN;
    }
  }
  if (unknown.empty()) return wfail(w, GFX_E_STATE, "no expression takes part in an assembly");
  if (nonlinear && (trees.size() != 1 || !trees[0].same1))
    return wfail(w, GFX_E_UNSUPPORTED, "the Neo-Hookean term must be the workspace's only expression "
                                       "(derivation order 2)");
  w->trees = trees;
  w->unknown = unknown;
  w->nonlinear = nonlinear;
  return GFX_OK;
}

int setup(gfx_workspace *w) {
  int r;
  if ((r = compile(w))) return r;
  const WsVar &u = w->vars.at(w->unknown);
  if (u.kind != VAR_FEM) return wfail(w, GFX_E_INVALID, "the unknown must be a fem variable");
  for (auto &t : w->trees) {
    if (!t.e2.empty() && (r = plan_for(w, t.p2, t.mim, &t.plan))) return r;
    if (!t.same1 && !t.e1.empty() && (r = plan_for(w, t.p1, t.mim, &t.plan1))) return r;
  }
  if (w->ndof != u.mf->ndof || !w->d_u) {
    if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
    wfree(w->d_u); wfree(w->d_V);
    const size_t bytes = 8 * (size_t)std::max<int64_t>(u.mf->ndof, 1);
    if (hipMalloc((void **)&w->d_u, bytes) != hipSuccess || hipMalloc((void **)&w->d_V, bytes) != hipSuccess)
      return wfail(w, GFX_E_OOM, "workspace vectors");
    w->ndof = u.mf->ndof;
  }
  w->dirty = false;
  w->coefs_dirty = false;
  w->state_dirty = !w->nonlinear;  // a nonlinear plan received the state at creation
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
  (void)hipSetDevice(w->ctx->device);
  drop_plans(w);
  wfree(w->d_u); wfree(w->d_V);
  for (auto &kv : w->vars)
    if (kv.second.mf) unref(kv.second.mf);
  for (auto &e : w->exprs) unref(e.mim);
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

int gfx_ws_set_values_device(gfx_workspace *w, const char *name, const double *d_V) {
  if (!w || !name || !d_V) return GFX_E_INVALID;
  auto it = w->vars.find(name);
  if (it == w->vars.end()) return wfail(w, GFX_E_INVALID, std::string("unknown name ") + name);
  if (it->second.kind != VAR_FEM) return wfail(w, GFX_E_INVALID, "device values are for the fem variable");
  int r;
  if (w->dirty && (r = setup(w))) return r;
  if (name != w->unknown) return wfail(w, GFX_E_INVALID, std::string(name) + " is not the unknown");
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  // stream-ordered: the next assembly reads it on the same stream
  WHIP(w, hipMemcpyAsync(w->d_u, d_V, 8 * (size_t)w->ndof, hipMemcpyDeviceToDevice, w->ctx->stream));
  if (w->nonlinear && (r = gfx_plan_set_state_device(w->trees[0].plan, w->d_u))) return r;
  w->state_dirty = false;  // the host copy is no longer the value; the device one is
  w->res_valid = false;
  return GFX_OK;
}

int gfx_ws_add_expression(gfx_workspace *w, const char *expr, gfx_mesh_im *mim, int der_order) {
  if (!w || !expr || !mim) return GFX_E_INVALID;
  if (der_order != 1 && der_order != 2)
    return wfail(w, GFX_E_UNSUPPORTED, "derivation order must be 1 or 2");
  if (mim->mesh->ctx != w->ctx) return wfail(w, GFX_E_INVALID, "mesh_im of another context");
  ref(mim);
  w->exprs.push_back({expr, mim, der_order, -1});
  w->dirty = true;
  return GFX_OK;
}

int gfx_ws_set_form(gfx_workspace *w, int form, int precision) {
  if (!w) return GFX_E_INVALID;
  w->form = form;
  w->precision = precision;
  for (auto &t : w->trees)
    for (gfx_plan *p : {t.plan, t.plan1})
      if (p) {
        int r = gfx_plan_set_form(p, form, precision);
        if (r) return r;
      }
  return GFX_OK;
}

int gfx_ws_compile(gfx_workspace *w) {
  if (!w) return GFX_E_INVALID;
  int r = compile(w);
  if (r == GFX_OK) w->dirty = true;  // the plans are made by the next assembly
  return r;
}

int gfx_ws_set_assembled_matrix(gfx_workspace *w, int orientation, int64_t n, const int64_t *d_ptr,
                                const int32_t *d_ind, double *d_vals) {
  if (!w) return GFX_E_INVALID;
  if (!d_ptr) {
    w->ext = false;
    return GFX_OK;
  }
  if (!d_ind || !d_vals || n < 0 || (orientation != GFX_CSR && orientation != GFX_CSC))
    return wfail(w, GFX_E_INVALID, "gfx_ws_set_assembled_matrix: invalid argument");
  w->ext = true;
  w->ext_orient = orientation;
  w->ext_n = n;
  w->ext_ptr = d_ptr;
  w->ext_ind = d_ind;
  w->ext_vals = d_vals;
  return GFX_OK;
}

// ga_workspace::assembly(order) (src/getfem_generic_assembly_workspace.cc:
// 791-936).  order 2: the tangent matrix (into the plan's CSR for one tree,
// the workspace's union matrix for several, or added into the caller's
// matrix after set_assembled_matrix).  order 1: the vector of the first
// derivative at the unknown's current value; for the Neo-Hookean term the
// element residual gathered per node, for the linear terms (whose first
// derivative is bilinear in u and the test function) K1 u with K1 the matrix
// of the order-1 tree.
int gfx_ws_assembly(gfx_workspace *w, int order) {
  if (!w) return GFX_E_INVALID;
  if (order != 1 && order != 2)
    return wfail(w, GFX_E_UNSUPPORTED, "only order-1 (vector) and order-2 (matrix) assembly are supported");
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  int r;
  if (w->dirty) {
    if ((r = setup(w))) return r;
  } else if (w->coefs_dirty) {
    // new coefficient values: the plans re-derive their thresholds and
    // patterns at their next assembly
    for (auto &t : w->trees) {
      if ((r = rebind(w, t.plan, t.p2))) return r;
      if ((r = rebind(w, t.plan1, t.p1))) return r;
      t.valid2 = t.valid1 = false;
    }
    w->coefs_dirty = false;
    w->res_valid = false;
  }
  const WsVar &u = w->vars.at(w->unknown);
  hipStream_t s = w->ctx->stream;
  if (w->nonlinear) {
    // One element pass gives the tangent and the residual (the element
    // kernel writes both; the matrix and vector gathers follow), so a
    // Newton iteration's assembly(2) + assembly(1) runs it once.
    gfx_plan *p = w->trees[0].plan;
    if (w->state_dirty) {
      if ((r = gfx_plan_set_state(p, u.V.data()))) return r;
      w->state_dirty = false;
      w->res_valid = false;
    }
    w->V_out = p->d_rhs;
    if (order == 1) {
      if (w->res_valid) return GFX_OK;
      if ((r = gfx_plan_assemble_residual(p))) return r;
      w->res_valid = true;
      return GFX_OK;
    }
    if ((r = gfx_plan_assemble(p))) return r;
    w->res_valid = true;
    if (w->ext) return gfx_plan_add_to_matrix(p, w->ext_orient, w->ext_n, w->ext_ptr, w->ext_ind, w->ext_vals);
    return GFX_OK;
  }
  if (order == 2) {
    std::vector<WsTree *> mt = matrix_trees(w);
    for (WsTree *t : mt) {
      if ((r = gfx_plan_assemble(t->plan))) return r;
      t->valid2 = true;
    }
    if (w->ext) {
      for (WsTree *t : mt)
        if ((r = gfx_plan_add_to_matrix(t->plan, w->ext_orient, w->ext_n, w->ext_ptr, w->ext_ind, w->ext_vals)))
          return r;
    } else if (mt.size() > 1) {
      if ((r = build_union(w))) return r;
      WHIP(w, launch_fill_f64(s, w->nnz, w->d_v, 0.0));
      for (WsTree *t : mt) WHIP(w, launch_scatter_add(s, t->plan->nnz, t->d_map, t->plan->d_vals, w->d_v));
    }
    return GFX_OK;
  }
  // order 1, linear terms: sum over the trees of K1 u
  if (w->state_dirty) {
    WHIP(w, hipMemcpyAsync(w->d_u, u.V.data(), 8 * (size_t)w->ndof, hipMemcpyHostToDevice, s));
    WHIP(w, hipStreamSynchronize(s));
    w->state_dirty = false;
  }
  WHIP(w, launch_fill_f64(s, w->ndof, w->d_V, 0.0));
  for (auto &t : w->trees) {
    if (t.e1.empty()) continue;
    gfx_plan *p = t.same1 ? t.plan : t.plan1;
    bool &valid = t.same1 ? t.valid2 : t.valid1;
    if (!valid) {
      if ((r = gfx_plan_assemble(p))) return r;
      valid = true;
    }
    if ((r = gfx_plan_spmv_add(p, w->d_u, w->d_V))) return r;
  }
  w->V_out = w->d_V;
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

// The workspace's own assembled matrix: the single tree's plan CSR, or the
// union matrix of several trees.
int gfx_ws_device_csr(gfx_workspace *w, int64_t *nrows, int64_t *nnz, const int64_t **d_indptr,
                      const int32_t **d_indices, double **d_vals) {
  if (!w) return GFX_E_INVALID;
  std::vector<WsTree *> mt = matrix_trees(w);
  if (mt.empty()) return wfail(w, GFX_E_STATE, "no matrix assembled yet");
  if (w->ext) return wfail(w, GFX_E_STATE, "the assembled matrix is the caller's (set_assembled_matrix)");
  if (mt.size() == 1) {
    gfx_plan *p = mt[0]->plan;
    if (nrows) *nrows = p->ndof;
    if (nnz) *nnz = p->nnz;
    return gfx_plan_device_csr(p, d_indptr, d_indices, d_vals);
  }
  if (w->nnz < 0) return wfail(w, GFX_E_STATE, "no matrix assembled yet");
  if (nrows) *nrows = w->ndof;
  if (nnz) *nnz = w->nnz;
  if (d_indptr) *d_indptr = w->d_ip;
  if (d_indices) *d_indices = w->d_ix;
  if (d_vals) *d_vals = w->d_v;
  return GFX_OK;
}

int gfx_ws_matrix_dims(gfx_workspace *w, int64_t *nrows, int64_t *nnz) {
  return gfx_ws_device_csr(w, nrows, nnz, nullptr, nullptr, nullptr);
}

int gfx_ws_export_csr(gfx_workspace *w, int64_t *h_indptr, int32_t *h_indices, double *h_vals) {
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  double *v = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, &v);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  WHIP(w, hipStreamSynchronize(w->ctx->stream));
  if (h_indptr) WHIP(w, hipMemcpy(h_indptr, ip, 8 * (size_t)(n + 1), hipMemcpyDeviceToHost));
  if (h_indices) WHIP(w, hipMemcpy(h_indices, ix, 4 * (size_t)nnz, hipMemcpyDeviceToHost));
  if (h_vals) WHIP(w, hipMemcpy(h_vals, v, 8 * (size_t)nnz, hipMemcpyDeviceToHost));
  return GFX_OK;
}

int gfx_ws_export_csc(gfx_workspace *w, int64_t *h_colptr, uint32_t *h_rows, double *h_vals) {
  int64_t n = 0, nnz = 0;
  int r = gfx_ws_matrix_dims(w, &n, &nnz);
  if (r) return r;
  if (!h_colptr) return GFX_E_INVALID;
  std::vector<int64_t> ip((size_t)n + 1);
  std::vector<int32_t> ix((size_t)nnz);
  std::vector<double> v(h_vals ? (size_t)nnz : 0);
  if ((r = gfx_ws_export_csr(w, ip.data(), ix.data(), h_vals ? v.data() : nullptr))) return r;
  // transpose (counting sort by column keeps the rows sorted)
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) cnt[ix[k] + 1]++;
  for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
  std::copy(cnt.begin(), cnt.end(), h_colptr);
  for (int64_t row = 0; row < n; ++row)
    for (int64_t k = ip[row]; k < ip[row + 1]; ++k) {
      const int64_t d = cnt[ix[k]]++;
      if (h_rows) h_rows[d] = (uint32_t)row;
      if (h_vals) h_vals[d] = v[k];
    }
  return GFX_OK;
}

static void describe_params(const GwflKernelParams &p, std::string &d) {
  d = p.describe + " unknown=" + p.unknown;
  const char *names3[3] = {"alpha", "beta", "gamma"};
  const char *names[3] = {"a/lambda", "mu", "c"};
  for (int s = 0; s < 3; ++s) {
    if (p.coef[s].terms.empty()) continue;
    d += std::string(" ") + (p.term == GFX_TERM_ISOTROPIC ? names3[s] : names[s]) + "=";
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

// "name:kind:n;..." declarations of the host-only compile entry points
static bool parse_decls(const char *decls, std::map<std::string, GwflVar> &gv) {
  std::string ds(decls);
  size_t b = 0;
  while (b < ds.size()) {
    size_t e = ds.find(';', b);
    if (e == std::string::npos) e = ds.size();
    const std::string item = ds.substr(b, e - b);
    b = e + 1;
    if (item.empty()) continue;
    const size_t c1 = item.find(':'), c2 = item.find(':', c1 + 1);
    if (c1 == std::string::npos || c2 == std::string::npos) return false;
    GwflVar g;
    g.name = item.substr(0, c1);
    const std::string kind = item.substr(c1 + 1, c2 - c1 - 1);
    g.qdim = std::atoi(item.c_str() + c2 + 1);
    g.is_fem = kind == "var" || kind == "p0";
    g.is_constant = kind != "var";
    g.per_element = kind == "p0";
    gv[g.name] = g;
  }
  return true;
}

int gfx_gwfl_order(const char *expr, const char *decls, int *order) {
  if (!expr || !decls || !order) return GFX_E_INVALID;
  std::map<std::string, GwflVar> gv;
  if (!parse_decls(decls, gv)) return GFX_E_INVALID;
  std::string err;
  *order = gwfl_expression_order(expr, gv, err);
  return *order < 0 ? GFX_E_UNSUPPORTED : GFX_OK;
}

int gfx_gwfl_compile(const char *exprs, const char *decls, char *out, int64_t len) {
  if (!exprs || !decls || !out || len <= 0) return GFX_E_INVALID;
  std::map<std::string, GwflVar> gv;
  if (!parse_decls(decls, gv)) return GFX_E_INVALID;
  std::vector<std::string> ex;
  std::string es(exprs);
  size_t b = 0;
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
  std::vector<WsTree *> mt = matrix_trees(w);
  if (mt.empty()) return wfail(w, GFX_E_STATE, "no assembly done yet");
  if (mt.size() > 1)
    return wfail(w, GFX_E_STATE, "the workspace has several trees (mesh_ims): use the gfx_ws_* matrix calls");
  *out = mt[0]->plan;
  return GFX_OK;
}

int gfx_ws_nb_trees(gfx_workspace *w) { return w ? (int)matrix_trees(w).size() : -1; }

int gfx_ws_describe(gfx_workspace *w, char *buf, int64_t len) {
  if (!w || !buf || len <= 0) return GFX_E_INVALID;
  std::string d;
  for (auto &t : w->trees) {
    std::string x;
    describe_params(t.e2.empty() ? t.p1 : t.p2, x);
    d += (d.empty() ? "" : "; ") + x;
  }
  std::strncpy(buf, d.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return GFX_OK;
}

}  // extern "C"
