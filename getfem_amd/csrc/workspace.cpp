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
  double *d_val = nullptr;      // values given on the device (polynomial route)
  bool dev = false;             // d_val, not V, holds the current values
};

struct WsExpr {
  std::string expr;
  gfx_mesh_im *mim = nullptr;
  int der_order = 2;
  int order = -1;  // test-function order, set by the compile stage
};

struct ScProg {
  GwflScalarProgram prog;
  int vf = 0;                   // order 1: the field of the variable
  int nm = 0;
  double *d_coef = nullptr;
  int32_t *d_off = nullptr, *d_sym = nullptr, *d_out = nullptr;
  int32_t *d_mptr = nullptr;    // order 2: monomials of each output (sorted by output)
};

struct ScBlock {                // an order-2 block: test variable x trial variable
  std::string tv, sv;
  ScProg p;
  int vt = 0, vs = 0;           // their fields
};

struct ScField {
  std::string name;
  gfx_mesh_fem *mf = nullptr;
  int off = 0;
  double *d_U = nullptr, *d_phi = nullptr, *d_gref = nullptr;
};

struct ScVar {                  // a fem variable's order-1 block
  std::string name;
  ScProg p;
  int32_t *d_inc_ptr = nullptr;
  uint32_t *d_inc = nullptr;
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

struct ScalarRoute {
  uint64_t version = ~0ull;
  int err_code = GFX_OK;        // the route's build result at `version`
  std::string err;
  gfx_mesh_im *mim = nullptr;
  std::vector<ScField> fields;
  std::vector<std::string> consts;
  int const_off = 0, nsym = 0;
  double *d_consts = nullptr, *d_qw = nullptr;
  double *d_gN = nullptr;       // hexahedra: the geometric transformation's gradients at the points
  ScProg p0;
  std::vector<ScVar> vars;
  double *d_elem = nullptr, *d_part = nullptr, *d_sum = nullptr, *d_V = nullptr;
  int *d_err = nullptr;
  int64_t nvec = 0;
  // order 2 (polymat.hip): the blocks, or why they do not compile
  std::vector<ScBlock> blocks;
  std::string err2;
  double *d_pelem = nullptr, *d_thr = nullptr;
  int64_t *d_cnt = nullptr, *d_off = nullptr, *d_scan = nullptr, *d_nnz = nullptr;
  // kept entries (keys, values and their sorted copies), sort and run work
  int64_t cap = 0;
  uint64_t *d_keys = nullptr, *d_keys2 = nullptr;
  double *d_vals = nullptr, *d_vals2 = nullptr;
  int64_t *d_work = nullptr;
  void *d_sort = nullptr;
  size_t sort_cap = 0;
  // the assembled matrix: nvec x nvec CSR
  int64_t *d_ip = nullptr;
  int32_t *d_ix = nullptr;
  double *d_v = nullptr;
  int64_t nnz = -1;
};

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
  // the polynomial route (scalar.hip, polymat.hip): potentials, derivatives
  // and tangents the matrix kernels do not cover (data fields, P0 or several
  // variables, blocks between mesh_fems, pulled-back coordinates, hexahedra)
  uint64_t version = 0;         // registrations and expressions; the route rebuilds on change
  bool mat_fail = false;        // the order-2 compile failed at this version
  ScalarRoute *sc = nullptr;
  int64_t vec_n = 0;            // size of the last assembled vector
  double potential = 0.0;       // the last assembly(0)
  bool has_potential = false;
  // set_assembled_matrix: a caller-owned device matrix the assembly adds into
  bool ext = false;
  int ext_orient = GFX_CSR;
  int64_t ext_n = 0;
  const int64_t *ext_ptr = nullptr;
  const int32_t *ext_ind = nullptr;
  double *ext_vals = nullptr;
  // solves and dof constraints on the workspace's matrix
  double *d_cg = nullptr, *d_go = nullptr, *d_gm = nullptr;
  uint8_t *d_fixed = nullptr;
  int *d_cerr = nullptr;
  int64_t cg_n = 0, con_n = 0;
  size_t gm_size = 0;
  IluData ilu;                  // GMRES's ILU preconditioner
  int64_t ilu_n = -1, ilu_nnz = -1;
  bool ilu_ready = false;       // the factor belongs to the current matrix (cleared by every assembly)
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
    if (e.order <= 1 && e.der_order >= 1) t->e1.push_back((int)i);
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

// ------------------------------------------------ the polynomial route (orders 0, 1, 2)
void sc_free_prog(ScProg &p) {
  wfree(p.d_coef); wfree(p.d_off); wfree(p.d_sym); wfree(p.d_out); wfree(p.d_mptr);
  p.nm = 0;
}

void sc_release(ScalarRoute *sc) {
  if (!sc) return;
  for (auto &f : sc->fields) { wfree(f.d_U); wfree(f.d_phi); wfree(f.d_gref); }
  sc_free_prog(sc->p0);
  for (auto &v : sc->vars) { sc_free_prog(v.p); wfree(v.d_inc_ptr); wfree(v.d_inc); }
  wfree(sc->d_consts); wfree(sc->d_qw); wfree(sc->d_gN); wfree(sc->d_elem); wfree(sc->d_part); wfree(sc->d_sum);
  wfree(sc->d_V); wfree(sc->d_err);
  for (auto &b : sc->blocks) sc_free_prog(b.p);
  wfree(sc->d_pelem); wfree(sc->d_thr); wfree(sc->d_cnt); wfree(sc->d_off); wfree(sc->d_scan);
  wfree(sc->d_nnz); wfree(sc->d_keys); wfree(sc->d_keys2); wfree(sc->d_vals); wfree(sc->d_vals2);
  wfree(sc->d_work); wfree(sc->d_sort); wfree(sc->d_ip); wfree(sc->d_ix); wfree(sc->d_v);
  sc->blocks.clear(); sc->err2.clear();
  sc->cap = 0; sc->sort_cap = 0; sc->nnz = -1;
  sc->fields.clear(); sc->vars.clear(); sc->consts.clear();
  sc->p0 = ScProg();
  sc->mim = nullptr;
}

std::map<std::string, GwflVar> gwfl_vars(const gfx_workspace *w) {
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
  return gv;
}

template <typename T>
hipError_t upload(T **d, const std::vector<T> &h) {
  if (hipMalloc((void **)d, sizeof(T) * std::max<size_t>(h.size(), 1)) != hipSuccess) return hipErrorOutOfMemory;
  if (h.empty()) return hipSuccess;
  return hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

// Compile the potential and every fem variable's derivative into programs,
// lay out the symbol table (fields, then fixed-size constants) and upload
// the programs, the fields' base tables and the incidences of each
// variable's mesh_fem.  Cached per registration/expression version.
int sc_build(gfx_workspace *w) {
  ScalarRoute *sc = w->sc;
  if (sc->version == w->version)
    return sc->err_code == GFX_OK ? GFX_OK : wfail(w, sc->err_code, sc->err);
  sc_release(sc);
  sc->version = w->version;
  sc->err_code = GFX_OK;
  auto bail = [&](int code, const std::string &m) {
    sc->err_code = code;
    sc->err = m;
    return wfail(w, code, m);
  };
  if (w->exprs.empty()) return bail(GFX_E_STATE, "no expression");
  gfx_mesh_im *mim = w->exprs[0].mim;
  for (auto &e : w->exprs)
    if (e.mim != mim) return bail(GFX_E_UNSUPPORTED, "the polynomial route takes expressions on one mesh_im");
  gfx_mesh *m = mim->mesh;
  if (m->gt != GT_TET_P1 && m->gt != GT_HEX_Q1)
    return bail(GFX_E_UNSUPPORTED, "the polynomial route needs GT_PK(3,1) or GT_QK(3,1) elements");
  sc->mim = mim;
  const std::map<std::string, GwflVar> gv = gwfl_vars(w);
  std::vector<std::string> all, deriv, deriv2;  // derived at least once (der_order >= 1), twice (2)
  for (auto &e : w->exprs) {
    all.push_back(e.expr);
    if (e.der_order >= 1) deriv.push_back(e.expr);
    if (e.der_order == 2) deriv2.push_back(e.expr);
  }
  std::string err;
  if (!gwfl_compile_scalar(all, gv, 0, "", sc->p0.prog, err)) return bail(GFX_E_UNSUPPORTED, err);
  for (auto &kv : w->vars) {
    if (kv.second.kind != VAR_FEM) continue;
    ScVar v;
    v.name = kv.first;
    if (!gwfl_compile_scalar(deriv, gv, 1, kv.first, v.p.prog, err)) return bail(GFX_E_UNSUPPORTED, err);
    sc->vars.push_back(v);
  }
  // order 2: one block per (test variable, trial variable) with terms; a
  // block that does not compile leaves orders 0 and 1 usable
  if (!deriv2.empty())
    for (auto &kt : w->vars) {
      if (kt.second.kind != VAR_FEM || !sc->err2.empty()) continue;
      for (auto &ks : w->vars) {
        if (ks.second.kind != VAR_FEM) continue;
        ScBlock b;
        b.tv = kt.first;
        b.sv = ks.first;
        if (!gwfl_compile_scalar(deriv2, gv, 2, b.tv, b.p.prog, err, b.sv)) {
          sc->err2 = err;
          sc->blocks.clear();
          break;
        }
        if (b.p.prog.monos.empty()) continue;
        std::stable_sort(b.p.prog.monos.begin(), b.p.prog.monos.end(),
                         [](const GwflScalarProgram::Mono &x, const GwflScalarProgram::Mono &y) {
                           return x.out < y.out;
                         });
        sc->blocks.push_back(b);
      }
    }
  // the fields and constants the programs read, and the variables
  std::vector<std::string> fnames;
  auto need_field = [&](const std::string &n) {
    if (std::find(fnames.begin(), fnames.end(), n) == fnames.end()) fnames.push_back(n);
  };
  auto scan = [&](const GwflScalarProgram &p) -> bool {
    for (const std::string &sy : p.syms) {
      const size_t c = sy.find(':');
      const std::string var = sy.substr(0, c);
      if (var == "c") {
        const std::string n = sy.substr(2);
        auto it = w->vars.find(n);
        if (it == w->vars.end()) { err = "unknown name " + n; return false; }
        if (it->second.kind == CONST_FIXED) {
          if (std::find(sc->consts.begin(), sc->consts.end(), n) == sc->consts.end()) sc->consts.push_back(n);
        } else {
          need_field(n);
        }
      } else {
        need_field(var);
      }
    }
    return true;
  };
  if (!scan(sc->p0.prog)) return bail(GFX_E_UNSUPPORTED, err);
  for (auto &v : sc->vars) {
    if (!scan(v.p.prog)) return bail(GFX_E_UNSUPPORTED, err);
    if (!v.p.prog.monos.empty()) need_field(v.name);
  }
  for (auto &b : sc->blocks) {
    if (!scan(b.p.prog)) return bail(GFX_E_UNSUPPORTED, err);
    need_field(b.tv);
    need_field(b.sv);
  }
  if (fnames.empty()) need_field(sc->vars.empty() ? w->vars.begin()->first : sc->vars[0].name);
  if ((int)fnames.size() > kScalarMaxFields)
    return bail(GFX_E_UNSUPPORTED, "the polynomial route reads at most 4 fem fields");
  int off = 0;
  const IntegrationRule &rule = mim->rule;
  const int nq = rule.n();
  for (const std::string &n : fnames) {
    const WsVar &v = w->vars.at(n);
    if (v.kind == CONST_FIXED || !v.mf) return bail(GFX_E_UNSUPPORTED, n + " is not a fem field");
    if (v.mf->mesh != m) return bail(GFX_E_UNSUPPORTED, n + " lives on another mesh");
    if (v.mf->fem.gt != m->gt) return bail(GFX_E_UNSUPPORTED, n + ": an element of another shape than the mesh's");
    ScField f;
    f.name = n;
    f.mf = v.mf;
    f.off = off;
    off += 4 * v.mf->qdim;
    const int nd = v.mf->fem.nd;
    std::vector<double> phi((size_t)nq * nd), gref((size_t)nq * nd * 3);
    for (int q = 0; q < nq; ++q) fem_eval(v.mf->fem, &rule.pts[3 * (size_t)q], &phi[(size_t)q * nd],
                                          &gref[(size_t)q * nd * 3]);
    if (upload(&f.d_phi, phi) != hipSuccess || upload(&f.d_gref, gref) != hipSuccess ||
        hipMalloc((void **)&f.d_U, 8 * (size_t)std::max<int64_t>(v.mf->ndof, 1)) != hipSuccess) {
      sc->fields.push_back(f);
      return bail(GFX_E_OOM, "polynomial route tables");
    }
    sc->fields.push_back(f);
  }
  sc->const_off = off;
  for (const std::string &n : sc->consts) {
    if (w->vars.at(n).V.size() != 1) return bail(GFX_E_UNSUPPORTED, "constant " + n + " is not a scalar");
    ++off;
  }
  sc->nsym = off;
  if (off > kScalarMaxSyms) return bail(GFX_E_UNSUPPORTED, "too many symbols for the polynomial route");
  auto field_index = [&](const std::string &n) {
    return (int)(std::find(fnames.begin(), fnames.end(), n) - fnames.begin());
  };
  // the programs in table slots
  auto map_prog = [&](ScProg &p) -> int {
    std::vector<int> slot;
    for (const std::string &sy : p.prog.syms) {
      const size_t c = sy.find(':');
      const std::string var = sy.substr(0, c);
      if (var == "c") {
        const std::string n = sy.substr(2);
        auto ci = std::find(sc->consts.begin(), sc->consts.end(), n);
        slot.push_back(ci != sc->consts.end() ? sc->const_off + (int)(ci - sc->consts.begin())
                                              : sc->fields[field_index(n)].off);
        continue;
      }
      const ScField &f = sc->fields[field_index(var)];
      const int q = f.mf->qdim;
      const size_t us = sy.find('_', c);
      const std::string tag = sy.substr(c + 1, us - c - 1), comp = sy.substr(us + 1);
      const int a = q == 1 ? 0 : comp[0] - '0';
      slot.push_back(tag == "V" ? f.off + a : f.off + q + 3 * a + (comp.back() - '0'));
    }
    std::vector<double> coef;
    std::vector<int32_t> offs{0}, sym, out;
    for (auto &mo : p.prog.monos) {
      coef.push_back(mo.coef);
      for (int x : mo.s) sym.push_back(slot[(size_t)x]);
      offs.push_back((int32_t)sym.size());
      out.push_back(mo.out);
    }
    p.nm = (int)coef.size();
    if (upload(&p.d_coef, coef) != hipSuccess || upload(&p.d_off, offs) != hipSuccess ||
        upload(&p.d_sym, sym) != hipSuccess || upload(&p.d_out, out) != hipSuccess)
      return bail(GFX_E_OOM, "polynomial route programs");
    return GFX_OK;
  };
  int r;
  if ((r = map_prog(sc->p0))) return r;
  int64_t pelem = 0;
  for (auto &b : sc->blocks) {
    if ((r = map_prog(b.p))) return r;
    b.vt = field_index(b.tv);
    b.vs = field_index(b.sv);
    const gfx_mesh_fem *ft = sc->fields[b.vt].mf, *fs = sc->fields[b.vs].mf;
    if (ft->fem.nd > kPolyMaxNd || fs->fem.nd > kPolyMaxNd ||
        (int64_t)ft->fem.nd * ft->qdim * fs->fem.nd * fs->qdim > kPolyMaxEntries || b.p.prog.nout > kPolyMaxOut)
      return bail(GFX_E_UNSUPPORTED, "the order-2 block " + b.tv + " x " + b.sv +
                                         " is larger than the polynomial route's element matrices");
    std::vector<int32_t> mptr((size_t)b.p.prog.nout + 1, 0);
    for (auto &mo : b.p.prog.monos) mptr[(size_t)mo.out + 1]++;
    for (size_t o = 0; o + 1 < mptr.size(); ++o) mptr[o + 1] += mptr[o];
    if (upload(&b.p.d_mptr, mptr) != hipSuccess) return bail(GFX_E_OOM, "order-2 route programs");
    pelem = std::max(pelem, m->n_cv * (int64_t)ft->fem.nd * ft->qdim * fs->fem.nd * fs->qdim);
  }
  if (!sc->blocks.empty() &&
      (hipMalloc((void **)&sc->d_pelem, 8 * (size_t)std::max<int64_t>(pelem, 1)) != hipSuccess ||
       hipMalloc((void **)&sc->d_thr, 8 * (size_t)std::max<int64_t>(m->n_cv, 1)) != hipSuccess ||
       hipMalloc((void **)&sc->d_cnt, 8 * (size_t)std::max<int64_t>(m->n_cv, 1)) != hipSuccess ||
       hipMalloc((void **)&sc->d_off, 8 * (size_t)std::max<int64_t>(m->n_cv, 1)) != hipSuccess ||
       hipMalloc((void **)&sc->d_scan, 8 * (size_t)std::max<int64_t>(scan_tmp_elems(m->n_cv), 1)) != hipSuccess))
    return bail(GFX_E_OOM, "order-2 route buffers");
  int64_t elem = m->n_cv, nvec = 0;
  hipStream_t s = w->ctx->stream;
  for (auto &v : sc->vars) {
    if ((r = map_prog(v.p))) return r;
    const WsVar &wv = w->vars.at(v.name);
    nvec = std::max(nvec, wv.first + wv.mf->ndof);
    if (v.p.nm == 0) continue;
    v.p.vf = field_index(v.name);
    const int nd = wv.mf->fem.nd, Q = wv.mf->qdim;
    elem = std::max(elem, m->n_cv * nd * Q);
    const int64_t nn = wv.mf->ndof / Q;
    int32_t *tmp = nullptr;
    if (hipMalloc((void **)&v.d_inc_ptr, 4 * (size_t)(nn + 1)) != hipSuccess ||
        hipMalloc((void **)&v.d_inc, 4 * (size_t)std::max<int64_t>(m->n_cv * nd, 1)) != hipSuccess ||
        hipMalloc((void **)&tmp, 4 * (size_t)(2 * nn + scan_tmp_elems(nn))) != hipSuccess) {
      wfree(tmp);
      return bail(GFX_E_OOM, "polynomial route incidences");
    }
    hipError_t e = launch_incidences(s, m->n_cv, nd, Q, wv.mf->d_dofs, nn, v.d_inc_ptr, v.d_inc, tmp);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    wfree(tmp);
    if (e != hipSuccess) return bail(GFX_E_HIP, std::string("incidences: ") + hipGetErrorString(e));
  }
  sc->nvec = nvec;
  if (m->gt == GT_HEX_Q1) {
    std::vector<double> gN((size_t)nq * m->nv * 3), Nv((size_t)m->nv);
    for (int q = 0; q < nq; ++q) geotrans_eval(m->gt, &rule.pts[3 * (size_t)q], Nv.data(), &gN[(size_t)q * m->nv * 3]);
    if (upload(&sc->d_gN, gN) != hipSuccess) return bail(GFX_E_OOM, "polynomial route geometry");
  }
  if (upload(&sc->d_qw, rule.w) != hipSuccess ||
      hipMalloc((void **)&sc->d_elem, 8 * (size_t)std::max<int64_t>(elem, 1)) != hipSuccess ||
      hipMalloc((void **)&sc->d_part, 8 * (size_t)scalar_partials()) != hipSuccess ||
      hipMalloc((void **)&sc->d_sum, 8) != hipSuccess ||
      hipMalloc((void **)&sc->d_V, 8 * (size_t)std::max<int64_t>(nvec, 1)) != hipSuccess ||
      hipMalloc((void **)&sc->d_err, 4) != hipSuccess ||
      hipMalloc((void **)&sc->d_consts, 8 * std::max<size_t>(sc->consts.size(), 1)) != hipSuccess)
    return bail(GFX_E_OOM, "polynomial route buffers");
  return GFX_OK;
}

// the current values of a variable on the host (a device copy is read back)
int host_values(gfx_workspace *w, const std::string &n, std::vector<double> &out) {
  const WsVar &v = w->vars.at(n);
  out = v.V;
  if (v.dev) {
    WHIP(w, hipStreamSynchronize(w->ctx->stream));
    WHIP(w, hipMemcpy(out.data(), v.d_val, 8 * out.size(), hipMemcpyDeviceToHost));
  }
  return GFX_OK;
}

// variables that Inv(...) expanded around zero must be zero
bool sc_zero_ok(gfx_workspace *w, const GwflScalarProgram &p) {
  for (const std::string &z : p.zero_vars) {
    std::vector<double> V;
    if (host_values(w, z, V)) return false;
    for (double x : V)
      if (x != 0.0) {
        wfail(w, GFX_E_UNSUPPORTED, "Inv(...) of " + z + " is expanded around " + z + " = 0, but " + z +
                                        " is not zero");
        return false;
      }
  }
  return true;
}

// the fields' values (from the host, or device to device), the constants,
// the error flag cleared
int sc_upload(gfx_workspace *w) {
  ScalarRoute *sc = w->sc;
  hipStream_t s = w->ctx->stream;
  for (auto &f : sc->fields) {
    const WsVar &v = w->vars.at(f.name);
    if (v.dev)
      WHIP(w, hipMemcpyAsync(f.d_U, v.d_val, 8 * v.V.size(), hipMemcpyDeviceToDevice, s));
    else
      WHIP(w, hipMemcpyAsync(f.d_U, v.V.data(), 8 * v.V.size(), hipMemcpyHostToDevice, s));
  }
  std::vector<double> cv;
  for (const std::string &n : sc->consts) cv.push_back(w->vars.at(n).V[0]);
  if (!cv.empty()) WHIP(w, hipMemcpyAsync(sc->d_consts, cv.data(), 8 * cv.size(), hipMemcpyHostToDevice, s));
  WHIP(w, hipMemsetAsync(sc->d_err, 0, 4, s));
  // the host buffers above are read before they can change
  WHIP(w, hipStreamSynchronize(s));
  return GFX_OK;
}

int sc_grow(gfx_workspace *w, int64_t need) {
  ScalarRoute *sc = w->sc;
  if (need <= sc->cap) return GFX_OK;
  const int64_t cap = std::max<int64_t>(need, 2 * sc->cap);
  hipStream_t s = w->ctx->stream;
  WHIP(w, hipStreamSynchronize(s));
  uint64_t *k = nullptr, *k2 = nullptr;
  double *v = nullptr, *v2 = nullptr, *vv = nullptr;
  int64_t *wk = nullptr;
  int32_t *ix = nullptr;
  const size_t n8 = 8 * (size_t)cap;
  if (hipMalloc((void **)&k, n8) != hipSuccess || hipMalloc((void **)&k2, n8) != hipSuccess ||
      hipMalloc((void **)&v, n8) != hipSuccess || hipMalloc((void **)&v2, n8) != hipSuccess ||
      hipMalloc((void **)&vv, n8) != hipSuccess || hipMalloc((void **)&ix, 4 * (size_t)cap) != hipSuccess ||
      hipMalloc((void **)&wk, 8 * (size_t)(3 * cap + 1 + scan_tmp_elems(cap))) != hipSuccess) {
    wfree(k); wfree(k2); wfree(v); wfree(v2); wfree(vv); wfree(ix); wfree(wk);
    return wfail(w, GFX_E_OOM, "order-2 route entries");
  }
  // the entries emitted so far move with the buffers
  if (sc->cap > 0) {
    WHIP(w, hipMemcpy(k, sc->d_keys, 8 * (size_t)sc->cap, hipMemcpyDeviceToDevice));
    WHIP(w, hipMemcpy(v, sc->d_vals, 8 * (size_t)sc->cap, hipMemcpyDeviceToDevice));
  }
  wfree(sc->d_keys); wfree(sc->d_keys2); wfree(sc->d_vals); wfree(sc->d_vals2); wfree(sc->d_v);
  wfree(sc->d_ix); wfree(sc->d_work);
  sc->d_keys = k; sc->d_keys2 = k2; sc->d_vals = v; sc->d_vals2 = v2; sc->d_v = vv; sc->d_ix = ix;
  sc->d_work = wk;
  sc->cap = cap;
  sc->nnz = -1;
  return GFX_OK;
}

// assembly(2) on the polynomial route: every block's element matrices, their
// entries above the element threshold as keys in element order, then one
// sort and sum into the nvec x nvec matrix (tested dofs x trial dofs at the
// variables' intervals)
int sc_run2(gfx_workspace *w) {
  if (!w->sc) w->sc = new (std::nothrow) ScalarRoute();
  if (!w->sc) return wfail(w, GFX_E_OOM, "workspace");
  int r;
  if ((r = sc_build(w))) return r;
  ScalarRoute *sc = w->sc;
  if (!sc->err2.empty()) return wfail(w, GFX_E_UNSUPPORTED, sc->err2);
  gfx_mesh *m = sc->mim->mesh;
  hipStream_t s = w->ctx->stream;
  for (auto &b : sc->blocks)
    if (!sc_zero_ok(w, b.p.prog)) return GFX_E_UNSUPPORTED;
  if ((r = sc_upload(w))) return r;
  const int64_t n = sc->nvec;
  if (n > INT32_MAX) return wfail(w, GFX_E_UNSUPPORTED, "more than 2^31 dofs");
  if (!sc->d_ip) {
    if (hipMalloc((void **)&sc->d_ip, 8 * (size_t)(n + 1)) != hipSuccess ||
        hipMalloc((void **)&sc->d_nnz, 8) != hipSuccess)
      return wfail(w, GFX_E_OOM, "order-2 route matrix");
  }
  PolyMatArgs a{};
  a.n_cv = m->n_cv; a.pts = m->d_pts; a.cvx = m->d_cvx; a.nv = m->nv; a.gN = sc->d_gN;
  a.nq = sc->mim->rule.n(); a.qw = sc->d_qw;
  a.nf = (int)sc->fields.size();
  for (int i = 0; i < a.nf; ++i) {
    const ScField &f = sc->fields[i];
    a.f[i] = ScalarField{f.mf->d_dofs, f.d_U, f.d_phi, f.d_gref, f.mf->fem.nd, f.mf->qdim, f.off};
  }
  a.nconst = (int)sc->consts.size(); a.const_off = sc->const_off; a.consts = sc->d_consts;
  a.elem = sc->d_pelem; a.thr = sc->d_thr; a.cnt = sc->d_cnt; a.err = sc->d_err;
  int64_t total = 0;
  for (auto &b : sc->blocks) {
    a.vt = b.vt; a.vs = b.vs;
    a.nout = b.p.prog.nout;
    a.mcoef = b.p.d_coef; a.mptr = b.p.d_mptr; a.moff = b.p.d_off; a.msym = b.p.d_sym;
    WHIP(w, launch_poly_mat(s, a));
    WHIP(w, scan_exclusive_i64(s, sc->d_cnt, sc->d_off, m->n_cv, sc->d_scan));
    int64_t tail[2] = {0, 0};
    if (m->n_cv > 0) {
      WHIP(w, hipMemcpyAsync(&tail[0], sc->d_off + m->n_cv - 1, 8, hipMemcpyDeviceToHost, s));
      WHIP(w, hipMemcpyAsync(&tail[1], sc->d_cnt + m->n_cv - 1, 8, hipMemcpyDeviceToHost, s));
    }
    WHIP(w, hipStreamSynchronize(s));
    const int64_t nb = tail[0] + tail[1];
    if (total + nb > INT32_MAX) return wfail(w, GFX_E_UNSUPPORTED, "more than 2^31 kept entries");
    if ((r = sc_grow(w, total + nb))) return r;
    const WsVar &vt = w->vars.at(b.tv), &vs = w->vars.at(b.sv);
    PolyEmitArgs pe{};
    pe.n_cv = m->n_cv;
    pe.ndt = vt.mf->fem.nd; pe.nds = vs.mf->fem.nd; pe.qt = vt.mf->qdim; pe.qs = vs.mf->qdim;
    pe.dt = vt.mf->d_dofs; pe.ds = vs.mf->d_dofs;
    pe.first_t = vt.first; pe.first_s = vs.first; pe.ncols = n;
    pe.elem = sc->d_pelem; pe.thr = sc->d_thr; pe.off = sc->d_off;
    pe.keys = sc->d_keys + total; pe.vals = sc->d_vals + total;
    WHIP(w, launch_poly_emit(s, pe));
    total += nb;
  }
  if ((r = sc_grow(w, std::max<int64_t>(total, 1)))) return r;
  int end_bit = 1;
  while (end_bit < 64 && ((uint64_t)1 << end_bit) < (uint64_t)n * (uint64_t)n) ++end_bit;
  const size_t sb = poly_sort_bytes(total, end_bit);
  if (sb > sc->sort_cap) {
    WHIP(w, hipStreamSynchronize(s));
    wfree(sc->d_sort);
    if (hipMalloc(&sc->d_sort, sb) != hipSuccess) return wfail(w, GFX_E_OOM, "order-2 route sort");
    sc->sort_cap = sb;
  }
  WHIP(w, poly_to_csr(s, sc->d_sort, sc->sort_cap, sc->d_keys, sc->d_vals, sc->d_keys2, sc->d_vals2, total,
                      end_bit, n, n, sc->d_work, sc->d_ip, sc->d_ix, sc->d_v, sc->d_nnz));
  int flag = 0;
  int64_t nnz = 0;
  WHIP(w, hipMemcpyAsync(&nnz, sc->d_nnz, 8, hipMemcpyDeviceToHost, s));
  WHIP(w, hipMemcpyAsync(&flag, sc->d_err, 4, hipMemcpyDeviceToHost, s));
  WHIP(w, hipStreamSynchronize(s));
  if (flag & 1) return wfail(w, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
  sc->nnz = nnz;
  if (w->ext) {
    if (w->ext_n != n) return wfail(w, GFX_E_INVALID, "target matrix of another size");
    WHIP(w, hipMemsetAsync(sc->d_err, 0, 4, s));
    WHIP(w, launch_add_into(s, w->ext_orient == GFX_CSC, n, sc->d_ip, sc->d_ix, sc->d_v, w->ext_ptr, w->ext_ind,
                            w->ext_vals, sc->d_err));
    WHIP(w, hipMemcpyAsync(&flag, sc->d_err, 4, hipMemcpyDeviceToHost, s));
    WHIP(w, hipStreamSynchronize(s));
    if (flag & 8) return wfail(w, GFX_E_PATTERN, "the target matrix lacks entries of the assembled pattern");
  }
  return GFX_OK;
}

// assembly(0): the potential; assembly(1): the vector of every fem variable's
// derivative, at the variable's interval (first .. first + nb_dof)
int sc_run(gfx_workspace *w, int order) {
  if (!w->sc) w->sc = new (std::nothrow) ScalarRoute();
  if (!w->sc) return wfail(w, GFX_E_OOM, "workspace");
  int r;
  if ((r = sc_build(w))) return r;
  ScalarRoute *sc = w->sc;
  gfx_mesh *m = sc->mim->mesh;
  hipStream_t s = w->ctx->stream;
  if (order == 0 && !sc_zero_ok(w, sc->p0.prog)) return GFX_E_UNSUPPORTED;
  if (order == 1)
    for (auto &v : sc->vars)
      if (!sc_zero_ok(w, v.p.prog)) return GFX_E_UNSUPPORTED;
  if ((r = sc_upload(w))) return r;
  ScalarArgs a{};
  a.n_cv = m->n_cv; a.pts = m->d_pts; a.cvx = m->d_cvx; a.nv = m->nv; a.gN = sc->d_gN;
  a.nq = sc->mim->rule.n(); a.qw = sc->d_qw;
  a.nf = (int)sc->fields.size();
  for (int i = 0; i < a.nf; ++i) {
    const ScField &f = sc->fields[i];
    a.f[i] = ScalarField{f.mf->d_dofs, f.d_U, f.d_phi, f.d_gref, f.mf->fem.nd, f.mf->qdim, f.off};
  }
  a.nconst = (int)sc->consts.size(); a.const_off = sc->const_off; a.consts = sc->d_consts;
  a.elem = sc->d_elem; a.err = sc->d_err;
  auto set_prog = [&](const ScProg &p) {
    a.nout = p.prog.nout; a.nm = p.nm; a.vf = p.vf;
    a.mcoef = p.d_coef; a.moff = p.d_off; a.msym = p.d_sym; a.mout = p.d_out;
  };
  int flag = 0;
  if (order == 0) {
    set_prog(sc->p0);
    a.order = 0;
    a.vf = 0;
    WHIP(w, launch_scalar_elem(s, a));
    WHIP(w, launch_scalar_sum(s, m->n_cv, sc->d_elem, sc->d_part, sc->d_sum));
    double J = 0.0;
    WHIP(w, hipMemcpyAsync(&J, sc->d_sum, 8, hipMemcpyDeviceToHost, s));
    WHIP(w, hipMemcpyAsync(&flag, sc->d_err, 4, hipMemcpyDeviceToHost, s));
    WHIP(w, hipStreamSynchronize(s));
    if (flag & 1) return wfail(w, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
    w->potential = J;
    w->has_potential = true;
    return GFX_OK;
  }
  WHIP(w, launch_fill_f64(s, sc->nvec, sc->d_V, 0.0));
  for (auto &v : sc->vars) {
    if (v.p.nm == 0) continue;
    set_prog(v.p);
    a.order = 1;
    const WsVar &wv = w->vars.at(v.name);
    WHIP(w, launch_scalar_elem(s, a));
    WHIP(w, launch_gather_vec(s, wv.mf->ndof / wv.mf->qdim, wv.mf->fem.nd, wv.mf->qdim, v.d_inc_ptr, v.d_inc,
                              sc->d_elem, sc->d_V + wv.first));
  }
  WHIP(w, hipMemcpyAsync(&flag, sc->d_err, 4, hipMemcpyDeviceToHost, s));
  WHIP(w, hipStreamSynchronize(s));
  if (flag & 1) return wfail(w, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
  w->V_out = sc->d_V;
  w->vec_n = sc->nvec;
  return GFX_OK;
}

}  // namespace

extern "C" {

static void ilu_release(gfx_workspace *w);

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
  wfree(w->d_cg); wfree(w->d_go); wfree(w->d_fixed); wfree(w->d_cerr); wfree(w->d_gm);
  ilu_release(w);
  sc_release(w->sc);
  delete w->sc;
  for (auto &kv : w->vars) {
    wfree(kv.second.d_val);
    if (kv.second.mf) unref(kv.second.mf);
  }
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
  if (old != w->vars.end()) {
    wfree(old->second.d_val);
    if (old->second.mf) unref(old->second.mf);
  }
  w->vars[name] = v;
  w->dirty = true;
  w->mat_fail = false;
  ++w->version;
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
  it->second.dev = false;
  if (it->second.kind == VAR_FEM) w->state_dirty = true;
  else w->coefs_dirty = true;
  return GFX_OK;
}

int gfx_ws_set_values_device(gfx_workspace *w, const char *name, const double *d_V) {
  if (!w || !name || !d_V) return GFX_E_INVALID;
  auto it = w->vars.find(name);
  if (it == w->vars.end()) return wfail(w, GFX_E_INVALID, std::string("unknown name ") + name);
  if (it->second.kind != VAR_FEM) return wfail(w, GFX_E_INVALID, "device values are for fem variables");
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  int r;
  bool poly = w->mat_fail;
  if (!poly && w->dirty && (r = setup(w))) {
    if (r != GFX_E_UNSUPPORTED && r != GFX_E_STATE && r != GFX_E_INVALID) return r;
    poly = true;  // the polynomial route reads the variables' device copies
  }
  if (poly || name != w->unknown) {
    WsVar &v = it->second;
    if (!v.d_val && hipMalloc((void **)&v.d_val, 8 * std::max<size_t>(v.V.size(), 1)) != hipSuccess)
      return wfail(w, GFX_E_OOM, "device values");
    WHIP(w, hipMemcpyAsync(v.d_val, d_V, 8 * v.V.size(), hipMemcpyDeviceToDevice, w->ctx->stream));
    v.dev = true;
    return GFX_OK;
  }
  // stream-ordered: the next assembly reads it on the same stream
  WHIP(w, hipMemcpyAsync(w->d_u, d_V, 8 * (size_t)w->ndof, hipMemcpyDeviceToDevice, w->ctx->stream));
  if (w->nonlinear && (r = gfx_plan_set_state_device(w->trees[0].plan, w->d_u))) return r;
  w->state_dirty = false;  // the host copy is no longer the value; the device one is
  w->res_valid = false;
  return GFX_OK;
}

int gfx_ws_add_expression(gfx_workspace *w, const char *expr, gfx_mesh_im *mim, int der_order) {
  if (!w || !expr || !mim) return GFX_E_INVALID;
  if (der_order < 0 || der_order > 2)
    return wfail(w, GFX_E_UNSUPPORTED, "derivation order must be 0, 1 or 2");
  if (mim->mesh->ctx != w->ctx) return wfail(w, GFX_E_INVALID, "mesh_im of another context");
  ref(mim);
  w->exprs.push_back({expr, mim, der_order, -1});
  w->dirty = true;
  w->mat_fail = false;
  ++w->version;
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
  if (order == 2) w->ilu_ready = false;
  if (order < 0 || order > 2) return wfail(w, GFX_E_UNSUPPORTED, "assembly order must be 0, 1 or 2");
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  if (order == 0) return sc_run(w, 0);
  int r;
  // outside the matrix kernels' subset (data fields, P0 or several
  // variables, blocks between mesh_fems, pulled-back coordinates,
  // polynomial nonlinearities): the polynomial route
  if (w->mat_fail) return order == 1 ? sc_run(w, 1) : sc_run2(w);
  if (w->dirty) {
    if ((r = setup(w))) {
      if (r != GFX_E_UNSUPPORTED && r != GFX_E_STATE && r != GFX_E_INVALID) return r;
      const std::string merr = w->ctx->err;
      const int r1 = order == 1 ? sc_run(w, 1) : sc_run2(w);
      if (r1 == GFX_OK) {
        w->mat_fail = true;
        return GFX_OK;
      }
      w->ctx->err = merr + " (polynomial route: " + w->ctx->err + ")";
      return r;
    }
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
    w->vec_n = w->ndof;
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
  w->vec_n = w->ndof;
  return GFX_OK;
}

int gfx_ws_copy_assembled_vector_device(gfx_workspace *w, double *d_V) {
  if (!w || !d_V) return GFX_E_INVALID;
  if (!w->V_out) return wfail(w, GFX_E_STATE, "no order-1 assembly done yet");
  if (hipSetDevice(w->ctx->device) != hipSuccess ||
      hipMemcpyAsync(d_V, w->V_out, 8 * (size_t)w->vec_n, hipMemcpyDeviceToDevice, w->ctx->stream) !=
          hipSuccess)
    return wfail(w, GFX_E_HIP, "copy of the assembled vector");
  return GFX_OK;
}

int gfx_ws_assembled_potential(gfx_workspace *w, double *J) {
  if (!w || !J) return GFX_E_INVALID;
  if (!w->has_potential) return wfail(w, GFX_E_STATE, "no order-0 assembly done yet");
  *J = w->potential;
  return GFX_OK;
}

int gfx_ws_vector_size(gfx_workspace *w, int64_t *n) {
  if (!w || !n) return GFX_E_INVALID;
  if (!w->V_out) return wfail(w, GFX_E_STATE, "no order-1 assembly done yet");
  *n = w->vec_n;
  return GFX_OK;
}

int gfx_ws_assembled_vector(gfx_workspace *w, double *h_V) {
  if (!w || !h_V) return GFX_E_INVALID;
  if (!w->V_out) return wfail(w, GFX_E_STATE, "no order-1 assembly done yet");
  if (hipSetDevice(w->ctx->device) != hipSuccess ||
      hipStreamSynchronize(w->ctx->stream) != hipSuccess ||
      hipMemcpy(h_V, w->V_out, 8 * (size_t)w->vec_n, hipMemcpyDeviceToHost) != hipSuccess)
    return wfail(w, GFX_E_HIP, "copy of the assembled vector");
  return GFX_OK;
}

// The workspace's own assembled matrix: the single tree's plan CSR, or the
// union matrix of several trees.
int gfx_ws_device_csr(gfx_workspace *w, int64_t *nrows, int64_t *nnz, const int64_t **d_indptr,
                      const int32_t **d_indices, double **d_vals) {
  if (!w) return GFX_E_INVALID;
  if (w->mat_fail) {  // the polynomial route's matrix
    if (w->ext) return wfail(w, GFX_E_STATE, "the assembled matrix is the caller's (set_assembled_matrix)");
    if (!w->sc || w->sc->nnz < 0) return wfail(w, GFX_E_STATE, "no matrix assembled yet");
    if (nrows) *nrows = w->sc->nvec;
    if (nnz) *nnz = w->sc->nnz;
    if (d_indptr) *d_indptr = w->sc->d_ip;
    if (d_indices) *d_indices = w->sc->d_ix;
    if (d_vals) *d_vals = w->sc->d_v;
    return GFX_OK;
  }
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

// CG on the workspace's assembled matrix (diagonal preconditioner; the
// 3x3 block Jacobi of gfx_plan_solve_cg needs one vector field)
int gfx_ws_solve_cg(gfx_workspace *w, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                    double *rel_res) {
  if (!w || !d_b || !d_x || !iters || !rel_res || !(rtol > 0) || maxit < 0) return GFX_E_INVALID;
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  double *v = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, &v);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  if (w->cg_n != n) {
    WHIP(w, hipStreamSynchronize(w->ctx->stream));
    wfree(w->d_cg);
    if (hipMalloc((void **)&w->d_cg, 8 * cg_workspace_doubles(n)) != hipSuccess)
      return wfail(w, GFX_E_OOM, "CG vectors");
    w->cg_n = n;
  }
  WHIP(w, cg_solve(w->ctx->stream, n, nnz, ip, ix, v, d_b, d_x, rtol, maxit, 1, w->d_cg, iters, rel_res));
  return GFX_OK;
}

static void ilu_release(gfx_workspace *w) {
  IluData &U = w->ilu;
  wfree(U.lv.d_dpos); wfree(U.lv.d_fdeg); wfree(U.lv.d_bdeg); wfree(U.lv.d_flist); wfree(U.lv.d_blist);
  wfree(U.lv.d_counter); wfree(U.d_T); wfree(U.d_rows); wfree(U.d_F); wfree(U.d_y);
  U.lv.flev.clear(); U.lv.blev.clear();
  w->ilu_n = w->ilu_nnz = -1;
}

// gmm::ilu_precond of the workspace's matrix: the analysis (diagonal
// positions, level sets) is redone at every solve, since the polynomial
// route rebuilds its pattern with every assembly
static int ilu_prepare(gfx_workspace *w, int64_t n, int64_t nnz, const int64_t *ip, const int32_t *ix,
                       const double *v, int64_t *replaced_out = nullptr) {
  hipStream_t s = w->ctx->stream;
  IluData &U = w->ilu;
  if (w->ilu_n != n || w->ilu_nnz < nnz) {
    WHIP(w, hipStreamSynchronize(s));
    ilu_release(w);
    const size_t n1 = (size_t)std::max<int64_t>(n, 1), z1 = (size_t)std::max<int64_t>(nnz, 1);
    if (hipMalloc((void **)&U.lv.d_dpos, 8 * n1) != hipSuccess ||
        hipMalloc((void **)&U.lv.d_fdeg, 4 * n1) != hipSuccess ||
        hipMalloc((void **)&U.lv.d_bdeg, 4 * n1) != hipSuccess ||
        hipMalloc((void **)&U.lv.d_flist, 4 * n1) != hipSuccess ||
        hipMalloc((void **)&U.lv.d_blist, 4 * n1) != hipSuccess || hipMalloc((void **)&U.lv.d_counter, 4) != hipSuccess ||
        hipMalloc((void **)&U.d_T, 8 * z1) != hipSuccess || hipMalloc((void **)&U.d_rows, 8 * n1) != hipSuccess ||
        hipMalloc((void **)&U.d_F, 8 * z1) != hipSuccess || hipMalloc((void **)&U.d_y, 8 * n1) != hipSuccess) {
      ilu_release(w);
      return wfail(w, GFX_E_OOM, "ILU buffers");
    }
    w->ilu_n = n;
    w->ilu_nnz = nnz;
  }
  // gmm's ILU keeps the first 1000 entries of a row (gmm_precond_ilu.h:147)
  std::vector<int64_t> hp((size_t)n + 1);
  WHIP(w, hipMemcpyAsync(hp.data(), ip, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
  WHIP(w, hipStreamSynchronize(s));
  for (int64_t i = 0; i < n; ++i)
    if (hp[(size_t)i + 1] - hp[(size_t)i] > 1000)
      return wfail(w, GFX_E_UNSUPPORTED, "ILU: a row of more than 1000 entries (GetFEM truncates it)");
  if (!w->d_cerr && hipMalloc((void **)&w->d_cerr, 4) != hipSuccess) return wfail(w, GFX_E_OOM, "flag");
  int herr = 0;
  WHIP(w, ildlt_analyse(s, U.lv, n, 1, ip, ix, w->d_cerr, &herr));
  if (herr)
    return wfail(w, GFX_E_UNSUPPORTED, (herr & 1) ? "ILU: a row has no diagonal entry"
                                                  : "ILU: the matrix pattern is not symmetric");
  int64_t replaced = 0;
  WHIP(w, ilu_factor(s, U, ip, ix, v, w->d_cerr, &herr, &replaced));
  if (herr) return wfail(w, GFX_E_UNSUPPORTED, "ILU: the matrix pattern is not symmetric");
  if (replaced_out) *replaced_out = replaced;
  return GFX_OK;
}

// restarted GMRES on the workspace's assembled matrix: what
// default_linear_solver gives a model that is not coercive
// (linear_solver_gmres_preconditioned_ilu, src/getfem/getfem_model_solvers.h:
// 101-109, restart 500), with its ILU or the inverse diagonal.  The Krylov
// basis takes (restart + 1) n doubles: the restart is lowered to keep it
// within 4 GiB.
int gfx_ws_solve_gmres(gfx_workspace *w, int precond, const double *d_b, double *d_x, double rtol, int maxit,
                       int restart, int *iters, double *rel_res) {
  if (!w || !d_b || !d_x || !iters || !rel_res || !(rtol > 0) || maxit < 0 || restart < 1) return GFX_E_INVALID;
  if (precond != GFX_PRECOND_ILU && precond != GFX_PRECOND_DIAGONAL)
    return wfail(w, GFX_E_INVALID, "GMRES takes GFX_PRECOND_ILU or GFX_PRECOND_DIAGONAL");
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  double *v = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, &v);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  int m = std::max(1, std::min(restart, std::max(maxit, 1)));
  const int64_t cap = (int64_t(4) << 30) / 8 / std::max<int64_t>(n, 1) - 1;
  if (m > cap) m = (int)std::max<int64_t>(cap, 1);
  const size_t need = gmres_workspace_doubles(n, m);
  if (need > w->gm_size) {
    WHIP(w, hipStreamSynchronize(w->ctx->stream));
    wfree(w->d_gm);
    if (hipMalloc((void **)&w->d_gm, 8 * need) != hipSuccess) return wfail(w, GFX_E_OOM, "GMRES basis");
    w->gm_size = need;
  }
  if (precond == GFX_PRECOND_ILU) {
    if ((r = ilu_prepare(w, n, nnz, ip, ix, v))) return r;
    w->ilu_ready = true;
  }
  WHIP(w, gmres_solve(w->ctx->stream, n, nnz, ip, ix, v, d_b, d_x, rtol, maxit, m, w->d_gm,
                      precond == GFX_PRECOND_ILU ? &w->ilu : nullptr, iters, rel_res));
  return GFX_OK;
}

// GMRES's ILU on its own: factor the workspace's current matrix, then apply
int gfx_ws_ilu_factor(gfx_workspace *w, int64_t *replaced) {
  if (!w) return GFX_E_INVALID;
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  double *v = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, &v);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  if ((r = ilu_prepare(w, n, nnz, ip, ix, v, replaced))) return r;
  w->ilu_ready = true;
  return GFX_OK;
}

int gfx_ws_ilu_apply(gfx_workspace *w, const double *d_r, double *d_z) {
  if (!w || !d_r || !d_z) return GFX_E_INVALID;
  if (!w->ilu_ready) return wfail(w, GFX_E_STATE, "no ILU factor (gfx_ws_ilu_factor)");
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, nullptr);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  WHIP(w, ilu_apply(w->ctx->stream, w->ilu, ip, ix, d_r, d_z));
  return GFX_OK;
}

// model::assembly's Dirichlet conditions by simplification
// (src/getfem_models.cc:2806-2870) on the workspace's assembled matrix, as
// gfx_plan_dof_constraints does on a plan's
int gfx_ws_dof_constraints(gfx_workspace *w, int64_t n_fixed, const int32_t *h_dofs, const double *h_go,
                           const double *d_pr, double *d_rhs, int flags) {
  if (!w || n_fixed < 0 || (n_fixed && (!h_dofs || !h_go))) return GFX_E_INVALID;
  int64_t n = 0, nnz = 0;
  const int64_t *ip = nullptr;
  const int32_t *ix = nullptr;
  double *v = nullptr;
  int r = gfx_ws_device_csr(w, &n, &nnz, &ip, &ix, &v);
  if (r) return r;
  if (hipSetDevice(w->ctx->device) != hipSuccess) return wfail(w, GFX_E_HIP, "hipSetDevice");
  hipStream_t s = w->ctx->stream;
  std::vector<uint8_t> fixed((size_t)n, 0);
  std::vector<double> go((size_t)n, 0.0);
  for (int64_t k = 0; k < n_fixed; ++k) {
    if (h_dofs[k] < 0 || h_dofs[k] >= n) return wfail(w, GFX_E_INVALID, "constrained dof out of range");
    fixed[h_dofs[k]] = 1;
    go[h_dofs[k]] = h_go[k];
  }
  if (w->con_n != n) {
    WHIP(w, hipStreamSynchronize(s));
    wfree(w->d_fixed); wfree(w->d_go); wfree(w->d_cerr);
    if (hipMalloc((void **)&w->d_fixed, std::max<size_t>((size_t)n, 1)) != hipSuccess ||
        hipMalloc((void **)&w->d_go, 8 * std::max<size_t>((size_t)n, 1)) != hipSuccess ||
        hipMalloc((void **)&w->d_cerr, 4) != hipSuccess)
      return wfail(w, GFX_E_OOM, "constraint buffers");
    w->con_n = n;
  }
  WHIP(w, hipMemcpyAsync(w->d_fixed, fixed.data(), (size_t)n, hipMemcpyHostToDevice, s));
  WHIP(w, hipMemcpyAsync(w->d_go, go.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s));
  WHIP(w, hipMemsetAsync(w->d_cerr, 0, 4, s));
  WHIP(w, launch_dof_constraints(s, n, ip, ix, v, w->d_fixed, w->d_go, d_pr, d_rhs, flags, w->d_cerr));
  int flag = 0;
  WHIP(w, hipMemcpyAsync(&flag, w->d_cerr, 4, hipMemcpyDeviceToHost, s));
  WHIP(w, hipStreamSynchronize(s));
  if (flag & 4) return wfail(w, GFX_E_PATTERN, "a constrained dof has no diagonal entry in the pattern");
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
    // var: fem variable; p0var: fem variable on a P0 mesh_fem; p0: P0 fem
    // constant; data: fem constant of another fem; anything else: fixed size
    g.is_fem = kind == "var" || kind == "p0" || kind == "p0var" || kind == "data";
    g.is_constant = kind != "var" && kind != "p0var";
    g.per_element = kind == "p0" || kind == "p0var";
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

int gfx_gwfl_compile_scalar(const char *exprs, const char *decls, int order, const char *var, char *out,
                            int64_t len) {
  if (!exprs || !decls || !out || len <= 0 || (order >= 1 && !var)) return GFX_E_INVALID;
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
  GwflScalarProgram prog;
  std::string err, d;
  // order 2: var names the test and the trial variable, "u,p"
  std::string v1 = var ? var : "", v2;
  if (order == 2) {
    const size_t c = v1.find(',');
    if (c == std::string::npos) return GFX_E_INVALID;
    v2 = v1.substr(c + 1);
    v1 = v1.substr(0, c);
  }
  const bool ok = gwfl_compile_scalar(ex, gv, order, v1, prog, err, v2);
  if (ok) {
    d = "nout=" + std::to_string(prog.nout) + " zero=";
    for (size_t k = 0; k < prog.zero_vars.size(); ++k) d += (k ? "," : "") + prog.zero_vars[k];
    d += "\nsyms=";
    for (size_t k = 0; k < prog.syms.size(); ++k) d += (k ? " " : "") + prog.syms[k];
    char num[40];
    for (const auto &m : prog.monos) {
      std::snprintf(num, sizeof num, "%.17g", m.coef);
      d += "\n" + std::to_string(m.out) + " " + num;
      for (int x : m.s) d += " " + std::to_string(x);
    }
  } else {
    d = err;
  }
  if ((int64_t)d.size() >= len) {
    std::strncpy(out, "output buffer too small", (size_t)len - 1);
    out[len - 1] = 0;
    return GFX_E_INVALID;
  }
  std::strncpy(out, d.c_str(), (size_t)len - 1);
  out[len - 1] = 0;
  return ok ? GFX_OK : GFX_E_UNSUPPORTED;
}

int gfx_ws_plan(gfx_workspace *w, gfx_plan **out) {
  if (!w || !out) return GFX_E_INVALID;
  if (w->mat_fail) return wfail(w, GFX_E_STATE, "the polynomial route holds the workspace's matrix: no plan");
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
