// C ABI of libgfx: contexts, meshes, mesh_fems, integration methods and
// order-2 assembly plans.  See include/gfx.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gfx_internal.h"

using namespace gfx;

namespace {

int fail(gfx_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(gfx_ctx *c, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(c, e == hipErrorOutOfMemory ? GFX_E_OOM : GFX_E_HIP, m);
}

#define HIPCHK(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hipfail((ctx), e_, #call); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void **)p, n * sizeof(T));
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void il_free(gfx::IldltData &L) {
  dfree(L.d_dpos); dfree(L.d_fdeg); dfree(L.d_bdeg); dfree(L.d_flist); dfree(L.d_blist); dfree(L.d_counter);
  dfree(L.d_F); dfree(L.d_work); dfree(L.d_cg);
  L = gfx::IldltData();
}

int set_device(gfx_ctx *c) {
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hipfail(c, e, "hipSetDevice");
  return GFX_OK;
}

void release(gfx_ctx *c) {
  (void)hipSetDevice(c->device);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

void release(gfx_mesh *m) {
  gfx_ctx *c = m->ctx;
  (void)hipSetDevice(c->device);
  dfree(m->d_pts); dfree(m->d_cvx);
  delete m;
  unref(c);
}

void release(gfx_mesh_fem *mf) {
  gfx_mesh *m = mf->mesh;
  (void)hipSetDevice(m->ctx->device);
  dfree(mf->d_dofs);
  delete mf;
  unref(m);
}

void release(gfx_mesh_im *mim) {
  gfx_mesh *m = mim->mesh;
  delete mim;
  unref(m);
}

template <typename T>
void close_or_mark(T *o) {
  o->closing = true;
  if (o->refs == 0) release(o);
}

}  // namespace

void gfx::ref(gfx_ctx *c) { ++c->refs; }
void gfx::unref(gfx_ctx *c) { if (--c->refs == 0 && c->closing) release(c); }
void gfx::ref(gfx_mesh *m) { ++m->refs; }
void gfx::unref(gfx_mesh *m) { if (--m->refs == 0 && m->closing) release(m); }
void gfx::ref(gfx_mesh_fem *mf) { ++mf->refs; }
void gfx::unref(gfx_mesh_fem *mf) { if (--mf->refs == 0 && mf->closing) release(mf); }
void gfx::ref(gfx_mesh_im *mim) { ++mim->refs; }
void gfx::unref(gfx_mesh_im *mim) { if (--mim->refs == 0 && mim->closing) release(mim); }

extern "C" {

const char *gfx_version(void) { return "libgfx 0.1 (gfx950)"; }

int gfx_integration_rule(const char *im_name, int *npts, double *h_pts, double *h_w) {
  if (!im_name || !npts) return GFX_E_INVALID;
  IntegrationRule r;
  if (!integration_rule(im_name, r)) return GFX_E_UNSUPPORTED;
  *npts = r.n();
  if (h_pts) std::copy(r.pts.begin(), r.pts.end(), h_pts);
  if (h_w) std::copy(r.w.begin(), r.w.end(), h_w);
  return GFX_OK;
}

int gfx_fem_eval(const char *fem_name, const double *x, int *nd, double *h_val, double *h_grad) {
  if (!fem_name || !x) return GFX_E_INVALID;
  FemDesc f;
  if (!fem_descriptor(fem_name, f)) return GFX_E_UNSUPPORTED;
  if (nd) *nd = f.nd;
  std::vector<double> v(f.nd), g(3 * (size_t)f.nd);
  fem_eval(f, x, v.data(), g.data());
  if (h_val) std::copy(v.begin(), v.end(), h_val);
  if (h_grad) std::copy(g.begin(), g.end(), h_grad);
  return GFX_OK;
}

int gfx_enumerate_dofs(const char *geotrans, const char *fem_name, int qdim, int64_t n_pts,
                       const double *h_pts, int64_t n_cv, const int32_t *h_cvx, int32_t *h_out,
                       int64_t *ndof) {
  if (!geotrans || !fem_name || !h_cvx || !h_out || !ndof || qdim < 1) return GFX_E_INVALID;
  const int gt = geotrans_descriptor(geotrans);
  FemDesc f;
  if (gt < 0 || !fem_descriptor(fem_name, f)) return GFX_E_UNSUPPORTED;
  const int nv = geotrans_nb_vertices(gt);
  for (int64_t i = 0; i < n_cv * nv; ++i)
    if (h_cvx[i] < 0 || h_cvx[i] >= n_pts) return GFX_E_INVALID;
  std::vector<int32_t> dofs;
  std::string err;
  const int64_t n = enumerate_dofs(gt, n_pts, h_pts, n_cv, h_cvx, f, qdim, dofs, err);
  if (n < 0) return GFX_E_UNSUPPORTED;
  std::copy(dofs.begin(), dofs.end(), h_out);
  *ndof = n;
  return GFX_OK;
}

int gfx_ctx_create(int device, gfx_ctx **out) {
  if (!out) return GFX_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GFX_E_HIP;
  gfx_ctx *c = new (std::nothrow) gfx_ctx();
  if (!c) return GFX_E_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return GFX_E_HIP;
  }
  c->own_stream = true;
  // set-up of the allocation-free compliance entry point: the address of this
  // device's copy of its static scratch (kept per device; a failure here is
  // reported by that entry point, which retries the lookup)
  double *scr = nullptr;
  (void)compliance_scratch(device, &scr);
  *out = c;
  return GFX_OK;
}

int gfx_ctx_destroy(gfx_ctx *c) {
  if (!c || c->closing) return GFX_E_INVALID;
  close_or_mark(c);
  return GFX_OK;
}

int gfx_ctx_set_stream(gfx_ctx *c, void *s) {
  if (!c) return GFX_E_INVALID;
  if (set_device(c)) return GFX_E_HIP;
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  if (s) {
    c->stream = (hipStream_t)s;
    c->own_stream = false;
  } else {
    if (hipStreamCreate(&c->stream) != hipSuccess) return fail(c, GFX_E_HIP, "hipStreamCreate");
    c->own_stream = true;
  }
  return GFX_OK;
}

void *gfx_ctx_stream(gfx_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gfx_ctx_synchronize(gfx_ctx *c) {
  if (!c) return GFX_E_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GFX_OK;
}

const char *gfx_ctx_last_error(gfx_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ----------------------------------------------------------------- mesh
int gfx_mesh_create(gfx_ctx *c, const char *geotrans, int64_t n_pts, const double *h_pts,
                    int64_t n_cv, const int32_t *h_cvx, gfx_mesh **out) {
  if (!c || !out || !geotrans || n_pts < 0 || n_cv < 0 || (n_pts && !h_pts) || (n_cv && !h_cvx))
    return fail(c, GFX_E_INVALID, "gfx_mesh_create: invalid argument");
  *out = nullptr;
  const int gt = geotrans_descriptor(geotrans);
  if (gt < 0) return fail(c, GFX_E_UNSUPPORTED, std::string("unsupported geometric transformation ") + geotrans);
  const int nv = geotrans_nb_vertices(gt);
  if (n_cv >= (int64_t(1) << 27)) return fail(c, GFX_E_INVALID, "too many elements (limit 2^27)");
  for (int64_t i = 0; i < n_cv * nv; ++i)
    if (h_cvx[i] < 0 || h_cvx[i] >= n_pts)
      return fail(c, GFX_E_INVALID, "convex references a point out of range");
  if (set_device(c)) return GFX_E_HIP;
  gfx_mesh *m = new (std::nothrow) gfx_mesh();
  if (!m) return fail(c, GFX_E_OOM, "host allocation");
  m->ctx = c; m->gt = gt; m->nv = nv; m->n_pts = n_pts; m->n_cv = n_cv;
  m->h_pts.assign(h_pts, h_pts + 3 * n_pts);
  m->h_cvx.assign(h_cvx, h_cvx + nv * n_cv);
  hipError_t e = dalloc(&m->d_pts, 3 * (size_t)n_pts);
  if (e == hipSuccess) e = dalloc(&m->d_cvx, (size_t)nv * n_cv);
  if (e == hipSuccess && n_pts) e = hipMemcpy(m->d_pts, h_pts, 24 * (size_t)n_pts, hipMemcpyHostToDevice);
  if (e == hipSuccess && n_cv) e = hipMemcpy(m->d_cvx, h_cvx, 4 * (size_t)nv * n_cv, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dfree(m->d_pts); dfree(m->d_cvx); delete m;
    return hipfail(c, e, "gfx_mesh_create upload");
  }
  ref(c);
  *out = m;
  return GFX_OK;
}

int gfx_mesh_set_points(gfx_mesh *m, const double *h_pts) {
  if (!m || !h_pts) return GFX_E_INVALID;
  gfx_ctx *c = m->ctx;
  if (set_device(c)) return GFX_E_HIP;
  std::copy(h_pts, h_pts + 3 * m->n_pts, m->h_pts.begin());
  HIPCHK(c, hipMemcpyAsync(m->d_pts, h_pts, 24 * (size_t)m->n_pts, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  m->version++;  // plans on this mesh re-derive their thresholds and pattern
  return GFX_OK;
}

int gfx_mesh_destroy(gfx_mesh *m) {
  if (!m || m->closing) return GFX_E_INVALID;
  close_or_mark(m);
  return GFX_OK;
}

// ------------------------------------------------------------- mesh_fem
static int mf_finish(gfx_mesh *m, gfx_mesh_fem *mf, gfx_mesh_fem **out) {
  gfx_ctx *c = m->ctx;
  hipError_t e = dalloc(&mf->d_dofs, mf->h_dofs.size());
  if (e == hipSuccess && !mf->h_dofs.empty())
    e = hipMemcpy(mf->d_dofs, mf->h_dofs.data(), 4 * mf->h_dofs.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dfree(mf->d_dofs); delete mf;
    return hipfail(c, e, "gfx_mesh_fem upload");
  }
  ref(m);
  *out = mf;
  return GFX_OK;
}

int gfx_mesh_fem_create(gfx_mesh *m, const char *fem_name, int qdim, gfx_mesh_fem **out) {
  if (!m || !fem_name || !out || qdim < 1 || qdim > 3) return fail(m ? m->ctx : nullptr, GFX_E_INVALID, "gfx_mesh_fem_create: invalid argument");
  gfx_ctx *c = m->ctx;
  *out = nullptr;
  gfx_mesh_fem *mf = new (std::nothrow) gfx_mesh_fem();
  if (!mf) return fail(c, GFX_E_OOM, "host allocation");
  mf->mesh = m; mf->qdim = qdim;
  if (!fem_descriptor(fem_name, mf->fem)) {
    delete mf;
    return fail(c, GFX_E_UNSUPPORTED, std::string("unsupported finite element ") + fem_name);
  }
  std::string err;
  mf->ndof = enumerate_dofs(m->gt, m->n_pts, m->h_pts.data(), m->n_cv, m->h_cvx.data(), mf->fem,
                            qdim, mf->h_dofs, err);
  if (mf->ndof < 0) { delete mf; return fail(c, GFX_E_UNSUPPORTED, err); }
  if (set_device(c)) { delete mf; return GFX_E_HIP; }
  return mf_finish(m, mf, out);
}

int gfx_mesh_fem_create_from_dofs(gfx_mesh *m, const char *fem_name, int qdim,
                                  const int32_t *h_basic_dofs, int64_t ndof, gfx_mesh_fem **out) {
  if (!m || !fem_name || !out || qdim < 1 || qdim > 3 || !h_basic_dofs || ndof < 0)
    return fail(m ? m->ctx : nullptr, GFX_E_INVALID, "gfx_mesh_fem_create_from_dofs: invalid argument");
  gfx_ctx *c = m->ctx;
  *out = nullptr;
  gfx_mesh_fem *mf = new (std::nothrow) gfx_mesh_fem();
  if (!mf) return fail(c, GFX_E_OOM, "host allocation");
  mf->mesh = m; mf->qdim = qdim; mf->ndof = ndof;
  if (!fem_descriptor(fem_name, mf->fem)) {
    delete mf;
    return fail(c, GFX_E_UNSUPPORTED, std::string("unsupported finite element ") + fem_name);
  }
  const size_t n = (size_t)m->n_cv * mf->fem.nd;
  for (size_t i = 0; i < n; ++i)
    if (h_basic_dofs[i] < 0 || h_basic_dofs[i] + qdim > ndof || h_basic_dofs[i] % qdim) {
      delete mf;
      return fail(c, GFX_E_INVALID, "basic dof out of range or not a multiple of qdim");
    }
  mf->h_dofs.assign(h_basic_dofs, h_basic_dofs + n);
  if (set_device(c)) { delete mf; return GFX_E_HIP; }
  return mf_finish(m, mf, out);
}

int64_t gfx_mesh_fem_nb_dof(gfx_mesh_fem *mf) { return mf ? mf->ndof : -1; }
int gfx_mesh_fem_nb_dof_per_element(gfx_mesh_fem *mf) { return mf ? mf->fem.nd : -1; }

int gfx_mesh_fem_basic_dofs(gfx_mesh_fem *mf, int32_t *h_out) {
  if (!mf || !h_out) return GFX_E_INVALID;
  std::copy(mf->h_dofs.begin(), mf->h_dofs.end(), h_out);
  return GFX_OK;
}

int gfx_mesh_fem_dof_nodes(gfx_mesh_fem *mf, double *h_out) {
  if (!mf || !h_out) return GFX_E_INVALID;
  const gfx_mesh *m = mf->mesh;
  const int nd = mf->fem.nd, nv = m->nv;
  std::vector<double> N(nv);
  std::vector<std::vector<double>> Nn(nd, std::vector<double>(nv));
  for (int i = 0; i < nd; ++i) geotrans_eval(m->gt, &mf->fem.nodes[3 * i], Nn[i].data(), nullptr);
  for (int64_t e = 0; e < m->n_cv; ++e)
    for (int i = 0; i < nd; ++i) {
      const int64_t node = mf->h_dofs[(size_t)e * nd + i] / mf->qdim;
      double x[3] = {0, 0, 0};
      for (int v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) x[k] += Nn[i][v] * m->h_pts[3 * (size_t)m->h_cvx[(size_t)e * nv + v] + k];
      for (int k = 0; k < 3; ++k) h_out[3 * node + k] = x[k];
    }
  return GFX_OK;
}

int gfx_mesh_fem_destroy(gfx_mesh_fem *mf) {
  if (!mf || mf->closing) return GFX_E_INVALID;
  close_or_mark(mf);
  return GFX_OK;
}

// -------------------------------------------------------------- mesh_im
int gfx_mesh_im_create(gfx_mesh *m, const char *im_name, gfx_mesh_im **out) {
  if (!m || !im_name || !out) return GFX_E_INVALID;
  *out = nullptr;
  gfx_mesh_im *mim = new (std::nothrow) gfx_mesh_im();
  if (!mim) return fail(m->ctx, GFX_E_OOM, "host allocation");
  mim->mesh = m;
  if (!integration_rule(im_name, mim->rule)) {
    delete mim;
    return fail(m->ctx, GFX_E_UNSUPPORTED, std::string("unsupported integration method ") + im_name);
  }
  if (mim->rule.gt != m->gt) {
    delete mim;
    return fail(m->ctx, GFX_E_INVALID, "integration method does not match the mesh element");
  }
  ref(m);
  *out = mim;
  return GFX_OK;
}

int gfx_mesh_im_nb_points(gfx_mesh_im *mim) { return mim ? mim->rule.n() : -1; }

int gfx_mesh_im_points(gfx_mesh_im *mim, double *h_pts, double *h_w) {
  if (!mim) return GFX_E_INVALID;
  if (h_pts) std::copy(mim->rule.pts.begin(), mim->rule.pts.end(), h_pts);
  if (h_w) std::copy(mim->rule.w.begin(), mim->rule.w.end(), h_w);
  return GFX_OK;
}

int gfx_mesh_im_destroy(gfx_mesh_im *mim) {
  if (!mim || mim->closing) return GFX_E_INVALID;
  close_or_mark(mim);
  return GFX_OK;
}

// ---------------------------------------------------------------- plans
static int plan_upload_coef(gfx_plan *p, int slot, const gfx_coef *cf) {
  gfx_ctx *c = p->ctx;
  if (cf->kind != GFX_COEF_CONST && cf->kind != GFX_COEF_PER_ELEMENT)
    return fail(c, GFX_E_INVALID, "unknown coefficient kind");
  p->coef[slot] = *cf;
  p->coef[slot].h_values = nullptr;
  p->coef_min[slot] = cf->value;
  dfree(p->d_coef[slot]);
  if (cf->kind == GFX_COEF_PER_ELEMENT) {
    if (!cf->h_values) return fail(c, GFX_E_INVALID, "per-element coefficient without values");
    const int64_t n = p->mf->mesh->n_cv;
    p->coef_min[slot] = n ? *std::min_element(cf->h_values, cf->h_values + n) : 0.0;
    HIPCHK(c, dalloc(&p->d_coef[slot], (size_t)n));
    HIPCHK(c, hipMemcpy(p->d_coef[slot], cf->h_values, 8 * (size_t)n, hipMemcpyHostToDevice));
  }
  p->values_changed = true;  // thresholds and pattern are re-derived at the next assembly
  return GFX_OK;
}

// number of coefficients of a term
static int term_coefs(int term) {
  return term == GFX_TERM_LAPLACE ? 1 : (term == GFX_TERM_ISOTROPIC ? 3 : 2);
}

int gfx_plan_destroy(gfx_plan *p) {
  if (!p) return GFX_E_INVALID;
  (void)hipSetDevice(p->ctx->device);
  for (hipStream_t st : p->aux) (void)hipStreamSynchronize(st);
  for (hipStream_t st : p->aux) (void)hipStreamDestroy(st);
  for (hipEvent_t e : p->join) (void)hipEventDestroy(e);
  if (p->fork) (void)hipEventDestroy(p->fork);
  dfree(p->d_coef[0]); dfree(p->d_coef[1]); dfree(p->d_coef[2]);
  dfree(p->d_gref); dfree(p->d_qw); dfree(p->d_R);
  dfree(p->d_rec); dfree(p->d_thr); dfree(p->d_keep);
  dfree(p->d_gref32); dfree(p->d_qw32); dfree(p->d_R32); dfree(p->d_rec32); dfree(p->d_Rq); dfree(p->d_E32);
  dfree(p->d_E64);
  dfree(p->d_inc_ptr); dfree(p->d_inc); dfree(p->d_slot); dfree(p->d_nslots); dfree(p->d_order); dfree(p->d_desc);
  dfree(p->d_node_ok);
  dfree(p->d_nbr); dfree(p->d_mask);
  dfree(p->d_indptr); dfree(p->d_indices); dfree(p->d_vals); dfree(p->d_err);
  dfree(p->d_gN); dfree(p->d_E); dfree(p->d_res); dfree(p->d_U); dfree(p->d_rhs); dfree(p->d_cg);
  dfree(p->d_fixed); dfree(p->d_go);
  il_free(p->il);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  gfx_ctx *c = p->ctx;
  gfx_mesh_fem *mf = p->mf;
  gfx_mesh_im *mim = p->mim;
  delete p;
  unref(mim); unref(mf); unref(c);
  return GFX_OK;
}

// The reference tensor R_ij = sum_q w_q gradref_i gradref_j^T of an exact
// rule is rational: R = N / L with small integers N (P1/IM_TETRAHEDRON(1):
// L = 6, |N| <= 1; P2/(5): L = 30, |N| <= 8; P3/(8): |N| < 2^15).  Stored as
// int8 (16-byte blocks) or int16 (32-byte blocks) it costs the patch kernel's
// lanes 16 or 32 bytes per pair block instead of 72; the integers convert to
// floating point exactly and 1/L is folded into the element's scale, so the
// result differs from the FP64 table only by rounding.  rk = 0 when R has no
// such form (a rule that does not integrate the products exactly).
static void detect_int_reference(const std::vector<double> &R, int nblk, int &rk, double &L,
                                 std::vector<uint8_t> &out) {
  static const int kL[] = {1,    2,    3,    4,    5,    6,    8,    9,     10,    12,    15,
                           16,   18,   20,   24,   30,   36,   40,   45,    48,    60,    72,
                           80,   90,   96,   120,  144,  180,  240,  280,   315,   360,   420,
                           480,  504,  560,  630,  720,  840,  1008, 1120,  1260,  1440,  1680,
                           2016, 2240, 2520, 2880, 3360, 4032, 5040, 6720,  7560,  10080, 15120,
                           20160, 30240, 40320, 60480, 80640, 120960, 181440, 362880};
  rk = 0;
  L = 1.0;
  for (int Lc : kL) {
    double nmax = 0.0;
    bool ok = true;
    for (double r : R) {
      const double n = r * Lc, rn = std::nearbyint(n);
      if (std::fabs(n - rn) > 1e-9 * std::max(1.0, std::fabs(n))) { ok = false; break; }
      nmax = std::max(nmax, std::fabs(rn));
    }
    if (!ok) continue;
    if (nmax > 32767.0) return;
    rk = nmax <= 127.0 ? 1 : 2;
    L = Lc;
    const int bs = rk == 1 ? 16 : 32;
    out.assign((size_t)nblk * bs, 0);
    for (int b = 0; b < nblk; ++b)
      for (int t = 0; t < 9; ++t) {
        const int n = (int)std::nearbyint(R[(size_t)b * 9 + t] * Lc);
        if (rk == 1) out[(size_t)b * 16 + t] = (uint8_t)(int8_t)n;
        else {
          const uint16_t u = (uint16_t)(int16_t)n;
          std::memcpy(&out[(size_t)b * 32 + 2 * t], &u, 2);
        }
      }
    return;
  }
}

int gfx_plan_create(gfx_mesh_fem *mf, gfx_mesh_im *mim, int term, const gfx_coef *coefs,
                    int n_coefs, gfx_plan **out) {
  if (!mf || !mim || !out || mf->closing || mim->closing) return GFX_E_INVALID;
  *out = nullptr;
  gfx_ctx *c = mf->mesh->ctx;
  if (mim->mesh != mf->mesh) return fail(c, GFX_E_INVALID, "mesh_fem and mesh_im on different meshes");
  const gfx_mesh *m = mf->mesh;
  const int Q = mf->qdim, nd = mf->fem.nd;
  const bool eb = term == GFX_TERM_NEOHOOKEAN;
  if (eb) {
    if (m->gt != GT_HEX_Q1 || mf->fem.kind != FEM_QK || mf->fem.k != 2 || Q != 3)
      return fail(c, GFX_E_UNSUPPORTED, "the Neo-Hookean plan needs FEM_QK(3,2), qdim 3, on GT_QK(3,1)");
  } else {
    if (m->gt != GT_TET_P1 || mf->fem.kind != FEM_PK)
      return fail(c, GFX_E_UNSUPPORTED, "this plan kind needs FEM_PK(3,k) on GT_PK(3,1)");
    if (!((term == GFX_TERM_LAPLACE && (Q == 1 || Q == 3)) ||
          ((term == GFX_TERM_ISO_ELAST || term == GFX_TERM_ISOTROPIC) && Q == 3)))
      return fail(c, GFX_E_UNSUPPORTED, "term / qdim combination not supported");
    if (!(nd == 4 || nd == 10 || nd == 20) || (term == GFX_TERM_LAPLACE && Q == 3 && nd == 20))
      return fail(c, GFX_E_UNSUPPORTED, "element degree not supported for this term");
  }
  if (term < GFX_TERM_LAPLACE || term > GFX_TERM_ISOTROPIC)
    return fail(c, GFX_E_UNSUPPORTED, "unknown term");
  const int need = term_coefs(term);
  if (n_coefs != need || !coefs) return fail(c, GFX_E_INVALID, "wrong number of coefficients");
  if (set_device(c)) return GFX_E_HIP;
  gfx_plan *p = new (std::nothrow) gfx_plan();
  if (!p) return fail(c, GFX_E_OOM, "host allocation");
  p->ctx = c; p->mf = mf; p->mim = mim; p->term = term;
  ref(c); ref(mf); ref(mim);  // released by gfx_plan_destroy (also on the failure paths below)
  p->nd = nd; p->qdim = Q; p->nloc = nd * Q; p->nq = mim->rule.n(); p->nv = m->nv;
  p->kwords = (p->nloc * p->nloc + 31) / 32;
  p->mwpi = (nd * Q * Q + 31) / 32;
  p->ndof = mf->ndof;
  p->n_nodes = mf->ndof / Q;
  for (int s = 0; s < 3; ++s) {
    p->coef[s].kind = GFX_COEF_CONST;
    p->coef[s].value = 0.0;
  }
  for (int s = 0; s < need; ++s) {
    int r = plan_upload_coef(p, s, &coefs[s]);
    if (r) { gfx_plan_destroy(p); return r; }
  }
  // reference tables: gradients at the points and R_ij = sum_q w_q g_i g_j^T
  const int nq = p->nq;
  std::vector<double> gref((size_t)nq * nd * 3), val(nd), R((size_t)nd * nd * 9, 0.0);
  for (int q = 0; q < nq; ++q) fem_eval(mf->fem, &mim->rule.pts[3 * q], val.data(), &gref[(size_t)q * nd * 3]);
  for (int q = 0; q < nq; ++q) {
    const double w = mim->rule.w[q];
    const double *g = &gref[(size_t)q * nd * 3];
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) R[((size_t)i * nd + j) * 9 + 3 * a + b] += w * g[3 * i + a] * g[3 * j + b];
  }
  p->eb = eb;
  std::vector<double> gN((size_t)nq * m->nv * 3), Nv(m->nv);
  for (int q = 0; q < nq; ++q) geotrans_eval(m->gt, &mim->rule.pts[3 * q], Nv.data(), &gN[(size_t)q * m->nv * 3]);
  hipError_t e = dalloc(&p->d_gref, gref.size());
  if (e == hipSuccess) e = dalloc(&p->d_qw, (size_t)nq);
  if (e == hipSuccess && eb) e = dalloc(&p->d_gN, gN.size());
  if (e == hipSuccess && eb) e = hipMemcpy(p->d_gN, gN.data(), 8 * gN.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && eb) e = dalloc(&p->d_E, (size_t)m->n_cv * p->nloc * p->nloc);
  if (e == hipSuccess && eb) e = dalloc(&p->d_res, (size_t)m->n_cv * p->nloc);
  if (e == hipSuccess && eb) e = dalloc(&p->d_U, (size_t)p->ndof);
  if (e == hipSuccess && eb) e = dalloc(&p->d_rhs, (size_t)p->ndof);
  if (e == hipSuccess && eb) e = hipMemset(p->d_U, 0, 8 * (size_t)p->ndof);
  if (e == hipSuccess) e = dalloc(&p->d_R, R.size());
  if (e == hipSuccess) e = dalloc(&p->d_rec, (size_t)kRecStride * m->n_cv);
  if (e == hipSuccess) e = dalloc(&p->d_thr, (size_t)m->n_cv);
  if (e == hipSuccess) e = dalloc(&p->d_err, 4);
  if (e == hipSuccess) e = hipMemcpy(p->d_gref, gref.data(), 8 * gref.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_qw, mim->rule.w.data(), 8 * (size_t)nq, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->d_R, R.data(), 8 * R.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->d_err, 0, 4);
  if (!eb) {  // integer reference tensor, if R has one
    std::vector<uint8_t> rq;
    detect_int_reference(R, nd * nd, p->rk, p->r_L, rq);
    if (p->rk && e == hipSuccess) e = dalloc(&p->d_Rq, rq.size());
    if (p->rk && e == hipSuccess) e = hipMemcpy(p->d_Rq, rq.data(), rq.size(), hipMemcpyHostToDevice);
  }
  if (!eb) {  // FP32 copies of the tables for the FP32 contraction form
    std::vector<float> gref32(gref.begin(), gref.end()), R32(R.begin(), R.end());
    std::vector<float> qw32(mim->rule.w.begin(), mim->rule.w.end());
    if (e == hipSuccess) e = dalloc(&p->d_gref32, gref32.size());
    if (e == hipSuccess) e = dalloc(&p->d_qw32, qw32.size());
    if (e == hipSuccess) e = dalloc(&p->d_R32, R32.size());
    if (e == hipSuccess) e = hipMemcpy(p->d_gref32, gref32.data(), 4 * gref32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_qw32, qw32.data(), 4 * qw32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_R32, R32.data(), 4 * R32.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) { gfx_plan_destroy(p); return hipfail(c, e, "gfx_plan_create"); }
  *out = p;
  return GFX_OK;
}

int gfx_plan_set_coef(gfx_plan *p, int slot, const gfx_coef *cf) {
  if (!p || !cf || slot < 0 || slot >= term_coefs(p->term)) return GFX_E_INVALID;
  if (set_device(p->ctx)) return GFX_E_HIP;
  return plan_upload_coef(p, slot, cf);
}

int gfx_plan_set_form(gfx_plan *p, int form, int precision) {
  if (!p) return GFX_E_INVALID;
  if ((form != GFX_FORM_FACTORIZED && form != GFX_FORM_PERQP) || precision < GFX_PREC_FP64 ||
      precision > GFX_PREC_FP32_MFMA)
    return fail(p->ctx, GFX_E_INVALID, "unknown form or precision");
  if (precision >= GFX_PREC_BF16X2 &&
      (p->eb || p->term != GFX_TERM_ISO_ELAST || p->qdim != 3 || (p->nd != 10 && p->nd != 20)))
    return fail(p->ctx, GFX_E_UNSUPPORTED,
                "the element-batched matrix-core forms need isotropic elasticity on a P2 or P3 vector field "
                "of an affine tet mesh");
  p->form = form;
  p->precision = precision;
  return GFX_OK;
}

// Element records for the pattern build (FP64, no thresholds yet) or for an
// assembly (with thresholds; FP32 records for the FP32 contraction form).
static int plan_records(gfx_plan *p, bool with_thr) {
  const gfx_mesh *m = p->mf->mesh;
  gfx_ctx *c = p->ctx;
  const bool fp32 = with_thr && p->precision == GFX_PREC_FP32;
  if (fp32 && !p->d_rec32) HIPCHK(c, dalloc(&p->d_rec32, (size_t)kRecStride * m->n_cv));
  // coefficient c of the record: mu for elasticity (its Laplacian part), the
  // third coefficient of the isotropic term, none for the Laplacian
  const int sc = p->term == GFX_TERM_ISOTROPIC ? 2 : (p->term == GFX_TERM_ISO_ELAST ? 1 : -1);
  auto dptr = [&](int k) -> const double * {
    return k >= 0 && p->coef[k].kind == GFX_COEF_PER_ELEMENT ? p->d_coef[k] : nullptr;
  };
  HIPCHK(c, launch_records_tet(c->stream, m->n_cv, m->d_pts, m->d_cvx, dptr(0), p->coef[0].value,
                               dptr(1), p->coef[1].value, dptr(sc), sc >= 0 ? p->coef[sc].value : 0.0,
                               with_thr ? p->d_thr : nullptr, p->rk ? 1.0 / p->r_L : 1.0, p->d_rec,
                               fp32 ? p->d_rec32 : nullptr, p->d_err));
  return GFX_OK;
}

// the element kernel; with_keep (FP64): it also writes the keep bits
static int plan_eb_elements(gfx_plan *p, int precision, bool with_keep = false) {
  const gfx_mesh *m = p->mf->mesh;
  gfx_ctx *c = p->ctx;
  EbArgs a{};
  a.n_cv = m->n_cv; a.pts = m->d_pts; a.cvx = m->d_cvx; a.dofs = p->mf->d_dofs; a.U = p->d_U;
  a.ca = p->coef[0].kind == GFX_COEF_PER_ELEMENT ? p->d_coef[0] : nullptr;
  a.cb = p->coef[1].kind == GFX_COEF_PER_ELEMENT ? p->d_coef[1] : nullptr;
  a.ca_c = p->coef[0].value; a.cb_c = p->coef[1].value;
  a.gN = p->d_gN; a.gref = p->d_gref; a.qw = p->d_qw; a.nq = p->nq;
  a.E = p->d_E; a.res = p->d_res; a.thr = p->d_thr; a.err = p->d_err;
  a.keep = with_keep ? p->d_keep : nullptr; a.kwords = p->kwords;
  HIPCHK(c, launch_nh_element(c->stream, a, precision));
  return GFX_OK;
}

static int check_err_flag(gfx_plan *p) {
  gfx_ctx *c = p->ctx;
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, p->d_err, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag & 1) return fail(c, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
  if (flag & 2) return fail(c, GFX_E_PATTERN, "a node patch exceeds the kernel limits");
  return GFX_OK;
}

// Group the nodes by patch size (number of neighbour slots).  The assembly
// kernels reserve LDS for the largest patch of a launch, so one launch per
// group lets the many small patches run at the occupancy their own size
// allows.  Size classes with few nodes join the next class up.
static int plan_bins(gfx_plan *p, const std::vector<uint16_t> &ns, const std::vector<int32_t> &iptr) {
  gfx_ctx *c = p->ctx;
  const int64_t nn = p->n_nodes;
  static const int kCaps[] = {8,   16,  24,  32,   48,   64,   96,   128,  192,   256,   384,
                              512, 768, 1024, 1536, 2048, 3072, 4096, 8192, 16384, 32767};
  constexpr int kClasses = sizeof(kCaps) / sizeof(kCaps[0]);
  constexpr int64_t kMinBin = 16384;  // nodes; below this a launch is not worth its overhead
  std::vector<int> cls((size_t)nn);
  std::vector<int64_t> cnt(kClasses, 0);
  std::vector<int> cmax(kClasses, 0);
  for (int64_t n = 0; n < nn; ++n) {
    const int v = ns[n] & 0x7fff;
    int k = 0;
    while (kCaps[k] < v) ++k;
    cls[n] = k;
    cnt[k]++;
    cmax[k] = std::max(cmax[k], v);
  }
  int last = -1;
  for (int k = 0; k < kClasses; ++k)
    if (cnt[k]) last = k;
  std::vector<int> group(kClasses, 0);
  p->bins.clear();
  int64_t pending = 0, off = 0;
  int gmax = 0;
  for (int k = 0; k <= last; ++k) {
    group[k] = (int)p->bins.size();
    pending += cnt[k];
    gmax = std::max(gmax, cmax[k]);
    if (pending && (pending >= kMinBin || k == last)) {
      p->bins.push_back({off, pending, std::max(gmax, 1), 1});
      off += pending;
      pending = 0;
      gmax = 0;
    }
  }
  // patches per wavefront: small patches are packed so that one wave's
  // incidences fill about three passes of its lanes
  std::vector<int> bmax(p->bins.size(), 0);
  for (int64_t n = 0; n < nn; ++n) {
    const int b = group[cls[n]];
    bmax[b] = std::max(bmax[b], iptr[n + 1] - iptr[n]);
  }
  const int lanes = (64 / p->nd) * p->nd;
  for (size_t b = 0; b < p->bins.size(); ++b) {
    const int per = std::max(1, bmax[b]) * p->nd;
    if (p->eb)  // gather workgroups stream ~4 incidences each
      p->bins[b].npw = std::max(1, std::min({4, 4 / std::max(1, bmax[b]), 64 / std::max(1, p->max_inc)}));
    else
      p->bins[b].npw = std::max(1, std::min(4, 3 * lanes / per));
  }
  std::vector<int64_t> cur(p->bins.size());
  for (size_t b = 0; b < p->bins.size(); ++b) cur[b] = p->bins[b].off;
  std::vector<int32_t> order((size_t)nn);
  for (int64_t n = 0; n < nn; ++n) order[cur[group[cls[n]]]++] = (int32_t)n;
  dfree(p->d_order);
  if (nn) {
    HIPCHK(c, dalloc(&p->d_order, (size_t)nn));
    HIPCHK(c, hipMemcpy(p->d_order, order.data(), 4 * (size_t)nn, hipMemcpyHostToDevice));
  }
  return GFX_OK;
}

// Patch plans: nodes in node order (which is spatial order: first-encounter
// numbering over the element order) are packed into waves of up to 4 node
// patches while their incidences fill about three passes of the lanes and
// their LDS stays within kWaveLds bytes, so that LDS does not bound the
// occupancy below what the kernel's registers allow.  One launch then sweeps
// the mesh in order and the element records of a region are read from HBM
// about once.  Nodes whose patch alone exceeds the budget get one wave each,
// in launches by size class.  d_order receives 4 node ids per wave (-1: an
// empty slot): the large classes first, then the packed waves, one bin each.
static int plan_waves(gfx_plan *p, const std::vector<uint16_t> &ns, const std::vector<int32_t> &iptr) {
  gfx_ctx *c = p->ctx;
  const int64_t nn = p->n_nodes;
  const int Q = p->qdim, half = p->nd / 2;
  constexpr int kWaveLds = 5120;  // measured: 5 KB beats 4, 6 and 8 KB (P2, 64^3 x 6)
  const int inc_cap = 3 * (64 / (p->nd >= 20 ? p->nd / 2 : p->nd));  // three passes (kernel: lanes per incidence)
  auto need = [&](int64_t n, int &acc, int &flat, int &inc) {
    const int s = ns[n] & 0x7fff;
    const bool full = (ns[n] & 0x8000) != 0;
    inc = iptr[n + 1] - iptr[n];
    acc = Q * Q * s;
    flat = (full ? 0 : (Q * Q * s + 31) / 32) + inc * (1 + half);
  };
  // large patches: size classes by LDS bytes (each launch reserves LDS for
  // its own largest patch), one node per wave
  static const int kLargeCaps[] = {8192, 12288, 16384, 24576, 32768, 49152, 65536};
  constexpr int kLargeClasses = sizeof(kLargeCaps) / sizeof(kLargeCaps[0]);
  std::vector<int32_t> packed, large[kLargeClasses];
  gfx_plan::Bin bm{0, 0, 0, 4}, bl[kLargeClasses];
  for (auto &b : bl) b = gfx_plan::Bin{0, 0, 0, 1};
  int cnt = 0, sinc = 0, sacc = 0, sflat = 0;
  auto close = [&]() {
    if (!cnt) return;
    for (; cnt < 4; ++cnt) packed.push_back(-1);
    bm.acc_cap = std::max(bm.acc_cap, sacc);
    bm.flat_cap = std::max(bm.flat_cap, sflat);
    bm.inc_max = std::max(bm.inc_max, sinc);
    bm.count++;
    cnt = sinc = sacc = sflat = 0;
  };
  for (int64_t n = 0; n < nn; ++n) {
    int acc, flat, inc;
    need(n, acc, flat, inc);
    const int bytes = 8 * acc + 4 * flat;
    if (bytes > kWaveLds) {
      int k = 0;
      while (k < kLargeClasses - 1 && kLargeCaps[k] < bytes) ++k;
      if (bytes > kLargeCaps[k]) return fail(c, GFX_E_PATTERN, "node patch too large for the assembly kernel");
      large[k].push_back((int32_t)n);
      for (int t = 1; t < 4; ++t) large[k].push_back(-1);
      bl[k].acc_cap = std::max(bl[k].acc_cap, acc);
      bl[k].flat_cap = std::max(bl[k].flat_cap, flat);
      bl[k].inc_max = std::max(bl[k].inc_max, inc);
      bl[k].count++;
      continue;
    }
    if (cnt && (cnt == 4 || sinc + inc > inc_cap || 8 * (sacc + acc) + 4 * (sflat + flat) > kWaveLds)) close();
    packed.push_back((int32_t)n);
    cnt++;
    sinc += inc;
    sacc += acc;
    sflat += flat;
  }
  close();
  // the large patches are launched first, largest first (their waves are the
  // longest); the packed launch fills the GPU around them
  p->bins.clear();
  std::vector<int32_t> all;
  int64_t off = 0;
  for (int k = kLargeClasses - 1; k >= 0; --k) {
    if (!bl[k].count) continue;
    bl[k].off = off;
    off += bl[k].count;
    p->bins.push_back(bl[k]);
    all.insert(all.end(), large[k].begin(), large[k].end());
  }
  if (bm.count) {
    bm.off = off;
    p->bins.push_back(bm);
    all.insert(all.end(), packed.begin(), packed.end());
  }
  packed.swap(all);
  p->n_order = (int64_t)packed.size();
  dfree(p->d_order);
  if (!packed.empty()) {
    HIPCHK(c, dalloc(&p->d_order, packed.size()));
    HIPCHK(c, hipMemcpy(p->d_order, packed.data(), 4 * packed.size(), hipMemcpyHostToDevice));
  }
  return GFX_OK;
}

int gfx_plan_build_pattern(gfx_plan *p, int64_t *nnz_out) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->pinned) return fail(c, GFX_E_STATE, "the pattern is pinned (callers hold its CSR arrays)");
  if (set_device(c)) return GFX_E_HIP;
  const gfx_mesh *m = p->mf->mesh;
  const int64_t n_cv = m->n_cv, nn = p->n_nodes;
  const int nd = p->nd, Q = p->qdim;
  hipStream_t s = c->stream;
  int r;
  HIPCHK(c, hipMemsetAsync(p->d_err, 0, 4, s));
  if (!p->eb && (r = plan_records(p, false))) return r;  // element-batched: with the keep bits below
  // incidences
  dfree(p->d_inc_ptr); dfree(p->d_inc); dfree(p->d_slot); dfree(p->d_nslots); dfree(p->d_order); dfree(p->d_desc);
  dfree(p->d_node_ok);
  dfree(p->d_nbr); dfree(p->d_mask); dfree(p->d_indptr); dfree(p->d_indices); dfree(p->d_vals);
  dfree(p->d_keep);
  p->nnz = -1;
  p->bins.clear();
  p->n_inc = n_cv * nd;
  if (p->n_inc >= INT32_MAX) return fail(c, GFX_E_INVALID, "too many element-node incidences");
  if (n_cv == 0 || nn == 0) {  // no element: an empty matrix (all rows empty)
    HIPCHK(c, dalloc(&p->d_indptr, (size_t)p->ndof + 1));
    HIPCHK(c, hipMemsetAsync(p->d_indptr, 0, 8 * ((size_t)p->ndof + 1), s));
    HIPCHK(c, dalloc(&p->d_indices, 1));
    HIPCHK(c, dalloc(&p->d_vals, 1));
    if (p->eb && p->ndof) HIPCHK(c, hipMemsetAsync(p->d_rhs, 0, 8 * (size_t)p->ndof, s));
    HIPCHK(c, hipStreamSynchronize(s));
    p->nnz = 0;
    p->pts_version = m->version;
    p->values_changed = false;
    p->pattern_id++;
    if (nnz_out) *nnz_out = 0;
    return GFX_OK;
  }
  int32_t *tmp32 = nullptr;
  HIPCHK(c, dalloc(&p->d_inc_ptr, (size_t)nn + 1));
  HIPCHK(c, dalloc(&p->d_inc, (size_t)p->n_inc));
  HIPCHK(c, dalloc(&tmp32, (size_t)(2 * nn + scan_tmp_elems(nn))));
  hipError_t he = launch_incidences(s, n_cv, nd, Q, p->mf->d_dofs, nn, p->d_inc_ptr, p->d_inc, tmp32);
  if (he != hipSuccess) { dfree(tmp32); return hipfail(c, he, "incidences"); }
  // largest patch (LDS sort size) on the host
  std::vector<int32_t> iptr((size_t)nn + 1);
  he = hipMemcpyAsync(iptr.data(), p->d_inc_ptr, 4 * ((size_t)nn + 1), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  dfree(tmp32);
  if (he != hipSuccess) return hipfail(c, he, "incidence copy");
  int maxinc = 0;
  for (int64_t n = 0; n < nn; ++n) maxinc = std::max(maxinc, iptr[n + 1] - iptr[n]);
  p->max_cand = maxinc * nd;
  p->max_inc = maxinc;
  if (p->max_cand > 4096)
    return fail(c, GFX_E_PATTERN, "a node belongs to too many elements for the patch kernels");
  // thresholds and keep bits
  HIPCHK(c, dalloc(&p->d_keep, (size_t)n_cv * p->kwords));
  PatternArgs a{};
  a.n_cv = n_cv; a.n_nodes = nn; a.nd = nd; a.qdim = Q; a.nloc = p->nloc;
  a.kwords = p->kwords; a.mwpi = p->mwpi; a.dofs = p->mf->d_dofs; a.rec = p->d_rec;
  a.gref = p->d_gref; a.qw = p->d_qw; a.nq = p->nq; a.thr = p->d_thr; a.keep = p->d_keep;
  a.inc_ptr = p->d_inc_ptr; a.inc = p->d_inc; a.err = p->d_err;
  if (p->eb) {
    if ((r = plan_eb_elements(p, GFX_PREC_FP64, true))) return r;
  } else {
    HIPCHK(c, launch_keep(s, p->term, a));
  }
  // patches (the patch kernel addresses presence words and slot pairs with
  // 32-bit word offsets)
  if ((int64_t)p->n_inc * std::max(p->mwpi, nd / 2) >= (int64_t(1) << 32))
    return fail(c, GFX_E_UNSUPPORTED, "mesh too large for one plan (presence/slot maps beyond 2^32 words)");
  HIPCHK(c, dalloc(&p->d_slot, (size_t)p->n_inc * nd));
  HIPCHK(c, dalloc(&p->d_nslots, (size_t)nn));
  HIPCHK(c, dalloc(&p->d_nbr, (size_t)p->n_inc * nd));
  HIPCHK(c, dalloc(&p->d_mask, (size_t)p->n_inc * p->mwpi));
  int64_t *rowcnt = nullptr, *tmp64 = nullptr;
  HIPCHK(c, dalloc(&rowcnt, (size_t)p->ndof));
  HIPCHK(c, dalloc(&p->d_indptr, (size_t)p->ndof + 1));
  a.slot = p->d_slot; a.nslots = p->d_nslots; a.nbr = p->d_nbr; a.mask = p->d_mask;
  a.rowcnt = rowcnt; a.indptr = p->d_indptr;
  he = launch_patch_build(s, a, p->max_cand);
  if (he == hipSuccess) he = dalloc(&tmp64, (size_t)scan_tmp_elems(p->ndof));
  if (he == hipSuccess) he = scan_exclusive_i64(s, rowcnt, p->d_indptr, p->ndof, tmp64);
  int64_t nnz = 0;
  if (he == hipSuccess)
    he = hipMemcpyAsync(&nnz, p->d_indptr + p->ndof, 8, hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  dfree(rowcnt); dfree(tmp64);
  if (he != hipSuccess) return hipfail(c, he, "patch build");
  if ((r = check_err_flag(p))) return r;
  // nslots maximum for the assembly kernel's LDS
  std::vector<uint16_t> ns((size_t)nn);
  HIPCHK(c, hipMemcpy(ns.data(), p->d_nslots, 2 * (size_t)nn, hipMemcpyDeviceToHost));
  p->max_nslots = 0;
  for (int64_t n = 0; n < nn; ++n) p->max_nslots = std::max<int>(p->max_nslots, ns[n] & 0x7fff);
  if ((size_t)Q * Q * p->max_nslots * 8 > 64 * 1024)
    return fail(c, GFX_E_PATTERN, "node patch too large for the assembly kernel");
  if ((r = p->eb ? plan_bins(p, ns, iptr) : plan_waves(p, ns, iptr))) return r;
  if (p->eb) p->n_order = nn;
  dfree(p->d_desc);
  HIPCHK(c, dalloc(&p->d_desc, (size_t)p->n_order));
  HIPCHK(c, launch_fill_desc(s, p->n_order, Q, p->d_order, p->d_inc_ptr, p->d_nslots, p->d_indptr, p->d_desc));
  HIPCHK(c, dalloc(&p->d_indices, (size_t)nnz));
  HIPCHK(c, dalloc(&p->d_vals, (size_t)nnz));
  a.indices = p->d_indices;
  HIPCHK(c, launch_fill_columns(s, a));
  HIPCHK(c, hipStreamSynchronize(s));
  dfree(p->d_nbr);
  // the keep bits stay: re-validation after a change of values compares
  // the pattern they give with this one
  p->nnz = nnz;
  p->pts_version = m->version;
  p->values_changed = false;
  p->pattern_id++;
  if (nnz_out) *nnz_out = nnz;
  return GFX_OK;
}

// The values behind the stored pattern changed (coefficients, node positions
// or state): re-derive the per-element thresholds and keep bits at the new
// values, as GetFEM's per-assembly threshold does, and compare the presence
// bits they give with the stored ones.  Equal: the pattern stands and the
// new thresholds are used.  Different: the pattern is rebuilt (refused while
// pinned).  A degenerate element is reported here, like GetFEM's "Non
// invertible matrix" at every assembly.  For element-batched plans the FP64
// element matrices computed here serve the assembly that follows (*eb_fresh).
static int plan_validate(gfx_plan *p, bool *eb_fresh) {
  gfx_ctx *c = p->ctx;
  const gfx_mesh *m = p->mf->mesh;
  hipStream_t s = c->stream;
  *eb_fresh = false;
  if (p->nnz == 0) {  // no element: nothing to re-derive
    p->pts_version = m->version;
    p->values_changed = false;
    return GFX_OK;
  }
  int r;
  HIPCHK(c, hipMemsetAsync(p->d_err, 0, 4, s));
  if (!p->d_keep) HIPCHK(c, dalloc(&p->d_keep, (size_t)m->n_cv * p->kwords));
  PatternArgs a{};
  a.n_cv = m->n_cv; a.n_nodes = p->n_nodes; a.nd = p->nd; a.qdim = p->qdim; a.nloc = p->nloc;
  a.kwords = p->kwords; a.mwpi = p->mwpi; a.dofs = p->mf->d_dofs; a.rec = p->d_rec;
  a.gref = p->d_gref; a.qw = p->d_qw; a.nq = p->nq; a.thr = p->d_thr; a.keep = p->d_keep;
  a.inc_ptr = p->d_inc_ptr; a.inc = p->d_inc; a.slot = p->d_slot; a.nslots = p->d_nslots;
  a.mask = p->d_mask; a.err = p->d_err;
  if (!p->d_node_ok) HIPCHK(c, dalloc(&p->d_node_ok, (size_t)std::max<int64_t>(p->n_nodes, 1)));
  a.node_ok = p->d_node_ok;
  if (p->eb) {
    if ((r = plan_eb_elements(p, GFX_PREC_FP64, true))) return r;
  } else {
    if ((r = plan_records(p, false))) return r;
    HIPCHK(c, launch_keep(s, p->term, a));
  }
  // nodes whose rows keep every entry are confirmed by one pass; the exact
  // check runs only when some node is not
  HIPCHK(c, launch_mask_fast(s, a));
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, p->d_err, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (flag & 16) {
    HIPCHK(c, launch_mask_exact(s, a, p->max_nslots));
    HIPCHK(c, hipMemcpyAsync(&flag, p->d_err, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
  }
  if (flag & 1) return fail(c, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
  if (flag & 4) {
    if (p->pinned)
      return fail(c, GFX_E_STATE, "the new values change the stored pattern, which is pinned "
                                  "(callers hold its CSR arrays): unpin it and assemble again");
    int64_t nnz = 0;
    if ((r = gfx_plan_build_pattern(p, &nnz))) return r;
    *eb_fresh = p->eb;  // the rebuild ran the FP64 element kernel at the current values
    return GFX_OK;
  }
  p->pts_version = m->version;
  p->values_changed = false;
  *eb_fresh = p->eb;
  return GFX_OK;
}

int gfx_plan_assemble(gfx_plan *p) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  int r;
  bool eb_fresh = false;  // element-batched: E and thresholds already computed in FP64
  if (p->values_changed || p->pts_version != p->mf->mesh->version)
    if ((r = plan_validate(p, &eb_fresh))) return r;
  hipEvent_t *ev = nullptr;
  if (p->timing) {
    const size_t need = 3 * (size_t)(p->n_timed + 1);
    while (p->ev.size() < need) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      p->ev.push_back(e);
    }
    ev = &p->ev[3 * (size_t)p->n_timed];
    p->n_timed++;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
  }
  if (!(eb_fresh && p->precision == GFX_PREC_FP64))
    if ((r = p->eb ? plan_eb_elements(p, p->precision) : plan_records(p, true))) return r;
  const bool bdb = !p->eb && p->precision >= GFX_PREC_BF16X2;
  const bool bdb64 = bdb && p->precision == GFX_PREC_FP64_MFMA;
  if (bdb) {  // element matrices on the matrix cores, then the node-patch gather below
    const int64_t n_cv = p->mf->mesh->n_cv;
    // the kernels factor D through its square root: mu > 0 and 3 lambda + 2 mu > 0
    // for every element (sufficient: from the smallest lambda and mu)
    const double lmin = p->coef_min[0], mmin = p->coef_min[1];
    if (!(mmin > 0.0 && 3.0 * std::min(lmin, 0.0) + 2.0 * mmin > 0.0))
      return fail(c, GFX_E_UNSUPPORTED,
                  "the matrix-core forms need mu > 0 and 3*lambda + 2*mu > 0 on every element");
    const size_t ne = (size_t)n_cv * p->nloc * p->nloc;
    if (bdb64) {
      if (!p->d_E64) HIPCHK(c, dalloc(&p->d_E64, ne));
      HIPCHK(c, launch_elast_bdb_mfma_hp(c->stream, n_cv, p->nd, p->nq, p->d_rec, p->d_gref, p->d_qw, p->d_E64,
                                         nullptr));
    } else {
      if (!p->d_E32) HIPCHK(c, dalloc(&p->d_E32, ne));
      if (p->precision == GFX_PREC_BF16X2)
        HIPCHK(c, launch_elast_bdb_mfma(c->stream, n_cv, p->nd, p->nq, p->d_rec, p->d_gref, p->d_qw, p->d_E32));
      else
        HIPCHK(c, launch_elast_bdb_mfma_hp(c->stream, n_cv, p->nd, p->nq, p->d_rec, p->d_gref, p->d_qw, nullptr,
                                           p->d_E32));
    }
  }
  if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  PatchArgs a{};
  a.inc_ptr = p->d_inc_ptr; a.inc = p->d_inc; a.slot = p->d_slot;
  a.nslots = p->d_nslots; a.mask = p->d_mask; a.mwpi = p->mwpi; a.indptr = p->d_indptr;
  a.nd = p->nd; a.qdim = p->qdim; a.nq = p->nq; a.rec = p->d_rec; a.gref = p->d_gref;
  a.qw = p->d_qw; a.R = p->d_R; a.vals = p->d_vals; a.max_inc = p->max_inc;
  a.rec32 = p->d_rec32; a.gref32 = p->d_gref32; a.qw32 = p->d_qw32; a.R32 = p->d_R32;
  a.Rq = p->d_Rq; a.rk = p->rk;
  // The launches (patch plans: large-patch size classes and the packed
  // waves; element-batched plans: gather bins) write disjoint rows: they run
  // concurrently, each on its own stream (up to seven auxiliary streams forked
  // from and joined back into the context's stream by events, no host
  // synchronisation), which fills the tail of one launch with the next.
  const int nb = (int)p->bins.size();
  const int naux = std::min(nb - 1, 7);
  while ((int)p->aux.size() < naux) {
    hipStream_t st; hipEvent_t ej;
    HIPCHK(c, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&ej, hipEventDisableTiming));
    p->aux.push_back(st); p->join.push_back(ej);
  }
  if (naux > 0) {
    if (!p->fork) HIPCHK(c, hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(p->fork, c->stream));
    for (int i = 0; i < naux; ++i) HIPCHK(c, hipStreamWaitEvent(p->aux[i], p->fork, 0));
  }
  for (int k = 0; k < nb; ++k) {  // one launch per bin, spread over streams
    const gfx_plan::Bin &b = p->bins[k];
    hipStream_t st = k == 0 ? c->stream : p->aux[(k - 1) % naux];
    if (p->eb) {  // the element kernel wrote E in the form's precision
      a.n_nodes = b.count;
      a.order = p->d_order + b.off;
      a.desc = p->d_desc + b.off;
      a.npw = b.npw;
      HIPCHK(c, launch_gather_emat(st, a, p->d_E, p->precision == GFX_PREC_FP32, p->d_thr,
                                   b.npw * p->qdim * p->qdim * b.cap, b.npw * p->max_inc));
    } else if (bdb) {  // the packed waves' descriptor groups (4 slots, empty ones zero)
      a.n_nodes = 4 * b.count;
      a.desc = p->d_desc + 4 * b.off;
      a.npw = 4;
      HIPCHK(c, launch_gather_emat(st, a, bdb64 ? (const void *)p->d_E64 : (const void *)p->d_E32, !bdb64,
                                   p->d_thr, b.acc_cap, b.inc_max));
    } else {
      a.n_waves = b.count;
      a.desc = p->d_desc + 4 * b.off;
      a.lds_acc = b.acc_cap;
      a.lds_flat = b.flat_cap;
      HIPCHK(c, launch_assemble_patch(st, p->term, p->form, p->precision, a));
    }
  }
  for (int i = 0; i < naux; ++i) {
    HIPCHK(c, hipEventRecord(p->join[i], p->aux[i]));
    HIPCHK(c, hipStreamWaitEvent(c->stream, p->join[i], 0));
  }
  if (p->eb)
    HIPCHK(c, launch_gather_vec(c->stream, p->n_nodes, p->nd, p->qdim, p->d_inc_ptr, p->d_inc, p->d_res,
                                p->d_rhs));
  if (ev) HIPCHK(c, hipEventRecord(ev[2], c->stream));
  return GFX_OK;
}

int gfx_plan_set_state(gfx_plan *p, const double *h_U) {
  if (!p || !h_U) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->eb) return fail(c, GFX_E_STATE, "this plan has no state");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipMemcpyAsync(p->d_U, h_U, 8 * (size_t)p->ndof, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  p->values_changed = true;  // the element matrices, thresholds and pattern follow the state
  return GFX_OK;
}

int gfx_plan_set_state_device(gfx_plan *p, const double *d_U) {
  if (!p || !d_U) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->eb) return fail(c, GFX_E_STATE, "this plan has no state");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipMemcpyAsync(p->d_U, d_U, 8 * (size_t)p->ndof, hipMemcpyDeviceToDevice, c->stream));
  p->values_changed = true;
  return GFX_OK;
}

int gfx_plan_pin_pattern(gfx_plan *p, int pin) {
  if (!p) return GFX_E_INVALID;
  p->pinned = pin != 0;
  return GFX_OK;
}

int gfx_plan_pattern_id(gfx_plan *p, int64_t *id) {
  if (!p || !id) return GFX_E_INVALID;
  *id = p->pattern_id;
  return GFX_OK;
}

int gfx_plan_assemble_residual(gfx_plan *p) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->eb) return fail(c, GFX_E_STATE, "this plan has no residual");
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  int r;
  if ((r = plan_eb_elements(p, p->precision))) return r;
  HIPCHK(c, launch_gather_vec(c->stream, p->n_nodes, p->nd, p->qdim, p->d_inc_ptr, p->d_inc, p->d_res,
                              p->d_rhs));
  return GFX_OK;
}

int gfx_plan_add_to_matrix(gfx_plan *p, int orientation, int64_t n, const int64_t *d_ptr,
                           const int32_t *d_ind, double *d_vals) {
  if (!p || !d_ptr || !d_ind || !d_vals || (orientation != GFX_CSR && orientation != GFX_CSC))
    return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern and assemble first");
  if (n != p->ndof) return fail(c, GFX_E_INVALID, "target matrix of another size");
  if (set_device(c)) return GFX_E_HIP;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemsetAsync(p->d_err, 0, 4, s));
  HIPCHK(c, launch_add_into(s, orientation == GFX_CSC, p->ndof, p->d_indptr, p->d_indices, p->d_vals, d_ptr,
                            d_ind, d_vals, p->d_err));
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, p->d_err, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (flag & 8) return fail(c, GFX_E_PATTERN, "the target matrix lacks entries of the assembled pattern");
  return GFX_OK;
}

int gfx_plan_spmv_add(gfx_plan *p, const double *d_x, double *d_y) {
  if (!p || !d_x || !d_y) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, launch_spmv(c->stream, p->ndof, p->nnz, p->d_indptr, p->d_indices, p->d_vals, d_x, d_y, true));
  return GFX_OK;
}

int gfx_plan_spmv(gfx_plan *p, const double *d_x, double *d_y) {
  if (!p || !d_x || !d_y) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, launch_spmv(c->stream, p->ndof, p->nnz, p->d_indptr, p->d_indices, p->d_vals, d_x, d_y));
  return GFX_OK;
}

// ---- incomplete LDL^T (gmm::ildlt_precond) of the plan's matrix
static int il_prepare(gfx_plan *p) {
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern and assemble first");
  IldltData &L = p->il;
  const int64_t n = p->ndof;
  const int qb = (p->qdim == 3 && n % 3 == 0) ? 3 : 1;
  hipStream_t s = c->stream;
  int herr = 0;
  if (L.pattern_id != p->pattern_id || L.n != n || L.qb != qb) {
    il_free(L);
    const int64_t nu = n / qb;
    HIPCHK(c, dalloc(&L.d_dpos, (size_t)std::max<int64_t>(n, 1)));
    HIPCHK(c, dalloc(&L.d_fdeg, (size_t)std::max<int64_t>(nu, 1)));
    HIPCHK(c, dalloc(&L.d_bdeg, (size_t)std::max<int64_t>(nu, 1)));
    HIPCHK(c, dalloc(&L.d_flist, (size_t)std::max<int64_t>(nu, 1)));
    HIPCHK(c, dalloc(&L.d_blist, (size_t)std::max<int64_t>(nu, 1)));
    HIPCHK(c, dalloc(&L.d_counter, 1));
    HIPCHK(c, dalloc(&L.d_F, (size_t)std::max<int64_t>(p->nnz, 1)));
    HIPCHK(c, dalloc(&L.d_work, (size_t)(n / 1024 + 2)));
    // the longest row sizes the factorization's LDS
    std::vector<int64_t> ip((size_t)n + 1);
    HIPCHK(c, hipMemcpyAsync(ip.data(), p->d_indptr, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    int64_t cap = 0;
    for (int64_t i = 0; i < n; ++i) cap = std::max(cap, ip[(size_t)i + 1] - ip[(size_t)i]);
    if (44 * cap > 64 * 1024) {
      il_free(L);
      return fail(c, GFX_E_UNSUPPORTED, "ILDLT: a row of " + std::to_string(cap) + " entries exceeds the "
                                        "factorization's LDS (1489 entries)");
    }
    L.cap = (int)cap;
    HIPCHK(c, ildlt_analyse(s, L, n, qb, p->d_indptr, p->d_indices, p->d_err, &herr));
    if (herr) {
      il_free(L);
      return fail(c, GFX_E_UNSUPPORTED, (herr & 1) ? "ILDLT: a row has no diagonal entry"
                                                   : "ILDLT: the matrix pattern is not symmetric");
    }
    L.pattern_id = p->pattern_id;
  }
  HIPCHK(c, ildlt_factor(s, L, p->d_indptr, p->d_indices, p->d_vals, L.cap, p->d_err, &herr));
  if (herr) {
    L.factored = false;
    return fail(c, GFX_E_UNSUPPORTED, "ILDLT: a pivot that GetFEM replaces by 1 (|D_k| <= the running pivot "
                                      "bound, gmm_precond_ildlt.h:171-176); the device factorization does not");
  }
  L.factored = true;
  return GFX_OK;
}

int gfx_plan_ildlt_factor(gfx_plan *p, int64_t *n_levels_forward, int64_t *n_levels_backward) {
  if (!p) return GFX_E_INVALID;
  if (set_device(p->ctx)) return GFX_E_HIP;
  int r = il_prepare(p);
  if (r) return r;
  if (n_levels_forward) *n_levels_forward = (int64_t)p->il.flev.size() - 1;
  if (n_levels_backward) *n_levels_backward = (int64_t)p->il.blev.size() - 1;
  return GFX_OK;
}

int gfx_plan_ildlt_values(gfx_plan *p, double *h_F) {
  if (!p || !h_F) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->il.factored || p->il.pattern_id != p->pattern_id) return fail(c, GFX_E_STATE, "no ILDLT factor");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(h_F, p->il.d_F, 8 * (size_t)p->nnz, hipMemcpyDeviceToHost));
  return GFX_OK;
}

int gfx_plan_ildlt_apply(gfx_plan *p, const double *d_r, double *d_z) {
  if (!p || !d_r || !d_z) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->il.factored || p->il.pattern_id != p->pattern_id) return fail(c, GFX_E_STATE, "no ILDLT factor");
  if (set_device(c)) return GFX_E_HIP;
  IldltData &L = p->il;
  if (!L.d_cg) HIPCHK(c, dalloc(&L.d_cg, ildlt_cg_doubles(p->ndof)));
  HIPCHK(c, ildlt_apply(c->stream, L, p->d_indptr, p->d_indices, d_r, d_z, L.d_cg + 4 * p->ndof,
                        L.d_cg + 5 * p->ndof));
  return GFX_OK;
}

int gfx_plan_solve_pcg(gfx_plan *p, int precond, const double *d_b, double *d_x, double rtol, int maxit,
                       int *iters, double *rel_res) {
  if (!p || !d_b || !d_x || !iters || !rel_res || !(rtol > 0) || maxit < 0) return GFX_E_INVALID;
  if (precond == GFX_PRECOND_BLOCK_JACOBI) return gfx_plan_solve_cg(p, d_b, d_x, rtol, maxit, iters, rel_res);
  gfx_ctx *c = p->ctx;
  if (precond != GFX_PRECOND_ILDLT) return fail(c, GFX_E_INVALID, "unknown preconditioner");
  if (set_device(c)) return GFX_E_HIP;
  // the preconditioner is built from the current values at every solve, as
  // linear_solver_cg_preconditioned_ildlt does
  int r = il_prepare(p);
  if (r) return r;
  IldltData &L = p->il;
  if (!L.d_cg) HIPCHK(c, dalloc(&L.d_cg, ildlt_cg_doubles(p->ndof)));
  HIPCHK(c, ildlt_pcg(c->stream, L, p->nnz, p->d_indptr, p->d_indices, p->d_vals, d_b, d_x, rtol, maxit, L.d_cg,
                      iters, rel_res));
  return GFX_OK;
}

int gfx_plan_solve_cg(gfx_plan *p, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                      double *rel_res) {
  if (!p || !d_b || !d_x || !iters || !rel_res || !(rtol > 0) || maxit < 0) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern and assemble first");
  if (set_device(c)) return GFX_E_HIP;
  if (!p->d_cg) HIPCHK(c, dalloc(&p->d_cg, cg_workspace_doubles(p->ndof)));
  // vector fields: 3x3 block Jacobi (a node's coupled components), else the diagonal
  const int qb = (p->qdim == 3 && p->ndof % 3 == 0) ? 3 : 1;
  HIPCHK(c, cg_solve(c->stream, p->ndof, p->nnz, p->d_indptr, p->d_indices, p->d_vals, d_b, d_x, rtol, maxit,
                     qb, p->d_cg, iters, rel_res));
  return GFX_OK;
}

int gfx_plan_dof_constraints(gfx_plan *p, int64_t n_fixed, const int32_t *h_dofs, const double *h_go,
                             const double *d_pr, double *d_rhs, int flags) {
  if (!p || n_fixed < 0 || (n_fixed && (!h_dofs || !h_go))) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern and assemble first");
  if (set_device(c)) return GFX_E_HIP;
  const int64_t n = p->ndof;
  std::vector<uint8_t> fixed((size_t)n, 0);
  std::vector<double> go((size_t)n, 0.0);
  for (int64_t k = 0; k < n_fixed; ++k) {
    if (h_dofs[k] < 0 || h_dofs[k] >= n) return fail(c, GFX_E_INVALID, "constrained dof out of range");
    fixed[h_dofs[k]] = 1;
    go[h_dofs[k]] = h_go[k];
  }
  if (!p->d_fixed) HIPCHK(c, dalloc(&p->d_fixed, (size_t)n));
  if (!p->d_go) HIPCHK(c, dalloc(&p->d_go, (size_t)n));
  HIPCHK(c, hipMemcpyAsync(p->d_fixed, fixed.data(), (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(p->d_go, go.data(), 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(p->d_err, 0, 4, c->stream));
  HIPCHK(c, launch_dof_constraints(c->stream, n, p->d_indptr, p->d_indices, p->d_vals, p->d_fixed, p->d_go,
                                   d_pr, d_rhs, flags, p->d_err));
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, p->d_err, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag & 4) return fail(c, GFX_E_PATTERN, "a constrained dof has no diagonal entry in the pattern");
  return GFX_OK;
}

int gfx_plan_residual(gfx_plan *p, double *h_V) {
  if (!p || !h_V) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->eb) return fail(c, GFX_E_STATE, "this plan has no residual");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(h_V, p->d_rhs, 8 * (size_t)p->ndof, hipMemcpyDeviceToHost));
  return GFX_OK;
}

int gfx_plan_copy_residual_device(gfx_plan *p, double *d_V) {
  if (!p || !d_V) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->eb) return fail(c, GFX_E_STATE, "this plan has no residual");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipMemcpyAsync(d_V, p->d_rhs, 8 * (size_t)p->ndof, hipMemcpyDeviceToDevice, c->stream));
  return GFX_OK;
}

int gfx_plan_set_timing(gfx_plan *p, int enable) {
  if (!p) return GFX_E_INVALID;
  p->timing = enable != 0;
  p->n_timed = 0;
  return GFX_OK;
}

int gfx_plan_timing_summary(gfx_plan *p, int *n_calls, double *ms_geometry, double *ms_assemble) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (set_device(c)) return GFX_E_HIP;
  double g = 0.0, a = 0.0;
  for (int k = 0; k < p->n_timed; ++k) {
    hipEvent_t *ev = &p->ev[3 * (size_t)k];
    HIPCHK(c, hipEventSynchronize(ev[2]));
    float x = 0.f, y = 0.f;
    HIPCHK(c, hipEventElapsedTime(&x, ev[0], ev[1]));
    HIPCHK(c, hipEventElapsedTime(&y, ev[1], ev[2]));
    g += x; a += y;
  }
  if (n_calls) *n_calls = p->n_timed;
  if (ms_geometry) *ms_geometry = g;
  if (ms_assemble) *ms_assemble = a;
  p->n_timed = 0;
  return GFX_OK;
}

int gfx_plan_timing_calls(gfx_plan *p, int max_calls, int *n_calls, float *ms_geometry, float *ms_assemble) {
  if (!p || !n_calls || max_calls < 0) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (set_device(c)) return GFX_E_HIP;
  const int n = std::min(p->n_timed, max_calls);
  for (int k = 0; k < n; ++k) {
    hipEvent_t *ev = &p->ev[3 * (size_t)k];
    HIPCHK(c, hipEventSynchronize(ev[2]));
    float x = 0.f, y = 0.f;
    HIPCHK(c, hipEventElapsedTime(&x, ev[0], ev[1]));
    HIPCHK(c, hipEventElapsedTime(&y, ev[1], ev[2]));
    if (ms_geometry) ms_geometry[k] = x;
    if (ms_assemble) ms_assemble[k] = y;
  }
  *n_calls = n;
  p->n_timed = 0;
  return GFX_OK;
}

int gfx_plan_copy_csr_device(gfx_plan *p, int64_t *d_indptr, int32_t *d_indices, double *d_vals) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  hipStream_t s = c->stream;
  if (d_indptr)
    HIPCHK(c, hipMemcpyAsync(d_indptr, p->d_indptr, 8 * ((size_t)p->ndof + 1), hipMemcpyDeviceToDevice, s));
  if (d_indices)
    HIPCHK(c, hipMemcpyAsync(d_indices, p->d_indices, 4 * (size_t)p->nnz, hipMemcpyDeviceToDevice, s));
  if (d_vals) HIPCHK(c, hipMemcpyAsync(d_vals, p->d_vals, 8 * (size_t)p->nnz, hipMemcpyDeviceToDevice, s));
  return GFX_OK;
}

int gfx_plan_dims(gfx_plan *p, int64_t *nrows, int64_t *nnz) {
  if (!p) return GFX_E_INVALID;
  if (nrows) *nrows = p->ndof;
  if (nnz) *nnz = p->nnz;
  return GFX_OK;
}

int gfx_plan_export_csr(gfx_plan *p, int64_t *h_indptr, int32_t *h_indices, double *h_vals) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (p->nnz < 0) return fail(c, GFX_E_STATE, "build the pattern first");
  if (set_device(c)) return GFX_E_HIP;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h_indptr) HIPCHK(c, hipMemcpy(h_indptr, p->d_indptr, 8 * ((size_t)p->ndof + 1), hipMemcpyDeviceToHost));
  if (h_indices) HIPCHK(c, hipMemcpy(h_indices, p->d_indices, 4 * (size_t)p->nnz, hipMemcpyDeviceToHost));
  if (h_vals) HIPCHK(c, hipMemcpy(h_vals, p->d_vals, 8 * (size_t)p->nnz, hipMemcpyDeviceToHost));
  return GFX_OK;
}

int gfx_plan_export_csc(gfx_plan *p, int64_t *h_colptr, uint32_t *h_rows, double *h_vals) {
  if (!p || !h_colptr) return GFX_E_INVALID;
  const int64_t n = p->ndof, nnz = p->nnz;
  if (nnz < 0) return fail(p->ctx, GFX_E_STATE, "build the pattern first");
  std::vector<int64_t> ip((size_t)n + 1);
  std::vector<int32_t> ix((size_t)nnz);
  std::vector<double> v(h_vals ? (size_t)nnz : 0);
  int r = gfx_plan_export_csr(p, ip.data(), ix.data(), h_vals ? v.data() : nullptr);
  if (r) return r;
  // transpose (counting sort by column keeps rows sorted)
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

int gfx_plan_device_csr(gfx_plan *p, const int64_t **d_indptr, const int32_t **d_indices,
                        double **d_vals) {
  if (!p) return GFX_E_INVALID;
  if (p->nnz < 0) return fail(p->ctx, GFX_E_STATE, "build the pattern first");
  if (d_indptr) *d_indptr = p->d_indptr;
  if (d_indices) *d_indices = p->d_indices;
  if (d_vals) *d_vals = p->d_vals;
  return GFX_OK;
}

int gfx_plan_last_timings(gfx_plan *p, float *ms_geometry, float *ms_assemble) {
  if (!p) return GFX_E_INVALID;
  gfx_ctx *c = p->ctx;
  if (!p->timing || p->n_timed == 0) return fail(c, GFX_E_STATE, "no timed call recorded");
  if (set_device(c)) return GFX_E_HIP;
  hipEvent_t *ev = &p->ev[3 * (size_t)(p->n_timed - 1)];
  HIPCHK(c, hipEventSynchronize(ev[2]));
  float g = 0.f, a = 0.f;
  HIPCHK(c, hipEventElapsedTime(&g, ev[0], ev[1]));
  HIPCHK(c, hipEventElapsedTime(&a, ev[1], ev[2]));
  if (ms_geometry) *ms_geometry = g;
  if (ms_assemble) *ms_assemble = a;
  return GFX_OK;
}

}  // extern "C"

// ----------------------------------------------------------- compliance
extern "C" {

int gfx_compliance_destroy(gfx_compliance *c) {
  if (!c) return GFX_E_INVALID;
  (void)hipSetDevice(c->ctx->device);
  dfree(c->d_vinc_ptr); dfree(c->d_dinc_ptr); dfree(c->d_vinc); dfree(c->d_dinc);
  dfree(c->d_cX); dfree(c->d_cU); dfree(c->d_Jpart); dfree(c->d_err);
  gfx_ctx *ctx = c->ctx;
  gfx_mesh_fem *mf = c->mf;
  delete c;
  unref(mf); unref(ctx);
  return GFX_OK;
}

int gfx_compliance_create(gfx_mesh_fem *mf, gfx_mesh_im *mim, double lambda, double mu,
                          double emin, double penal, gfx_compliance **out) {
  if (!mf || !mim || !out) return GFX_E_INVALID;
  *out = nullptr;
  gfx_ctx *ctx = mf->mesh->ctx;
  const gfx_mesh *m = mf->mesh;
  if (m->gt != GT_TET_P1 || mf->fem.kind != FEM_PK || mf->fem.k != 1 || mf->qdim != 3)
    return fail(ctx, GFX_E_UNSUPPORTED, "compliance adjoint needs FEM_PK(3,1), qdim 3, GT_PK(3,1)");
  if (mim->mesh != m) return fail(ctx, GFX_E_INVALID, "mesh_im on another mesh");
  if (m->n_cv == 0) return fail(ctx, GFX_E_INVALID, "empty mesh");
  if (set_device(ctx)) return GFX_E_HIP;
  gfx_compliance *c = new (std::nothrow) gfx_compliance();
  if (!c) return fail(ctx, GFX_E_OOM, "host allocation");
  c->ctx = ctx; c->mf = mf; c->lam = lambda; c->mu = mu; c->emin = emin; c->penal = penal;
  ref(ctx); ref(mf);  // released by gfx_compliance_destroy
  c->n_cv = m->n_cv; c->n_pts = m->n_pts; c->ndof = mf->ndof;
  c->wsum = 0.0;
  for (double w : mim->rule.w) c->wsum += w;  // P1: the integrand is constant per element
  const int64_t nnode = mf->ndof / 3;
  int32_t *tmp = nullptr;
  const int64_t big = std::max(m->n_pts, nnode);
  hipError_t e = dalloc(&c->d_vinc_ptr, (size_t)m->n_pts + 1);
  if (e == hipSuccess) e = dalloc(&c->d_dinc_ptr, (size_t)nnode + 1);
  if (e == hipSuccess) e = dalloc(&c->d_vinc, (size_t)m->n_cv * 4);
  if (e == hipSuccess) e = dalloc(&c->d_dinc, (size_t)m->n_cv * 4);
  if (e == hipSuccess) e = dalloc(&c->d_cX, (size_t)m->n_cv * 12);
  if (e == hipSuccess) e = dalloc(&c->d_cU, (size_t)m->n_cv * 12);
  if (e == hipSuccess) e = dalloc(&c->d_Jpart, 2 * (size_t)compliance_nblocks(m->n_cv));
  if (e == hipSuccess) e = dalloc(&c->d_err, 1);
  if (e == hipSuccess) e = hipMemset(c->d_err, 0, 4);  // sticky: set by any evaluation that meets J = 0
  if (e == hipSuccess) e = dalloc(&tmp, (size_t)(2 * big + scan_tmp_elems(big)));
  if (e == hipSuccess)
    e = launch_incidences(ctx->stream, m->n_cv, 4, 1, m->d_cvx, m->n_pts, c->d_vinc_ptr, c->d_vinc, tmp);
  if (e == hipSuccess)
    e = launch_incidences(ctx->stream, m->n_cv, 4, 3, mf->d_dofs, nnode, c->d_dinc_ptr, c->d_dinc, tmp);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dfree(tmp);
  if (e != hipSuccess) {
    gfx_compliance_destroy(c);
    return hipfail(ctx, e, "gfx_compliance_create");
  }
  *out = c;
  return GFX_OK;
}

int gfx_compliance_status(gfx_compliance *c) {
  if (!c) return GFX_E_INVALID;
  gfx_ctx *ctx = c->ctx;
  if (set_device(ctx)) return GFX_E_HIP;
  int flag = 0;
  HIPCHK(ctx, hipMemcpyAsync(&flag, c->d_err, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (flag) {
    HIPCHK(ctx, hipMemsetAsync(c->d_err, 0, 4, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return fail(ctx, GFX_E_DEGENERATE_ELEMENT, "Non invertible matrix (degenerate element)");
  }
  return GFX_OK;
}

int gfx_compliance_eval(gfx_compliance *c, const double *d_X, const double *d_rho, const double *d_u,
                        double *d_J, double *d_dJ_drho, double *d_dJ_dX, double *d_dJ_du) {
  if (!c || !d_rho || !d_u || !d_J) return GFX_E_INVALID;
  gfx_ctx *ctx = c->ctx;
  if (set_device(ctx)) return GFX_E_HIP;
  hipStream_t s = ctx->stream;
  const gfx_mesh *m = c->mf->mesh;
  // kernels only (no memset): the call can be captured into a graph and replayed
  HIPCHK(ctx, launch_compliance(s, c->n_cv, d_X ? d_X : m->d_pts, m->d_cvx, c->mf->d_dofs, d_u, d_rho,
                                c->lam, c->mu, c->emin, c->penal, c->wsum, c->d_Jpart, d_J, d_dJ_drho,
                                d_dJ_dX ? c->d_cX : nullptr, d_dJ_du ? c->d_cU : nullptr, c->d_err));
  HIPCHK(ctx, launch_compliance_finish(s, c->n_cv, c->n_pts, c->d_vinc_ptr, c->d_vinc, c->d_cX, d_dJ_dX,
                                       c->ndof / 3, c->d_dinc_ptr, c->d_dinc, c->d_cU, d_dJ_du,
                                       c->d_Jpart, d_J));
  return GFX_OK;
}

int gfx_compliance_eval_noalloc(gfx_compliance *c, const double *d_X, const double *d_rho, const double *d_u,
                                double *d_J, double *d_dJ_drho, double *d_dJ_dX, double *d_dJ_du) {
  if (!c || !d_rho || !d_u || !d_J) return GFX_E_INVALID;
  gfx_ctx *ctx = c->ctx;
  if (set_device(ctx)) return GFX_E_HIP;
  double *scr = nullptr;
  if (compliance_scratch(ctx->device, &scr) != hipSuccess || !scr)
    return fail(ctx, GFX_E_HIP, "hipGetSymbolAddress of the compliance scratch failed");
  // [J block partials, padded to 256 bytes][dJ/dX contributions][dJ/du contributions]
  const int64_t jp = (2 * compliance_nblocks(c->n_cv) + 31) & ~int64_t(31);
  const int64_t per = 12 * c->n_cv;
  const int64_t need = jp + (d_dJ_dX ? per : 0) + (d_dJ_du ? per : 0);
  if (8 * need > (int64_t)kComplianceScratchBytes)
    return fail(ctx, GFX_E_INVALID,
                "gfx_compliance_eval_noalloc: " + std::to_string(c->n_cv) + " elements need " +
                    std::to_string(8 * need) + " bytes of scratch, the static scratch holds " +
                    std::to_string(kComplianceScratchBytes) + " (at most about " +
                    std::to_string(kComplianceScratchBytes / 192) + " elements with both node gradients)");
  double *Jpart = scr, *cX = d_dJ_dX ? scr + jp : nullptr;
  double *cU = d_dJ_du ? scr + jp + (d_dJ_dX ? per : 0) : nullptr;
  hipStream_t s = ctx->stream;
  const gfx_mesh *m = c->mf->mesh;
  HIPCHK(ctx, launch_compliance(s, c->n_cv, d_X ? d_X : m->d_pts, m->d_cvx, c->mf->d_dofs, d_u, d_rho, c->lam,
                                c->mu, c->emin, c->penal, c->wsum, Jpart, d_J, d_dJ_drho, cX, cU, c->d_err));
  HIPCHK(ctx, launch_compliance_finish(s, c->n_cv, c->n_pts, c->d_vinc_ptr, c->d_vinc, cX, d_dJ_dX, c->ndof / 3,
                                       c->d_dinc_ptr, c->d_dinc, cU, d_dJ_du, Jpart, d_J));
  return GFX_OK;
}

}  // extern "C"
