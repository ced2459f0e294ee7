// gfx_workspace.hpp -- header-only C++ facade with GetFEM's generic assembly
// API (getfem::ga_workspace, src/getfem/getfem_generic_assembly.h:262-597)
// over the C ABI of libgfx (gfx.h).  Errors are thrown as
// getfem_amd::gfx_error, a std::logic_error like GetFEM's gmm::gmm_error
// (src/gmm/gmm_except.h:55).
#ifndef GFX_WORKSPACE_HPP
#define GFX_WORKSPACE_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gfx.h"

namespace getfem_amd {

class gfx_error : public std::logic_error {
 public:
  gfx_error(int code, const std::string &msg)
      : std::logic_error("libgfx error " + std::to_string(code) + ": " + msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class context;
inline void check(int code, gfx_ctx *c) {
  if (code != GFX_OK) throw gfx_error(code, c ? gfx_ctx_last_error(c) : "");
}

class context {
 public:
  explicit context(int device = 0) {
    if (int r = gfx_ctx_create(device, &h_)) throw gfx_error(r, "cannot open device");
  }
  ~context() { if (h_) gfx_ctx_destroy(h_); }
  context(const context &) = delete;
  context &operator=(const context &) = delete;
  gfx_ctx *handle() const { return h_; }
  void synchronize() { check(gfx_ctx_synchronize(h_), h_); }

 private:
  gfx_ctx *h_ = nullptr;
};

// getfem::mesh built from points (3 per point) and convex point lists
class mesh {
 public:
  mesh(context &c, const std::string &geotrans, const std::vector<double> &pts,
       const std::vector<int32_t> &cvx)
      : ctx_(c.handle()) {
    const int nv = geotrans == "GT_QK(3,1)" ? 8 : 4;
    check(gfx_mesh_create(ctx_, geotrans.c_str(), (int64_t)pts.size() / 3, pts.data(),
                          (int64_t)cvx.size() / nv, cvx.data(), &h_),
          ctx_);
  }
  ~mesh() { if (h_) gfx_mesh_destroy(h_); }
  mesh(const mesh &) = delete;
  mesh &operator=(const mesh &) = delete;
  gfx_mesh *handle() const { return h_; }
  gfx_ctx *ctx() const { return ctx_; }

 private:
  gfx_ctx *ctx_;
  gfx_mesh *h_ = nullptr;
};

// getfem::mesh_fem: set_finite_element + set_qdim in one step
class mesh_fem {
 public:
  mesh_fem(const mesh &m, const std::string &fem, int qdim = 1) : ctx_(m.ctx()) {
    check(gfx_mesh_fem_create(m.handle(), fem.c_str(), qdim, &h_), ctx_);
  }
  ~mesh_fem() { if (h_) gfx_mesh_fem_destroy(h_); }
  mesh_fem(const mesh_fem &) = delete;
  mesh_fem &operator=(const mesh_fem &) = delete;
  gfx_mesh_fem *handle() const { return h_; }
  int64_t nb_dof() const { return gfx_mesh_fem_nb_dof(h_); }

 private:
  gfx_ctx *ctx_;
  gfx_mesh_fem *h_ = nullptr;
};

class mesh_im {
 public:
  mesh_im(const mesh &m, const std::string &im) : ctx_(m.ctx()) {
    check(gfx_mesh_im_create(m.handle(), im.c_str(), &h_), ctx_);
  }
  ~mesh_im() { if (h_) gfx_mesh_im_destroy(h_); }
  mesh_im(const mesh_im &) = delete;
  mesh_im &operator=(const mesh_im &) = delete;
  gfx_mesh_im *handle() const { return h_; }

 private:
  gfx_ctx *ctx_;
  gfx_mesh_im *h_ = nullptr;
};

// gmm::sub_interval(first, size)
struct sub_interval {
  int64_t first_, size_;
  sub_interval(int64_t f, int64_t s) : first_(f), size_(s) {}
  int64_t first() const { return first_; }
  int64_t size() const { return size_; }
};

struct csr_matrix {  // rows = Test_u, sorted columns = Test2_u
  int64_t n = 0;
  std::vector<int64_t> indptr;
  std::vector<int32_t> indices;
  std::vector<double> values;
};

struct csc_matrix {  // GetFEM's orientation: columns = trial dofs, sorted rows
  int64_t n = 0;
  std::vector<int64_t> colptr;
  std::vector<uint32_t> rows;
  std::vector<double> values;
};

class ga_workspace {
 public:
  explicit ga_workspace(context &c) : ctx_(c.handle()) { check(gfx_ws_create(ctx_, &h_), ctx_); }
  ~ga_workspace() { if (h_) gfx_ws_destroy(h_); }
  ga_workspace(const ga_workspace &) = delete;
  ga_workspace &operator=(const ga_workspace &) = delete;

  void add_fem_variable(const std::string &name, const mesh_fem &mf, const sub_interval &I,
                        const std::vector<double> &V) {
    check(gfx_ws_add_fem_variable(h_, name.c_str(), mf.handle(), I.first(), V.empty() ? nullptr : V.data()),
          ctx_);
  }
  void add_fem_constant(const std::string &name, const mesh_fem &mf, const std::vector<double> &V) {
    check(gfx_ws_add_fem_constant(h_, name.c_str(), mf.handle(), V.data()), ctx_);
  }
  void add_fixed_size_constant(const std::string &name, const std::vector<double> &V) {
    check(gfx_ws_add_fixed_size_constant(h_, name.c_str(), V.data(), (int64_t)V.size()), ctx_);
  }
  // region: -1 = all convexes (the only region supported)
  size_t add_expression(const std::string &expr, const mesh_im &mim, int region = -1,
                        size_t add_derivative_order = 2) {
    if (region != -1) throw gfx_error(GFX_E_UNSUPPORTED, "only the whole mesh region is supported");
    check(gfx_ws_add_expression(h_, expr.c_str(), mim.handle(), (int)add_derivative_order), ctx_);
    return 2;
  }
  void set_form(int form, int precision) { check(gfx_ws_set_form(h_, form, precision), ctx_); }
  void assembly(size_t order, bool condensation = false) {
    if (condensation) throw gfx_error(GFX_E_UNSUPPORTED, "condensation is not supported");
    check(gfx_ws_assembly(h_, (int)order), ctx_);
  }
  std::string describe() const {
    char buf[1024];
    check(gfx_ws_describe(h_, buf, sizeof buf), ctx_);
    return buf;
  }
  csr_matrix assembled_matrix_csr() const {
    gfx_plan *p = nullptr;
    check(gfx_ws_plan(h_, &p), ctx_);
    csr_matrix m;
    int64_t nnz = 0;
    check(gfx_plan_dims(p, &m.n, &nnz), ctx_);
    m.indptr.resize((size_t)m.n + 1);
    m.indices.resize((size_t)nnz);
    m.values.resize((size_t)nnz);
    check(gfx_plan_export_csr(p, m.indptr.data(), m.indices.data(), m.values.data()), ctx_);
    return m;
  }
  csc_matrix assembled_matrix() const {
    gfx_plan *p = nullptr;
    check(gfx_ws_plan(h_, &p), ctx_);
    csc_matrix m;
    int64_t nnz = 0;
    check(gfx_plan_dims(p, &m.n, &nnz), ctx_);
    m.colptr.resize((size_t)m.n + 1);
    m.rows.resize((size_t)nnz);
    m.values.resize((size_t)nnz);
    check(gfx_plan_export_csc(p, m.colptr.data(), m.rows.data(), m.values.data()), ctx_);
    return m;
  }
  // ga_workspace::assembled_vector() (.h:389-390), after assembly(1)
  std::vector<double> assembled_vector() const {
    gfx_plan *p = nullptr;
    check(gfx_ws_plan(h_, &p), ctx_);
    int64_t n = 0, nnz = 0;
    check(gfx_plan_dims(p, &n, &nnz), ctx_);
    std::vector<double> V((size_t)n);
    check(gfx_ws_assembled_vector(h_, V.data()), ctx_);
    return V;
  }
  void set_values(const std::string &name, const std::vector<double> &V) {
    check(gfx_ws_set_values(h_, name.c_str(), V.data()), ctx_);
  }

 private:
  gfx_ctx *ctx_;
  gfx_workspace *h_ = nullptr;
};

}  // namespace getfem_amd

#endif
