// Internal structures of libgfx: contexts, meshes, mesh_fems, integration
// methods and assembly plans, plus the launch wrappers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gfx.h"
#include "fem_ref.h"

// Lifetimes: every object counts the objects that point at it (a mesh its
// mesh_fems/mesh_ims, a context its meshes, plans, workspaces).  Destroying
// an object that is still referenced only marks it; it is released when the
// last referring object goes, so any destruction order is safe.
struct gfx_refcount {
  int refs = 0;
  bool closing = false;
};

struct gfx_ctx : gfx_refcount {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
};

struct gfx_mesh : gfx_refcount {
  gfx_ctx *ctx = nullptr;
  uint64_t version = 0;        // incremented by gfx_mesh_set_points
  int gt = 0;
  int nv = 4;
  int64_t n_pts = 0, n_cv = 0;
  double *d_pts = nullptr;     // n_pts*3
  int32_t *d_cvx = nullptr;    // n_cv*nv
  std::vector<double> h_pts;   // host copy (dof nodes, checks)
  std::vector<int32_t> h_cvx;
};

struct gfx_mesh_fem : gfx_refcount {
  gfx_mesh *mesh = nullptr;
  gfx::FemDesc fem;
  int qdim = 1;
  int64_t ndof = 0;
  std::vector<int32_t> h_dofs;  // n_cv*nd basic dofs
  int32_t *d_dofs = nullptr;
};

struct gfx_mesh_im : gfx_refcount {
  gfx_mesh *mesh = nullptr;
  gfx::IntegrationRule rule;
};

namespace gfx {

constexpr int kRecStride = 16;  // doubles per element record

// Per-element record written by the geometry kernel (affine tetrahedra):
// [0..8] B = K^{-T} row-major, [9] |J| a s, [10] |J| b s (s: scale of the
// stored reference tensor), [11] insertion threshold, [12] |J| a, [13] |J| b,
// [14] |J| c, [15] |J| c s; a, b, c are the term's coefficients (Laplace: a;
// elasticity: lambda, mu, mu; isotropic: alpha, beta, gamma).

// What a patch kernel needs first about its node, in one 16-byte load:
// the node's first CSR value, its incidence range and its slot count (bit 15:
// all rows full).
struct PatchDesc {
  int64_t rowbase;          // indptr[node * qdim]
  int32_t i0;               // inc_ptr[node]
  uint32_t ninc_ns;         // incidences (low 16 bits) | nslots (high 16 bits)
};

struct PatchArgs {
  int64_t n_nodes;
  const int32_t *order;     // node ids of this launch's bin
  const PatchDesc *desc;    // descriptors of the same nodes
  const int32_t *inc_ptr;   // n_nodes+1
  const uint32_t *inc;      // (element << 5) | local node
  const uint16_t *slot;     // per incidence, nd slots
  const uint16_t *nslots;   // per node
  const uint32_t *mask;     // presence bits, mwpi words per incidence
  int mwpi;
  const int64_t *indptr;    // CSR rows
  int nd, qdim, nq;
  const double *rec;        // element records
  const double *gref;       // nq*nd*3 reference gradients
  const double *qw;         // nq
  const double *R;          // nd*nd*9 weighted reference gradient products
  // the same in FP32, read by the FP32 contraction form
  const float *rec32, *gref32, *qw32, *R32;
  // the reference tensor as integers N with R = N / L (rk 1: int8 in 16-byte
  // blocks, 2: int16 in 32-byte blocks; 0: not available, R / R32 are used)
  const void *Rq;
  int rk;
  double *vals;             // CSR values
  int xcd_chunk;            // waves (patch kernel) or nodes (gather) per XCD in the block swizzle
  int max_inc;              // largest number of incidences of a node
  int npw;                  // gather kernel: node patches per workgroup (1..4)
  int lds_acc, lds_mask, lds_inc;  // gather kernel: dynamic LDS layout (set by its launcher)
  // patch kernel: desc holds n_waves groups of 4 descriptors (empty slots
  // zero); lds_acc and lds_flat are the largest accumulator (doubles) and
  // staging run (4-byte words) of any of the launch's waves
  int64_t n_waves;
  int lds_flat;
};

struct PatternArgs {
  int64_t n_cv, n_nodes;
  int nd, qdim, nloc, kwords, mwpi;
  const int32_t *dofs;
  const double *rec;
  const double *gref;
  const double *qw;
  int nq;
  double *thr;              // out: per element threshold
  uint32_t *keep;           // out: per element keep bits (row-major nloc x nloc)
  int32_t *inc_ptr;
  uint32_t *inc;
  uint16_t *slot;
  uint16_t *nslots;
  int32_t *nbr;
  uint32_t *mask;
  uint8_t *node_ok;         // pattern check: nodes confirmed full by their keep bits
  int64_t *rowcnt;          // per dof
  int64_t *indptr;
  int32_t *indices;
  int *err;                 // device error flag
};

struct EbArgs {
  int64_t n_cv;
  const double *pts;
  const int32_t *cvx;
  const int32_t *dofs;
  const double *U;          // state (ndof)
  const double *ca, *cb;    // per-element coefficients or null
  double ca_c, cb_c;
  const double *gN;         // nq*8*3 transformation gradients
  const double *gref;       // nq*nd*3 base gradients
  const double *qw;
  int nq;
  double *E;                // n_cv * nl * nl element matrices, node-blocked: K[(i,a),(j,b)] at
                            // ((3 i + a) 3 + b) nd + j (FP32 values in the FP32 form)
  double *res;              // n_cv * nl element residuals
  double *thr;              // n_cv thresholds
  int *err;
  uint32_t *keep;           // FP64 form: when set, the keep bits (row-major, kwords words per element)
  int kwords;
};

// Order-0 / order-1 polynomial programs (scalar.hip).  The symbol table of a
// thread holds, for every field f, its values at off..off+q-1 and gradients
// at off+q+3a+b, then the fixed-size constants from const_off.
constexpr int kScalarMaxFields = 4;
constexpr int kScalarMaxSyms = 56;
constexpr int kScalarMaxOut = 12;
struct ScalarField {
  const int32_t *dofs;      // n_cv * nd, dof of component 0 of each element node
  const double *U;          // the field's values
  const double *phi;        // nq * nd base values at the rule's points
  const double *gref;       // nq * nd * 3 reference gradients
  int nd, qdim, off;
};
struct ScalarArgs {
  int64_t n_cv;
  const double *pts;
  const int32_t *cvx;       // n_cv * nv vertices
  int nv;                   // 4: affine tetrahedra; 8: Q1 hexahedra
  const double *gN;         // hexahedra: nq * nv * 3 shape gradients at the rule's points
  int nq;
  const double *qw;
  int nf;
  ScalarField f[kScalarMaxFields];
  int nconst, const_off;
  const double *consts;
  int order;                // 0: potential, 1: derivative's element vectors
  int vf;                   // order 1: the field of the differentiated variable
  int nout, nm;
  const double *mcoef;      // program: nm monomials
  const int32_t *moff;      // nm + 1 offsets into msym
  const int32_t *msym;      // symbol table indices
  const int32_t *mout;      // output coefficient of each monomial
  double *elem;             // order 0: n_cv values; order 1: n_cv * nd * qdim
  int *err;
};

// Order-2 polynomial programs (polymat.hip): the element matrices of one
// block (test variable x trial variable), thresholded per element as
// ga_instruction_matrix_assembly_mf_mf does, then the kept entries of every
// block as (row, column) keys, sorted and summed in element order into a CSR.
constexpr int kPolyMaxOut = 144;       // 16 q q2 for q, q2 <= 3
constexpr int kPolyMaxNd = 32;         // base functions of one field (P3: 20, Q2: 27)
constexpr int kPolyMaxEntries = 4096;  // entries of one block's element matrix
struct PolyMatArgs {
  int64_t n_cv;
  const double *pts;
  const int32_t *cvx;        // n_cv * nv vertices
  int nv;                    // 4: affine tetrahedra; 8: Q1 hexahedra
  const double *gN;          // hexahedra: nq * nv * 3 shape gradients at the rule's points
  int nq;
  const double *qw;
  int nf;
  ScalarField f[kScalarMaxFields];
  int nconst, const_off;
  const double *consts;
  int vt, vs;                // fields of the test and of the trial variable
  int nout;
  const double *mcoef;       // monomials, sorted by output
  const int32_t *mptr;       // nout + 1: the monomials of output o
  const int32_t *moff;       // offsets into msym
  const int32_t *msym;
  double *elem;              // n_cv * nr * nc, row r = i qt + a, column c = j qs + b
  double *thr;               // n_cv: 1e-14 max|elem| (INFINITY: a zero element)
  int64_t *cnt;              // n_cv: entries above the threshold
  int *err;
};
struct PolyEmitArgs {
  int64_t n_cv;
  int ndt, nds, qt, qs;
  const int32_t *dt, *ds;    // dof tables of the two fields (component 0 of each node)
  int64_t first_t, first_s;  // intervals of the two variables
  int64_t ncols;             // key = row * ncols + column
  const double *elem, *thr;
  const int64_t *off;        // n_cv: first slot of each element's entries
  uint64_t *keys;
  double *vals;
};
// K = sum_v X_v grad(N_v) at point q (hexahedra) or the tetrahedron's edge
// matrix; J = det K and B = K^{-T} row-major; false where J = 0
__device__ __forceinline__ bool poly_geometry(const double *pts, const int32_t *cv, int nv, const double *gNq,
                                              double (&B)[9], double &aJ) {
  double k[9];
  if (!gNq) {
    const double *p0 = pts + 3 * (int64_t)cv[0];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) k[r + 3 * c] = pts[3 * (int64_t)cv[c + 1] + r] - p0[r];
  } else {
    for (int t = 0; t < 9; ++t) k[t] = 0.0;
    for (int v = 0; v < nv; ++v) {
      const double *x = pts + 3 * (int64_t)cv[v];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) k[r + 3 * c] += x[r] * gNq[3 * v + c];
    }
  }
  const double c0 = k[4] * k[8] - k[5] * k[7], c1 = k[5] * k[6] - k[3] * k[8], c2 = k[3] * k[7] - k[4] * k[6];
  const double J = k[0] * c0 + k[1] * c1 + k[2] * c2;
  if (J == 0.0) return false;
  B[0] = c0 / J; B[1] = (k[2] * k[7] - k[1] * k[8]) / J; B[2] = (k[1] * k[5] - k[2] * k[4]) / J;
  B[3] = c1 / J; B[4] = (k[0] * k[8] - k[2] * k[6]) / J; B[5] = (k[2] * k[3] - k[0] * k[5]) / J;
  B[6] = c2 / J; B[7] = (k[1] * k[6] - k[0] * k[7]) / J; B[8] = (k[0] * k[4] - k[1] * k[3]) / J;
  aJ = fabs(J);
  return true;
}
hipError_t launch_poly_mat(hipStream_t s, const PolyMatArgs &a);
hipError_t launch_poly_emit(hipStream_t s, const PolyEmitArgs &a);
// temporary storage of the stable key sort of n entries (bytes)
size_t poly_sort_bytes(int64_t n, int end_bit);
// sort (stable, by key) then sum equal keys in their order: n entries ->
// CSR with nrows rows (ip, ix, v: at least n entries); *d_nnz receives the
// entry count on the device.  work: 3 n + 1 + scan_tmp_elems(n) int64.
hipError_t poly_to_csr(hipStream_t s, void *tmp, size_t tmp_bytes, uint64_t *keys, double *vals, uint64_t *keys2,
                       double *vals2, int64_t n, int end_bit, int64_t nrows, int64_t ncols, int64_t *work,
                       int64_t *ip, int32_t *ix, double *v, int64_t *d_nnz);

// reference counting of the objects above (api.cpp)
void ref(gfx_ctx *c);
void unref(gfx_ctx *c);
void ref(gfx_mesh *m);
void unref(gfx_mesh *m);
void ref(gfx_mesh_fem *mf);
void unref(gfx_mesh_fem *mf);
void ref(gfx_mesh_im *mim);
void unref(gfx_mesh_im *mim);

// kernel launch wrappers (kernels.hip)
hipError_t launch_records_tet(hipStream_t s, int64_t n_cv, const double *pts,
                              const int32_t *cvx, const double *ca, double ca_c,
                              const double *cb, double cb_c, const double *cc, double cc_c,
                              const double *thr, double r_scale, double *rec, float *rec32,
                              int *err);
hipError_t launch_incidences(hipStream_t s, int64_t n_cv, int nd, int qdim,
                             const int32_t *dofs, int64_t n_nodes,
                             int32_t *inc_ptr, uint32_t *inc, int32_t *tmp);
hipError_t launch_keep(hipStream_t s, int term, const PatternArgs &a);
hipError_t launch_patch_build(hipStream_t s, const PatternArgs &a, int max_cand);
hipError_t launch_fill_columns(hipStream_t s, const PatternArgs &a);
// the pattern check against new keep bits: the thread-per-node pass (sets bit
// 16 of the error flag when some node needs the exact check), then the exact
// check (bit 4: the pattern changed)
hipError_t launch_mask_fast(hipStream_t s, const PatternArgs &a);
hipError_t launch_mask_exact(hipStream_t s, const PatternArgs &a, int max_nslots);
hipError_t launch_fill_desc(hipStream_t s, int64_t n, int qdim, const int32_t *order, const int32_t *inc_ptr,
                            const uint16_t *nslots, const int64_t *indptr, PatchDesc *desc);
hipError_t launch_assemble_patch(hipStream_t s, int term, int form, int prec, const PatchArgs &a);
hipError_t scan_exclusive_i64(hipStream_t s, const int64_t *in, int64_t *out,
                              int64_t n, int64_t *tmp);
hipError_t scan_exclusive_i32(hipStream_t s, const int32_t *in, int32_t *out,
                              int64_t n, int32_t *tmp);
int64_t scan_tmp_elems(int64_t n);

// kernels_eb.hip
hipError_t launch_nh_element(hipStream_t s, const EbArgs &a, int precision);
hipError_t launch_elast_bdb_mfma(hipStream_t s, int64_t n_cv, int nd, int nq, const double *rec,
                                 const double *gref, const double *qw, float *E);
hipError_t launch_elast_bdb_mfma_hp(hipStream_t s, int64_t n_cv, int nd, int nq, const double *rec,
                                    const double *gref, const double *qw, double *E64, float *E32);
hipError_t launch_gather_emat(hipStream_t s, const PatchArgs &a, const void *E, bool e_fp32, const double *thr,
                              int acc_cap, int group_inc);
hipError_t launch_gather_vec(hipStream_t s, int64_t n_nodes, int nd, int Q, const int32_t *inc_ptr,
                             const uint32_t *inc, const double *res, double *V);

// scalar.hip
hipError_t launch_scalar_elem(hipStream_t s, const ScalarArgs &a);
hipError_t launch_scalar_sum(hipStream_t s, int64_t n, const double *x, double *part, double *out);
int64_t scalar_partials();

// spmv.hip
hipError_t launch_spmv(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr,
                       const int32_t *indices, const double *vals, const double *x, double *y,
                       bool acc = false);
hipError_t launch_add_into(hipStream_t s, bool csc, int64_t n, const int64_t *ip, const int32_t *ix,
                           const double *v, const int64_t *tp, const int32_t *ti, double *tv, int *err);
hipError_t launch_union_count(hipStream_t s, int64_t n, const int64_t *ipa, const int32_t *ixa,
                              const int64_t *ipb, const int32_t *ixb, int64_t *cnt);
hipError_t launch_union_fill(hipStream_t s, int64_t n, const int64_t *ipa, const int32_t *ixa,
                             const int64_t *ipb, const int32_t *ixb, const int64_t *ipu, int32_t *ixu,
                             int64_t *mapa, int64_t *mapb);
hipError_t launch_scatter_add(hipStream_t s, int64_t nnz, const int64_t *map, const double *v, double *out);
hipError_t launch_gather_i64(hipStream_t s, int64_t n, const int64_t *map, const int64_t *src, int64_t *dst);
hipError_t launch_fill_f64(hipStream_t s, int64_t n, double *p, double v);
hipError_t launch_dof_constraints(hipStream_t s, int64_t n, const int64_t *indptr, const int32_t *indices,
                                  double *vals, const uint8_t *fixed, const double *go, const double *pr,
                                  double *rhs, int flags, int *err);
size_t cg_workspace_doubles(int64_t n);
size_t gmres_workspace_doubles(int64_t n, int m);
struct IluData;
hipError_t gmres_solve(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                       const double *vals, const double *b, double *x, double rtol, int maxit, int m,
                       double *work, const IluData *ilu, int *iters, double *rel_res);
hipError_t cg_solve(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                    const double *vals, const double *b, double *x, double rtol, int maxit, int qb, double *work,
                    int *iters, double *rel_res);

// adjoint.hip
int64_t compliance_nblocks(int64_t n_cv);
// the allocation-free compliance entry point's static scratch (adjoint.hip):
// its size, and the address of the given device's copy (looked up once per
// device with that device current; later calls return the kept address)
constexpr size_t kComplianceScratchBytes = size_t(64) << 20;
hipError_t compliance_scratch(int device, double **ptr);
hipError_t launch_compliance(hipStream_t s, int64_t n_cv, const double *X, const int32_t *cvx,
                             const int32_t *dofs, const double *u, const double *rho, double lam,
                             double mu, double emin, double penal, double wsum, double *Jpart,
                             double *J, double *drho, double *cX, double *cU, int *err);
hipError_t launch_compliance_finish(hipStream_t s, int64_t n_cv, int64_t nx, const int32_t *xptr,
                                    const uint32_t *xinc, const double *cX, double *outX, int64_t nu,
                                    const int32_t *uptr, const uint32_t *uinc, const double *cU,
                                    double *outU, const double *Jpart, double *J);

}  // namespace gfx

struct gfx_compliance {
  gfx_ctx *ctx = nullptr;
  gfx_mesh_fem *mf = nullptr;
  double lam = 0, mu = 0, emin = 0, penal = 3, wsum = 0;
  int64_t n_cv = 0, n_pts = 0, ndof = 0;
  int32_t *d_vinc_ptr = nullptr, *d_dinc_ptr = nullptr;  // point / dof-node incidences
  uint32_t *d_vinc = nullptr, *d_dinc = nullptr;
  double *d_cX = nullptr, *d_cU = nullptr, *d_Jpart = nullptr;
  int *d_err = nullptr;
};

namespace gfx {
// Incomplete LDL^T of a plan's matrix (ildlt.hip): diagonal positions, level
// sets of units of qb rows (forward: host offsets flev into d_flist;
// backward: blev into d_blist), and the factor F aligned with the CSR.
struct IldltData {
  int64_t n = 0;
  int qb = 1;
  int cap = 0;                    // longest row
  int64_t pattern_id = -1;        // pattern the analysis belongs to
  bool factored = false;
  int64_t *d_dpos = nullptr;
  int32_t *d_fdeg = nullptr, *d_bdeg = nullptr, *d_flist = nullptr, *d_blist = nullptr, *d_counter = nullptr;
  double *d_F = nullptr;          // nnz
  double *d_work = nullptr;       // pivot-check block maxima
  double *d_cg = nullptr;         // conjugate gradient vectors
  std::vector<int64_t> flev, blev;
};
hipError_t ildlt_analyse(hipStream_t s, IldltData &L, int64_t n, int qb, const int64_t *ip, const int32_t *ix,
                         int *d_err, int *h_err);
hipError_t ildlt_factor(hipStream_t s, IldltData &L, const int64_t *ip, const int32_t *ix, const double *av,
                        int cap, int *d_err, int *h_err);
hipError_t ildlt_apply(hipStream_t s, const IldltData &L, const int64_t *ip, const int32_t *ix, const double *r,
                       double *z, double *y, double *w);
size_t ildlt_cg_doubles(int64_t n);
hipError_t ildlt_pcg(hipStream_t s, const IldltData &L, int64_t nnz, const int64_t *ip, const int32_t *ix,
                     const double *av, const double *b, double *x, double rtol, int maxit, double *work,
                     int *iters, double *rel_res);
// gmm::ilu_precond of GetFEM's column-major matrix (ildlt.hip): the
// analysis of the ILDLT (qb = 1: diagonal positions, level sets) plus the
// transpose positions, the factor and a work vector.
struct IluData {
  IldltData lv;
  int64_t *d_T = nullptr;         // nnz: position of (j, i) for the entry (i, j)
  int64_t *d_rows = nullptr;      // n: rows whose pivot gmm replaces by 1
  double *d_F = nullptr;          // nnz
  double *d_y = nullptr;          // n
  int64_t pattern_nnz = -1;
};
hipError_t ilu_factor(hipStream_t s, IluData &U, const int64_t *ip, const int32_t *ix, const double *av,
                      int *d_err, int *h_err, int64_t *replaced);
hipError_t ilu_apply(hipStream_t s, const IluData &U, const int64_t *ip, const int32_t *ix, const double *r,
                     double *z);
}  // namespace gfx

struct gfx_plan {
  gfx_ctx *ctx = nullptr;
  gfx_mesh_fem *mf = nullptr;
  gfx_mesh_im *mim = nullptr;
  int term = GFX_TERM_LAPLACE;
  int nd = 0, qdim = 1, nloc = 0, nq = 0, nv = 4;
  int kwords = 0, mwpi = 0;
  gfx_coef coef[3];
  double coef_min[3] = {0.0, 0.0, 0.0};  // smallest value of each coefficient (host copy)
  double *d_coef[3] = {nullptr, nullptr, nullptr};
  // The stored pattern is GetFEM's for the values it was built from.  After
  // the coefficients, the node positions (mesh version) or the state change,
  // the next assembly re-derives the thresholds and keep bits and rebuilds
  // the pattern if its presence bits differ (GetFEM thresholds every element
  // matrix at every assembly, compile_and_exec.cc:5441-5453).
  uint64_t pts_version = 0;  // mesh version the thresholds were derived from
  bool values_changed = false;
  bool pinned = false;       // callers hold the CSR arrays: a rebuild is refused
  int64_t pattern_id = 0;    // incremented by every pattern (re)build
  double *d_gref = nullptr, *d_qw = nullptr, *d_R = nullptr;
  double *d_rec = nullptr, *d_thr = nullptr;
  // FP32 contraction form: the tables and the element records in FP32
  float *d_gref32 = nullptr, *d_qw32 = nullptr, *d_R32 = nullptr, *d_rec32 = nullptr;
  float *d_E32 = nullptr;  // element matrices of the bf16x2 / FP32 matrix-core forms (allocated on first use)
  double *d_E64 = nullptr; // element matrices of the FP64 matrix-core form (allocated on first use)
  // R = N / r_L with small integers N (exact quadrature of polynomials): rk 1
  // int8 / 2 int16 blocks in d_Rq, 0 if R has no such form
  int rk = 0;
  double r_L = 1.0;
  uint8_t *d_Rq = nullptr;
  uint32_t *d_keep = nullptr;
  int64_t n_nodes = 0, n_inc = 0;
  int32_t *d_inc_ptr = nullptr;
  uint32_t *d_inc = nullptr;
  uint16_t *d_slot = nullptr, *d_nslots = nullptr;
  uint8_t *d_node_ok = nullptr;   // pattern check scratch (n_nodes)
  int32_t *d_nbr = nullptr;
  uint32_t *d_mask = nullptr;
  int max_nslots = 0, max_cand = 0, max_inc = 0;
  // element-batched plans: nodes grouped by patch size so each gather launch
  // reserves LDS for its own largest patch (d_order holds the bins back to
  // back); patch plans: 4 node ids per wave (-1: empty slot)
  int32_t *d_order = nullptr;
  gfx::PatchDesc *d_desc = nullptr;  // per position of d_order
  // gather bins: node positions [off, off+count), largest nslots cap, npw
  // patches per workgroup; patch launches: wave groups [off, off+count) of
  // d_desc / d_order (4 entries each), LDS caps acc_cap doubles + flat_cap words
  struct Bin { int64_t off, count; int cap, npw; int acc_cap = 0, flat_cap = 0, inc_max = 0; };
  std::vector<Bin> bins;
  int64_t n_order = 0;       // entries of d_order / d_desc
  int64_t ndof = 0, nnz = -1;
  int64_t *d_indptr = nullptr;
  int32_t *d_indices = nullptr;
  double *d_vals = nullptr;
  int *d_err = nullptr;
  int form = GFX_FORM_FACTORIZED;
  int precision = GFX_PREC_FP64;
  // element-batched plans (non-affine / nonlinear): buffered element matrices
  bool eb = false;
  double *d_gN = nullptr, *d_E = nullptr, *d_res = nullptr, *d_U = nullptr, *d_rhs = nullptr;
  double *d_cg = nullptr;  // conjugate gradient work vectors (allocated on first solve)
  gfx::IldltData il;       // incomplete LDL^T preconditioner (built on first use)
  uint8_t *d_fixed = nullptr;  // dof constraints: mask and prescribed values
  double *d_go = nullptr;
  std::vector<hipStream_t> aux;  // streams the patch-size bins run on beside ctx->stream
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  bool timing = false;
  std::vector<hipEvent_t> ev;  // 3 events per timed call
  int n_timed = 0;
};
