// Sparse operations on the assembled CSR (rows = test dofs, sorted columns):
// the product y = K x, dof constraints (Dirichlet by simplification) and a
// Jacobi-preconditioned conjugate gradient solve.
//
// SpMV: G lanes of a wavefront share one row; each lane walks the row with
// stride G, then the G partial sums are combined by shuffles.  G is picked
// from the mean row length so that short rows (P1) do not leave most lanes
// idle.
//
// CG: every scalar of the recurrence (alpha, beta, r.z, r.r) lives in device
// memory and is produced by fixed-order block reductions, so an iteration
// issues kernels only and the result does not depend on scheduling.  The host
// reads r.r once per iteration for the stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gfx_internal.h"

namespace gfx {
namespace {

constexpr int kSpmvThreads = 256;
constexpr int kRedThreads = 256;   // threads of a reduction block
constexpr int kRedBlocks = 1024;   // partial sums per dot product (fixed: deterministic order)

template <int G, bool ACC>
__global__ __launch_bounds__(kSpmvThreads) void k_spmv(int64_t n, const int64_t *__restrict__ indptr,
                                                        const int32_t *__restrict__ indices,
                                                        const double *__restrict__ vals,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ y) {
  const int64_t row = ((int64_t)blockIdx.x * kSpmvThreads + threadIdx.x) / G;
  const int lane = threadIdx.x % G;
  double s = 0.0;
  if (row < n) {
    const int64_t b = indptr[row], e = indptr[row + 1];
    for (int64_t k = b + lane; k < e; k += G) s += vals[k] * x[indices[k]];
  }
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
  if (row < n && lane == 0) y[row] = ACC ? y[row] + s : s;
}

template <int G>
hipError_t spmv_g(hipStream_t s, int64_t n, const int64_t *indptr, const int32_t *indices,
                  const double *vals, const double *x, double *y, bool acc) {
  const int64_t threads = n * G;
  const unsigned grid = (unsigned)((threads + kSpmvThreads - 1) / kSpmvThreads);
  if (acc) k_spmv<G, true><<<grid, kSpmvThreads, 0, s>>>(n, indptr, indices, vals, x, y);
  else k_spmv<G, false><<<grid, kSpmvThreads, 0, s>>>(n, indptr, indices, vals, x, y);
  return hipGetLastError();
}

// ------------------------------------------- adding into another matrix
// The plan's CSR (rows = test dofs) added into a caller matrix whose pattern
// holds it.  CSR target: one thread per row walks both sorted rows (the
// target's row is a superset).  CSC target (GetFEM's col_matrix orientation:
// column = trial dof, sorted rows = test dofs): each entry (r, c) is found by
// a binary search in the target's column c.  An entry the target lacks sets
// bit 8 of the error flag (the value is dropped).
__global__ void k_add_into_csr(int64_t n, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                               const double *__restrict__ v, const int64_t *__restrict__ tip,
                               const int32_t *__restrict__ tix, double *__restrict__ tv, int *__restrict__ err) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int64_t t = tip[row];
  const int64_t te = tip[row + 1];
  for (int64_t k = ip[row]; k < ip[row + 1]; ++k) {
    const int32_t c = ix[k];
    while (t < te && tix[t] < c) ++t;
    if (t < te && tix[t] == c) tv[t] += v[k];
    else atomicOr(err, 8);
  }
}

__global__ void k_add_into_csc(int64_t n, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                               const double *__restrict__ v, const int64_t *__restrict__ tcp,
                               const int32_t *__restrict__ tri, double *__restrict__ tv, int *__restrict__ err) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  for (int64_t k = ip[row]; k < ip[row + 1]; ++k) {
    const int32_t c = ix[k];
    int64_t lo = tcp[c], hi = tcp[c + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (tri[mid] < row) lo = mid + 1;
      else hi = mid;
    }
    if (lo < tcp[c + 1] && tri[lo] == row) tv[lo] += v[k];
    else atomicOr(err, 8);
  }
}

// union of two CSR patterns with the same rows: per-row merge counts, then
// the merged columns and, for every entry of either input, its position in
// the union
__global__ void k_union_count(int64_t n, const int64_t *__restrict__ ipa, const int32_t *__restrict__ ixa,
                              const int64_t *__restrict__ ipb, const int32_t *__restrict__ ixb,
                              int64_t *__restrict__ cnt) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int64_t a = ipa[row], ae = ipa[row + 1], b = ipb[row], be = ipb[row + 1], m = 0;
  while (a < ae || b < be) {
    const int32_t ca = a < ae ? ixa[a] : INT32_MAX, cb = b < be ? ixb[b] : INT32_MAX;
    a += ca <= cb;
    b += cb <= ca;
    ++m;
  }
  cnt[row] = m;
}

__global__ void k_union_fill(int64_t n, const int64_t *__restrict__ ipa, const int32_t *__restrict__ ixa,
                             const int64_t *__restrict__ ipb, const int32_t *__restrict__ ixb,
                             const int64_t *__restrict__ ipu, int32_t *__restrict__ ixu,
                             int64_t *__restrict__ mapa, int64_t *__restrict__ mapb) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int64_t a = ipa[row], ae = ipa[row + 1], b = ipb[row], be = ipb[row + 1], u = ipu[row];
  while (a < ae || b < be) {
    const int32_t ca = a < ae ? ixa[a] : INT32_MAX, cb = b < be ? ixb[b] : INT32_MAX;
    const int32_t c = ca < cb ? ca : cb;
    ixu[u] = c;
    if (ca == c) mapa[a++] = u;
    if (cb == c) mapb[b++] = u;
    ++u;
  }
}

// out[map[k]] += v[k] (map injective: no two entries share a target)
__global__ void k_scatter_add(int64_t nnz, const int64_t *__restrict__ map, const double *__restrict__ v,
                              double *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) out[map[k]] += v[k];
}

// dst[k] = src[map[k]] (composition of position maps)
__global__ void k_gather_i64(int64_t n, const int64_t *__restrict__ map, const int64_t *__restrict__ src,
                             int64_t *__restrict__ dst) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[map[k]];
}

__global__ void k_fill_f64(int64_t n, double *__restrict__ p, double v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// ---------------------------------------------------------------- CG
struct CgScal {
  double rz, rr, bb, pq, alpha, beta;
};

__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// inverse diagonal (0 where the diagonal is absent or zero: that dof is left alone)
__global__ void k_inv_diag(int64_t n, const int64_t *__restrict__ indptr,
                           const int32_t *__restrict__ indices, const double *__restrict__ vals,
                           double *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t lo = indptr[i], hi = indptr[i + 1];
  double d = 0.0;
  while (lo < hi) {  // sorted columns: binary search for column i
    const int64_t mid = (lo + hi) / 2;
    const int32_t c = indices[mid];
    if (c == i) { d = vals[mid]; break; }
    if (c < i) lo = mid + 1;
    else hi = mid;
  }
  dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
}

// 3x3 block Jacobi for vector fields (dof = 3 node + component): the inverse
// of each node's diagonal block; a singular block (e.g. rows cleared by a
// Dirichlet simplification) falls back to the inverse diagonal of that node
__global__ void k_inv_block3(int64_t nn, const int64_t *__restrict__ indptr,
                             const int32_t *__restrict__ indices, const double *__restrict__ vals,
                             double *__restrict__ minv) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nn) return;
  double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < 3; ++a) {
    const int64_t row = 3 * u + a;
    int64_t lo = indptr[row], hi = indptr[row + 1];
    while (lo < hi) {  // first column >= 3u (sorted columns)
      const int64_t mid = (lo + hi) / 2;
      if (indices[mid] < 3 * u) lo = mid + 1;
      else hi = mid;
    }
    for (int64_t k = lo; k < indptr[row + 1] && indices[k] < 3 * u + 3; ++k) M[a][indices[k] - 3 * u] = vals[k];
  }
  const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1], c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2],
               c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
  double *o = minv + 9 * u;
  const double scale = fabs(M[0][0]) + fabs(M[1][1]) + fabs(M[2][2]);
  if (det != 0.0 && fabs(det) > 1e-300 && fabs(det) > 1e-14 * scale * scale * scale) {
    const double id = 1.0 / det;
    o[0] = c00 * id; o[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
    o[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
    o[3] = c01 * id; o[4] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
    o[5] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
    o[6] = c02 * id; o[7] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
    o[8] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
  } else {
    for (int t = 0; t < 9; ++t) o[t] = 0.0;
    for (int a = 0; a < 3; ++a) o[4 * a] = M[a][a] != 0.0 ? 1.0 / M[a][a] : 0.0;
  }
}

// z = M^-1 r on one unit: a dof (QB = 1, inverse diagonal) or a node's three
// dofs (QB = 3, inverse 3x3 block)
template <int QB>
__device__ __forceinline__ void precond(const double *__restrict__ minv, int64_t u, const double (&r)[QB],
                                        double (&z)[QB]) {
  if constexpr (QB == 1) {
    z[0] = minv[u] * r[0];
  } else {
    const double *M = minv + 9 * u;
    for (int a = 0; a < 3; ++a) z[a] = M[3 * a] * r[0] + M[3 * a + 1] * r[1] + M[3 * a + 2] * r[2];
  }
}

// r = b - r (r holds K x on entry), z = M^-1 r, p = z; partials of r.z, r.r, b.b
template <int QB>
__global__ __launch_bounds__(kRedThreads) void k_cg_start(int64_t nu, const double *__restrict__ b,
                                                          const double *__restrict__ minv,
                                                          double *__restrict__ r, double *__restrict__ z,
                                                          double *__restrict__ p, double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  double rz = 0, rr = 0, bb = 0;
  for (int64_t u = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; u < nu; u += (int64_t)kRedBlocks * kRedThreads) {
    double ri[QB], zi[QB];
    for (int a = 0; a < QB; ++a) ri[a] = b[QB * u + a] - r[QB * u + a];
    precond<QB>(minv, u, ri, zi);
    for (int a = 0; a < QB; ++a) {
      const int64_t i = QB * u + a;
      r[i] = ri[a]; z[i] = zi[a]; p[i] = zi[a];
      rz += ri[a] * zi[a]; rr += ri[a] * ri[a]; bb += b[i] * b[i];
    }
  }
  rz = block_sum(rz, sh); rr = block_sum(rr, sh); bb = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz; part[kRedBlocks + blockIdx.x] = rr; part[2 * kRedBlocks + blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_dot(int64_t n, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kRedBlocks * kRedThreads)
    s += x[i] * y[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sum `count` partial arrays of kRedBlocks entries in a fixed order;
// mode 0: start (rz, rr, bb), 1: pq -> alpha, 2: (rz_new, rr) -> beta
__global__ __launch_bounds__(kRedThreads) void k_cg_scalars(const double *__restrict__ part, int mode,
                                                            CgScal *__restrict__ sc) {
  __shared__ double sh[kRedThreads];
  double v[3] = {0, 0, 0};
  const int count = mode == 1 ? 1 : (mode == 0 ? 3 : 2);
  for (int k = 0; k < count; ++k) {
    double s = 0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) s += part[k * kRedBlocks + i];
    v[k] = block_sum(s, sh);
  }
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    sc->rz = v[0]; sc->rr = v[1]; sc->bb = v[2];
  } else if (mode == 1) {
    sc->pq = v[0];
    sc->alpha = v[0] != 0.0 ? sc->rz / v[0] : 0.0;
  } else {
    sc->beta = sc->rz != 0.0 ? v[0] / sc->rz : 0.0;
    sc->rz = v[0]; sc->rr = v[1];
  }
}

// x += alpha p, r -= alpha q, z = M^-1 r; partials of r.z and r.r
template <int QB>
__global__ __launch_bounds__(kRedThreads) void k_cg_update(int64_t nu, const CgScal *__restrict__ sc,
                                                           const double *__restrict__ p,
                                                           const double *__restrict__ q,
                                                           const double *__restrict__ minv,
                                                           double *__restrict__ x, double *__restrict__ r,
                                                           double *__restrict__ z, double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  const double alpha = sc->alpha;
  double rz = 0, rr = 0;
  for (int64_t u = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; u < nu; u += (int64_t)kRedBlocks * kRedThreads) {
    double ri[QB], zi[QB];
    for (int a = 0; a < QB; ++a) {
      const int64_t i = QB * u + a;
      x[i] += alpha * p[i];
      ri[a] = r[i] - alpha * q[i];
    }
    precond<QB>(minv, u, ri, zi);
    for (int a = 0; a < QB; ++a) {
      const int64_t i = QB * u + a;
      r[i] = ri[a]; z[i] = zi[a];
      rz += ri[a] * zi[a]; rr += ri[a] * ri[a];
    }
  }
  rz = block_sum(rz, sh); rr = block_sum(rr, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = rz; part[kRedBlocks + blockIdx.x] = rr; }
}

// p = z + beta p
__global__ void k_cg_direction(int64_t n, const CgScal *__restrict__ sc, const double *__restrict__ z,
                               double *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + sc->beta * p[i];
}

unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------- dof constraints
// fixed[i] = 1 for a constrained dof, go[i] its prescribed value.  One
// thread per row; the row's entries are read before any is cleared, so the
// symmetric linear right-hand side correction uses the assembled K.
__global__ void k_dof_constraints(int64_t n, const int64_t *__restrict__ indptr,
                                  const int32_t *__restrict__ indices, double *__restrict__ vals,
                                  const uint8_t *__restrict__ fixed, const double *__restrict__ go,
                                  const double *__restrict__ pr, double *__restrict__ rhs, int flags,
                                  int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool mat = flags & GFX_BUILD_MATRIX, vec = (flags & GFX_BUILD_RHS) && rhs;
  const bool lin = flags & GFX_MODEL_LINEAR, sym = flags & GFX_MODEL_SYMMETRIC;
  const int64_t b = indptr[i], e = indptr[i + 1];
  if (fixed[i]) {
    if (vec) rhs[i] = lin ? go[i] : go[i] - (pr ? pr[i] : 0.0);
    if (mat) {
      bool diag = false;
      for (int64_t k = b; k < e; ++k) {
        const bool d = indices[k] == i;
        vals[k] = d ? 1.0 : 0.0;
        diag |= d;
      }
      if (!diag) atomicOr(err, 4);
    }
    return;
  }
  if (vec && lin && sym) {  // rhs -= K(:, fixed) go
    double s = 0.0;
    for (int64_t k = b; k < e; ++k)
      if (fixed[indices[k]]) s += vals[k] * go[indices[k]];
    rhs[i] -= s;
  }
  if (mat && sym)
    for (int64_t k = b; k < e; ++k)
      if (fixed[indices[k]]) vals[k] = 0.0;
}

}  // namespace

hipError_t launch_dof_constraints(hipStream_t s, int64_t n, const int64_t *indptr, const int32_t *indices,
                                  double *vals, const uint8_t *fixed, const double *go, const double *pr,
                                  double *rhs, int flags, int *err) {
  if (n == 0) return hipSuccess;
  k_dof_constraints<<<nb256(n), 256, 0, s>>>(n, indptr, indices, vals, fixed, go, pr, rhs, flags, err);
  return hipGetLastError();
}

hipError_t launch_spmv(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr,
                       const int32_t *indices, const double *vals, const double *x, double *y, bool acc) {
  if (n == 0) return hipSuccess;
  const double mean = (double)nnz / (double)n;
  if (mean <= 12) return spmv_g<4>(s, n, indptr, indices, vals, x, y, acc);
  if (mean <= 24) return spmv_g<8>(s, n, indptr, indices, vals, x, y, acc);
  if (mean <= 64) return spmv_g<16>(s, n, indptr, indices, vals, x, y, acc);
  return spmv_g<32>(s, n, indptr, indices, vals, x, y, acc);
}

hipError_t launch_add_into(hipStream_t s, bool csc, int64_t n, const int64_t *ip, const int32_t *ix,
                           const double *v, const int64_t *tp, const int32_t *ti, double *tv, int *err) {
  if (n == 0) return hipSuccess;
  if (csc) k_add_into_csc<<<nb256(n), 256, 0, s>>>(n, ip, ix, v, tp, ti, tv, err);
  else k_add_into_csr<<<nb256(n), 256, 0, s>>>(n, ip, ix, v, tp, ti, tv, err);
  return hipGetLastError();
}

hipError_t launch_union_count(hipStream_t s, int64_t n, const int64_t *ipa, const int32_t *ixa,
                              const int64_t *ipb, const int32_t *ixb, int64_t *cnt) {
  if (n == 0) return hipSuccess;
  k_union_count<<<nb256(n), 256, 0, s>>>(n, ipa, ixa, ipb, ixb, cnt);
  return hipGetLastError();
}

hipError_t launch_union_fill(hipStream_t s, int64_t n, const int64_t *ipa, const int32_t *ixa,
                             const int64_t *ipb, const int32_t *ixb, const int64_t *ipu, int32_t *ixu,
                             int64_t *mapa, int64_t *mapb) {
  if (n == 0) return hipSuccess;
  k_union_fill<<<nb256(n), 256, 0, s>>>(n, ipa, ixa, ipb, ixb, ipu, ixu, mapa, mapb);
  return hipGetLastError();
}

hipError_t launch_scatter_add(hipStream_t s, int64_t nnz, const int64_t *map, const double *v, double *out) {
  if (nnz == 0) return hipSuccess;
  k_scatter_add<<<nb256(nnz), 256, 0, s>>>(nnz, map, v, out);
  return hipGetLastError();
}

hipError_t launch_gather_i64(hipStream_t s, int64_t n, const int64_t *map, const int64_t *src, int64_t *dst) {
  if (n == 0) return hipSuccess;
  k_gather_i64<<<nb256(n), 256, 0, s>>>(n, map, src, dst);
  return hipGetLastError();
}

hipError_t launch_fill_f64(hipStream_t s, int64_t n, double *p, double v) {
  if (n == 0) return hipSuccess;
  k_fill_f64<<<nb256(n), 256, 0, s>>>(n, p, v);
  return hipGetLastError();
}

// preconditioner storage (3 n: room for the 3x3 blocks) + r, z, p, q + partials
size_t cg_workspace_doubles(int64_t n) { return 7 * (size_t)n + 3 * kRedBlocks + 8; }

// ------------------------------------------------------------------ GMRES
namespace {

constexpr int kGmBlocks = 256;  // partial sums per dot product (fixed order)
constexpr int kGmRows = 8;      // basis vectors per block of the multi-dot

// a row without diagonal (or a zero one) is left unscaled
__global__ void k_gm_fix_dinv(int64_t n, double *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && dinv[i] == 0.0) dinv[i] = 1.0;
}

// y = dinv x
__global__ void k_gm_scale_diag(int64_t n, const double *__restrict__ dinv, const double *x, double *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = dinv[i] * x[i];
}

// w = b - w
__global__ void k_gm_sub(int64_t n, const double *__restrict__ b, double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = b[i] - w[i];
}

// partials of V_k . w for the basis vectors k0 .. k0 + kGmRows - 1 (k < nk)
__global__ __launch_bounds__(kRedThreads) void k_gm_multidot(int64_t n, int nk, const double *__restrict__ V,
                                                             const double *__restrict__ w,
                                                             double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  const int k0 = blockIdx.y * kGmRows;
  double acc[kGmRows];
  for (int r = 0; r < kGmRows; ++r) acc[r] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kGmBlocks * kRedThreads) {
    const double wi = w[i];
    for (int r = 0; r < kGmRows; ++r)
      if (k0 + r < nk) acc[r] += V[(int64_t)(k0 + r) * n + i] * wi;
  }
  for (int r = 0; r < kGmRows; ++r) {
    const double t = block_sum(acc[r], sh);  // uniform: every thread takes part
    if (threadIdx.x == 0 && k0 + r < nk) part[(int64_t)(k0 + r) * kGmBlocks + blockIdx.x] = t;
  }
}

// h_k (+)= sum of the partials of k, in a fixed order
__global__ __launch_bounds__(kRedThreads) void k_gm_sum(int nk, const double *__restrict__ part,
                                                        double *__restrict__ h, int accumulate) {
  __shared__ double sh[kRedThreads];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < kGmBlocks; i += kRedThreads) s += part[(int64_t)k * kGmBlocks + i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && k < nk) h[k] = accumulate ? h[k] + s : s;
}

// w -= sum_k h_k V_k
__global__ void k_gm_project(int64_t n, int nk, const double *__restrict__ V, const double *__restrict__ h,
                             double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = w[i];
  for (int k = 0; k < nk; ++k) t -= h[k] * V[(int64_t)k * n + i];
  w[i] = t;
}

// out = w * (1 / a)  with a = *anorm (0 leaves zeros)
__global__ void k_gm_normalize(int64_t n, const double *__restrict__ anorm, const double *__restrict__ w,
                               double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = *anorm;
  if (i < n) out[i] = a != 0.0 ? w[i] * (1.0 / a) : 0.0;
}

// x += sum_k y_k V_k
__global__ void k_gm_combine(int64_t n, int nk, const double *__restrict__ V, const double *__restrict__ y,
                             double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t = x[i];
  for (int k = 0; k < nk; ++k) t += y[k] * V[(int64_t)k * n + i];
  x[i] = t;
}

// ||w||^2 partials then the norm into *out
__global__ __launch_bounds__(kRedThreads) void k_gm_norm(int64_t n, const double *__restrict__ w,
                                                         double *__restrict__ part) {
  __shared__ double sh[kRedThreads];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (int64_t)kGmBlocks * kRedThreads)
    s += w[i] * w[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(kRedThreads) void k_gm_norm_final(const double *__restrict__ part,
                                                               double *__restrict__ out) {
  __shared__ double sh[kRedThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < kGmBlocks; i += kRedThreads) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// Givens rotation zeroing b in (a, b), as gmm::Givens_rotation
void givens(double a, double b, double &c, double &s) {
  if (b == 0.0) { c = 1.0; s = 0.0; return; }
  if (std::fabs(b) > std::fabs(a)) {
    const double t = a / b;
    s = 1.0 / std::sqrt(1.0 + t * t);
    c = s * t;
  } else {
    const double t = b / a;
    c = 1.0 / std::sqrt(1.0 + t * t);
    s = c * t;
  }
}
void rotate(double &x, double &y, double c, double s) {
  const double t = c * x + s * y;
  y = c * y - s * x;
  x = t;
}

}  // namespace

size_t gmres_workspace_doubles(int64_t n, int m) {
  return (size_t)(m + 1) * (size_t)n + 3 * (size_t)n + (size_t)(m + 2) * kGmBlocks + 2 * (size_t)(m + 2) + 8;
}

// gmm::gmres (src/gmm/gmm_solver_gmres.h:85-157): restarted, left
// preconditioned (gmm::ilu_precond, or the inverse diagonal), the stopping test on the
// preconditioned residual relative to ||M b||, a restart that makes no
// progress ten times in a row ends the solve.  The basis is orthogonalised by
// classical Gram-Schmidt applied twice (two batched passes per step, where
// gmm runs modified Gram-Schmidt one vector at a time): the same subspace,
// with the same orthogonality to working precision.  The Hessenberg matrix,
// its Givens rotations and the triangular solve live on the host.
hipError_t gmres_solve(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                       const double *vals, const double *b, double *x, double rtol, int maxit, int m,
                       double *work, const IluData *ilu, int *iters, double *rel_res) {
  *iters = 0;
  *rel_res = 0.0;
  if (n == 0) return hipSuccess;
  if (m < 1) return hipErrorInvalidValue;
  double *V = work, *dinv = V + (size_t)(m + 1) * n, *w = dinv + n, *u = w + n;
  double *part = u + n, *hd = part + (size_t)(m + 2) * kGmBlocks, *yd = hd + (m + 2);
  double *anorm = yd + (m + 2);
  hipError_t e;
  const unsigned nb = nb256(n);
  auto norm = [&](const double *v, double &out) -> hipError_t {
    k_gm_norm<<<kGmBlocks, kRedThreads, 0, s>>>(n, v, part);
    k_gm_norm_final<<<1, kRedThreads, 0, s>>>(part, anorm);
    hipError_t er = hipMemcpyAsync(&out, anorm, 8, hipMemcpyDeviceToHost, s);
    if (er == hipSuccess) er = hipStreamSynchronize(s);
    return er;
  };
  k_inv_diag<<<nb, 256, 0, s>>>(n, indptr, indices, vals, dinv);
  k_gm_fix_dinv<<<nb, 256, 0, s>>>(n, dinv);
  // out = M^-1 in (in == out allowed): the ILU solves or the inverse diagonal
  auto prec = [&](const double *in, double *out) -> hipError_t {
    if (ilu) return ilu_apply(s, *ilu, indptr, indices, in, out);
    k_gm_scale_diag<<<nb, 256, 0, s>>>(n, dinv, in, out);
    return hipGetLastError();
  };
  // ||M b||
  if ((e = prec(b, w)) != hipSuccess) return e;
  double bn = 0.0;
  if ((e = norm(w, bn)) != hipSuccess) return e;
  if (bn == 0.0) {
    if ((e = hipMemsetAsync(x, 0, 8 * (size_t)n, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  auto residual = [&](double &beta) -> hipError_t {  // w = M (b - A x), beta = ||w||
    hipError_t er = launch_spmv(s, n, nnz, indptr, indices, vals, x, w, false);
    if (er != hipSuccess) return er;
    k_gm_sub<<<nb, 256, 0, s>>>(n, b, w);
    if ((er = prec(w, w)) != hipSuccess) return er;
    return norm(w, beta);
  };
  double beta = 0.0;
  if ((e = residual(beta)) != hipSuccess) return e;
  double beta_old = beta;
  int blocked = 0, it = 0;
  std::vector<double> H((size_t)(m + 1) * m), cr(m + 1), sr(m + 1), g(m + 1), h(m + 2);
  while (beta / bn > rtol && it < maxit) {
    k_gm_normalize<<<nb, 256, 0, s>>>(n, anorm, w, V);  // anorm holds beta
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int i = 0;
    do {
      // u = M A V_i, orthogonalised against V_0..V_i twice
      if ((e = launch_spmv(s, n, nnz, indptr, indices, vals, V + (size_t)i * n, u, false)) != hipSuccess) return e;
      if ((e = prec(u, u)) != hipSuccess) return e;
      const int nk = i + 1;
      const dim3 grid(kGmBlocks, (unsigned)((nk + kGmRows - 1) / kGmRows));
      for (int pass = 0; pass < 2; ++pass) {
        k_gm_multidot<<<grid, kRedThreads, 0, s>>>(n, nk, V, u, part);
        k_gm_sum<<<nk, kRedThreads, 0, s>>>(nk, part, pass == 0 ? hd : yd, 0);
        k_gm_project<<<nb, 256, 0, s>>>(n, nk, V, pass == 0 ? hd : yd, u);
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
      double a = 0.0;
      if ((e = norm(u, a)) != hipSuccess) return e;
      if ((e = hipMemcpy(h.data(), hd, 8 * (size_t)nk, hipMemcpyDeviceToHost)) != hipSuccess) return e;
      std::vector<double> h2((size_t)nk);
      if ((e = hipMemcpy(h2.data(), yd, 8 * (size_t)nk, hipMemcpyDeviceToHost)) != hipSuccess) return e;
      double *col = &H[(size_t)i * (m + 1)];  // column i of H (column-major)
      for (int k = 0; k < nk; ++k) col[k] = h[k] + h2[k];
      col[i + 1] = a;
      k_gm_normalize<<<nb, 256, 0, s>>>(n, anorm, u, V + (size_t)(i + 1) * n);
      for (int k = 0; k < i; ++k) rotate(col[k], col[k + 1], cr[k], sr[k]);
      givens(col[i], col[i + 1], cr[i], sr[i]);
      rotate(col[i], col[i + 1], cr[i], sr[i]);
      rotate(g[i], g[i + 1], cr[i], sr[i]);
      ++i;
      ++it;
    } while (std::fabs(g[i]) / bn > rtol && i < m && it < maxit);
    // y = H(0:i, 0:i)^{-1} g, x += V y
    std::vector<double> y(g.begin(), g.begin() + i);
    for (int r = i - 1; r >= 0; --r) {
      for (int c = r + 1; c < i; ++c) y[r] -= H[(size_t)c * (m + 1) + r] * y[c];
      y[r] /= H[(size_t)r * (m + 1) + r];
    }
    if ((e = hipMemcpyAsync(yd, y.data(), 8 * (size_t)i, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    k_gm_combine<<<nb, 256, 0, s>>>(n, i, V, yd, x);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;  // y is read before it changes
    beta_old = std::min(beta, beta_old);
    if ((e = residual(beta)) != hipSuccess) return e;
    if (i < m - 1 || beta_old <= beta) ++blocked;
    else blocked = 0;
    if (blocked > 10) break;
  }
  *iters = it;
  *rel_res = beta / bn;
  return hipSuccess;
}


hipError_t cg_solve(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr, const int32_t *indices,
                    const double *vals, const double *b, double *x, double rtol, int maxit, int qb,
                    double *work, int *iters, double *rel_res) {
  *iters = 0;
  *rel_res = 0.0;
  if (n == 0) return hipSuccess;
  if (qb != 1 && (qb != 3 || n % 3 != 0)) return hipErrorInvalidValue;
  double *minv = work, *r = work + 3 * n, *z = work + 4 * n, *p = work + 5 * n, *q = work + 6 * n;
  double *part = work + 7 * n;
  CgScal *sc = reinterpret_cast<CgScal *>(part + 3 * kRedBlocks);
  const int64_t nu = n / qb;
  hipError_t e;
  if (qb == 3) k_inv_block3<<<nb256(nu), 256, 0, s>>>(nu, indptr, indices, vals, minv);
  else k_inv_diag<<<nb256(n), 256, 0, s>>>(n, indptr, indices, vals, minv);
  if ((e = launch_spmv(s, n, nnz, indptr, indices, vals, x, r, false)) != hipSuccess) return e;
  if (qb == 3) k_cg_start<3><<<kRedBlocks, kRedThreads, 0, s>>>(nu, b, minv, r, z, p, part);
  else k_cg_start<1><<<kRedBlocks, kRedThreads, 0, s>>>(nu, b, minv, r, z, p, part);
  k_cg_scalars<<<1, kRedThreads, 0, s>>>(part, 0, sc);
  CgScal h{};
  if ((e = hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const double bn = sqrt(h.bb);
  if (bn == 0.0) {  // b = 0: x = 0 is the solution
    if ((e = hipMemsetAsync(x, 0, 8 * (size_t)n, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  double res = sqrt(h.rr) / bn;
  int it = 0;
  while (res > rtol && it < maxit) {
    if ((e = launch_spmv(s, n, nnz, indptr, indices, vals, p, q, false)) != hipSuccess) return e;
    k_dot<<<kRedBlocks, kRedThreads, 0, s>>>(n, p, q, part);
    k_cg_scalars<<<1, kRedThreads, 0, s>>>(part, 1, sc);
    if (qb == 3) k_cg_update<3><<<kRedBlocks, kRedThreads, 0, s>>>(nu, sc, p, q, minv, x, r, z, part);
    else k_cg_update<1><<<kRedBlocks, kRedThreads, 0, s>>>(nu, sc, p, q, minv, x, r, z, part);
    k_cg_scalars<<<1, kRedThreads, 0, s>>>(part, 2, sc);
    k_cg_direction<<<nb256(n), 256, 0, s>>>(n, sc, z, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&h, sc, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    ++it;
    res = sqrt(h.rr) / bn;
    if (h.pq <= 0.0) break;  // K not positive definite along p
  }
  *iters = it;
  *rel_res = res;
  return hipSuccess;
}

}  // namespace gfx
