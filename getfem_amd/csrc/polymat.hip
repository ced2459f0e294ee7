// The order-2 polynomial route of the workspace: tangent matrices of the
// integrands the matrix kernels do not cover (several fem variables, blocks
// between two mesh_fems, data fields, polynomial nonlinearities), as
// ga_workspace::assembly(2) (src/getfem_generic_assembly_workspace.cc:791-936)
// runs them with the two-mesh_fem matrix instruction
// ga_instruction_matrix_assembly_mf_mf
// (src/getfem_generic_assembly_compile_and_exec.cc:5074-5189) and
// add_elem_matrix (:4853-4936).
//
// k_poly_mat   one workgroup (one wavefront) per element and block: at each
//              integration point the lanes fill the symbol table (values and
//              gradients of the fields the program reads), evaluate the
//              program's outputs (one lane per output, its monomials in a
//              fixed order), then add w_q |J| times the bilinear form of the
//              two fields' base functions to their entries of the element
//              matrix held in LDS.  Then the element's threshold
//              1e-14 max|elem| (a zero element adds nothing) and its count
//              of entries above it.
// k_poly_emit  the entries above the threshold as (row, column) keys at the
//              element's offset (the exclusive scan of the counts), in
//              element order.
// poly_to_csr  a stable radix sort by key keeps the element order among equal
//              keys, so each CSR value is summed over the elements in
//              increasing order, the order add_elem_matrix adds them in.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "gfx_internal.h"

namespace gfx {
namespace {

constexpr int kPm = 64;  // one wavefront
constexpr int kPoly1d = 256;

__device__ __forceinline__ double pm_wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int pm_wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ __launch_bounds__(kPm) void k_poly_mat(PolyMatArgs a) {
  __shared__ double T[kScalarMaxSyms];
  __shared__ double O[kPolyMaxOut];
  __shared__ double bt[kPolyMaxNd][4], bs[kPolyMaxNd][4];  // phi, real gradient
  __shared__ double E[kPolyMaxEntries];
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  // geometry, the same on every lane: once for an affine tetrahedron, at
  // every point for a hexahedron (a degenerate point leaves the element,
  // uniformly)
  const int32_t *cv = a.cvx + (int64_t)a.nv * e;
  double B[9], aJ = 0.0;
  bool ok = poly_geometry(a.pts, cv, a.nv, a.gN, B, aJ);
  for (int q = 1; ok && a.gN && q < a.nq; ++q) {  // every hexahedron point checked first
    double Bq[9], aq;
    ok = poly_geometry(a.pts, cv, a.nv, a.gN + (int64_t)q * a.nv * 3, Bq, aq);
  }
  if (!ok) {
    if (t == 0) {
      atomicOr(a.err, 1);
      a.cnt[e] = 0;
      a.thr[e] = INFINITY;
    }
    return;
  }
  const ScalarField &ft = a.f[a.vt], &fs = a.f[a.vs];
  const int qt = ft.qdim, qs = fs.qdim, nc = fs.nd * qs, ne = ft.nd * qt * nc, w4 = 4 * qs;
  for (int c = t; c < a.nconst; c += kPm) T[a.const_off + c] = a.consts[c];
  for (int x = t; x < ne; x += kPm) E[x] = 0.0;
  for (int q = 0; q < a.nq; ++q) {
    if (a.gN && q > 0) poly_geometry(a.pts, cv, a.nv, a.gN + (int64_t)q * a.nv * 3, B, aJ);
    // symbol table: lane -> (field, component)
    int kc = t;
    for (int fi = 0; fi < a.nf; ++fi) {
      const ScalarField &f = a.f[fi];
      if (kc >= 0 && kc < f.qdim) {
        const int32_t *dofs = f.dofs + e * f.nd;
        double v = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
        for (int i = 0; i < f.nd; ++i) {
          const double u = f.U[(int64_t)dofs[i] + kc];
          const double *g = f.gref + ((int64_t)q * f.nd + i) * 3;
          v += f.phi[q * f.nd + i] * u;
          r0 += g[0] * u; r1 += g[1] * u; r2 += g[2] * u;
        }
        T[f.off + kc] = v;
        for (int b = 0; b < 3; ++b)
          T[f.off + f.qdim + 3 * kc + b] = B[3 * b] * r0 + B[3 * b + 1] * r1 + B[3 * b + 2] * r2;
      }
      kc -= f.qdim;
    }
    // base functions of the two fields at the point (nd <= 32)
    if (t < ft.nd) {
      const double *h = ft.gref + ((int64_t)q * ft.nd + t) * 3;
      bt[t][0] = ft.phi[q * ft.nd + t];
      for (int b = 0; b < 3; ++b) bt[t][1 + b] = B[3 * b] * h[0] + B[3 * b + 1] * h[1] + B[3 * b + 2] * h[2];
    }
    if (t >= 32 && t - 32 < fs.nd) {
      const int j = t - 32;
      const double *h = fs.gref + ((int64_t)q * fs.nd + j) * 3;
      bs[j][0] = fs.phi[q * fs.nd + j];
      for (int b = 0; b < 3; ++b) bs[j][1 + b] = B[3 * b] * h[0] + B[3 * b + 1] * h[1] + B[3 * b + 2] * h[2];
    }
    __syncthreads();
    for (int o = t; o < a.nout; o += kPm) {
      double s = 0.0;
      for (int m = a.mptr[o]; m < a.mptr[o + 1]; ++m) {
        double v = a.mcoef[m];
        for (int x = a.moff[m]; x < a.moff[m + 1]; ++x) v *= T[a.msym[x]];
        s += v;
      }
      O[o] = s;
    }
    __syncthreads();
    const double wJ = a.qw[q] * aJ;
    for (int x = t; x < ne; x += kPm) {
      const int r = x / nc, c = x - r * nc;
      const int i = r / qt, ca = r - i * qt, j = c / qs, cb = c - j * qs;
      double s = 0.0;
      for (int p = 0; p < 4; ++p) {  // slot of the test function: value, then gradient b
        const double *Or = O + (p == 0 ? ca : qt + 3 * ca + p - 1) * w4;
        const double u = Or[cb] * bs[j][0] + Or[qs + 3 * cb] * bs[j][1] + Or[qs + 3 * cb + 1] * bs[j][2] +
                         Or[qs + 3 * cb + 2] * bs[j][3];
        s += bt[i][p] * u;
      }
      E[x] += wJ * s;
    }
    __syncthreads();
  }
  double mx = 0.0;
  for (int x = t; x < ne; x += kPm) mx = fmax(mx, fabs(E[x]));
  mx = pm_wave_max(mx);
  const double th = mx > 0.0 ? 1e-14 * mx : INFINITY;
  double *dst = a.elem + e * ne;
  int n = 0;
  for (int x = t; x < ne; x += kPm) {
    dst[x] = E[x];
    n += fabs(E[x]) > th;
  }
  n = pm_wave_sum(n);
  if (t == 0) {
    a.thr[e] = th;
    a.cnt[e] = n;
  }
}

__global__ __launch_bounds__(kPm) void k_poly_emit(PolyEmitArgs a) {
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  const int nc = a.nds * a.qs, ne = a.ndt * a.qt * nc;
  const double th = a.thr[e];
  const double *E = a.elem + e * ne;
  int64_t pos = a.off[e];
  for (int x0 = 0; x0 < ne; x0 += kPm) {
    const int x = x0 + t;
    const double v = x < ne ? E[x] : 0.0;
    const bool keep = x < ne && fabs(v) > th;
    const uint64_t bal = __ballot(keep);
    if (keep) {
      const int r = x / nc, c = x - r * nc;
      const int i = r / a.qt, j = c / a.qs;
      const int64_t row = a.first_t + a.dt[e * a.ndt + i] + (r - i * a.qt);
      const int64_t col = a.first_s + a.ds[e * a.nds + j] + (c - j * a.qs);
      const int64_t p = pos + __popcll(bal & ((1ull << t) - 1ull));
      a.keys[p] = (uint64_t)(row * a.ncols + col);
      a.vals[p] = v;
    }
    pos += __popcll(bal);
  }
}

// head[k] = 1 where a run of equal keys starts
__global__ void k_poly_head(int64_t n, const uint64_t *__restrict__ key, int64_t *__restrict__ head) {
  const int64_t k = (int64_t)blockIdx.x * kPoly1d + threadIdx.x;
  if (k < n) head[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// run u starts at start[u]; its column; the run count
__global__ void k_poly_starts(int64_t n, const uint64_t *__restrict__ key, const int64_t *__restrict__ head,
                              const int64_t *__restrict__ pos, int64_t ncols, int64_t *__restrict__ start,
                              int32_t *__restrict__ ix, int64_t *__restrict__ nnz) {
  const int64_t k = (int64_t)blockIdx.x * kPoly1d + threadIdx.x;
  if (k >= n) return;
  if (head[k]) {
    start[pos[k]] = k;
    ix[pos[k]] = (int32_t)(key[k] % (uint64_t)ncols);
  }
  if (k == n - 1) {
    const int64_t u = pos[k] + head[k];
    start[u] = n;
    *nnz = u;
  }
}

// each value summed over its run in order (the element order)
__global__ void k_poly_sum(int64_t n, const int64_t *__restrict__ start, const int64_t *__restrict__ nnz,
                           const double *__restrict__ vals, double *__restrict__ v) {
  const int64_t u = (int64_t)blockIdx.x * kPoly1d + threadIdx.x;
  if (u >= n || u >= *nnz) return;
  const int64_t b = start[u], e = start[u + 1];
  double s = vals[b];
  for (int64_t k = b + 1; k < e; ++k) s += vals[k];
  v[u] = s;
}

// row pointers: run u opens the rows after the previous run's row up to its own
__global__ void k_poly_indptr(int64_t n, const uint64_t *__restrict__ key, const int64_t *__restrict__ start,
                              const int64_t *__restrict__ nnz, int64_t nrows, int64_t ncols,
                              int64_t *__restrict__ ip) {
  const int64_t u = (int64_t)blockIdx.x * kPoly1d + threadIdx.x;
  const int64_t m = *nnz;
  if (u >= n || u >= m) return;
  const int64_t r = (int64_t)(key[start[u]] / (uint64_t)ncols);
  const int64_t rp = u == 0 ? -1 : (int64_t)(key[start[u - 1]] / (uint64_t)ncols);
  for (int64_t rr = rp + 1; rr <= r; ++rr) ip[rr] = u;
  if (u == m - 1)
    for (int64_t rr = r + 1; rr <= nrows; ++rr) ip[rr] = m;
}

__global__ void k_poly_zero_ip(int64_t n, int64_t *__restrict__ ip, int64_t *__restrict__ nnz) {
  const int64_t k = (int64_t)blockIdx.x * kPoly1d + threadIdx.x;
  if (k < n) ip[k] = 0;
  if (k == 0) *nnz = 0;
}

unsigned pm_blocks(int64_t n) { return (unsigned)((n + kPoly1d - 1) / kPoly1d); }

}  // namespace

hipError_t launch_poly_mat(hipStream_t s, const PolyMatArgs &a) {
  if (a.n_cv == 0) return hipSuccess;
  if (a.nf < 1 || a.nf > kScalarMaxFields || a.vt < 0 || a.vt >= a.nf || a.vs < 0 || a.vs >= a.nf ||
      a.nout > kPolyMaxOut || a.f[a.vt].nd > kPolyMaxNd || a.f[a.vs].nd > kPolyMaxNd ||
      (int64_t)a.f[a.vt].nd * a.f[a.vt].qdim * a.f[a.vs].nd * a.f[a.vs].qdim > kPolyMaxEntries)
    return hipErrorInvalidValue;
  int lanes = 0;
  for (int i = 0; i < a.nf; ++i) lanes += a.f[i].qdim;
  if (lanes > kPm) return hipErrorInvalidValue;
  k_poly_mat<<<(unsigned)a.n_cv, kPm, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_poly_emit(hipStream_t s, const PolyEmitArgs &a) {
  if (a.n_cv == 0) return hipSuccess;
  k_poly_emit<<<(unsigned)a.n_cv, kPm, 0, s>>>(a);
  return hipGetLastError();
}

size_t poly_sort_bytes(int64_t n, int end_bit) {
  size_t bytes = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                         (const double *)nullptr, (double *)nullptr, (int)n, 0, end_bit) !=
      hipSuccess)
    return 0;
  return bytes;
}

hipError_t poly_to_csr(hipStream_t s, void *tmp, size_t tmp_bytes, uint64_t *keys, double *vals, uint64_t *keys2,
                       double *vals2, int64_t n, int end_bit, int64_t nrows, int64_t ncols, int64_t *work,
                       int64_t *ip, int32_t *ix, double *v, int64_t *d_nnz) {
  if (n == 0) {
    k_poly_zero_ip<<<pm_blocks(nrows + 1), kPoly1d, 0, s>>>(nrows + 1, ip, d_nnz);
    return hipGetLastError();
  }
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, vals, vals2, (int)n, 0,
                                                    end_bit, s);
  if (e != hipSuccess) return e;
  int64_t *head = work, *pos = work + n, *start = work + 2 * n, *stmp = work + 3 * n + 1;
  k_poly_head<<<pm_blocks(n), kPoly1d, 0, s>>>(n, keys2, head);
  if ((e = scan_exclusive_i64(s, head, pos, n, stmp)) != hipSuccess) return e;
  k_poly_starts<<<pm_blocks(n), kPoly1d, 0, s>>>(n, keys2, head, pos, ncols, start, ix, d_nnz);
  k_poly_sum<<<pm_blocks(n), kPoly1d, 0, s>>>(n, start, d_nnz, vals2, v);
  k_poly_indptr<<<pm_blocks(n), kPoly1d, 0, s>>>(n, keys2, start, d_nnz, nrows, ncols, ip);
  return hipGetLastError();
}

}  // namespace gfx
