// Incomplete LDL^T preconditioner on the device: gmm::ildlt_precond
// (src/gmm/gmm_precond_ildlt.h:136-196), the preconditioner GetFEM's default
// solver gives coercive problems (linear_solver_cg_preconditioned_ildlt,
// src/getfem/getfem_model_solvers.h:90-98, :652-655), and its application
// gmm::mult(P, r, z) (:200-206) inside gmm::cg (src/gmm/gmm_solver_cg.h:84-118).
//
// The factor lives in one array F aligned with the plan's CSR (symmetric
// pattern, sorted columns): F[diag(i)] = D_i, F at (i, j > i) = U_ij / D_i
// and at (j, i) the same value (L_ji = U_ij / D_i), so both triangular
// solves are row gathers.
//
// Rows are processed in units of qb consecutive rows (the qdim components of
// a node, one after another by one wave or thread), and units in level sets
// of the dependency graph of the lower triangle (forward: factorization and
// forward solve) or of the upper triangle (backward solve), computed on the
// device by frontier sweeps.  Every value is formed with the reference's
// operations in the reference's order, one rounding per product and per
// difference (no fused multiply-add), so F, the forward and the backward
// solves equal the sequential CPU algorithm bit for bit:
//   row r, entry j >= r: a_rj minus, for the lower entries (r, k) in column
//   order, (L_rk D_k) U_kj / D_k where row k holds j;
//   forward:  y_r = b_r - sum over lower (r, k) in column order of y_k L_rk;
//   w_r = y_r / D_r;
//   backward: z_r = w_r - sum over upper (r, j) in column order of F_rj z_j.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gfx_internal.h"

namespace gfx {

namespace {

constexpr int kIlThreads = 64;
constexpr int kIlFactorBytes = 44;  // LDS bytes per row entry of the factorization kernel

// one rounding per product and per difference: no contraction into fma
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double sub_rn(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}

// first position in [lo, hi) whose column is >= c (sorted columns)
__device__ __forceinline__ int64_t lower_bound_col(const int32_t *__restrict__ ix, int64_t lo, int64_t hi,
                                                   int32_t c) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ix[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// diagonal positions; err bit 1: a row without its diagonal entry
__global__ void k_il_diag(int64_t n, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                          int64_t *__restrict__ dpos, int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = lower_bound_col(ix, ip[i], ip[i + 1], (int32_t)i);
  const bool ok = p < ip[i + 1] && ix[p] == i;
  dpos[i] = ok ? p : -1;
  if (!ok) atomicOr(err, 1);
}

// err bit 2: an entry (i, j) whose mirror (j, i) is absent (the pattern is not symmetric)
__global__ void k_il_symmetric(int64_t n, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                               int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t p = ip[i]; p < ip[i + 1]; ++p) {
    const int32_t j = ix[p];
    const int64_t q = lower_bound_col(ix, ip[j], ip[j + 1], (int32_t)i);
    if (q >= ip[j + 1] || ix[q] != i) { atomicOr(err, 2); return; }
  }
}

// dependency counts of unit u (rows qb u .. qb u + qb - 1): entries left of
// the unit's first row (forward) and right of its last row (backward)
__global__ void k_il_degrees(int64_t nu, int qb, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                             int32_t *__restrict__ fdeg, int32_t *__restrict__ bdeg) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  const int64_t r0 = qb * u, r1 = r0 + qb;
  int32_t f = 0, b = 0;
  for (int64_t r = r0; r < r1; ++r) {
    f += (int32_t)(lower_bound_col(ix, ip[r], ip[r + 1], (int32_t)r0) - ip[r]);
    b += (int32_t)(ip[r + 1] - lower_bound_col(ix, ip[r], ip[r + 1], (int32_t)r1));
  }
  fdeg[u] = f;
  bdeg[u] = b;
}

__global__ void k_il_seed(int64_t nu, const int32_t *__restrict__ deg, int32_t *__restrict__ list,
                          int32_t *__restrict__ count) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < nu && deg[u] == 0) list[atomicAdd(count, 1)] = (int32_t)u;
}

// One frontier step: every unit of the current level releases one count of
// each unit it is a dependency of (forward: the entries right of the unit,
// which by symmetry are the dependents' entries left of theirs; backward:
// the entries left of the unit); a unit whose count reaches zero joins the
// next level.  The set of a level is fixed; the order inside it is not, and
// does not matter (its units are independent).
template <bool FWD>
__global__ void k_il_frontier(int32_t m, const int32_t *__restrict__ cur, int qb, const int64_t *__restrict__ ip,
                              const int32_t *__restrict__ ix, int32_t *__restrict__ deg,
                              int32_t *__restrict__ next, int32_t *__restrict__ count) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t w = cur[t], r0 = qb * w, r1 = r0 + qb;
  for (int64_t r = r0; r < r1; ++r) {
    const int64_t b = FWD ? lower_bound_col(ix, ip[r], ip[r + 1], (int32_t)r1) : ip[r];
    const int64_t e = FWD ? ip[r + 1] : lower_bound_col(ix, ip[r], ip[r + 1], (int32_t)r0);
    for (int64_t p = b; p < e; ++p) {
      const int32_t v = ix[p] / qb;
      if (atomicSub(&deg[v], 1) == 1) next[atomicAdd(count, 1)] = v;
    }
  }
}

// Factorization of one level: one wavefront per unit, its rows in order.
// Row r: acc[t] = a at the t-th entry of r from the diagonal on (LDS).  The
// lanes first look up, for every lower entry (r, k) at once, zz = L_rk D_k and
// where row k's part from column r on starts and ends; then, in column order
// of k, that part of row k is staged in LDS (columns and values) and each lane
// finds its column there and subtracts zz U_kj / D_k.  Then D_r = acc[0], and
// the entries right of the diagonal are divided by D_r and written at (r, j)
// and at their mirrors (j, r).  Dynamic LDS: kIlFactorBytes per entry of the
// longest row (cap).
__global__ __launch_bounds__(kIlThreads) void k_il_factor(int32_t m, const int32_t *__restrict__ units, int qb,
                                                         const int64_t *__restrict__ ip,
                                                         const int32_t *__restrict__ ix,
                                                         const double *__restrict__ av,
                                                         const int64_t *__restrict__ dpos, int cap,
                                                         double *__restrict__ F) {
  extern __shared__ double il_lds[];
  double *acc = il_lds, *sval = il_lds + cap, *szz = il_lds + 2 * cap;
  int64_t *sb = reinterpret_cast<int64_t *>(il_lds + 3 * cap), *se = sb + cap;
  int32_t *scol = reinterpret_cast<int32_t *>(se + cap);
  if ((int32_t)blockIdx.x >= m) return;
  const int lane = threadIdx.x;
  const int64_t u = units[blockIdx.x];
  for (int64_t r = qb * u; r < qb * u + qb; ++r) {
    const int64_t d = dpos[r], e = ip[r + 1], s0r = ip[r];
    const int nup = (int)(e - d);  // diagonal + entries right of it
    const int nl = (int)(d - s0r);  // entries left of the diagonal
    for (int t = lane; t < nup; t += kIlThreads) acc[t] = av[d + t];
    for (int t = lane; t < nl; t += kIlThreads) {
      const int32_t k = ix[s0r + t];
      szz[t] = mul_rn(F[s0r + t], F[dpos[k]]);  // L_rk D_k
      // row k from column r on: r is in row k (symmetric pattern)
      sb[t] = lower_bound_col(ix, dpos[k], ip[k + 1], (int32_t)r);
      se[t] = ip[k + 1];
    }
    __syncthreads();
    for (int q = 0; q < nl; ++q) {
      const double zz = szz[q];
      const int64_t s0 = sb[q];
      const int ns = (int)(se[q] - s0);
      for (int t = lane; t < ns; t += kIlThreads) {
        scol[t] = ix[s0 + t];
        sval[t] = F[s0 + t];
      }
      __syncthreads();
      for (int t = lane; t < nup; t += kIlThreads) {
        const int32_t j = ix[d + t];
        int lo = 0, hi = ns;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (scol[mid] < j) lo = mid + 1;
          else hi = mid;
        }
        if (lo < ns && scol[lo] == j) acc[t] = sub_rn(acc[t], mul_rn(zz, sval[lo]));
      }
      __syncthreads();
    }
    const double D = acc[0];
    for (int t = lane; t < nup; t += kIlThreads) {
      if (t == 0) {
        F[d] = D;
      } else {
        const double v = acc[t] / D;
        const int32_t j = ix[d + t];
        F[d + t] = v;
        F[lower_bound_col(ix, ip[j], dpos[j], (int32_t)r)] = v;  // (j, r)
      }
    }
    __syncthreads();  // the next row of the unit reads these values
  }
}

// Triangular solves of one level: one wavefront per unit, its rows in order
// (forward) or in reverse (backward).  The lanes form the row's products in
// parallel (one rounding each) into LDS, then lane 0 subtracts them in column
// order, so the chain of differences is the reference's; the row's result is
// in global memory before the next row of the unit reads it.  Dynamic LDS:
// cap doubles.
//   forward:  y_r = b_r - sum_{(r,k), k < r} y_k L_rk,  w_r = y_r / D_r
//   backward: z_r = w_r - sum_{(r,j), j > r} F_rj z_j
__global__ __launch_bounds__(kIlThreads) void k_il_forward(int32_t m, const int32_t *__restrict__ units, int qb,
                                                          const int64_t *__restrict__ ip,
                                                          const int32_t *__restrict__ ix,
                                                          const int64_t *__restrict__ dpos,
                                                          const double *__restrict__ F,
                                                          const double *__restrict__ b, double *__restrict__ y,
                                                          double *__restrict__ w) {
  extern __shared__ double il_lds[];
  if ((int32_t)blockIdx.x >= m) return;
  const int lane = threadIdx.x;
  const int64_t u = units[blockIdx.x];
  for (int64_t r = qb * u; r < qb * u + qb; ++r) {
    const int64_t s0 = ip[r], d = dpos[r];
    const int nl = (int)(d - s0);
    for (int t = lane; t < nl; t += kIlThreads) il_lds[t] = mul_rn(y[ix[s0 + t]], F[s0 + t]);
    __syncthreads();
    if (lane == 0) {
      double s = b[r];
      for (int t = 0; t < nl; ++t) s = sub_rn(s, il_lds[t]);
      y[r] = s;
      w[r] = s / F[d];
    }
    __syncthreads();  // y_r before the unit's next row reads it
  }
}

__global__ __launch_bounds__(kIlThreads) void k_il_backward(int32_t m, const int32_t *__restrict__ units, int qb,
                                                           const int64_t *__restrict__ ip,
                                                           const int32_t *__restrict__ ix,
                                                           const int64_t *__restrict__ dpos,
                                                           const double *__restrict__ F,
                                                           const double *__restrict__ w, double *__restrict__ z) {
  extern __shared__ double il_lds[];
  if ((int32_t)blockIdx.x >= m) return;
  const int lane = threadIdx.x;
  const int64_t u = units[blockIdx.x];
  for (int64_t r = qb * u + qb - 1; r >= qb * u; --r) {
    const int64_t d = dpos[r];
    const int nu = (int)(ip[r + 1] - d - 1);
    for (int t = lane; t < nu; t += kIlThreads) il_lds[t] = mul_rn(F[d + 1 + t], z[ix[d + 1 + t]]);
    __syncthreads();
    if (lane == 0) {
      double s = w[r];
      for (int t = 0; t < nu; ++t) s = sub_rn(s, il_lds[t]);
      z[r] = s;
    }
    __syncthreads();
  }
}

// pivot check of gmm (:171-176): pivot k is replaced when |D_k| <= maxp_k,
// maxp_0 = eps |a_00|, maxp_{k+1} = max(maxp_k, min(eps |D_k|, 1)).  With
// c_k = min(eps |D_k|, 1), maxp_k = max(eps |a_00|, max_{k' < k} c_k'): block
// maxima of c, then each block checks its pivots against the running
// maximum (two launches; the host combines the block maxima in between).
constexpr int kPivBlock = 1024;
__global__ void k_il_pivot_max(int64_t n, const int64_t *__restrict__ dpos, const double *__restrict__ F,
                               double *__restrict__ bmax) {
  __shared__ double sh[256];
  const int64_t b0 = (int64_t)blockIdx.x * kPivBlock;
  double mx = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < min(n, b0 + kPivBlock); i += 256)
    mx = fmax(mx, fmin(fabs(F[dpos[i]]) * 2.220446049250313e-16, 1.0));
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = sh[0];
}

// pre[b]: the running maximum before block b (exclusive, including eps |a_00|);
// err bit 8: a pivot GetFEM would replace by 1
__global__ void k_il_pivot_check(int64_t n, const int64_t *__restrict__ dpos, const double *__restrict__ F,
                                 const double *__restrict__ pre, int *__restrict__ err) {
  if (threadIdx.x != 0) return;
  const int64_t b0 = (int64_t)blockIdx.x * kPivBlock;
  double maxp = pre[blockIdx.x];
  for (int64_t i = b0; i < min(n, b0 + kPivBlock); ++i) {
    const double z = fabs(F[dpos[i]]);
    if (z <= maxp) { atomicOr(err, 8); return; }
    maxp = fmax(maxp, fmin(z * 2.220446049250313e-16, 1.0));
  }
}

// ---- conjugate gradient pieces (gmm::cg with a separate preconditioner pass)
constexpr int kIlRedThreads = 256, kIlRedBlocks = 1024;

__global__ __launch_bounds__(kIlRedThreads) void k_il_dot2(int64_t n, const double *__restrict__ a,
                                                          const double *__restrict__ b,
                                                          const double *__restrict__ c,
                                                          const double *__restrict__ d,
                                                          double *__restrict__ part) {
  __shared__ double s1[kIlRedThreads], s2[kIlRedThreads];
  double x = 0.0, y = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kIlRedThreads + threadIdx.x; i < n;
       i += (int64_t)kIlRedBlocks * kIlRedThreads) {
    x += a[i] * b[i];
    if (c) y += c[i] * d[i];
  }
  s1[threadIdx.x] = x;
  s2[threadIdx.x] = y;
  __syncthreads();
  for (int o = kIlRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[kIlRedBlocks + blockIdx.x] = s2[0];
  }
}

// fixed-order sum of the partials: out[0] = sum a.b and, when two, out[1] = sum c.d
__global__ __launch_bounds__(kIlRedThreads) void k_il_sum2(const double *__restrict__ part, double *__restrict__ out,
                                                          int two) {
  __shared__ double s1[kIlRedThreads], s2[kIlRedThreads];
  double x = 0.0, y = 0.0;
  for (int i = threadIdx.x; i < kIlRedBlocks; i += kIlRedThreads) {
    x += part[i];
    y += part[kIlRedBlocks + i];
  }
  s1[threadIdx.x] = x;
  s2[threadIdx.x] = y;
  __syncthreads();
  for (int o = kIlRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    if (two) out[1] = s2[0];
  }
}

// r = b - r (r holds K x)
__global__ void k_il_resid(int64_t n, const double *__restrict__ b, double *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

// p = z + beta p with beta = rho / rho_1 (sc[0] = rho, sc[2] = rho_1); first: p = z
__global__ void k_il_direction(int64_t n, const double *__restrict__ sc, const double *__restrict__ z,
                               double *__restrict__ p, int first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = first ? z[i] : z[i] + (sc[0] / sc[2]) * p[i];
}

// a = rho / (q.p) (sc[3] = q.p); x += a p; r -= a q
__global__ void k_il_step(int64_t n, const double *__restrict__ sc, const double *__restrict__ p,
                          const double *__restrict__ q, double *__restrict__ x, double *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = sc[0] / sc[3];
  x[i] += a * p[i];
  r[i] -= a * q[i];
}

__global__ void k_il_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

unsigned il_blocks(int64_t n, int th) { return (unsigned)((n + th - 1) / th); }

}  // namespace

// Level sets of units: host-side offsets (levels + 1 entries) into the device
// list of units.  The frontier loop synchronises the stream once per level.
static hipError_t il_levels(hipStream_t s, int64_t nu, int qb, const int64_t *ip, const int32_t *ix, bool fwd,
                            int32_t *deg, int32_t *list, int32_t *counter, std::vector<int64_t> &lev) {
  hipError_t e;
  lev.assign(1, 0);
  if ((e = hipMemsetAsync(counter, 0, 4, s)) != hipSuccess) return e;
  k_il_seed<<<il_blocks(nu, 256), 256, 0, s>>>(nu, deg, list, counter);
  int32_t done = 0, m = 0;
  if ((e = hipMemcpyAsync(&m, counter, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  while (m > 0) {
    lev.push_back(done + m);
    const int32_t cur = done;
    done += m;
    if (done >= nu) break;
    if ((e = hipMemsetAsync(counter, 0, 4, s)) != hipSuccess) return e;
    if (fwd)
      k_il_frontier<true><<<il_blocks(m, 256), 256, 0, s>>>(m, list + cur, qb, ip, ix, deg, list + done, counter);
    else
      k_il_frontier<false><<<il_blocks(m, 256), 256, 0, s>>>(m, list + cur, qb, ip, ix, deg, list + done, counter);
    if ((e = hipMemcpyAsync(&m, counter, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  }
  return done == nu ? hipSuccess : hipErrorInvalidValue;  // a cycle cannot happen for a symmetric pattern
}

hipError_t ildlt_analyse(hipStream_t s, IldltData &L, int64_t n, int qb, const int64_t *ip, const int32_t *ix,
                         int *d_err, int *h_err) {
  hipError_t e;
  *h_err = 0;
  L.n = n;
  L.qb = qb;
  const int64_t nu = n / qb;
  if ((e = hipMemsetAsync(d_err, 0, 4, s)) != hipSuccess) return e;
  k_il_diag<<<il_blocks(n, 256), 256, 0, s>>>(n, ip, ix, L.d_dpos, d_err);
  k_il_symmetric<<<il_blocks(n, 256), 256, 0, s>>>(n, ip, ix, d_err);
  k_il_degrees<<<il_blocks(nu, 256), 256, 0, s>>>(nu, qb, ip, ix, L.d_fdeg, L.d_bdeg);
  if ((e = hipMemcpyAsync(h_err, d_err, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (*h_err) return hipSuccess;
  if ((e = il_levels(s, nu, qb, ip, ix, true, L.d_fdeg, L.d_flist, L.d_counter, L.flev)) != hipSuccess) return e;
  if ((e = il_levels(s, nu, qb, ip, ix, false, L.d_bdeg, L.d_blist, L.d_counter, L.blev)) != hipSuccess) return e;
  return hipSuccess;
}

hipError_t ildlt_factor(hipStream_t s, IldltData &L, const int64_t *ip, const int32_t *ix, const double *av,
                        int cap, int *d_err, int *h_err) {
  hipError_t e;
  *h_err = 0;
  const size_t lds = (size_t)cap * kIlFactorBytes;
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  for (size_t l = 0; l + 1 < L.flev.size(); ++l) {
    const int32_t m = (int32_t)(L.flev[l + 1] - L.flev[l]);
    k_il_factor<<<(unsigned)m, kIlThreads, lds, s>>>(m, L.d_flist + L.flev[l], L.qb, ip, ix, av, L.d_dpos, cap,
                                                     L.d_F);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // pivots: block maxima, exclusive running maxima on the host, block checks
  const int64_t nb = (L.n + kPivBlock - 1) / kPivBlock;
  if (nb == 0) return hipSuccess;
  k_il_pivot_max<<<(unsigned)nb, 256, 0, s>>>(L.n, L.d_dpos, L.d_F, L.d_work);
  std::vector<double> bm((size_t)nb), pre((size_t)nb);
  double a00 = 0.0;  // row 0's first entry is its diagonal (present, sorted columns)
  if ((e = hipMemcpyAsync(bm.data(), L.d_work, 8 * (size_t)nb, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(&a00, av, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (a00 == 0.0) {  // GetFEM sets U_00 = 1 first: not reproduced here
    *h_err = 8;
    return hipSuccess;
  }
  double run = fabs(a00) * 2.220446049250313e-16;
  for (int64_t b = 0; b < nb; ++b) {
    pre[(size_t)b] = run;
    run = std::max(run, bm[(size_t)b]);
  }
  if ((e = hipMemcpyAsync(L.d_work, pre.data(), 8 * (size_t)nb, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(d_err, 0, 4, s)) != hipSuccess) return e;
  k_il_pivot_check<<<(unsigned)nb, 64, 0, s>>>(L.n, L.d_dpos, L.d_F, L.d_work, d_err);
  if ((e = hipMemcpyAsync(h_err, d_err, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t ildlt_apply(hipStream_t s, const IldltData &L, const int64_t *ip, const int32_t *ix, const double *r,
                       double *z, double *y, double *w) {
  const size_t lds = 8 * (size_t)std::max(L.cap, 1);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  for (size_t l = 0; l + 1 < L.flev.size(); ++l) {
    const int32_t m = (int32_t)(L.flev[l + 1] - L.flev[l]);
    k_il_forward<<<(unsigned)m, kIlThreads, lds, s>>>(m, L.d_flist + L.flev[l], L.qb, ip, ix, L.d_dpos, L.d_F, r,
                                                       y, w);
  }
  for (size_t l = 0; l + 1 < L.blev.size(); ++l) {
    const int32_t m = (int32_t)(L.blev[l + 1] - L.blev[l]);
    k_il_backward<<<(unsigned)m, kIlThreads, lds, s>>>(m, L.d_blist + L.blev[l], L.qb, ip, ix, L.d_dpos, L.d_F, w,
                                                        z);
  }
  return hipGetLastError();
}

size_t ildlt_cg_doubles(int64_t n) { return 6 * (size_t)n + 2 * kIlRedBlocks + 8; }

// ------------------------------------------------------------------- ILU(0)
// gmm::ilu_precond (src/gmm/gmm_precond_ilu.h:104-219) of GetFEM's matrix,
// which is column-major: the preconditioner factors A^T by rows (invert =
// true) and gmm::mult (:221-231) applies it as the lower solve with U^T, then
// the unit upper solve with L^T (src/gmm/gmm_tri_solve.h:75-89, :45-57).
// With a symmetric pattern, A^T's row i has A's row i's columns: F is aligned
// with A's CSR, F at (i, j) = L_ij (j < i) or U_ij (j >= i) of A^T, and T at
// (i, j) is the position of (j, i).  One thread per row, rows of one level
// set at a time; every value is formed with the reference's operations in
// its order (no fused multiply-add), so F and the solves equal the
// sequential algorithm bit for bit.
namespace {

__global__ void k_ilu_tpos(int64_t n, const int64_t *__restrict__ ip, const int32_t *__restrict__ ix,
                           int64_t *__restrict__ T, const double *__restrict__ av, double *__restrict__ F,
                           int *__restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t p = ip[i]; p < ip[i + 1]; ++p) {
    const int32_t j = ix[p];
    const int64_t q = lower_bound_col(ix, ip[j], ip[j + 1], (int32_t)i);
    if (q == ip[j + 1] || ix[q] != i) {
      atomicOr(err, 2);
      T[p] = p;
      F[p] = 0.0;
      continue;
    }
    T[p] = q;
    F[p] = av[q];  // A^T
  }
}

__global__ void k_ilu_diag(int64_t n, const int64_t *__restrict__ dpos, const double *__restrict__ F,
                           double *__restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = F[dpos[i]];
}

// the diagonals gmm replaces by 1 before their row is eliminated
__global__ void k_ilu_unit(int64_t m, const int64_t *__restrict__ rows, const int64_t *__restrict__ dpos,
                           double *__restrict__ F) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) F[dpos[rows[k]]] = 1.0;
}

// row i: for its lower entries k in column order, the multiplier L_ik /= U_kk,
// then every later entry (i, c) of the row present in row k: -= L_ik U_kc
__global__ void k_ilu_factor(int32_t m, const int32_t *__restrict__ list, const int64_t *__restrict__ ip,
                             const int32_t *__restrict__ ix, const int64_t *__restrict__ dpos,
                             double *__restrict__ F) {
  const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= m) return;
  const int64_t i = list[u], e = ip[i + 1];
  for (int64_t p = ip[i]; p < e && ix[p] < i; ++p) {
    const int32_t k = ix[p];
    const double mult = F[p] / F[dpos[k]];
    F[p] = mult;
    int64_t r = p + 1;
    for (int64_t q = dpos[k] + 1; q < ip[k + 1]; ++q) {
      const int32_t c = ix[q];
      while (r < e && ix[r] < c) ++r;
      if (r < e && ix[r] == c) F[r] = sub_rn(F[r], mul_rn(mult, F[q]));
    }
  }
}

// y_c = (r_c - sum over j < c, ascending, of y_j U_jc) / U_cc
__global__ void k_ilu_forward(int32_t m, const int32_t *__restrict__ list, const int64_t *__restrict__ ip,
                              const int32_t *__restrict__ ix, const int64_t *__restrict__ dpos,
                              const int64_t *__restrict__ T, const double *__restrict__ F,
                              const double *__restrict__ r, double *__restrict__ y) {
  const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= m) return;
  const int64_t c = list[u];
  double x = r[c];
  for (int64_t p = ip[c]; p < ip[c + 1] && ix[p] < c; ++p) x = sub_rn(x, mul_rn(y[ix[p]], F[T[p]]));
  y[c] = x / F[dpos[c]];
}

// z_i = y_i - sum over j > i, descending, of z_j L_ji (unit diagonal)
__global__ void k_ilu_backward(int32_t m, const int32_t *__restrict__ list, const int64_t *__restrict__ ip,
                               const int32_t *__restrict__ ix, const int64_t *__restrict__ T,
                               const double *__restrict__ F, const double *__restrict__ y, double *__restrict__ z) {
  const int32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= m) return;
  const int64_t i = list[u];
  double x = y[i];
  for (int64_t p = ip[i + 1] - 1; p >= ip[i] && ix[p] > i; --p) x = sub_rn(x, mul_rn(z[ix[p]], F[T[p]]));
  z[i] = x;
}

}  // namespace

hipError_t ilu_factor(hipStream_t s, IluData &U, const int64_t *ip, const int32_t *ix, const double *av,
                      int *d_err, int *h_err, int64_t *replaced) {
  hipError_t e;
  *h_err = 0;
  *replaced = 0;
  const int64_t n = U.lv.n;
  if ((e = hipMemsetAsync(d_err, 0, 4, s)) != hipSuccess) return e;
  k_ilu_tpos<<<il_blocks(n, 256), 256, 0, s>>>(n, ip, ix, U.d_T, av, U.d_F, d_err);
  // the pivots gmm replaces by 1: decided on the raw diagonal of each row in
  // order, against the running bound max_k min(|U_kk| eps, 1) (:181-196)
  std::vector<double> diag((size_t)n);
  k_ilu_diag<<<il_blocks(n, 256), 256, 0, s>>>(n, U.lv.d_dpos, U.d_F, U.d_y);
  if ((e = hipMemcpyAsync(diag.data(), U.d_y, 8 * (size_t)n, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(h_err, d_err, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (*h_err) return hipSuccess;
  const double eps = 2.220446049250313e-16;
  std::vector<int64_t> unit;
  double maxp = std::fabs(diag[0]) * eps;
  if (diag[0] == 0.0) unit.push_back(0);
  for (int64_t i = 1; i < n; ++i) {
    double d = diag[(size_t)i];
    if (std::fabs(d) <= maxp) {
      d = 1.0;
      unit.push_back(i);
    }
    maxp = std::max(maxp, std::min(std::fabs(d) * eps, 1.0));
  }
  *replaced = (int64_t)unit.size();
  if (!unit.empty()) {
    if ((e = hipMemcpyAsync(U.d_rows, unit.data(), 8 * unit.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
      return e;
    k_ilu_unit<<<il_blocks((int64_t)unit.size(), 256), 256, 0, s>>>((int64_t)unit.size(), U.d_rows, U.lv.d_dpos,
                                                                      U.d_F);
  }
  for (size_t l = 0; l + 1 < U.lv.flev.size(); ++l) {
    const int32_t m = (int32_t)(U.lv.flev[l + 1] - U.lv.flev[l]);
    k_ilu_factor<<<il_blocks(m, 64), 64, 0, s>>>(m, U.lv.d_flist + U.lv.flev[l], ip, ix, U.lv.d_dpos, U.d_F);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipStreamSynchronize(s);  // the host vectors above stay alive until here
}

hipError_t ilu_apply(hipStream_t s, const IluData &U, const int64_t *ip, const int32_t *ix, const double *r,
                     double *z) {
  for (size_t l = 0; l + 1 < U.lv.flev.size(); ++l) {
    const int32_t m = (int32_t)(U.lv.flev[l + 1] - U.lv.flev[l]);
    k_ilu_forward<<<il_blocks(m, 64), 64, 0, s>>>(m, U.lv.d_flist + U.lv.flev[l], ip, ix, U.lv.d_dpos, U.d_T,
                                                   U.d_F, r, U.d_y);
  }
  for (size_t l = 0; l + 1 < U.lv.blev.size(); ++l) {
    const int32_t m = (int32_t)(U.lv.blev[l + 1] - U.lv.blev[l]);
    k_ilu_backward<<<il_blocks(m, 64), 64, 0, s>>>(m, U.lv.d_blist + U.lv.blev[l], ip, ix, U.d_T, U.d_F, U.d_y, z);
  }
  return hipGetLastError();
}

// gmm::cg with the ILDLT preconditioner: r = b - K x, z = P r, rho = z.r,
// p = z; while |r| > rtol |b|: (after the first step: z = P r, rho = z.r,
// p = z + rho / rho_1 p), q = K p, a = rho / q.p, x += a p, r -= a q.
// work: r, z, p, q, y, w (n each), the partials and 8 scalars.
hipError_t ildlt_pcg(hipStream_t s, const IldltData &L, int64_t nnz, const int64_t *ip, const int32_t *ix,
                     const double *av, const double *b, double *x, double rtol, int maxit, double *work,
                     int *iters, double *rel_res) {
  const int64_t n = L.n;
  *iters = 0;
  *rel_res = 0.0;
  if (n == 0) return hipSuccess;
  double *r = work, *z = work + n, *p = work + 2 * n, *q = work + 3 * n, *y = work + 4 * n, *w = work + 5 * n;
  double *part = work + 6 * n, *sc = part + 2 * kIlRedBlocks;  // sc: rho, |r|^2, rho_1, q.p, |b|^2
  hipError_t e;
  const unsigned gb = il_blocks(n, 256);
  // |b|
  k_il_dot2<<<kIlRedBlocks, kIlRedThreads, 0, s>>>(n, b, b, nullptr, nullptr, part);
  k_il_sum2<<<1, kIlRedThreads, 0, s>>>(part, sc + 4, 0);
  if ((e = launch_spmv(s, n, nnz, ip, ix, av, x, r, false)) != hipSuccess) return e;
  k_il_resid<<<gb, 256, 0, s>>>(n, b, r);
  if ((e = ildlt_apply(s, L, ip, ix, r, z, y, w)) != hipSuccess) return e;
  k_il_dot2<<<kIlRedBlocks, kIlRedThreads, 0, s>>>(n, z, r, r, r, part);
  k_il_sum2<<<1, kIlRedThreads, 0, s>>>(part, sc, 1);  // rho, |r|^2
  double h[5];
  if ((e = hipMemcpyAsync(h, sc, 40, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const double bn = sqrt(h[4]);
  if (bn == 0.0) {
    k_il_copy<<<gb, 256, 0, s>>>(n, b, x);  // b = 0: x = 0
    return hipStreamSynchronize(s);
  }
  double res = sqrt(h[1]) / bn;
  int it = 0;
  while (res > rtol && it < maxit) {
    if (it > 0) {
      k_il_copy<<<1, 1, 0, s>>>(1, sc, sc + 2);  // rho_1 = rho
      if ((e = ildlt_apply(s, L, ip, ix, r, z, y, w)) != hipSuccess) return e;
      k_il_dot2<<<kIlRedBlocks, kIlRedThreads, 0, s>>>(n, z, r, nullptr, nullptr, part);
      k_il_sum2<<<1, kIlRedThreads, 0, s>>>(part, sc, 0);  // rho
    }
    k_il_direction<<<gb, 256, 0, s>>>(n, sc, z, p, it == 0);
    if ((e = launch_spmv(s, n, nnz, ip, ix, av, p, q, false)) != hipSuccess) return e;
    k_il_dot2<<<kIlRedBlocks, kIlRedThreads, 0, s>>>(n, q, p, nullptr, nullptr, part);
    k_il_sum2<<<1, kIlRedThreads, 0, s>>>(part, sc + 3, 0);  // q.p
    k_il_step<<<gb, 256, 0, s>>>(n, sc, p, q, x, r);
    k_il_dot2<<<kIlRedBlocks, kIlRedThreads, 0, s>>>(n, r, r, nullptr, nullptr, part);
    k_il_sum2<<<1, kIlRedThreads, 0, s>>>(part, sc + 1, 0);  // |r|^2
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h, sc, 40, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    ++it;
    res = sqrt(h[1]) / bn;
    if (!(h[3] > 0.0)) break;  // K not positive definite along p
  }
  *iters = it;
  *rel_res = res;
  return hipSuccess;
}

}  // namespace gfx
