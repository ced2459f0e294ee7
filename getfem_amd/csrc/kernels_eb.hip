// Element-batched assembly (non-affine elements / nonlinear material):
// a workgroup computes a whole element matrix, the matrices go to an HBM
// buffer (row-major, rows = Test_u), and a node-patch gather adds the rows
// each node owns into its CSR rows, with GetFEM's per-element insertion
// threshold (compile_and_exec.cc:5441-5453).  Used for the compressible
// Neo-Hookean (Bonet) tangent and residual on Q2 hexahedra:
//   GWFL ((Id+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(Grad_u,[lambda;mu]))):Grad_Test_u
//   (src/getfem_nonlinear_elasticity.cc:2319-2320), law :612-696, derivative
//   wrapper :1799-1827.
#include <hip/hip_runtime.h>

#include "gfx_internal.h"

namespace gfx {

namespace {

constexpr int kEbThreads = 256;
constexpr int kNd = 27, kNl = 81, kNv = 8;
constexpr int kPerThread = (kNl * kNl + kEbThreads - 1) / kEbThreads;  // 26

// One workgroup per element.  Per integration point q:
//   K = G grad(N)(q), J, B = K^{-T}   (trilinear transformation, per point)
//   g_i = B grad(phi_i)(q);  H = sum_i U_i (x) g_i;  F = I + H
//   S = mu (I - C^-1) + lambda/2 ln(I3) C^-1,  A = dS/dE
//   D[(a,c),(b,n)] = delta_ab S[n][c] + sum_km F[a][k] A[k c m n] F[b][m]
//   T[(a,c)][(j,b)] = w|J| sum_n D[(a,c),(b,n)] g_j[n]
//   Ke[(i,a),(j,b)] += sum_c g_i[c] T[(a,c)][(j,b)];  re[(i,a)] += w|J| (F S g_i)_a
__global__ __launch_bounds__(kEbThreads) void k_nh_element(EbArgs a) {
  __shared__ double X[kNv][3], Ue[kNd][3], g[kNd][3];
  __shared__ double Kg[9], Bm[9], sc[4];  // sc: |J| w, lnI3
  __shared__ double H[9], F[9], S[9], Ci[9], Cm[9], P[9];
  __shared__ double A[81], D[81];
  __shared__ double T[9 * kNl];
  __shared__ double red[kEbThreads];
  const int64_t e = blockIdx.x;
  if (e >= a.n_cv) return;
  const int t = threadIdx.x;
  if (t < kNv * 3) X[t / 3][t % 3] = a.pts[3 * (int64_t)a.cvx[kNv * e + t / 3] + t % 3];
  if (t < kNl) Ue[t / 3][t % 3] = a.U[(int64_t)a.dofs[kNd * e + t / 3] + t % 3];
  const double lam = a.ca ? a.ca[e] : a.ca_c, mu = a.cb ? a.cb[e] : a.cb_c;
  double Ke[kPerThread];
  for (int k = 0; k < kPerThread; ++k) Ke[k] = 0.0;
  double re = 0.0;
  __syncthreads();
  for (int q = 0; q < a.nq; ++q) {
    // geometry
    if (t < 9) {
      const int r = t / 3, c = t % 3;
      const double *gn = a.gN + (int64_t)q * kNv * 3;
      double s = X[0][r] * gn[c];
      for (int v = 1; v < kNv; ++v) s += X[v][r] * gn[3 * v + c];
      Kg[r + 3 * c] = s;  // column-major K(r,c)
    }
    __syncthreads();
    if (t == 0) {
      const double *k = Kg;
      const double a0 = k[4] * k[8] - k[5] * k[7];
      const double a1 = k[5] * k[6] - k[3] * k[8];
      const double a2 = k[3] * k[7] - k[4] * k[6];
      const double J = k[0] * a0 + k[1] * a1 + k[2] * a2;
      if (J == 0.0) atomicOr(a.err, 1);
      Bm[0] = a0 / J; Bm[3] = a1 / J; Bm[6] = a2 / J;  // row-major B(r,c) = Bm[3r+c]
      Bm[1] = (k[2] * k[7] - k[1] * k[8]) / J;
      Bm[4] = (k[0] * k[8] - k[2] * k[6]) / J;
      Bm[7] = (k[1] * k[6] - k[0] * k[7]) / J;
      Bm[2] = (k[1] * k[5] - k[2] * k[4]) / J;
      Bm[5] = (k[2] * k[3] - k[0] * k[5]) / J;
      Bm[8] = (k[0] * k[4] - k[1] * k[3]) / J;
      sc[0] = fabs(J) * a.qw[q];
    }
    __syncthreads();
    if (t < kNl) {
      const int i = t / 3, c = t % 3;
      const double *gr = a.gref + ((int64_t)q * kNd + i) * 3;
      g[i][c] = Bm[3 * c] * gr[0] + Bm[3 * c + 1] * gr[1] + Bm[3 * c + 2] * gr[2];
    }
    __syncthreads();
    if (t < 9) {
      const int r = t / 3, c = t % 3;
      double s = 0.0;
      for (int i = 0; i < kNd; ++i) s += Ue[i][r] * g[i][c];
      H[t] = s;
      F[t] = s + (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    if (t == 0) {
      double C[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
      const double c0 = C[4] * C[8] - C[5] * C[7];
      const double c1 = C[5] * C[6] - C[3] * C[8];
      const double c2 = C[3] * C[7] - C[4] * C[6];
      const double I3 = C[0] * c0 + C[1] * c1 + C[2] * c2;
      Ci[0] = c0 / I3; Ci[3] = c1 / I3; Ci[6] = c2 / I3;
      Ci[1] = (C[2] * C[7] - C[1] * C[8]) / I3;
      Ci[4] = (C[0] * C[8] - C[2] * C[6]) / I3;
      Ci[7] = (C[1] * C[6] - C[0] * C[7]) / I3;
      Ci[2] = (C[1] * C[5] - C[2] * C[4]) / I3;
      Ci[5] = (C[2] * C[3] - C[0] * C[5]) / I3;
      Ci[8] = (C[0] * C[4] - C[1] * C[3]) / I3;
      const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                          F[2] * (F[3] * F[7] - F[4] * F[6]);
      const double l3 = log(I3);
      sc[1] = l3;
      for (int i = 0; i < 9; ++i) {
        S[i] = mu * ((i % 4 == 0 ? 1.0 : 0.0) - Ci[i]) + 0.5 * lam * l3 * Ci[i];
        if (detF <= 0.0) S[i] += 1e200 * C[i];
        Cm[i] = C[i];
      }
    }
    __syncthreads();
    if (t < 81) {  // A[i j k l] = dS_ij / dE_kl
      const int i = t / 27, j = (t / 9) % 3, k = (t / 3) % 3, l = t % 3;
      const double dCi = -0.5 * (Ci[3 * i + k] * Ci[3 * l + j] + Ci[3 * i + l] * Ci[3 * k + j]);
      A[t] = 2.0 * (-mu * dCi + 0.5 * lam * (Ci[3 * k + l] * Ci[3 * i + j] + sc[1] * dCi));
    } else if (t < 90) {
      const int r = (t - 81) / 3, c = (t - 81) % 3;
      P[t - 81] = F[3 * r] * S[c] + F[3 * r + 1] * S[3 + c] + F[3 * r + 2] * S[6 + c];
    }
    __syncthreads();
    const double wq = sc[0];
    if (t < 81) {
      // D[(a,c),(b,n)] with t = (a*3+c)*9 + (b*3+n)
      const int ac = t / 9, bn = t % 9;
      const int aa = ac / 3, c = ac % 3, bb = bn / 3, n = bn % 3;
      double s = (aa == bb) ? S[3 * n + c] : 0.0;
      for (int k = 0; k < 3; ++k) {
        double in = 0.0;
        for (int m = 0; m < 3; ++m) in += A[27 * k + 9 * c + 3 * m + n] * F[3 * bb + m];
        s += F[3 * aa + k] * in;
      }
      D[t] = s;
      // residual row (i, a) = t
      const int i = t / 3, ar = t % 3;
      re += wq * (P[3 * ar] * g[i][0] + P[3 * ar + 1] * g[i][1] + P[3 * ar + 2] * g[i][2]);
    }
    __syncthreads();
    for (int x = t; x < 9 * kNl; x += kEbThreads) {
      const int ac = x / kNl, jb = x % kNl;
      const int j = jb / 3, bb = jb % 3;
      const double *d = D + 9 * ac + 3 * bb;
      T[x] = wq * (d[0] * g[j][0] + d[1] * g[j][1] + d[2] * g[j][2]);
    }
    __syncthreads();
    for (int k = 0; k < kPerThread; ++k) {
      const int idx = t + kEbThreads * k;
      if (idx < kNl * kNl) {
        const int r = idx / kNl, c = idx % kNl;
        const int i = r / 3, ar = r % 3;
        const double *tt = T + (3 * ar) * kNl + c;
        Ke[k] += g[i][0] * tt[0] + g[i][1] * tt[kNl] + g[i][2] * tt[2 * kNl];
      }
    }
    __syncthreads();
  }
  // threshold: 1e-14 * max|Ke| (an all-zero element is skipped: thr = inf)
  double mx = 0.0;
  for (int k = 0; k < kPerThread; ++k)
    if (t + kEbThreads * k < kNl * kNl) mx = fmax(mx, fabs(Ke[k]));
  red[t] = mx;
  __syncthreads();
  for (int o = kEbThreads / 2; o > 0; o >>= 1) {
    if (t < o) red[t] = fmax(red[t], red[t + o]);
    __syncthreads();
  }
  double *E = a.E + (int64_t)e * kNl * kNl;
  for (int k = 0; k < kPerThread; ++k) {
    const int idx = t + kEbThreads * k;
    if (idx < kNl * kNl) E[idx] = Ke[k];
  }
  if (t < kNl) a.res[(int64_t)e * kNl + t] = re;
  if (t == 0) a.thr[e] = red[0] > 0.0 ? 1e-14 * red[0] : INFINITY;
}

// keep bits of element matrices held in the buffer: bit r*nl + c
__global__ __launch_bounds__(256) void k_keep_emat(int64_t n_cv, int nl, int kwords,
                                                   const double *__restrict__ E,
                                                   const double *__restrict__ thr,
                                                   uint32_t *__restrict__ keep) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_cv * kwords) return;
  const int64_t e = w / kwords;
  const int word = (int)(w - e * kwords);
  const double th = thr[e];
  const double *Ee = E + e * (int64_t)nl * nl;
  uint32_t m = 0;
  for (int b = 0; b < 32; ++b) {
    const int bit = word * 32 + b;
    if (bit < nl * nl && fabs(Ee[bit]) > th) m |= 1u << b;
  }
  keep[w] = m;
}

__device__ __forceinline__ bool mbit(const uint32_t *m, int b) { return (m[b >> 5] >> (b & 31)) & 1u; }

// node-patch gather of the rows of buffered element matrices
template <int ND, int Q>
__global__ __launch_bounds__(64) void k_gather_emat(PatchArgs a, const double *__restrict__ E,
                                                    const double *__restrict__ thr) {
  extern __shared__ double acc[];
  const int64_t wg = blockIdx.x;
  const int64_t node = (wg & 7) * (int64_t)a.xcd_chunk + (wg >> 3);
  if (node >= a.n_nodes) return;
  constexpr int NL = ND * Q;
  const int lane = threadIdx.x;
  const int i0 = a.inc_ptr[node], ninc = a.inc_ptr[node + 1] - i0;
  const int ns = a.nslots[node];
  for (int t = lane; t < Q * Q * ns; t += 64) acc[t] = 0.0;
  __syncthreads();
  constexpr int PER = 64 / ND;
  const int kk = lane / ND, j = lane - kk * ND;
  for (int k0 = 0; k0 < ninc; k0 += PER) {
    const int k = k0 + kk;
    if (kk < PER && k < ninc) {
      const uint32_t inc = a.inc[i0 + k];
      const int64_t e = inc >> 5;
      const int i = inc & 31;
      const double th = thr[e];
      const double *Ee = E + e * (int64_t)NL * NL;
      const int s = a.slot[(int64_t)(i0 + k) * ND + j];
      for (int x = 0; x < Q; ++x)
        for (int y = 0; y < Q; ++y) {
          const double v = Ee[(i * Q + x) * NL + j * Q + y];
          if (fabs(v) > th) atomicAdd(&acc[(x * ns + s) * Q + y], v);
        }
    }
  }
  __syncthreads();
  const uint32_t *m = a.mask + (int64_t)i0 * a.mwpi;
  for (int x = 0; x < Q; ++x) {
    int64_t base = a.indptr[node * Q + x];
    const int off = x * ns * Q, len = ns * Q;
    for (int c0 = 0; c0 < len; c0 += 64) {
      const int d = c0 + lane;
      const bool pr = d < len && mbit(m, off + d);
      const uint64_t bal = __ballot(pr);
      if (pr) a.vals[base + __popcll(bal & ((1ull << lane) - 1ull))] = acc[off + d];
      base += __popcll(bal);
    }
  }
}

// V[node*Q + x] = sum over incidences (e, i) of res[(e*nd + i)*Q + x]
__global__ __launch_bounds__(256) void k_gather_vec(int64_t n_nodes, int nd, int Q,
                                                    const int32_t *__restrict__ inc_ptr,
                                                    const uint32_t *__restrict__ inc,
                                                    const double *__restrict__ res,
                                                    double *__restrict__ V) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  for (int x = 0; x < Q; ++x) {
    double s = 0.0;
    for (int k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
      const uint32_t c = inc[k];
      s += res[((int64_t)(c >> 5) * nd + (c & 31)) * Q + x];
    }
    V[n * Q + x] = s;
  }
}

}  // namespace

hipError_t launch_nh_element(hipStream_t s, const EbArgs &a) {
  if (a.n_cv == 0) return hipSuccess;
  k_nh_element<<<(unsigned)a.n_cv, kEbThreads, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_keep_emat(hipStream_t s, int64_t n_cv, int nl, int kwords, const double *E,
                            const double *thr, uint32_t *keep) {
  const int64_t n = n_cv * kwords;
  if (n == 0) return hipSuccess;
  k_keep_emat<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n_cv, nl, kwords, E, thr, keep);
  return hipGetLastError();
}

hipError_t launch_gather_emat(hipStream_t s, const PatchArgs &a0, const double *E, const double *thr,
                              int max_nslots) {
  if (a0.n_nodes == 0) return hipSuccess;
  PatchArgs a = a0;
  a.xcd_chunk = (int)((a.n_nodes + 7) / 8);
  const unsigned grid = (unsigned)(8 * (int64_t)a.xcd_chunk);
  const size_t lds = (size_t)a.qdim * a.qdim * max_nslots * sizeof(double);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  if (a.nd == 27 && a.qdim == 3) k_gather_emat<27, 3><<<grid, 64, lds, s>>>(a, E, thr);
  else if (a.nd == 8 && a.qdim == 3) k_gather_emat<8, 3><<<grid, 64, lds, s>>>(a, E, thr);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_gather_vec(hipStream_t s, int64_t n_nodes, int nd, int Q, const int32_t *inc_ptr,
                             const uint32_t *inc, const double *res, double *V) {
  if (n_nodes == 0) return hipSuccess;
  k_gather_vec<<<(unsigned)((n_nodes + 255) / 256), 256, 0, s>>>(n_nodes, nd, Q, inc_ptr, inc, res, V);
  return hipGetLastError();
}

}  // namespace gfx
