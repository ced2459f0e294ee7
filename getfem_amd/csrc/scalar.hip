// The order-0 / order-1 route of the workspace: potentials and their
// derivatives (ga_workspace::assembly(0 | 1),
// src/getfem_generic_assembly_workspace.cc:791-936, with the potential and
// vector assembly instructions of
// src/getfem_generic_assembly_compile_and_exec.cc:4669-4735) for integrands
// the compile stage turned into polynomial programs.
//
// k_scalar_elem: one thread per element (affine tetrahedra, or Q1 hexahedra
// with the geometry at every point).  At
// each integration point the thread fills its row of a symbol table in LDS
// (values and gradients of the fem fields the program reads, then the
// fixed-size constants), evaluates the program's monomials into its output
// coefficients and integrates them: order 0, the potential's element value;
// order 1, the element vector of the differentiated variable,
//   elem[i, a] += w_q |J| (c_a phi_i + sum_b c_(q + 3a + b) (B grad phi_i)_b).
// The program is the same for every thread (uniform loads).  Element vectors
// go to a buffer and are summed per dof over the node's incidences (no
// atomics); element values are summed by a fixed-order two-stage reduction,
// so both results are deterministic.
#include <hip/hip_runtime.h>

#include "gfx_internal.h"

namespace gfx {
namespace {

constexpr int kScThreads = 64;
constexpr int kScRedThreads = 256;
constexpr int kScRedBlocks = 256;

__global__ __launch_bounds__(kScThreads) void k_scalar_elem(ScalarArgs a) {
  __shared__ double tab[kScThreads][kScalarMaxSyms + 1];
  __shared__ double oc[kScThreads][kScalarMaxOut + 1];
  const int64_t e = (int64_t)blockIdx.x * kScThreads + threadIdx.x;
  if (e >= a.n_cv) return;  // no barrier below
  double *T = tab[threadIdx.x];
  double *O = oc[threadIdx.x];
  // geometry: once for an affine tetrahedron, at every point for a hexahedron
  const int32_t *cv = a.cvx + (int64_t)a.nv * e;
  double B[9], aJ = 0.0;
  bool ok = poly_geometry(a.pts, cv, a.nv, a.gN, B, aJ);
  for (int q = 1; ok && a.gN && q < a.nq; ++q) {
    double Bq[9], aq;
    ok = poly_geometry(a.pts, cv, a.nv, a.gN + (int64_t)q * a.nv * 3, Bq, aq);
  }
  if (!ok) {
    atomicOr(a.err, 1);
    return;
  }
  for (int c = 0; c < a.nconst; ++c) T[a.const_off + c] = a.consts[c];
  const ScalarField &vf = a.f[a.vf];
  const int nl = a.order == 1 ? vf.nd * vf.qdim : 0;
  double *el = a.order == 1 ? a.elem + e * nl : nullptr;
  for (int t = 0; t < nl; ++t) el[t] = 0.0;
  double acc = 0.0;
  for (int q = 0; q < a.nq; ++q) {
    if (a.gN && q > 0) poly_geometry(a.pts, cv, a.nv, a.gN + (int64_t)q * a.nv * 3, B, aJ);
    const double wJ = a.qw[q] * aJ;
    for (int fi = 0; fi < a.nf; ++fi) {
      const ScalarField &f = a.f[fi];
      const int Q = f.qdim;
      const int32_t *dofs = f.dofs + e * f.nd;
      for (int c = 0; c < Q; ++c) {
        double v = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
        for (int i = 0; i < f.nd; ++i) {
          const double u = f.U[(int64_t)dofs[i] + c];
          const double *g = f.gref + ((int64_t)q * f.nd + i) * 3;
          v += f.phi[q * f.nd + i] * u;
          r0 += g[0] * u; r1 += g[1] * u; r2 += g[2] * u;
        }
        T[f.off + c] = v;
        for (int b = 0; b < 3; ++b) T[f.off + Q + 3 * c + b] = B[3 * b] * r0 + B[3 * b + 1] * r1 + B[3 * b + 2] * r2;
      }
    }
    for (int o = 0; o < a.nout; ++o) O[o] = 0.0;
    for (int m = 0; m < a.nm; ++m) {
      double v = a.mcoef[m];
      for (int s = a.moff[m]; s < a.moff[m + 1]; ++s) v *= T[a.msym[s]];
      O[a.mout[m]] += v;
    }
    if (a.order == 0) {
      acc += wJ * O[0];
      continue;
    }
    const int Q = vf.qdim;
    for (int i = 0; i < vf.nd; ++i) {
      const double phi = vf.phi[q * vf.nd + i];
      const double *h = vf.gref + ((int64_t)q * vf.nd + i) * 3;
      double g[3];
      for (int b = 0; b < 3; ++b) g[b] = B[3 * b] * h[0] + B[3 * b + 1] * h[1] + B[3 * b + 2] * h[2];
      for (int c = 0; c < Q; ++c) {
        const double *oc3 = O + Q + 3 * c;
        el[i * Q + c] += wJ * (O[c] * phi + oc3[0] * g[0] + oc3[1] * g[1] + oc3[2] * g[2]);
      }
    }
  }
  if (a.order == 0) a.elem[e] = acc;
}

__device__ __forceinline__ double sc_block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kScRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  return sh[0];
}

// partial sums of x over a fixed grid (each block a fixed stride of x), then
// one block sums the partials: the order never depends on scheduling
__global__ __launch_bounds__(kScRedThreads) void k_sc_partial(int64_t n, const double *__restrict__ x,
                                                              double *__restrict__ part) {
  __shared__ double sh[kScRedThreads];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kScRedThreads + threadIdx.x; i < n;
       i += (int64_t)kScRedBlocks * kScRedThreads)
    s += x[i];
  s = sc_block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kScRedThreads) void k_sc_final(const double *__restrict__ part,
                                                            double *__restrict__ out) {
  __shared__ double sh[kScRedThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < kScRedBlocks; i += kScRedThreads) s += part[i];
  s = sc_block_sum(s, sh);
  if (threadIdx.x == 0) *out = s;
}

}  // namespace

int64_t scalar_partials() { return kScRedBlocks; }

hipError_t launch_scalar_elem(hipStream_t s, const ScalarArgs &a) {
  if (a.n_cv == 0) return hipSuccess;
  if (a.nf < 1 || a.nf > kScalarMaxFields || a.nout > kScalarMaxOut || a.vf < 0 || a.vf >= a.nf)
    return hipErrorInvalidValue;
  const unsigned grid = (unsigned)((a.n_cv + kScThreads - 1) / kScThreads);
  k_scalar_elem<<<grid, kScThreads, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_scalar_sum(hipStream_t s, int64_t n, const double *x, double *part, double *out) {
  k_sc_partial<<<kScRedBlocks, kScRedThreads, 0, s>>>(n, x, part);
  k_sc_final<<<1, kScRedThreads, 0, s>>>(part, out);
  return hipGetLastError();
}

}  // namespace gfx
