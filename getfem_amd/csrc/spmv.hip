// y = K x over the assembled CSR (rows = test dofs, sorted columns).  G lanes
// of a wavefront share one row: each lane walks the row with stride G, then
// the G partial sums are combined by shuffles.  G is picked from the mean row
// length so that short rows (P1) do not leave most lanes idle.
#include <hip/hip_runtime.h>

#include "gfx_internal.h"

namespace gfx {
namespace {

constexpr int kSpmvThreads = 256;

template <int G>
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
  if (row < n && lane == 0) y[row] = s;
}

template <int G>
hipError_t spmv_g(hipStream_t s, int64_t n, const int64_t *indptr, const int32_t *indices,
                  const double *vals, const double *x, double *y) {
  const int64_t threads = n * G;
  const unsigned grid = (unsigned)((threads + kSpmvThreads - 1) / kSpmvThreads);
  k_spmv<G><<<grid, kSpmvThreads, 0, s>>>(n, indptr, indices, vals, x, y);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_spmv(hipStream_t s, int64_t n, int64_t nnz, const int64_t *indptr,
                       const int32_t *indices, const double *vals, const double *x, double *y) {
  if (n == 0) return hipSuccess;
  const double mean = (double)nnz / (double)n;
  if (mean <= 12) return spmv_g<4>(s, n, indptr, indices, vals, x, y);
  if (mean <= 24) return spmv_g<8>(s, n, indptr, indices, vals, x, y);
  if (mean <= 64) return spmv_g<16>(s, n, indptr, indices, vals, x, y);
  return spmv_g<32>(s, n, indptr, indices, vals, x, y);
}

}  // namespace gfx
