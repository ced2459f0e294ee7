// Adjoint of the SIMP compliance J(rho, X) = u^T K(rho, X) u for P1 vector
// fields on affine tetrahedra, E_e = emin + rho_e^p (1 - emin):
//   J_e        = E_e |T| W(H),  W = lambda tr(H)^2 + 2 mu eps:eps, H = grad u
//   dJ/drho_e  = E'(rho_e) |T| W
//   dJ/dx_v    = sum_e E_e |T| (W I - H^T P) g_v,   P = dW/dH (shape derivative)
//   dJ/du_v    = sum_e E_e |T| P g_v   (= 2 K u)
// In GWFL terms this is the order-1 assembly of the order-0 potential with
// rho (and the pulled-back coordinates) as variables
// (src/getfem_generic_assembly_workspace.cc:520-541, :832-847).
// One thread per element writes per-(element, vertex) contributions; the
// node sums are gathers over node -> element incidences (deterministic, no
// floating-point atomics), and J is a fixed-order two-stage reduction.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gfx_internal.h"

namespace gfx {

namespace {

constexpr int kRedThreads = 256;

// (s, c) += (b, bc) with Knuth's TwoSum on the leading parts
__device__ __forceinline__ void two_sum_into(double &s, double &c, double b, double bc) {
  const double t = s + b;
  const double bb = t - s;
  const double err = (s - (t - bb)) + (b - bb);
  s = t;
  c += err + bc;
}

__global__ __launch_bounds__(256) void k_compliance_elem(
    int64_t n_cv, const double *__restrict__ X, const int32_t *__restrict__ cvx,
    const int32_t *__restrict__ dofs, const double *__restrict__ u,
    const double *__restrict__ rho, double lam, double mu, double emin, double penal,
    double wsum, double *__restrict__ Jpart, double *__restrict__ drho,
    double *__restrict__ cX, double *__restrict__ cU, int *__restrict__ err) {
  __shared__ double red[kRedThreads], redc[kRedThreads];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double Je = 0.0;
  if (e < n_cv) {
    const int4 v = *reinterpret_cast<const int4 *>(cvx + 4 * e);
    const int vv[4] = {v.x, v.y, v.z, v.w};
    double x[4][3];
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r) x[i][r] = X[3 * (int64_t)vv[i] + r];
    double k[9];  // column-major K(r,c)
    for (int r = 0; r < 3; ++r) {
      k[r] = x[1][r] - x[0][r];
      k[r + 3] = x[2][r] - x[0][r];
      k[r + 6] = x[3][r] - x[0][r];
    }
    const double a0 = k[4] * k[8] - k[5] * k[7];
    const double a1 = k[5] * k[6] - k[3] * k[8];
    const double a2 = k[3] * k[7] - k[4] * k[6];
    const double J = k[0] * a0 + k[1] * a1 + k[2] * a2;
    if (J == 0.0) atomicOr(err, 1);
    double B[3][3];  // B = K^{-T}
    B[0][0] = a0 / J; B[1][0] = a1 / J; B[2][0] = a2 / J;
    B[0][1] = (k[2] * k[7] - k[1] * k[8]) / J;
    B[1][1] = (k[0] * k[8] - k[2] * k[6]) / J;
    B[2][1] = (k[1] * k[6] - k[0] * k[7]) / J;
    B[0][2] = (k[1] * k[5] - k[2] * k[4]) / J;
    B[1][2] = (k[2] * k[3] - k[0] * k[5]) / J;
    B[2][2] = (k[0] * k[4] - k[1] * k[3]) / J;
    // P1 reference gradients: -1,-1,-1 / e1 / e2 / e3
    double g[4][3];
    for (int r = 0; r < 3; ++r) {
      g[1][r] = B[r][0];
      g[2][r] = B[r][1];
      g[3][r] = B[r][2];
      g[0][r] = -(B[r][0] + B[r][1] + B[r][2]);
    }
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < 4; ++i) {
      const int64_t d = dofs[4 * e + i];
      for (int a = 0; a < 3; ++a) {
        const double ua = u[d + a];
        for (int b = 0; b < 3; ++b) H[a][b] += ua * g[i][b];
      }
    }
    const double tr = H[0][0] + H[1][1] + H[2][2];
    double W = lam * tr * tr, P[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double eps = 0.5 * (H[a][b] + H[b][a]);
        W += 2.0 * mu * eps * eps;
        P[a][b] = 4.0 * mu * eps + (a == b ? 2.0 * lam * tr : 0.0);
      }
    const double w = fabs(J) * wsum;
    const double r = rho[e];
    const double E = emin + pow(r, penal) * (1.0 - emin);
    const double dE = penal * pow(r, penal - 1.0) * (1.0 - emin);
    Je = E * w * W;
    if (drho) drho[e] = dE * w * W;
    // M = W I - H^T P
    double M[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        M[a][b] = (a == b ? W : 0.0) - (H[0][a] * P[0][b] + H[1][a] * P[1][b] + H[2][a] * P[2][b]);
    const double s = E * w;
    for (int i = 0; i < 4; ++i)
      for (int a = 0; a < 3; ++a) {
        if (cX) cX[(4 * e + i) * 3 + a] = s * (M[a][0] * g[i][0] + M[a][1] * g[i][1] + M[a][2] * g[i][2]);
        if (cU)  // P = dW/dH already carries the factor 2 of the quadratic form
          cU[(4 * e + i) * 3 + a] = s * (P[a][0] * g[i][0] + P[a][1] * g[i][1] + P[a][2] * g[i][2]);
      }
  }
  // compensated (TwoSum) tree reduction: J is accurate to a few ulps so that
  // finite differences of J resolve its derivatives
  red[threadIdx.x] = Je;
  redc[threadIdx.x] = 0.0;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) two_sum_into(red[threadIdx.x], redc[threadIdx.x], red[threadIdx.x + o],
                                           redc[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Jpart[2 * blockIdx.x] = red[0];
    Jpart[2 * blockIdx.x + 1] = redc[0];
  }
}

// fixed-order compensated sum of the block partials (value, compensation)
__device__ __forceinline__ void sum_partials_block(const double *__restrict__ part, int64_t n,
                                                   double *__restrict__ out) {
  __shared__ double red[kRedThreads], redc[kRedThreads];
  double s = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kRedThreads) two_sum_into(s, c, part[2 * i], part[2 * i + 1]);
  red[threadIdx.x] = s;
  redc[threadIdx.x] = c;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) two_sum_into(red[threadIdx.x], redc[threadIdx.x], red[threadIdx.x + o],
                                           redc[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] + redc[0];
}

// out[node*3 + a] = sum over incidences (e, i) of c[(4e+i)*3 + a]
__device__ __forceinline__ void gather3_node(int64_t n, int64_t n_nodes, const int32_t *__restrict__ inc_ptr,
                                             const uint32_t *__restrict__ inc,
                                             const double *__restrict__ c, double *__restrict__ out) {
  if (n >= n_nodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
    const uint32_t x = inc[k];
    const int64_t idx = ((int64_t)(x >> 5) * 4 + (x & 31)) * 3;
    s0 += c[idx];
    s1 += c[idx + 1];
    s2 += c[idx + 2];
  }
  out[3 * n] = s0;
  out[3 * n + 1] = s1;
  out[3 * n + 2] = s2;
}

// The three reductions after the element kernel are independent, so they run
// as one launch: blocks [0, nbx) gather dJ/dX per mesh point, the next nbu
// blocks gather dJ/du per dof node, and the last block sums J.
__global__ __launch_bounds__(kRedThreads) void k_compliance_finish(
    int64_t nbx, int64_t nx, const int32_t *__restrict__ xptr, const uint32_t *__restrict__ xinc,
    const double *__restrict__ cX, double *__restrict__ outX, int64_t nbu, int64_t nu,
    const int32_t *__restrict__ uptr, const uint32_t *__restrict__ uinc, const double *__restrict__ cU,
    double *__restrict__ outU, const double *__restrict__ Jpart, int64_t nbJ, double *__restrict__ J) {
  int64_t b = blockIdx.x;
  if (b < nbx) {
    gather3_node(b * kRedThreads + threadIdx.x, nx, xptr, xinc, cX, outX);
    return;
  }
  b -= nbx;
  if (b < nbu) {
    gather3_node(b * kRedThreads + threadIdx.x, nu, uptr, uinc, cU, outU);
    return;
  }
  sum_partials_block(Jpart, nbJ, J);
}

}  // namespace

// The allocation-free entry point's scratch: the block partials of J and the
// per-(element, vertex) contributions to dJ/dX and dJ/du, in one file-scope
// device array of kComplianceScratchBytes.  Every GPU the library runs on
// holds its own copy until the process ends; host code reaches a device's
// copy only through the address hipGetSymbolAddress returns with that device
// current (the array's name is a host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[kComplianceScratchBytes / 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices];
}  // namespace

hipError_t compliance_scratch(int device, double **ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double *>(p);
  }
  *ptr = g_scratch[device];
  return hipSuccess;
}

int64_t compliance_nblocks(int64_t n_cv) { return (n_cv + 255) / 256; }

hipError_t launch_compliance(hipStream_t s, int64_t n_cv, const double *X, const int32_t *cvx,
                             const int32_t *dofs, const double *u, const double *rho, double lam,
                             double mu, double emin, double penal, double wsum, double *Jpart,
                             double *J, double *drho, double *cX, double *cU, int *err) {
  const int64_t nb = compliance_nblocks(n_cv);
  if (nb == 0) return hipErrorInvalidValue;
  k_compliance_elem<<<(unsigned)nb, 256, 0, s>>>(n_cv, X, cvx, dofs, u, rho, lam, mu, emin, penal,
                                                 wsum, Jpart, drho, cX, cU, err);
  return hipGetLastError();
}

hipError_t launch_compliance_finish(hipStream_t s, int64_t n_cv, int64_t nx, const int32_t *xptr,
                                    const uint32_t *xinc, const double *cX, double *outX, int64_t nu,
                                    const int32_t *uptr, const uint32_t *uinc, const double *cU,
                                    double *outU, const double *Jpart, double *J) {
  // a gather runs only when its output is requested (outX / outU non-null)
  const int64_t nbx = outX ? (nx + kRedThreads - 1) / kRedThreads : 0;
  const int64_t nbu = outU ? (nu + kRedThreads - 1) / kRedThreads : 0;
  k_compliance_finish<<<(unsigned)(nbx + nbu + 1), kRedThreads, 0, s>>>(
      nbx, nx, xptr, xinc, cX, outX, nbu, nu, uptr, uinc, cU, outU, Jpart, compliance_nblocks(n_cv), J);
  return hipGetLastError();
}

}  // namespace gfx
