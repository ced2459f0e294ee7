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

#include <type_traits>
#include <utility>

#include "gfx_internal.h"

namespace gfx {

namespace {

constexpr int kEbThreads = 256;
constexpr int kNd = 27, kNl = 81, kNv = 8;

// One workgroup per element.
// Phase A, every integration point at once (FP64):
//   K_q = G grad(N)(q), J_q, B_q = K_q^{-T}, wJ_q = |J_q| w_q     (trilinear map)
//   g[q][i] = B_q grad(phi_i)(q);  H_q = sum_i U_i (x) g[q][i];  F_q = I + H_q
//   S_q = mu (I - C^-1) + lambda/2 ln(I3) C^-1,  A = dS/dE  (Bonet law)
//   re[(i,a)] = sum_q wJ_q (F_q S_q g[q][i])_a
// Phase B, the tangent K[(i,a),(j,b)] = sum_q g[q][i][c] D_q[(a,c),(b,n)]
// g[q][j][n] with D_q[(a,c),(b,n)] = wJ_q (delta_ab S[n][c] + F[a][k]
// A[k c m n] F[b][m]) (the GWFL derivative of the residual):
//   in the spatial form (h = F^{-T} g, see k_nh_element): per node pair and
//   point, two outer products of spatial gradients and a dot product of
//   material ones, as products on the matrix cores in FP64 (FP64 form) or
//   FP32 (FP32 form), upper tiles only (nh_phase_b_mfma).
constexpr int kNq = 27;
constexpr int kKeepWords = (kNl * kNl + 31) / 32;  // keep-bit words per element (206)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// maximum over the workgroup's threads: shuffles within each wavefront, then
// the wavefronts' maxima through LDS (one barrier instead of a tree)
__device__ __forceinline__ double block_max(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int t = threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < kEbThreads / 64; ++w) m = fmax(m, red[w]);
  return m;
}

// Lane layout of the 16x16x4 matrix-core products (v_mfma_f64_16x16x4_f64,
// v_mfma_f32_16x16x4_f32): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k
// (both instructions); D[i][j] in lane j + 16 (i / 4), element i % 4 for the
// FP32 one, but in lane j + 16 (i % 4), element i / 4 for the FP64 one
// (BdbMfma<T>::row gives the row of lane group g and element r).
template <typename T>
struct BdbMfma;
template <>
struct BdbMfma<double> {
  typedef __attribute__((__vector_size__(4 * sizeof(double)))) double v4;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * r + g; }
};
template <>
struct BdbMfma<float> {
  typedef f32x4 v4;
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// Phase B in the spatial form (see k_nh_element) on the matrix cores, in T
// (FP64: v_mfma_f64_16x16x4_f64; FP32: v_mfma_f32_16x16x4_f32, exact
// products and FP32 sums).  With H_c[i][q] = h_qic (node i, point q), the
// (a, b) component block of K over the node pairs is
//   K_ab = H_a diag(alpha) H_b^T + H_b diag(beta) H_a^T + delta_ab K3,
//   K3 = sum_c G_c M_c^T,  G_c[i][q] = g_qic,  M_c[j][q] = (gamma_q g_qj + pen_q C_q g_qj)_c,
// so both outer-product terms of a block go into one accumulator and no
// value moves between lanes.  A tile is a 16 x 16 node tile (I, J) in
// {0, 1}^2 (27 nodes and 5 zero pads) of one component pair (a, b); the k
// axis is the 27 points and one zero (7 steps of 4).  hC ([q][c][32 nodes],
// zero pads) holds h as the fragments read it: one LDS read per value.
// K is symmetric: with P = 2 a + I and Q = 2 b + J only the 21 tiles P <= Q
// are formed (the others are their transposes), and of K3 the node tiles
// (0, 0), (0, 1), (1, 1).  Wavefront w < 3 forms K3 tile w and keeps it in
// its accumulators beside the three tiles (a, a) of the same node tiles, to
// which it is added, and two of the twelve tiles with a < b; wavefront 3
// forms the other six (91, 91, 91, 84 products).  The tiles stay in the
// accumulators until phase A's arrays and hC are free; then E is staged in
// LDS over them (node rows 0-15, then 16-26), each tile written with its
// transpose, and each staged run, contiguous in E, is streamed by the
// workgroup.
constexpr int kNhQ = 28, kNhN = 32;

// tile N (0..20): wavefront N / 5 (N < 15) takes N % 5 = 0..2 as the (a, a)
// tiles of its K3 node tile and 3..4 as with-a<b tiles 2 (N / 5) + (N % 5 - 3);
// N >= 15 are the with-a<b tiles 6..11 of wavefront 3.  With-a<b tile n:
// (a, b) = (0,1), (0,2), (1,2) for n / 4 = 0, 1, 2, (I, J) = (n % 4 / 2, n % 2).
template <int N>
struct NhTile {
  static constexpr bool diag = N < 15 && N % 5 < 3;  // an (a, a) tile: K3 added
  static constexpr int w = N < 15 ? N / 5 : 3;
  static constexpr int n = N < 15 ? 2 * w + (N % 5 - 3) : 6 + (N - 15);  // with-a<b index (when !diag)
  static constexpr int a = diag ? N % 5 : (n / 4 == 2 ? 1 : 0);
  static constexpr int b = diag ? N % 5 : (n / 4 == 0 ? 1 : 2);
  static constexpr int I = diag ? w / 2 : n % 4 / 2;
  static constexpr int J = diag ? (w + 1) / 2 : n % 2;
  static constexpr int P = 2 * a + I, Q = 2 * b + J;
  static_assert(P <= Q, "upper tiles only");
};
static_assert(NhTile<0>::P == 0 && NhTile<0>::Q == 0 && NhTile<5>::P == 0 && NhTile<5>::Q == 1 &&
              NhTile<14>::a == 0 && NhTile<14>::b == 2 && NhTile<14>::I == 0 && NhTile<14>::J == 1 &&
              NhTile<20>::a == 1 && NhTile<20>::b == 2 && NhTile<20>::I == 1 && NhTile<20>::J == 1,
              "tile table");

// one k-step of tile N: f[c][I] = h at this lane's point, component c, node 16 I + lr
template <typename T, int N>
__device__ __forceinline__ void nh_w_mma(typename BdbMfma<T>::v4 &acc, T al, T be, const T (&f)[3][2]) {
  typedef NhTile<N> D;
  acc = BdbMfma<T>::mma(al * f[D::a][D::I], f[D::b][D::J], acc);
  acc = BdbMfma<T>::mma(be * f[D::b][D::I], f[D::a][D::J], acc);
}
template <typename T, int T0, int... U>
__device__ __forceinline__ void nh_w_step(std::integer_sequence<int, U...>, typename BdbMfma<T>::v4 (&acc)[6], T al,
                                          T be, const T (&f)[3][2]) {
  (nh_w_mma<T, T0 + U>(acc[U], al, be, f), ...);
}

// the products of tiles T0 .. T0 + NT - 1 (accumulators acc[0 .. NT - 1])
template <typename T, int T0, int NT>
__device__ __forceinline__ void nh_w_tiles(int lr, int lg, const T (*hC)[3][kNhN], const T (&al)[kNhQ / 4],
                                           const T (&be)[kNhQ / 4], typename BdbMfma<T>::v4 (&acc)[6]) {
#pragma unroll
  for (int u = 0; u < 6; ++u) acc[u] = typename BdbMfma<T>::v4{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int st = 0; st < kNhQ / 4; ++st) {
    T f[3][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int I = 0; I < 2; ++I) f[c][I] = hC[4 * st + lg][c][16 * I + lr];
    nh_w_step<T, T0>(std::make_integer_sequence<int, NT>{}, acc, al[st], be[st], f);
  }
}

// tile N's entries in node rows [n0, n1) of E into the stage (E's
// node-blocked layout, entry ((3 (i - n0) + a) 3 + b) 27 + j for K[(i,a),(j,b)]),
// K3 added to (a, a) tiles, the transpose written for P < Q.  Branch-free:
// a lane whose entry is outside the slab (or a zero pad) writes its own slot
// of the trash row behind the stage instead.
template <typename T, int N>
__device__ __forceinline__ void nh_w_store1(int lr, int lg, int n0, int n1, int trash,
                                            const typename BdbMfma<T>::v4 &acc,
                                            const typename BdbMfma<T>::v4 &k3, T *stage) {
  typedef NhTile<N> D;
  const int j = 16 * D::J + lr;
  const bool jv = j < kNd, jin = jv && j >= n0 && j < n1;
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int i = 16 * D::I + BdbMfma<T>::row(lg, x);
    const bool iv = i < kNd, iin = iv && i >= n0 && i < n1;
    const T v = D::diag ? acc[x] + k3[x] : acc[x];
    stage[iin && jv ? ((3 * (i - n0) + D::a) * 3 + D::b) * kNd + j : trash] = v;
    if (D::P != D::Q) stage[jin && iv ? ((3 * (j - n0) + D::b) * 3 + D::a) * kNd + i : trash] = v;
  }
}
template <typename T, int T0, int... U>
__device__ __forceinline__ void nh_w_store_seq(std::integer_sequence<int, U...>, int lr, int lg, int n0, int n1,
                                               int trash, const typename BdbMfma<T>::v4 (&acc)[6],
                                               const typename BdbMfma<T>::v4 &k3, T *stage) {
  (nh_w_store1<T, T0 + U>(lr, lg, n0, n1, trash, acc[U], k3, stage), ...);
}
template <typename T, int T0, int NT>
__device__ __forceinline__ void nh_w_store(int lr, int lg, int n0, int n1, int trash,
                                           const typename BdbMfma<T>::v4 (&acc)[6],
                                           const typename BdbMfma<T>::v4 &k3, T *stage) {
  nh_w_store_seq<T, T0>(std::make_integer_sequence<int, NT>{}, lr, lg, n0, n1, trash, acc, k3, stage);
}

// the stage aliases phase A's arrays and hC: written only after the
// workgroup barrier that follows the last read of them
// this lane's largest |entry| of tiles T0 .. T0 + NT - 1 (K3 added to the
// (a, a) tiles; the zero pads add nothing)
template <typename T, int N>
__device__ __forceinline__ double nh_w_max1(const typename BdbMfma<T>::v4 &acc, const typename BdbMfma<T>::v4 &k3) {
  double m = 0.0;
#pragma unroll
  for (int x = 0; x < 4; ++x) m = fmax(m, (double)fabs(NhTile<N>::diag ? acc[x] + k3[x] : acc[x]));
  return m;
}
template <typename T, int T0, int... U>
__device__ __forceinline__ double nh_w_max_seq(std::integer_sequence<int, U...>,
                                               const typename BdbMfma<T>::v4 (&acc)[6],
                                               const typename BdbMfma<T>::v4 &k3) {
  double m = 0.0;
  ((m = fmax(m, nh_w_max1<T, T0 + U>(acc[U], k3))), ...);
  return m;
}
template <typename T, int T0, int NT>
__device__ __forceinline__ double nh_w_max(const typename BdbMfma<T>::v4 (&acc)[6],
                                           const typename BdbMfma<T>::v4 &k3) {
  return nh_w_max_seq<T, T0>(std::make_integer_sequence<int, NT>{}, acc, k3);
}

template <typename T, bool KEEP>
__device__ __forceinline__ void nh_phase_b_mfma(int t, int64_t e, const double (*coef)[2], const double (*Mq)[9],
                                                const T (*hC)[3][kNhN], T *stage, uint32_t *skeep, double *red,
                                                const EbArgs &a) {
  typedef typename BdbMfma<T>::v4 v4;
  // the wavefront index as a scalar: the branches on it are uniform
  const int w = __builtin_amdgcn_readfirstlane(t / 64), l = t % 64, lr = l % 16, lg = l / 16;
  // K3 node tile (w / 2, (w + 1) / 2) for wavefronts 0-2, in the spatial
  // gradients: K3[i][j] = sum_q h_qi^T M_q h_qj, k = (c, q): A = h_qic,
  // B = (M_q h_qj)_c (zero pads in hC make the pads' products zero)
  v4 k3 = v4{T(0), T(0), T(0), T(0)};
  if (w < 3) {
    const int i = 16 * (w / 2) + lr, j = 16 * ((w + 1) / 2) + lr;
#pragma unroll
    for (int st = 0; st < kNhQ / 4; ++st) {
      const int q = 4 * st + lg;
      const double *M = Mq[min(q, kNq - 1)];
      const T h0 = hC[q][0][j], h1 = hC[q][1][j], h2 = hC[q][2][j];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T m = (T)M[3 * c] * h0 + (T)M[3 * c + 1] * h1 + (T)M[3 * c + 2] * h2;
        k3 = BdbMfma<T>::mma(hC[q][c][i], m, k3);
      }
    }
  }
  // this lane's alpha_q and beta_q at q = 4 st + lg (hC is zero at q = 27)
  T al[kNhQ / 4], be[kNhQ / 4];
#pragma unroll
  for (int st = 0; st < kNhQ / 4; ++st) {
    const int qc = min(4 * st + lg, kNq - 1);
    al[st] = (T)coef[qc][0];
    be[st] = (T)coef[qc][1];
  }
  v4 acc[6];
  if (w == 0) nh_w_tiles<T, 0, 5>(lr, lg, hC, al, be, acc);
  else if (w == 1) nh_w_tiles<T, 5, 5>(lr, lg, hC, al, be, acc);
  else if (w == 2) nh_w_tiles<T, 10, 5>(lr, lg, hC, al, be, acc);
  else nh_w_tiles<T, 15, 6>(lr, lg, hC, al, be, acc);
  // threshold 1e-14 max|K| (all-zero element skipped): over the streamed
  // entries, or, when the FP64 pass also sets the keep bits while it streams
  // E (pattern build, re-validation), from the accumulators before any entry
  // is written
  constexpr bool kb = KEEP;
  double mx = 0.0, th = 0.0;
  if constexpr (kb) {
    if (w == 0) mx = nh_w_max<T, 0, 5>(acc, k3);
    else if (w == 1) mx = nh_w_max<T, 5, 5>(acc, k3);
    else if (w == 2) mx = nh_w_max<T, 10, 5>(acc, k3);
    else mx = nh_w_max<T, 15, 6>(acc, k3);
    const double mxe = block_max(mx, red);
    th = mxe > 0.0 ? 1e-14 * mxe : INFINITY;
    if (t == 0) a.thr[e] = th;
    for (int x = t; x < a.kwords; x += kEbThreads) skeep[x] = 0u;
  }
  T *E = reinterpret_cast<T *>(a.E) + e * kNl * kNl;
  // two stages: node rows 0-15, then 16-26; the trash row (a slot per lane)
  // behind the larger one.  (FP32 would fit one stage of all 27 node rows;
  // that variant gave wrong values with this compiler at -O3, with a fill of
  // the stage before the tile stores or at -O1 it did not, and the cause was
  // not found: both precisions use the two stages, which are checked.)
  constexpr int kSlabs = 2;
  const int trash = 3 * 16 * kNl + l;
#pragma unroll
  for (int sl = 0; sl < kSlabs; ++sl) {
    const int n0 = 16 * sl, n1 = sl == 1 ? kNd : 16;
    __syncthreads();  // the last reads of what the stage overlays (or of the previous slab) are done
    if (w == 0) nh_w_store<T, 0, 5>(lr, lg, n0, n1, trash, acc, k3, stage);
    else if (w == 1) nh_w_store<T, 5, 5>(lr, lg, n0, n1, trash, acc, k3, stage);
    else if (w == 2) nh_w_store<T, 10, 5>(lr, lg, n0, n1, trash, acc, k3, stage);
    else nh_w_store<T, 15, 6>(lr, lg, n0, n1, trash, acc, k3, stage);
    __syncthreads();
    const int len = 3 * (n1 - n0) * kNl;
    T *dst = E + 3 * n0 * kNl;
    for (int x = t; x < len; x += kEbThreads) {
      const T v = stage[x];
      __builtin_nontemporal_store(v, dst + x);
      if constexpr (!kb) mx = fmax(mx, (double)fabs(v));
    }
    if constexpr (kb) {
      // keep bits, row-major (r, c) = (3 i + a, 3 j + b): one thread per
      // 32-bit word of this slab's rows reads its entries from the stage
      // (staged at ((3 i + a) 3 + b) 27 + j); the word shared with the other
      // slab gets its two halves from the same thread, a barrier apart
      const int b0 = 3 * n0 * kNl, b1 = 3 * n1 * kNl;
      const int w0 = b0 >> 5, w1 = (b1 + 31) >> 5;
      // two lanes per word (16 bits each, lanes 2m and 2m + 1: both in the
      // loop, since the bound is even), combined by a lane exchange
      for (int id = t; id < 2 * (w1 - w0); id += kEbThreads) {
        const int wd = w0 + (id >> 1), k0 = 16 * (id & 1);
        uint32_t bits = 0u;
        for (int k = k0; k < k0 + 16; ++k) {
          const int bit = 32 * wd + k;
          if (bit < b0 || bit >= b1) continue;
          const int r = bit / kNl, c = bit - r * kNl, j = c / 3;
          const int x = (3 * r + (c - 3 * j)) * kNd + j - b0;  // the slab starts at entry b0
          if (fabs((double)stage[x]) > th) bits |= 1u << k;
        }
        bits |= __shfl_xor(bits, 1);
        if ((id & 1) == 0) skeep[wd] |= bits;
      }
    }
  }
  if constexpr (!kb) {
    const double mxe = block_max(mx, red);
    if (t == 0) a.thr[e] = mxe > 0.0 ? 1e-14 * mxe : INFINITY;
  } else {
    __syncthreads();
    for (int x = t; x < a.kwords; x += kEbThreads) a.keep[e * a.kwords + x] = skeep[x];
  }
}

// phase A's per-element arrays; the material gradients gq (phase A) and
// then the spatial ones hC (phase B) share their space; phase B stages E
// over all of it
struct NhPhaseA {
  double X[kNv][3], Ue[kNd][3];
  double Bq[kNq][9], wJ[kNq], Fq[kNq][9], Sq[kNq][9], Cq[kNq][9], cf[kNq][2], Mq[kNq][9];
};
template <typename T>
union NhGrad {
  double gq[kNq][kNd][3];
  T hC[kNhQ][3][kNhN];
};
template <typename T>
union NhLds {
  struct {
    NhPhaseA A;
    NhGrad<T> G;
  } P;
  T stage[3 * 16 * kNl + 64];  // the slab and the trash row
};

template <typename Real, bool KEEP>
__global__ __launch_bounds__(kEbThreads, 4) void k_nh_element(EbArgs a) {
  __shared__ NhLds<Real> L;
  double(&X)[kNv][3] = L.P.A.X;
  double(&Ue)[kNd][3] = L.P.A.Ue;
  double(&Bq)[kNq][9] = L.P.A.Bq;
  double(&wJ)[kNq] = L.P.A.wJ;
  double(&Fq)[kNq][9] = L.P.A.Fq;
  double(&Sq)[kNq][9] = L.P.A.Sq;
  double(&Cq)[kNq][9] = L.P.A.Cq;
  double(&cf)[kNq][2] = L.P.A.cf;
  double(&Mq)[kNq][9] = L.P.A.Mq;
  double(&gq)[kNq][kNd][3] = L.P.G.gq;
  Real(&hC)[kNhQ][3][kNhN] = L.P.G.hC;  // phase B: the spatial gradients in the fragments' layout
  __shared__ double red[kEbThreads];
  __shared__ uint32_t skeep[KEEP ? kKeepWords : 1];  // the element's keep bits
  const int64_t e = blockIdx.x;
  if (e >= a.n_cv || a.nq != kNq) return;
  const int t = threadIdx.x;
  if (t < kNv * 3) X[t / 3][t % 3] = a.pts[3 * (int64_t)a.cvx[kNv * e + t / 3] + t % 3];
  if (t < kNl) Ue[t / 3][t % 3] = a.U[(int64_t)a.dofs[kNd * e + t / 3] + t % 3];
  const double lam = a.ca ? a.ca[e] : a.ca_c, mu = a.cb ? a.cb[e] : a.cb_c;
  __syncthreads();
  // A1: geometry per point
  if (t < kNq) {
    const double *gn = a.gN + (int64_t)t * kNv * 3;
    double k[9];  // column-major K(r,c)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double s = X[0][r] * gn[c];
        for (int v = 1; v < kNv; ++v) s += X[v][r] * gn[3 * v + c];
        k[r + 3 * c] = s;
      }
    const double a0 = k[4] * k[8] - k[5] * k[7];
    const double a1 = k[5] * k[6] - k[3] * k[8];
    const double a2 = k[3] * k[7] - k[4] * k[6];
    const double J = k[0] * a0 + k[1] * a1 + k[2] * a2;
    if (J == 0.0) atomicOr(a.err, 1);
    const double rJ = 1.0 / J;
    double *B = Bq[t];  // row-major B(r,c)
    B[0] = a0 * rJ; B[3] = a1 * rJ; B[6] = a2 * rJ;
    B[1] = (k[2] * k[7] - k[1] * k[8]) * rJ;
    B[4] = (k[0] * k[8] - k[2] * k[6]) * rJ;
    B[7] = (k[1] * k[6] - k[0] * k[7]) * rJ;
    B[2] = (k[1] * k[5] - k[2] * k[4]) * rJ;
    B[5] = (k[2] * k[3] - k[0] * k[5]) * rJ;
    B[8] = (k[0] * k[4] - k[1] * k[3]) * rJ;
    wJ[t] = fabs(J) * a.qw[t];
  }
  __syncthreads();
  // A2: real gradients
  for (int x = t; x < kNq * kNd; x += kEbThreads) {
    const int q = x / kNd, i = x % kNd;
    const double *h = a.gref + ((int64_t)q * kNd + i) * 3;
    const double h0 = h[0], h1 = h[1], h2 = h[2];
    const double *B = Bq[q];
#pragma unroll
    for (int c = 0; c < 3; ++c) gq[q][i][c] = B[3 * c] * h0 + B[3 * c + 1] * h1 + B[3 * c + 2] * h2;
  }
  __syncthreads();
  // A3: deformation gradient
  for (int x = t; x < kNq * 9; x += kEbThreads) {
    const int q = x / 9, r = (x % 9) / 3, c = x % 3;
    double s = 0.0;
    for (int i = 0; i < kNd; ++i) s += Ue[i][r] * gq[q][i][c];
    Fq[q][3 * r + c] = s + (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  // A4: the law at every point
  if (t < kNq) {
    const double *F = Fq[t];
    double C[9], Ci[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    const double c0 = C[4] * C[8] - C[5] * C[7];
    const double c1 = C[5] * C[6] - C[3] * C[8];
    const double c2 = C[3] * C[7] - C[4] * C[6];
    const double I3 = C[0] * c0 + C[1] * c1 + C[2] * c2;
    const double rI3 = 1.0 / I3;
    Ci[0] = c0 * rI3; Ci[3] = c1 * rI3; Ci[6] = c2 * rI3;
    Ci[1] = (C[2] * C[7] - C[1] * C[8]) * rI3;
    Ci[4] = (C[0] * C[8] - C[2] * C[6]) * rI3;
    Ci[7] = (C[1] * C[6] - C[0] * C[7]) * rI3;
    Ci[2] = (C[1] * C[5] - C[2] * C[4]) * rI3;
    Ci[5] = (C[2] * C[3] - C[0] * C[5]) * rI3;
    Ci[8] = (C[0] * C[4] - C[1] * C[3]) * rI3;
    const double f0 = F[4] * F[8] - F[5] * F[7], f1 = F[5] * F[6] - F[3] * F[8], f2 = F[3] * F[7] - F[4] * F[6];
    const double detF = F[0] * f0 + F[1] * f1 + F[2] * f2;
    const double ln3 = log(I3);
    for (int i = 0; i < 9; ++i) {
      double sv = mu * ((i % 4 == 0 ? 1.0 : 0.0) - Ci[i]) + 0.5 * lam * ln3 * Ci[i];
      if (detF <= 0.0) sv += 1e200 * C[i];
      Sq[t][i] = sv;
    }
    // the spatial form's inputs of phase B: F^{-1} (row-major, into Cq) and
    // the point's coefficients alpha = wJ lambda, beta = wJ (mu - lambda ln J),
    // gamma = wJ mu and the penalty pen = 1e200 wJ where det F <= 0
    const double rF = 1.0 / detF;
    double *Fi = Cq[t];
    Fi[0] = f0 * rF; Fi[3] = f1 * rF; Fi[6] = f2 * rF;
    Fi[1] = (F[2] * F[7] - F[1] * F[8]) * rF;
    Fi[4] = (F[0] * F[8] - F[2] * F[6]) * rF;
    Fi[7] = (F[1] * F[6] - F[0] * F[7]) * rF;
    Fi[2] = (F[1] * F[5] - F[2] * F[4]) * rF;
    Fi[5] = (F[2] * F[3] - F[0] * F[5]) * rF;
    Fi[8] = (F[0] * F[4] - F[1] * F[3]) * rF;
    const double w = wJ[t];
    cf[t][0] = w * lam;
    cf[t][1] = w * (mu - 0.5 * lam * ln3);
    // the node term's kernel in the spatial gradients: with g = F^T h,
    // gamma g_i . g_j + pen (F g_i) . (F g_j) = h_i^T M h_j, M = gamma b + pen b^2,
    // b = F F^T (symmetric)
    const double ga = w * mu, pen = detF <= 0.0 ? w * 1e200 : 0.0;
    double b[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) b[3 * r + c] = F[3 * r] * F[3 * c] + F[3 * r + 1] * F[3 * c + 1] + F[3 * r + 2] * F[3 * c + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double m = ga * b[3 * r + c];
        if (pen != 0.0) m += pen * (b[3 * r] * b[c] + b[3 * r + 1] * b[3 + c] + b[3 * r + 2] * b[6 + c]);
        Mq[t][3 * r + c] = m;
      }
  }
  __syncthreads();
  // A5: the residual re[(i,a)] = sum_q sum_c P_q[a][c] g[q][i][c] with
  // P_q = wJ_q F_q S_q (first Piola stress times the weight, into Bq, which
  // is free after A2), as three partial sums over 9 points each
  if (t < kNq * 9) {
    const int q = t / 9, ar = (t % 9) / 3, c = t % 3;
    const double *F = Fq[q], *S = Sq[q];
    Bq[q][t % 9] = wJ[q] * (F[3 * ar] * S[c] + F[3 * ar + 1] * S[3 + c] + F[3 * ar + 2] * S[6 + c]);
  }
  __syncthreads();
  if (t < 3 * kNl) {
    const int grp = t / kNl, o = t % kNl, i = o / 3, ar = o % 3;
    double re = 0.0;
    for (int q = 9 * grp; q < 9 * grp + 9; ++q)
      re += Bq[q][3 * ar] * gq[q][i][0] + Bq[q][3 * ar + 1] * gq[q][i][1] + Bq[q][3 * ar + 2] * gq[q][i][2];
    red[t] = re;
  }
  __syncthreads();
  if (t < kNl) a.res[(int64_t)e * kNl + t] = red[t] + red[kNl + t] + red[2 * kNl + t];
  // Phase B in the spatial form.  With h_qi = F_q^{-T} g_qi (spatial
  // gradients), the push-forward of this law's material tangent is
  // J c = lambda 1 x 1 + (mu - lambda ln J)(two symmetric identities), and
  // F^{-1} b F^{-T} = I folds the stress term into mu g_qi . g_qj, so the
  // (i, j) block of D's contraction is, per point,
  //   wJ_q [lambda h_qi (x) h_qj + (mu - lambda ln J_q) h_qj (x) h_qi
  //         + (mu g_qi . g_qj + pen_q (F g_qi) . (F g_qj)) I]
  // (pen_q = 1e200 where det F <= 0: the derivative of the law's penalty
  // term 1e200 C added to S there).  Same matrix as the material form
  // D[(a,c),(b,n)] = wJ (delta_ab S_nc + F_ak A_kcmn F_bm) up to roundoff,
  // without forming D or D g (nh_phase_b_mfma).
  // hC[q][c][i] = h_qic in Real (zero for q = 27 and nodes i >= 27), over
  // gq: formed in registers, written once every thread has read gq
  constexpr int kHc = (kNhQ * kNhN + kEbThreads - 1) / kEbThreads;
  Real hv[kHc][3];
#pragma unroll
  for (int k = 0; k < kHc; ++k) {
    const int x = t + k * kEbThreads, q = x / kNhN, i = x % kNhN;
    hv[k][0] = hv[k][1] = hv[k][2] = Real(0);
    if (q < kNq && i < kNd) {
      const double *Fi = Cq[q], *g = gq[q][i];
      const double g0 = g[0], g1 = g[1], g2 = g[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) hv[k][c] = (Real)(Fi[c] * g0 + Fi[3 + c] * g1 + Fi[6 + c] * g2);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kHc; ++k) {
    const int x = t + k * kEbThreads, q = x / kNhN, i = x % kNhN;
    if (x < kNhQ * kNhN)
#pragma unroll
      for (int c = 0; c < 3; ++c) hC[q][c][i] = hv[k][c];
  }
  __syncthreads();
  nh_phase_b_mfma<Real, KEEP>(t, e, cf, Mq, hC, L.stage, skeep, red, a);
}

// node-patch gather of the rows of buffered element matrices: the Q rows of
// an incidence are one contiguous run of Q*NL values (FP64 or FP32); the runs
// of all the node's incidences are streamed by the whole workgroup at once
// (independent loads, LDS atomics), then the node's CSR rows are written once.
// TH threads per workgroup: 256 for the Q2 hexahedra (one node, long runs;
// four wavefronts owning two, two, two and three (row component, column
// component) pairs), 64 for the affine tets' packed wave groups (short
// runs, more groups in flight); wavefront w writes rows w, w + TH/64, ... of
// each node.
constexpr int kGatherThreadsQ2 = 256;

// npw (1..4) small patches of one bin share a workgroup: their incidence
// lists are concatenated and streamed together; patch k's accumulator starts
// at the prefix sum of Q*Q*ns over the patches before it.
constexpr int kGatherNpw = 4;

template <int ND, int Q, typename TE, int TH, int MI = 64>
__global__ __launch_bounds__(TH) void k_gather_emat(PatchArgs a, const TE *__restrict__ E,
                                                    const double *__restrict__ thr) {
  extern __shared__ double acc[];  // the npw accumulators, acc_cap doubles in all
  constexpr int kMaxInc = MI;      // incidences per workgroup (checked by the launcher)
  __shared__ int64_t sbase[kMaxInc];   // element-matrix run of each incidence
  __shared__ double sthr[kMaxInc];
  __shared__ uint16_t sslot[kMaxInc * ND];
  __shared__ int sacc[kMaxInc];        // accumulator offset and slot count of each incidence's patch
  __shared__ int sns[kMaxInc];
  constexpr int NL = ND * Q, RUN = Q * NL;
  const int64_t wg = blockIdx.x;
  const int npw = a.npw;
  const int64_t pos0 = ((wg & 7) * (int64_t)a.xcd_chunk + (wg >> 3)) * npw;
  if (pos0 >= a.n_nodes) return;
  const int t = threadIdx.x;
  PatchDesc dsc[kGatherNpw];
  int pre[kGatherNpw + 1];
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < kGatherNpw; ++k) {
    dsc[k] = PatchDesc{0, 0, 0u};
    if (k < npw && pos0 + k < a.n_nodes) dsc[k] = a.desc[pos0 + k];
    pre[k + 1] = pre[k] + (int)(dsc[k].ninc_ns & 0xffffu);
  }
  const int total_inc = pre[kGatherNpw];
  if (total_inc > kMaxInc) return;  // excluded by the launcher
  // staging, flattened over the workgroup's incidences (patch k of combined
  // incidence c: the number of patch starts at or below c): the incidence
  // words and slot maps are loaded together, then the thresholds
  auto patch_of = [&](int c) { return (int)(c >= pre[1]) + (int)(c >= pre[2]) + (int)(c >= pre[3]); };
  int gi0[kGatherNpw], gns[kGatherNpw], pre4[kGatherNpw], aoff[kGatherNpw];
  int apre = 0;  // accumulator of patch k at the prefix sum of Q*Q*ns
#pragma unroll
  for (int k = 0; k < kGatherNpw; ++k) {
    gi0[k] = dsc[k].i0;
    gns[k] = (int)(dsc[k].ninc_ns >> 16) & 0x7fff;
    pre4[k] = pre[k];
    aoff[k] = apre;
    apre += Q * Q * gns[k];
  }
  auto sel = [](const int (&v)[kGatherNpw], int k) {  // by value: no indexed array
    return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
  };
  // presence words of the patches that are not full (patch k's at the prefix
  // sum wpre[k]), staged in LDS behind the accumulators with the other staged
  // data, so that the write-out makes no global loads
  int wpre[kGatherNpw];
  int wtot = 0;
#pragma unroll
  for (int k = 0; k < kGatherNpw; ++k) {
    wpre[k] = wtot;
    wtot += (dsc[k].ninc_ns >> 31) ? 0 : (Q * Q * gns[k] + 31) / 32;
  }
  uint32_t *swords = reinterpret_cast<uint32_t *>(acc + a.lds_acc);
  uint32_t inc = 0;
  if (t < total_inc) {
    const int k = patch_of(t);
    inc = a.inc[sel(gi0, k) + (t - sel(pre4, k))];
  }
  for (int x = t; x < total_inc * ND; x += TH) {
    const int c = x / ND, k = patch_of(c);
    sslot[x] = a.slot[((int64_t)sel(gi0, k) + (c - sel(pre4, k))) * ND + (x - c * ND)];
  }
  for (int x = t; x < wtot; x += TH) {
    const int k = (int)(x >= wpre[1]) + (int)(x >= wpre[2]) + (int)(x >= wpre[3]);
    swords[x] = a.mask[(int64_t)sel(gi0, k) * a.mwpi + (x - sel(wpre, k))];
  }
  for (int x = t; x < apre; x += TH) acc[x] = 0.0;
  if (t < total_inc) {
    const int k = patch_of(t);
    const int64_t e = inc >> 5;
    sbase[t] = e * (int64_t)NL * NL + (int64_t)(inc & 31) * RUN;
    sthr[t] = thr[e];
    sacc[t] = sel(aoff, k);
    sns[t] = sel(gns, k);
  }
  __syncthreads();
  // the stream: incidence by incidence (uniform across the workgroup), each
  // thread owning the offsets t, t + TH, ... of the incidence's run, whose
  // (row, node, component) split is fixed; kI incidences' loads are issued
  // before any is used (offsets past the run load entry 0 and are dropped).
  // One wavefront (affine tets, runs of 90 or 180, 12-24 incidences per
  // group) loads kI incidences with clamped indices; 256 threads (Q2
  // hexahedra, runs of 243, often fewer incidences than kI) skip the loads
  // past the group's last incidence.
  // Which entries (row component x, column node j, column component y) of
  // a run a thread owns.  One wavefront (E row-major, run offset
  // x NL + j Q + y): t, t + 64, ...  Several wavefronts (E node-blocked, run
  // offset (x Q + y) ND + j, the Neo-Hookean element kernel's layout): each
  // owns whole (x, y) pairs, each ND consecutive values: 64 / ND pairs for
  // all but the last wavefront, which owns the rest in NV passes.  Every
  // accumulator entry (x, slot, y) is then added to by one wavefront only,
  // in incidence order, and no two lanes of one ds_add_f64 share an entry
  // (the nodes of an incidence have distinct slots): the sums do not depend
  // on timing.
  constexpr bool kPairs = TH > 64;
  constexpr int kW = TH / 64, kPpw = 64 / ND, kLast = Q * Q - kPpw * (kW - 1);
  static_assert(!kPairs || (kLast > 0 && kLast * ND <= 2 * 64), "the last wavefront's pairs fit two passes");
  constexpr int NV = kPairs ? (kLast * ND + 63) / 64 : (RUN + TH - 1) / TH;
  constexpr int kI = TH == 64 ? (NV >= 16 ? 1 : 16 / NV) : 8;
  int ov[NV], xv[NV], jv[NV], yv[NV];
  bool okv[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (kPairs) {  // E node-blocked: pair p's ND values are consecutive
      const int w = t / 64, L = t % 64 + 64 * v;  // this wavefront's value index
      const int p = kPpw * w + L / ND;
      okv[v] = w < kW - 1 ? L < kPpw * ND : L < kLast * ND;
      xv[v] = okv[v] ? p / Q : 0;
      yv[v] = okv[v] ? p % Q : 0;
      jv[v] = okv[v] ? L % ND : 0;
      ov[v] = okv[v] ? p * ND + jv[v] : 0;
    } else {
      const int o = t + v * TH;
      okv[v] = o < RUN;
      ov[v] = okv[v] ? o : 0;
      xv[v] = ov[v] / NL;
      const int c = ov[v] - xv[v] * NL;
      jv[v] = c / Q;
      yv[v] = c - jv[v] * Q;
    }
  }
  for (int c0 = 0; c0 < total_inc; c0 += kI) {
    TE val[kI][NV];
#pragma unroll
    for (int i = 0; i < kI; ++i) {
      if (TH == 64 || c0 + i < total_inc) {  // uniform
        const int64_t base = sbase[min(c0 + i, total_inc - 1)];
#pragma unroll
        for (int v = 0; v < NV; ++v) val[i][v] = __builtin_nontemporal_load(E + base + ov[v]);  // read once
      }
    }
#pragma unroll
    for (int i = 0; i < kI; ++i) {
      const int k = c0 + i;
      if (k >= total_inc) break;
      const double th = sthr[k];
      const int ns = sns[k], ak = sacc[k];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double x = (double)val[i][v];
        if (okv[v] && fabs(x) > th) atomicAdd(&acc[ak + (xv[v] * ns + sslot[k * ND + jv[v]]) * Q + yv[v]], x);
      }
    }
  }
  __syncthreads();
  // values are streamed past the caches (written once, never re-read here)
  const int w = t / 64, lane = t % 64;
#pragma unroll
  for (int k = 0; k < kGatherNpw; ++k) {
    if (k >= npw || pos0 + k >= a.n_nodes) break;
    const int nsraw = (int)(dsc[k].ninc_ns >> 16), ns = nsraw & 0x7fff;
    const double *ak = acc + aoff[k];
    if (nsraw & 0x8000) {  // all rows full: the node's Q rows are one contiguous run of the CSR
      double *dst = a.vals + dsc[k].rowbase;
      for (int d = t; d < Q * ns * Q; d += TH) __builtin_nontemporal_store(ak[d], dst + d);
      continue;
    }
    // wavefront x writes row x from the presence bits; the node's rows are
    // consecutive CSR rows, so row x starts after the present entries of
    // rows 0..x-1 (bits [0, x*ns*Q) of the node's presence words)
    const uint32_t *m = swords + wpre[k];
    for (int x = w; x < Q; x += TH / 64) {
      const int off = x * ns * Q, len = ns * Q;
      int cnt = 0;
      for (int b0 = 0; b0 < off; b0 += 64 * 32) {
        const int b = b0 + 32 * lane;  // this lane's word of the prefix
        if (b < off) {
          const uint32_t word = m[b >> 5];
          cnt += __popc(off - b >= 32 ? word : (word & ((1u << (off - b)) - 1u)));
        }
      }
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
      int64_t base = dsc[k].rowbase + cnt;
      // kOut chunks of 64 per round: presence words and values are read with
      // clamped indices before any ballot, so their latencies overlap
      constexpr int kOut = 8;
      for (int c0 = 0; c0 < len; c0 += kOut * 64) {
        uint32_t word[kOut];
        double val[kOut];
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
          const int dc = off + min(c0 + u * 64 + lane, len - 1);
          word[u] = m[dc >> 5];
          val[u] = ak[dc];
        }
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
          const int d = c0 + u * 64 + lane;
          const bool pr = (d < len) & (bool)((word[u] >> ((off + d) & 31)) & 1u);
          const uint64_t bal = __ballot(pr);
          if (pr) __builtin_nontemporal_store(val[u], a.vals + base + __popcll(bal & ((1ull << lane) - 1ull)));
          base += __popcll(bal);
        }
      }
    }
  }
}

// Element-batched isotropic elasticity on the matrix cores (affine tets, the
// bf16 split form): E = sum_q w_q |J| B_q^T D B_q with B_q the 6 x NL Voigt
// strain-displacement matrix at point q (engineering shears; rows xx, yy, zz,
// yz, xz, xy) and D the isotropic Voigt matrix (lambda + 2 mu on the normal
// diagonal, lambda off it, mu on the shears).  D = S S with the symmetric root
// S = sqrt(2 mu) I + c 1 1^T on the normal block, c = (sqrt(2 mu + 3 lambda)
// - sqrt(2 mu)) / 3, and sqrt(mu) on the shears (D positive definite: mu > 0,
// 3 lambda + 2 mu > 0).  With F_q = sqrt(|w_q| |J|) S B_q,
// E = sum_q sign(w_q) F_q^T F_q: one product E[m][n] = sum_k A[m][k] C[k][n]
// with A[m][k] = F[k][m], C[k][n] = sign F[k][n] and k = 8 q + s (s < 6; 6, 7
// zero), so one 32-wide k-step is 4 points.  E is symmetric: one wavefront
// computes one element's upper NT x NT tiles (16 x 16) and mirrors them.
// MFMA 16x16x32 bf16 lane layout: A[i][k] in lane i + 16 (k / 8), element
// k % 8; C[k][j] in lane j + 16 (k / 8), element k % 8; D[i][j] in lane
// j + 16 (i / 4), element i % 4.  So the A fragment of row tile a and the C
// fragment of column tile a are the same registers (up to the sign of w_q):
// the 8 rows of F at one point (q = 4 step + lane / 16) for the dof
// 16 a + lane % 16, built from the reference gradient of its shape function
// (the same for every element: staged in LDS once per workgroup) and the
// element's inverse Jacobian, then split into bf16 hi + lo; the products
// hi hi, hi lo and lo hi are accumulated in FP32 (lo lo is below 2^-16 of
// the leading term).  The workgroup's four wavefronts sweep the elements with
// a grid stride, the next element's record loaded while the current one is
// computed.  E is written row-major in FP32 (rows = test dofs (i, a)) through
// the wavefront's LDS block as whole 16-byte pieces.
__device__ __forceinline__ void split2(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

__device__ __forceinline__ bf16x8 bf16_xor(const bf16x8 &x, uint32_t m) {
  u32x4 u = __builtin_bit_cast(u32x4, x);
  u ^= m;
  return __builtin_bit_cast(bf16x8, u);
}

// rows of F at the point of the staged reference gradient h (h.w = the signed
// sqrt(|w_q|)) for dof m (m >= NL: padding, zero); on / off / sh: sqrt(|J|)
// times the normal-diagonal, normal-off-diagonal and shear factors of S
template <int NL>
__device__ __forceinline__ void bdb_factor(const float4 *hp, int m, const float (&Bm)[9], float on, float off,
                                           float sh, bf16x8 &hi, bf16x8 &lo) {
  const float4 h = hp[m < NL ? m / 3 : 0];
  const float gx = Bm[0] * h.x + Bm[1] * h.y + Bm[2] * h.z;
  const float gy = Bm[3] * h.x + Bm[4] * h.y + Bm[5] * h.z;
  const float gz = Bm[6] * h.x + Bm[7] * h.y + Bm[8] * h.z;
  const int ax = m % 3;
  const float r = m < NL ? fabsf(h.w) : 0.f;  // padding rows are zero
  const float ga = r * (ax == 0 ? gx : (ax == 1 ? gy : gz)), vn = on * ga, vo = off * ga;
  const float rs = sh * r;
  float v[8];
  v[0] = ax == 0 ? vn : vo;
  v[1] = ax == 1 ? vn : vo;
  v[2] = ax == 2 ? vn : vo;
  v[3] = rs * (ax == 1 ? gz : (ax == 2 ? gy : 0.f));  // yz
  v[4] = rs * (ax == 0 ? gz : (ax == 2 ? gx : 0.f));  // xz
  v[5] = rs * (ax == 0 ? gy : (ax == 1 ? gx : 0.f));  // xy
  v[6] = v[7] = 0.f;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    __bf16 a, b;
    split2(v[x], a, b);
    hi[x] = a;
    lo[x] = b;
  }
}

constexpr int kBdbMaxGroups = 8192;

// the k-steps of one element: fragments of the NT dof tiles at 4 points, then
// the upper tiles' three bf16 products; NEG: C = sign(w_q) F (one XOR)
template <int ND, bool NEG>
__device__ __forceinline__ void bdb_steps(const float4 *hs, int steps, int qq, int lr, const float (&Bm)[9],
                                          float on, float off, float sh,
                                          f32x4 (&acc)[((3 * ND + 15) / 16) * ((3 * ND + 15) / 16 + 1) / 2]) {
  constexpr int NL = 3 * ND, NT = (NL + 15) / 16;
  for (int k = 0; k < steps; ++k) {
    const float4 *hp = hs + (4 * k + qq) * ND;
    bf16x8 fh[NT], fl[NT];
#pragma unroll
    for (int a = 0; a < NT; ++a) bdb_factor<NL>(hp, 16 * a + lr, Bm, on, off, sh, fh[a], fl[a]);
    const uint32_t sgn = NEG && hp[0].w < 0.f ? 0x80008000u : 0u;
    int p = 0;
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const bf16x8 ch = NEG ? bf16_xor(fh[b], sgn) : fh[b];
      const bf16x8 cl = NEG ? bf16_xor(fl[b], sgn) : fl[b];
#pragma unroll
      for (int a = 0; a <= b; ++a, ++p) {
        acc[p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fl[a], ch, acc[p], 0, 0, 0);
        acc[p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[a], cl, acc[p], 0, 0, 0);
        acc[p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fh[a], ch, acc[p], 0, 0, 0);
      }
    }
  }
}

// rows of the element matrix staged at once: half of them when that half is a
// whole number of 16-byte pieces (P3: 30 x 60 floats), else all (P2)
__host__ __device__ constexpr int bdb_stage_rows(int nl) { return (nl / 2) * nl % 4 == 0 && nl % 2 == 0 ? nl / 2 : nl; }

template <int ND>
__global__ __launch_bounds__(256) void k_elast_bdb_mfma(int64_t n_cv, int nq, const double *__restrict__ rec,
                                                        const double *__restrict__ gref,
                                                        const double *__restrict__ qw, float *__restrict__ E) {
  constexpr int NL = 3 * ND, NT = (NL + 15) / 16, NP = NT * (NT + 1) / 2;
  constexpr int RC = bdb_stage_rows(NL), KC = NL / RC;
  static_assert(NL * NL % 4 == 0, "element matrices are whole float4 runs");
  extern __shared__ float4 hs[];  // [4 * steps][ND]: (reference gradient of phi_i at point q,
                                  // signed sqrt(|w_q|)), then one RC x NL float block per wavefront
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int steps = (nq + 3) / 4;
  float *es = reinterpret_cast<float *>(hs + 4 * steps * ND) + w * RC * NL;
  int neg = 0;
  for (int x = t; x < 4 * steps * ND; x += 256) {
    const int q = x / ND;
    float4 h{0.f, 0.f, 0.f, 0.f};
    if (q < nq) {
      const double *g = gref + (int64_t)x * 3;  // [q][i][3]
      const double wq = qw[q];
      h = float4{(float)g[0], (float)g[1], (float)g[2], (float)copysign(sqrt(fabs(wq)), wq)};
      neg |= wq < 0.0;
    }
    hs[x] = h;
  }
  const bool any_neg = __syncthreads_or(neg);  // a rule with negative weights
  const int lr = l % 16, qq = l / 16;
  const int64_t stride = 4 * (int64_t)gridDim.x;
  int64_t e = 4 * (int64_t)blockIdx.x + __builtin_amdgcn_readfirstlane(w);  // uniform: scalar record loads
  if (e >= n_cv) return;  // no barrier follows
  double nx[11];  // the record of the element in flight: B row-major, |J| lambda, |J| mu
#pragma unroll
  for (int x = 0; x < 9; ++x) nx[x] = rec[kRecStride * e + x];
#pragma unroll
  for (int x = 0; x < 2; ++x) nx[9 + x] = rec[kRecStride * e + 12 + x];
  for (; e < n_cv; e += stride) {
    float Bm[9];
#pragma unroll
    for (int x = 0; x < 9; ++x) Bm[x] = (float)nx[x];
    // S's factors in FP32 (the products are rounded to bf16 pairs anyway); the
    // record's coefficients carry |J|, and S(|J| lambda, |J| mu) = sqrt(|J|) S(lambda, mu)
    const float lam = (float)nx[9], mu = (float)nx[10];
    // c = (sqrt(2 mu + 3 lambda) - sqrt(2 mu)) / 3 without the cancellation
    const float s2m = sqrtf(2.f * mu), c = lam / (sqrtf(2.f * mu + 3.f * lam) + s2m);
    const float on = s2m + c, off = c, sh = sqrtf(mu);
    const int64_t en = min(e + stride, n_cv - 1);
#pragma unroll
    for (int x = 0; x < 9; ++x) nx[x] = rec[kRecStride * en + x];
#pragma unroll
    for (int x = 0; x < 2; ++x) nx[9 + x] = rec[kRecStride * en + 12 + x];
    f32x4 acc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (any_neg)  // uniform: only rules with a negative weight pay for the signs
      bdb_steps<ND, true>(hs, steps, qq, lr, Bm, on, off, sh, acc);
    else
      bdb_steps<ND, false>(hs, steps, qq, lr, Bm, on, off, sh, acc);
    // KC row chunks of RC rows, each staged whole and written as 16-byte pieces
#pragma unroll
    for (int h = 0; h < KC; ++h) {
      int p = 0;
#pragma unroll
      for (int b = 0; b < NT; ++b)
#pragma unroll
        for (int a = 0; a <= b; ++a, ++p) {
          const int col = 16 * b + lr;
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            const int row = 16 * a + 4 * qq + x;
            if (row < NL && col < NL) {
              if (row / RC == h) es[(row - RC * h) * NL + col] = acc[p][x];
              if (a != b && col / RC == h) es[(col - RC * h) * NL + row] = acc[p][x];
            }
          }
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const f32x4 *es4 = reinterpret_cast<const f32x4 *>(es);
      f32x4 *E4 = reinterpret_cast<f32x4 *>(E + e * NL * NL + RC * h * NL);
      for (int x = l; x < RC * NL / 4; x += 64) __builtin_nontemporal_store(es4[x], E4 + x);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // reads done before the next writes
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

hipError_t launch_nh_element(hipStream_t s, const EbArgs &a, int precision) {
  if (a.n_cv == 0) return hipSuccess;
  if (a.nq != kNq) return hipErrorInvalidValue;
  if (precision == GFX_PREC_FP32) {
    if (a.keep) return hipErrorInvalidValue;  // keep bits come from the FP64 pass
    k_nh_element<float, false><<<(unsigned)a.n_cv, kEbThreads, 0, s>>>(a);
  } else if (a.keep) {
    if (a.kwords != kKeepWords) return hipErrorInvalidValue;
    k_nh_element<double, true><<<(unsigned)a.n_cv, kEbThreads, 0, s>>>(a);
  } else {
    k_nh_element<double, false><<<(unsigned)a.n_cv, kEbThreads, 0, s>>>(a);
  }
  return hipGetLastError();
}

// Element-batched isotropic elasticity on the FP64 and FP32 matrix cores
// (v_mfma_f64_16x16x4_f64, v_mfma_f32_16x16x4_f32): the symmetric root
// factorization of the bf16 form above, E = sum_q sign(w_q) F_q^T F_q with
// F_q = sqrt(|w_q| |J|) S B_q, computed in T.  The k axis is the 6 Voigt rows
// of every point back to back (k = 6 q + s, no padding rows), 4 k per MFMA.
// Lane layout of the 16x16x4 products: A[i][k] in lane i + 16 k, C[k][j] in
// lane j + 16 k (both instructions); D[i][j] in lane j + 16 (i / 4), element
// i % 4 for v_mfma_f32_16x16x4_f32, but in lane j + 16 (i % 4), element i / 4
// for v_mfma_f64_16x16x4_f64 (BdbMfma<T>::row gives the row of a lane group g
// and element r).  So a lane holds one F value per dof tile and k-step, and
// the A fragment of row tile a is the C fragment of column tile a up to
// sign(w_q).  One wavefront per element (grid stride), only the upper tiles
// computed.  E (row-major, T) is stored from the accumulators: tile (a, b) as
// 16-value runs of its rows, its mirror (b, a) as 4-value runs per lane.
// F[s][m] of one point: Voigt row s (xx, yy, zz, yz, xz, xy) of S B_q for
// dof m = 3 i + ax, from the staged (reference gradient of phi_i, signed
// sqrt(|w_q|)) h; on / off / sh: the normal-diagonal, normal-off-diagonal and
// shear factors of S (with |J| folded in)
template <typename T>
__device__ __forceinline__ T bdb_fvalue(const T (&h)[4], int s, int ax, const T (&Bm)[9], T on, T off, T sh) {
  const T gx = Bm[0] * h[0] + Bm[1] * h[1] + Bm[2] * h[2];
  const T gy = Bm[3] * h[0] + Bm[4] * h[1] + Bm[5] * h[2];
  const T gz = Bm[6] * h[0] + Bm[7] * h[1] + Bm[8] * h[2];
  const T r = h[3] < T(0) ? -h[3] : h[3];
  T v;
  if (s < 3) {
    v = (s == ax ? on : off) * (ax == 0 ? gx : (ax == 1 ? gy : gz));
  } else {
    // shear rows: yz (s 3) couples y and z, xz (4) x and z, xy (5) x and y
    const int p = s == 3 ? 1 : 0, q = s == 5 ? 1 : 2;
    const T gp = p == 0 ? gx : gy, gq = q == 1 ? gy : gz;
    v = ax == p ? sh * gq : (ax == q ? sh * gp : T(0));
  }
  return r * v;
}

template <int ND, typename T>
__global__ __launch_bounds__(256) void k_elast_bdb_mfma_hp(int64_t n_cv, int nq, const double *__restrict__ rec,
                                                           const double *__restrict__ gref,
                                                           const double *__restrict__ qw, T *__restrict__ E) {
  typedef typename BdbMfma<T>::v4 v4;
  constexpr int NL = 3 * ND, NT = (NL + 15) / 16, NP = NT * (NT + 1) / 2;
  extern __shared__ double hs_raw[];  // [nq][ND][4] in T: reference gradient, signed sqrt(|w_q|)
  T *hs = reinterpret_cast<T *>(hs_raw);
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  int neg = 0;
  for (int x = t; x < nq * ND; x += 256) {
    const double *g = gref + (int64_t)x * 3;  // [q][i][3]
    const double wq = qw[x / ND];
    hs[4 * x] = (T)g[0];
    hs[4 * x + 1] = (T)g[1];
    hs[4 * x + 2] = (T)g[2];
    hs[4 * x + 3] = (T)copysign(sqrt(fabs(wq)), wq);
    neg |= wq < 0.0;
  }
  const bool any_neg = __syncthreads_or(neg);
  const int lr = l % 16, lg = l / 16;
  const int K = 6 * nq, steps = (K + 3) / 4;
  const int64_t stride = 4 * (int64_t)gridDim.x;
  for (int64_t e = 4 * (int64_t)blockIdx.x + __builtin_amdgcn_readfirstlane(w); e < n_cv; e += stride) {
    const double *re = rec + kRecStride * e;  // uniform: scalar loads
    T Bm[9];
#pragma unroll
    for (int x = 0; x < 9; ++x) Bm[x] = (T)re[x];
    // S(|J| lambda, |J| mu) = sqrt(|J|) S(lambda, mu)
    const T lam = (T)re[12], mu = (T)re[13];
    const T s2m = sqrt(T(2) * mu), c = lam / (sqrt(T(2) * mu + T(3) * lam) + s2m);
    const T on = s2m + c, off = c, sh = sqrt(mu);
    v4 acc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] = v4{T(0), T(0), T(0), T(0)};
    int q = lg / 6, sv = lg % 6;  // this lane's k = 4 step + lg as (point, Voigt row)
    for (int st = 0; st < steps; ++st) {
      const bool kin = q < nq;
      T f[NT];
#pragma unroll
      for (int a = 0; a < NT; ++a) {
        const int m = 16 * a + lr;
        T h[4] = {T(0), T(0), T(0), T(0)};
        if (kin && m < NL) {
          const T *hp = hs + 4 * (q * ND + m / 3);  // q < nq, m / 3 < ND here
          h[0] = hp[0]; h[1] = hp[1]; h[2] = hp[2]; h[3] = hp[3];
        }
        f[a] = bdb_fvalue<T>(h, sv, m % 3, Bm, on, off, sh);
      }
      const T sg = any_neg && hs[4 * (min(q, nq - 1) * ND) + 3] < T(0) ? T(-1) : T(1);
      int p = 0;
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        const T cb = sg * f[b];
#pragma unroll
        for (int a = 0; a <= b; ++a, ++p) acc[p] = BdbMfma<T>::mma(f[a], cb, acc[p]);
      }
      sv += 4;
      if (sv >= 6) { sv -= 6; ++q; }
    }
    T *Ee = E + e * (int64_t)(NL * NL);
    int p = 0;
#pragma unroll
    for (int b = 0; b < NT; ++b)
#pragma unroll
      for (int a = 0; a <= b; ++a, ++p) {
        const int col = 16 * b + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * a + BdbMfma<T>::row(lg, r);
          if (row < NL && col < NL) {
            __builtin_nontemporal_store(acc[p][r], Ee + row * NL + col);
            if (a != b) __builtin_nontemporal_store(acc[p][r], Ee + col * NL + row);
          }
        }
      }
  }
}

hipError_t launch_elast_bdb_mfma(hipStream_t s, int64_t n_cv, int nd, int nq, const double *rec,
                                 const double *gref, const double *qw, float *E) {
  if (n_cv == 0) return hipSuccess;
  // one float4 per (point, node), then the staged rows of one element matrix per wavefront
  const size_t lds = 16 * (size_t)(4 * ((nq + 3) / 4)) * nd + 4 * 4 * (size_t)(bdb_stage_rows(3 * nd) * 3 * nd);
  if (nq < 1 || lds > 64 * 1024) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)std::min<int64_t>((n_cv + 3) / 4, kBdbMaxGroups);  // 4 wavefronts each
  if (nd == 10) k_elast_bdb_mfma<10><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E);
  else if (nd == 20) k_elast_bdb_mfma<20><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// the FP64 (E64) or FP32 (E32) matrix-core form; the other pointer is null
hipError_t launch_elast_bdb_mfma_hp(hipStream_t s, int64_t n_cv, int nd, int nq, const double *rec,
                                    const double *gref, const double *qw, double *E64, float *E32) {
  if (n_cv == 0) return hipSuccess;
  const size_t lds = (size_t)nq * nd * 4 * (E64 ? sizeof(double) : sizeof(float));
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)std::min<int64_t>((n_cv + 3) / 4, kBdbMaxGroups);  // 4 wavefronts each
  if (E64) {
    if (nd == 10) k_elast_bdb_mfma_hp<10, double><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E64);
    else if (nd == 20) k_elast_bdb_mfma_hp<20, double><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E64);
    else return hipErrorInvalidValue;
  } else {
    if (nd == 10) k_elast_bdb_mfma_hp<10, float><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E32);
    else if (nd == 20) k_elast_bdb_mfma_hp<20, float><<<grid, 256, lds, s>>>(n_cv, nq, rec, gref, qw, E32);
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}


hipError_t launch_gather_emat(hipStream_t s, const PatchArgs &a0, const void *E, bool e_fp32, const double *thr,
                              int acc_cap, int group_inc) {
  // acc_cap: the largest sum of Q*Q*nslots over the patches of one workgroup
  // (doubles of LDS); group_inc: the largest number of incidences of one workgroup
  if (a0.n_nodes == 0) return hipSuccess;
  PatchArgs a = a0;
  if (a.npw < 1 || a.npw > kGatherNpw) return hipErrorInvalidValue;
  const int64_t groups = (a.n_nodes + a.npw - 1) / a.npw;
  a.xcd_chunk = (int)((groups + 7) / 8);
  const unsigned grid = (unsigned)(8 * (int64_t)a.xcd_chunk);
  // accumulators, then the presence words: at most ceil(Q*Q*ns/32) per patch,
  // so acc_cap/32 + kGatherNpw in all
  if (acc_cap < 0 || 8 * (size_t)acc_cap > 48 * 1024 || group_inc > 64) return hipErrorInvalidValue;
  a.lds_acc = acc_cap;
  const size_t lds = (size_t)acc_cap * sizeof(double) + 4 * (size_t)(acc_cap / 32 + kGatherNpw);
  const double *E64 = static_cast<const double *>(E);
  const float *E32 = static_cast<const float *>(E);
  if (a.nd == 27 && a.qdim == 3) {
    if (e_fp32) k_gather_emat<27, 3, float, kGatherThreadsQ2><<<grid, kGatherThreadsQ2, lds, s>>>(a, E32, thr);
    else k_gather_emat<27, 3, double, kGatherThreadsQ2><<<grid, kGatherThreadsQ2, lds, s>>>(a, E64, thr);
  } else if (a.nd == 10 && a.qdim == 3 && !e_fp32) {  // element-batched affine P2, FP64 matrix cores
    if (group_inc <= 32) k_gather_emat<10, 3, double, 64, 32><<<grid, 64, lds, s>>>(a, E64, thr);
    else k_gather_emat<10, 3, double, 64><<<grid, 64, lds, s>>>(a, E64, thr);
  } else if (a.nd == 20 && a.qdim == 3 && !e_fp32) {  // element-batched affine P3, FP64 matrix cores
    if (group_inc <= 32) k_gather_emat<20, 3, double, 64, 32><<<grid, 64, lds, s>>>(a, E64, thr);
    else k_gather_emat<20, 3, double, 64><<<grid, 64, lds, s>>>(a, E64, thr);
  } else if (a.nd == 10 && a.qdim == 3 && e_fp32) {  // element-batched affine P2 (bf16 split form)
    // groups of at most 32 incidences (the packed waves): half the static
    // LDS per group, so more groups fit on a CU
    if (group_inc <= 32) k_gather_emat<10, 3, float, 64, 32><<<grid, 64, lds, s>>>(a, E32, thr);
    else k_gather_emat<10, 3, float, 64><<<grid, 64, lds, s>>>(a, E32, thr);
  } else if (a.nd == 20 && a.qdim == 3 && e_fp32) {  // element-batched affine P3
    if (group_inc <= 32) k_gather_emat<20, 3, float, 64, 32><<<grid, 64, lds, s>>>(a, E32, thr);
    else k_gather_emat<20, 3, float, 64><<<grid, 64, lds, s>>>(a, E32, thr);
  }
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
