// HIP kernels of libgfx for gfx950 (MI355X, CDNA4).
//
// Data flow of one order-2 assembly (what ga_exec does element by element,
// src/getfem_generic_assembly_compile_and_exec.cc:8750-8869, reorganised
// for the GPU):
//   k_records_tet    per element, FP64: K = G grad(N), det K, B = K^{-T}
//                    (bgeot_geometric_trans.cc:270-413) and the element's
//                    coefficients -> one 128-byte record per element.
//   k_assemble_patch one wavefront per group of up to 4 mesh nodes
//                    ("patches"): for every element incident to a node,
//                    computes the rows of the element matrix that belong to
//                    the node's dofs, drops
//                    entries under the element's insertion threshold
//                    (compile_and_exec.cc:5441-5453), accumulates them in
//                    LDS and writes the node's CSR rows once.  Rows are owned
//                    by exactly one workgroup: no atomics in HBM, and every
//                    CSR value is written exactly once.
// Pattern build (once per plan): node -> element incidences by a counting
// sort, per-element thresholds and keep bits, then per node a segmented key
// sort of the neighbour nodes in LDS gives the sorted CSR columns.
#include <hip/hip_runtime.h>

#include "gfx_internal.h"

namespace gfx {

namespace {

constexpr int kWave = 64;
constexpr int kMaxCand = 4096;  // neighbour candidates per node (LDS sort size)
constexpr int kFullRows = 0x8000;  // flag in nslots: all of the node's rows are full

__device__ __forceinline__ int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// ------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------
// rec (FP64) or rec32 (FP32 contraction form, 64 bytes per element) is written
__global__ __launch_bounds__(256) void k_records_tet(
    int64_t n_cv, const double *__restrict__ pts, const int32_t *__restrict__ cvx,
    const double *__restrict__ ca, double ca_c, const double *__restrict__ cb,
    double cb_c, const double *__restrict__ cc, double cc_c, const double *__restrict__ thr,
    double r_scale, double *__restrict__ rec, float *__restrict__ rec32, int *__restrict__ err) {
  // The block's records are staged in LDS (rows padded by one 16-byte piece
  // against bank conflicts) and written back as one contiguous run: 16-byte
  // pieces of consecutive records from consecutive threads.
  __shared__ double2 stage[256 * 9];
  const int64_t e0 = (int64_t)blockIdx.x * 256, e = e0 + threadIdx.x;
  const int nblk = (int)min<int64_t>(256, n_cv - e0);
  if (e < n_cv) {
    const int4 v = *reinterpret_cast<const int4 *>(cvx + 4 * e);
    const double *p0 = pts + 3 * (int64_t)v.x, *p1 = pts + 3 * (int64_t)v.y;
    const double *p2 = pts + 3 * (int64_t)v.z, *p3 = pts + 3 * (int64_t)v.w;
    // column-major K(r,c) = k[r + 3c] = x_{c+1}[r] - x_0[r]  (G . grad N for
    // the P1 transformation)
    double k[9];
    for (int r = 0; r < 3; ++r) {
      const double x0 = p0[r];
      k[r] = p1[r] - x0;
      k[r + 3] = p2[r] - x0;
      k[r + 6] = p3[r] - x0;
    }
    const double a0 = k[4] * k[8] - k[5] * k[7];
    const double a1 = k[5] * k[6] - k[3] * k[8];
    const double a2 = k[3] * k[7] - k[4] * k[6];
    const double J = k[0] * a0 + k[1] * a1 + k[2] * a2;
    if (J == 0.0) { atomicOr(err, 1); }
    // Record: B = K^{-T} from cofactors, row-major B(r,c) = o[3r+c]; then the
    // factorized form's scales |J| a s and |J| b s (s: the scale of the stored
    // reference tensor, 1/L for an integer table R = N/L), the insertion
    // threshold, the per-point form's |J| a, |J| b, |J| c, and |J| c s
    double o[16];
    o[0] = a0 / J; o[1] = (k[2] * k[7] - k[1] * k[8]) / J;
    o[2] = (k[1] * k[5] - k[2] * k[4]) / J; o[3] = a1 / J;
    o[4] = (k[0] * k[8] - k[2] * k[6]) / J; o[5] = (k[2] * k[3] - k[0] * k[5]) / J;
    o[6] = a2 / J; o[7] = (k[1] * k[6] - k[0] * k[7]) / J;
    o[8] = (k[0] * k[4] - k[1] * k[3]) / J;
    const double aJ = fabs(J), va = ca ? ca[e] : ca_c, vb = cb ? cb[e] : cb_c;
    const double vc = cc ? cc[e] : cc_c;
    o[9] = aJ * va * r_scale; o[10] = aJ * vb * r_scale;
    o[11] = thr ? thr[e] : 0.0;
    o[12] = aJ * va; o[13] = aJ * vb; o[14] = aJ * vc; o[15] = aJ * vc * r_scale;
    if (rec32) {
      float4 *st4 = reinterpret_cast<float4 *>(stage + 9 * threadIdx.x);
      for (int t = 0; t < 4; ++t)
        st4[t] = make_float4((float)o[4 * t], (float)o[4 * t + 1], (float)o[4 * t + 2], (float)o[4 * t + 3]);
    } else {
      for (int t = 0; t < 8; ++t) stage[9 * threadIdx.x + t] = make_double2(o[2 * t], o[2 * t + 1]);
    }
  }
  __syncthreads();
  // 16-byte pieces: 8 per FP64 record, 4 per FP32 record
  const int per = rec32 ? 4 : 8;
  double2 *dst = rec32 ? reinterpret_cast<double2 *>(rec32 + (int64_t)kRecStride * e0)
                       : reinterpret_cast<double2 *>(rec + (int64_t)kRecStride * e0);
  for (int x = threadIdx.x; x < nblk * per; x += 256) {
    const int r = x / per, piece = x - r * per;
    dst[x] = stage[9 * r + piece];
  }
}

// ------------------------------------------------------------------------
// element pair blocks: rows (i, a), columns (j, b) of the element matrix
// ------------------------------------------------------------------------
// RK: format of the reference tensor R for the factorized form: 0 = Real
// (9 per block), 1 = int8 N with R = N/L (16-byte blocks), 2 = int16 N
// (32-byte blocks); the scale 1/L is folded into the record's |J| a s, |J| b s.
template <int RK>
__device__ __forceinline__ void ref_block_int(const void *__restrict__ R, int64_t blk, int (&n)[9]) {
  if constexpr (RK == 1) {
    const int4 w = reinterpret_cast<const int4 *>(R)[blk];
    const int ww[3] = {w.x, w.y, w.z};
    for (int t = 0; t < 9; ++t) n[t] = (int)(int8_t)(ww[t >> 2] >> (8 * (t & 3)));
  } else {
    const int4 w0 = reinterpret_cast<const int4 *>(R)[2 * blk];
    const int w1 = reinterpret_cast<const int *>(R)[8 * blk + 4];
    const int ww[5] = {w0.x, w0.y, w0.z, w0.w, w1};
    for (int t = 0; t < 9; ++t) n[t] = (int)(int16_t)(ww[t >> 1] >> (16 * (t & 1)));
  }
}

template <int Q, int TERM, int FORM, typename Real, int RK = 0>
__device__ __forceinline__ void pair_block(const Real *__restrict__ rec, int i, int j,
                                           int nd, int nq, const Real *__restrict__ gref,
                                           const Real *__restrict__ qw,
                                           const void *__restrict__ R, Real v[Q][Q]) {
  // Terms (rows (i, a), columns (j, b), g = B grad(phi) at a point):
  //   LAPLACE     a g_i.g_j delta_ab
  //   ISO_ELAST   lambda g_ia g_jb + mu g_ib g_ja + mu delta_ab g_i.g_j
  //   ISOTROPIC   alpha g_ia g_jb + beta g_ib g_ja + gamma delta_ab g_i.g_j
  Real B[9];
  for (int t = 0; t < 9; ++t) B[t] = (Real)rec[t];
  if (FORM == GFX_FORM_FACTORIZED) {
    // G = B R_ij B^T, R_ij = sum_q w_q gradref_i gradref_j^T
    const int64_t blk = (int64_t)i * nd + j;
    Real rr[9];
    if constexpr (RK == 0) {
      const Real *r = reinterpret_cast<const Real *>(R) + blk * 9;
      for (int t = 0; t < 9; ++t) rr[t] = r[t];
    } else {
      int n[9];
      ref_block_int<RK>(R, blk, n);
      for (int t = 0; t < 9; ++t) rr[t] = (Real)n[t];
    }
    const Real sa = rec[9], sb = rec[10];  // |J| a s, |J| b s
    Real T[3][3], G[3][3];
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        T[c][b] = rr[3 * c] * B[3 * b] + rr[3 * c + 1] * B[3 * b + 1] + rr[3 * c + 2] * B[3 * b + 2];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        G[a][b] = B[3 * a] * T[0][b] + B[3 * a + 1] * T[1][b] + B[3 * a + 2] * T[2][b];
    const Real tr = G[0][0] + G[1][1] + G[2][2];
    if (TERM == GFX_TERM_ISO_ELAST) {
      for (int a = 0; a < Q; ++a)
        for (int b = 0; b < Q; ++b)
          v[a][b] = sa * G[a][b] + sb * (G[b][a] + (a == b ? tr : Real(0)));
    } else if (TERM == GFX_TERM_ISOTROPIC) {
      const Real sc = rec[15];  // |J| c s
      for (int a = 0; a < Q; ++a)
        for (int b = 0; b < Q; ++b)
          v[a][b] = sa * G[a][b] + sb * G[b][a] + (a == b ? sc * tr : Real(0));
    } else {
      for (int a = 0; a < Q; ++a)
        for (int b = 0; b < Q; ++b) v[a][b] = (a == b) ? sa * tr : Real(0);
    }
  } else {
    const Real ca = rec[12], cb = rec[13];  // |J| a, |J| b
    const Real cc = TERM == GFX_TERM_ISOTROPIC ? rec[14] : cb;
    for (int a = 0; a < Q; ++a)
      for (int b = 0; b < Q; ++b) v[a][b] = Real(0);
    for (int q = 0; q < nq; ++q) {
      const Real *hi = gref + ((int64_t)q * nd + i) * 3;
      const Real *hj = gref + ((int64_t)q * nd + j) * 3;
      const Real h0 = (Real)hi[0], h1 = (Real)hi[1], h2 = (Real)hi[2];
      const Real k0 = (Real)hj[0], k1 = (Real)hj[1], k2 = (Real)hj[2];
      Real gi[3], gj[3];
      for (int a = 0; a < 3; ++a) {
        gi[a] = B[3 * a] * h0 + B[3 * a + 1] * h1 + B[3 * a + 2] * h2;
        gj[a] = B[3 * a] * k0 + B[3 * a + 1] * k1 + B[3 * a + 2] * k2;
      }
      const Real wq = (Real)qw[q];
      const Real d = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
      if (TERM == GFX_TERM_ISO_ELAST || TERM == GFX_TERM_ISOTROPIC) {
        for (int a = 0; a < Q; ++a)
          for (int b = 0; b < Q; ++b)
            v[a][b] += wq * (ca * gi[a] * gj[b] + cb * gi[b] * gj[a] + (a == b ? cc * d : Real(0)));
      } else {
        for (int a = 0; a < Q; ++a) v[a][a] += wq * (ca * d);
      }
    }
  }
}

// ------------------------------------------------------------------------
// incidences: node -> (element, local node), counting sort
// ------------------------------------------------------------------------
__global__ void k_inc_count(int64_t n, int nd, int qdim, const int32_t *__restrict__ dofs,
                            int32_t *__restrict__ cnt) {
  const int64_t t = gtid();
  if (t >= n) return;
  atomicAdd(&cnt[dofs[t] / qdim], 1);
}

__global__ void k_inc_fill(int64_t n, int nd, int qdim, const int32_t *__restrict__ dofs,
                           const int32_t *__restrict__ inc_ptr, int32_t *__restrict__ cursor,
                           uint32_t *__restrict__ inc) {
  const int64_t t = gtid();
  if (t >= n) return;
  const int node = dofs[t] / qdim;
  const int pos = inc_ptr[node] + atomicAdd(&cursor[node], 1);
  const int64_t e = t / nd;
  const int i = (int)(t - e * nd);
  inc[pos] = ((uint32_t)e << 5) | (uint32_t)i;
}

// deterministic order inside each node's list (element, then local index)
__global__ void k_inc_sort(int64_t n_nodes, const int32_t *__restrict__ inc_ptr,
                           uint32_t *__restrict__ inc) {
  const int64_t n = gtid();
  if (n >= n_nodes) return;
  const int b = inc_ptr[n], e = inc_ptr[n + 1];
  for (int i = b + 1; i < e; ++i) {
    const uint32_t x = inc[i];
    int j = i - 1;
    while (j >= b && inc[j] > x) { inc[j + 1] = inc[j]; --j; }
    inc[j + 1] = x;
  }
}

// ------------------------------------------------------------------------
// per-element threshold and keep bits (one wavefront per element)
// ------------------------------------------------------------------------
template <int Q, int TERM>
__global__ __launch_bounds__(256) void k_keep(PatternArgs a) {
  extern __shared__ uint32_t kbits[];  // 4 waves x kwords
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int64_t e = (int64_t)blockIdx.x * 4 + w;
  const bool live = e < a.n_cv;
  const int nd = a.nd, nl = a.nloc, npair = nd * nd;
  uint32_t *my = kbits + w * a.kwords;
  for (int t = lane; t < a.kwords; t += kWave) my[t] = 0u;
  double mx = 0.0;
  const double *rec = a.rec + (int64_t)kRecStride * (live ? e : 0);
  if (live) {
    for (int p = lane; p < npair; p += kWave) {
      double v[Q][Q];
      pair_block<Q, TERM, GFX_FORM_PERQP, double>(rec, p / nd, p % nd, nd, a.nq, a.gref, a.qw,
                                                  nullptr, v);
      for (int x = 0; x < Q; ++x)
        for (int y = 0; y < Q; ++y) mx = fmax(mx, fabs(v[x][y]));
    }
  }
  mx = wave_max(mx);
  const double thr = mx * 1e-14;
  __syncthreads();
  if (live && mx > 0.0) {  // an all-zero element is skipped (:5441-5442)
    for (int p = lane; p < npair; p += kWave) {
      const int i = p / nd, j = p % nd;
      double v[Q][Q];
      pair_block<Q, TERM, GFX_FORM_PERQP, double>(rec, i, j, nd, a.nq, a.gref, a.qw, nullptr, v);
      for (int x = 0; x < Q; ++x)
        for (int y = 0; y < Q; ++y)
          if (fabs(v[x][y]) > thr) {
            const int bit = (i * Q + x) * nl + (j * Q + y);
            atomicOr(&my[bit >> 5], 1u << (bit & 31));
          }
    }
  }
  __syncthreads();
  if (live) {
    if (lane == 0) a.thr[e] = (mx > 0.0) ? thr : INFINITY;
    for (int t = lane; t < a.kwords; t += kWave) a.keep[e * a.kwords + t] = my[t];
  }
}

// ------------------------------------------------------------------------
// node patches: sorted unique neighbour nodes, slots, presence bits
// ------------------------------------------------------------------------
__device__ void block_bitonic_sort(uint64_t *k, int P) {
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const uint64_t x = k[lo], y = k[hi];
        if ((x > y) == up) { k[lo] = y; k[hi] = x; }
      }
      __syncthreads();
    }
}

// Dynamic LDS, sized by the launcher for the largest candidate list pmax
// (a power of two): [keys: pmax u64][slots: pmax u16][presence: mwords u32].
__global__ __launch_bounds__(256) void k_patch_build(PatternArgs a, int pmax) {
  extern __shared__ uint64_t lds_pb[];
  uint64_t *keys = lds_pb;
  uint16_t *slotL = reinterpret_cast<uint16_t *>(keys + pmax);
  uint32_t *maskL = reinterpret_cast<uint32_t *>(slotL + pmax + (pmax & 1));
  __shared__ int32_t part[256];
  __shared__ int32_t rowc[3];
  const int64_t node = blockIdx.x;
  if (node >= a.n_nodes) return;
  const int nd = a.nd, Q = a.qdim, nl = a.nloc;
  const int i0 = a.inc_ptr[node], ninc = a.inc_ptr[node + 1] - i0;
  const int ncand = ninc * nd;
  if (ncand > pmax) { if (threadIdx.x == 0) atomicOr(a.err, 2); return; }
  int P = 1;
  while (P < ncand) P <<= 1;
  for (int t = threadIdx.x; t < P; t += blockDim.x) {
    uint64_t key = ~0ull;
    if (t < ncand) {
      const uint32_t inc = a.inc[i0 + t / nd];
      const int64_t e = inc >> 5;
      const int32_t nb = a.dofs[e * nd + (t % nd)] / Q;
      key = ((uint64_t)(uint32_t)nb << 32) | (uint32_t)t;
    }
    keys[t] = key;
  }
  if (threadIdx.x < 3) rowc[threadIdx.x] = 0;
  __syncthreads();
  block_bitonic_sort(keys, P);
  // unique flags -> slot numbers (block scan, 16 items per thread)
  const int per = (P + blockDim.x - 1) / blockDim.x;
  const int tb = threadIdx.x * per;
  int s = 0;
  for (int t = tb; t < tb + per && t < ncand; ++t)
    s += (t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32)) ? 1 : 0;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;  // exclusive prefix
  const int ns = part[blockDim.x - 1];
  for (int t = tb; t < tb + per && t < ncand; ++t) {
    const bool u = (t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32));
    if (u) {
      a.nbr[(int64_t)i0 * nd + run] = (int32_t)(keys[t] >> 32);
      ++run;
    }
    const int orig = (int)(keys[t] & 0xffffffffu);
    slotL[orig] = (uint16_t)(run - 1);
  }
  if (ns >= kFullRows) { if (threadIdx.x == 0) atomicOr(a.err, 2); return; }
  const int mw = (Q * Q * ns + 31) / 32;
  for (int t = threadIdx.x; t < mw; t += blockDim.x) maskL[t] = 0u;
  __syncthreads();
  for (int t = threadIdx.x; t < ncand; t += blockDim.x) a.slot[(int64_t)i0 * nd + t] = slotL[t];
  // presence: OR of the keep bits of the node's rows over its incidences
  const int per_inc = Q * nd * Q;
  for (int t = threadIdx.x; t < ninc * per_inc; t += blockDim.x) {
    const int k = t / per_inc, rem = t - k * per_inc;
    const int x = rem / (nd * Q), c = rem - x * (nd * Q);  // row component x, column c=(j,y)
    const uint32_t inc = a.inc[i0 + k];
    const int64_t e = inc >> 5;
    const int i = inc & 31;
    const int bit = (i * Q + x) * nl + c;
    if ((a.keep[e * a.kwords + (bit >> 5)] >> (bit & 31)) & 1u) {
      const int j = c / Q, y = c - j * Q;
      const int pos = (x * ns + slotL[k * nd + j]) * Q + y;
      atomicOr(&maskL[pos >> 5], 1u << (pos & 31));
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < mw; t += blockDim.x) {
    const uint32_t m = maskL[t];
    a.mask[(int64_t)i0 * a.mwpi + t] = m;
    // split the word's bits by row (row x spans bits [x*ns*Q, (x+1)*ns*Q))
    for (int x = 0; x < Q; ++x) {
      const int lo = x * ns * Q, hi = lo + ns * Q;
      const int wl = t * 32, wh = wl + 32;
      const int l = max(lo, wl), h = min(hi, wh);
      if (l < h) {
        uint32_t mm = m >> (l - wl);
        if (h - l < 32) mm &= (1u << (h - l)) - 1u;
        atomicAdd(&rowc[x], __popc(mm));
      }
    }
  }
  __syncthreads();
  // bit 15 of nslots: every (row component, slot, column component) is
  // present, so the node's rows are a plain copy of its dense accumulator
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int x = 0; x < Q; ++x) tot += rowc[x];
    a.nslots[node] = (uint16_t)(ns | (tot == Q * Q * ns ? kFullRows : 0));
  }
  if (threadIdx.x < Q) a.rowcnt[node * Q + threadIdx.x] = rowc[threadIdx.x];
}

// Re-validation of a built pattern after the values it was built from changed
// (coefficients, node positions, the state of a nonlinear plan): the
// presence words of every node are recomputed from the new keep bits through
// the stored slot maps (the slots depend on the topology only) and compared
// with the stored ones.  Any difference sets bit 4 of the error flag: the
// stored pattern is no longer GetFEM's and must be rebuilt.  Dynamic LDS:
// mwords u32 (the largest node's presence words).
// One thread per node: a node stored as full whose incidence rows have every
// keep bit set has every presence bit set again (the converse needs the exact
// check of k_mask_check, which the other nodes take).
__global__ __launch_bounds__(256) void k_mask_fast(PatternArgs a) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= a.n_nodes) return;
  const int Q = a.qdim, nl = a.nloc;
  bool ok = (a.nslots[node] & kFullRows) != 0;
  for (int k = a.inc_ptr[node]; ok && k < a.inc_ptr[node + 1]; ++k) {
    const uint32_t inc = a.inc[k];
    const uint32_t *kw = a.keep + (int64_t)(inc >> 5) * a.kwords;
    const int i = inc & 31;
    for (int x = 0; ok && x < Q; ++x) {
      const int b0 = (i * Q + x) * nl, b1 = b0 + nl;  // the row's bits [b0, b1)
      for (int b = b0; ok && b < b1;) {
        const int w = b >> 5, lo = b & 31, hi = min(32, b1 - (w << 5));
        const uint32_t m = (hi - lo == 32 ? 0xffffffffu : ((1u << (hi - lo)) - 1u)) << lo;
        ok = (kw[w] & m) == m;
        b = (w << 5) + hi;
      }
    }
  }
  a.node_ok[node] = ok ? 1 : 0;
  if (!ok) atomicOr(a.err, 16);  // some node needs the exact check
}

__global__ __launch_bounds__(128) void k_mask_check(PatternArgs a) {
  extern __shared__ uint32_t maskC[];
  const int64_t node = blockIdx.x;
  if (node >= a.n_nodes) return;
  const int nd = a.nd, Q = a.qdim, nl = a.nloc;
  const int i0 = a.inc_ptr[node], ninc = a.inc_ptr[node + 1] - i0;
  if (a.node_ok[node]) return;  // confirmed by k_mask_fast (uniform)
  const int ns = a.nslots[node] & (kFullRows - 1);
  const int mw = (Q * Q * ns + 31) / 32;
  for (int t = threadIdx.x; t < mw; t += blockDim.x) maskC[t] = 0u;
  __syncthreads();
  const int per_inc = Q * nd * Q;
  for (int t = threadIdx.x; t < ninc * per_inc; t += blockDim.x) {
    const int k = t / per_inc, rem = t - k * per_inc;
    const int x = rem / (nd * Q), c = rem - x * (nd * Q);
    const uint32_t inc = a.inc[i0 + k];
    const int64_t e = inc >> 5;
    const int i = inc & 31;
    const int bit = (i * Q + x) * nl + c;
    if ((a.keep[e * a.kwords + (bit >> 5)] >> (bit & 31)) & 1u) {
      const int j = c / Q, y = c - j * Q;
      const int pos = (x * ns + (int)a.slot[(int64_t)(i0 + k) * nd + j]) * Q + y;
      atomicOr(&maskC[pos >> 5], 1u << (pos & 31));
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < mw; t += blockDim.x)
    if (maskC[t] != a.mask[(int64_t)i0 * a.mwpi + t]) atomicOr(a.err, 4);
}

__device__ __forceinline__ bool mask_bit(const uint32_t *m, int b) {
  return (m[b >> 5] >> (b & 31)) & 1u;
}

__global__ __launch_bounds__(64) void k_fill_cols(PatternArgs a) {
  const int64_t node = blockIdx.x;
  if (node >= a.n_nodes) return;
  const int lane = threadIdx.x, Q = a.qdim, nd = a.nd;
  const int i0 = a.inc_ptr[node];
  const int ns = a.nslots[node] & (kFullRows - 1);
  const uint32_t *m = a.mask + (int64_t)i0 * a.mwpi;
  const int32_t *nb = a.nbr + (int64_t)i0 * nd;
  for (int x = 0; x < Q; ++x) {
    int64_t base = a.indptr[node * Q + x];
    const int off = x * ns * Q, len = ns * Q;
    for (int c0 = 0; c0 < len; c0 += kWave) {
      const int d = c0 + lane;
      const bool pr = d < len && mask_bit(m, off + d);
      const uint64_t bal = __ballot(pr);
      if (pr) {
        const int rank = __popcll(bal & ((1ull << lane) - 1ull));
        a.indices[base + rank] = nb[d / Q] * Q + (d % Q);
      }
      base += __popcll(bal);
    }
  }
}

__global__ void k_fill_desc(int64_t n, int qdim, const int32_t *__restrict__ order,
                            const int32_t *__restrict__ inc_ptr, const uint16_t *__restrict__ nslots,
                            const int64_t *__restrict__ indptr, PatchDesc *__restrict__ desc) {
  const int64_t p = gtid();
  if (p >= n) return;
  const int64_t node = order[p];
  PatchDesc d{0, 0, 0u};  // node -1: an empty slot of a wave's descriptor group
  if (node >= 0) {
    d.rowbase = indptr[node * qdim];
    d.i0 = inc_ptr[node];
    d.ninc_ns = (uint32_t)(inc_ptr[node + 1] - d.i0) | ((uint32_t)nslots[node] << 16);
  }
  desc[p] = d;
}

// ------------------------------------------------------------------------
// the assembly kernel: one workgroup (one wavefront) per group of node patches
// ------------------------------------------------------------------------
// A wavefront assembles up to kMaxNpw node patches at once, given by its
// group of kMaxNpw descriptors (empty slots are zero).  The host packs
// consecutive nodes (node order is spatial order) into a wave while their
// incidences fill about three passes of the lanes and their LDS fits the
// launch's budget, so one launch sweeps the mesh in order and the element
// records of a region are read from HBM about once.  The incidence lists of
// the wave's patches are concatenated: the lanes of one pass may serve
// different patches, and the chain of dependent loads (descriptors ->
// incidences -> element records) is paid once for all of them.
// Dynamic LDS: [accumulators: lds_acc doubles, patch k at the prefix sum of
// Q*Q*ns] then one flat run of lds_flat 4-byte words: the presence words of
// the patches that are not full, the incidences, and the slot maps as u16
// pairs (ND is even).  The run is filled by one round of independent loads
// (kStage words per lane in flight), so staging costs one memory round trip.
constexpr int kMaxNpw = 4;
constexpr int kStage = 6;

template <typename T>
__device__ __forceinline__ T pick4(const T (&v)[kMaxNpw], int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}

template <int ND, int Q, int TERM, int FORM, typename Real, int RK>
__global__ __launch_bounds__(kWave) void k_assemble_patch(PatchArgs a) {
  static_assert(ND % 2 == 0, "slot maps are staged as u16 pairs");
  extern __shared__ double lds[];
  const int64_t wg = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t w = (wg & 7) * (int64_t)a.xcd_chunk + (wg >> 3);  // XCD x: a contiguous run of waves
  if (w >= a.n_waves) return;
  const PatchDesc *wd = a.desc + w * kMaxNpw;
  int i0[kMaxNpw], ns[kMaxNpw], st[kMaxNpw], pre[kMaxNpw + 1], mst[kMaxNpw], mpre[kMaxNpw + 1];
  int aoff[kMaxNpw];
  bool full[kMaxNpw];
  int64_t rb[kMaxNpw];
  pre[0] = 0;
  mpre[0] = 0;
  int apre = 0;
#pragma unroll
  for (int k = 0; k < kMaxNpw; ++k) {
    const PatchDesc d = wd[k];
    const int nsraw = (int)(d.ninc_ns >> 16);
    i0[k] = d.i0;
    ns[k] = nsraw & (kFullRows - 1);
    full[k] = (nsraw & kFullRows) != 0;
    rb[k] = d.rowbase;
    st[k] = pre[k];
    pre[k + 1] = pre[k] + (int)(d.ninc_ns & 0xffffu);
    mst[k] = mpre[k];
    mpre[k + 1] = mpre[k] + (full[k] ? 0 : (Q * Q * ns[k] + 31) / 32);
    aoff[k] = apre;
    apre += Q * Q * ns[k];
  }
  const int total = pre[kMaxNpw];  // incidences of the wave's patches
  const int W = mpre[kMaxNpw];     // presence words of the wave's patches
  constexpr int HALF = ND / 2;
  const int nwords = W + total + total * HALF;
  // pair blocks per lane: 2 for large elements (P3: 6 incidences per pass
  // instead of 3, half the passes of a vertex patch), else 1
  constexpr int JB = ND >= 20 ? 2 : 1;
  constexpr int LPI = ND / JB;      // lanes per incidence
  constexpr int PER = kWave / LPI;  // incidences per pass
  double *acc = lds;
  uint32_t *smask = reinterpret_cast<uint32_t *>(lds + a.lds_acc);
  const uint32_t *sinc = smask + W;
  const uint16_t *sslot = reinterpret_cast<const uint16_t *>(sinc + total);
  const uint32_t *slot32 = reinterpret_cast<const uint32_t *>(a.slot);
  // patch of combined incidence c: k = number of patch starts at or below c
  auto patch_of = [&](int c) { return (int)(c >= pre[1]) + (int)(c >= pre[2]) + (int)(c >= pre[3]); };
  // The staged words form 12 runs (presence words, incidences, slot pairs of
  // patches 0..3), each a contiguous range of its source array: run q starts
  // at flat index rs[q] and maps t to source word ro[q] + t (mod 2^32).  The
  // source offset of t is ro[0] + t plus the steps ro[q] - ro[q-1] of every
  // run start at or below t (a sum of selects: no indexed array, so the
  // table stays in scalar registers; empty runs add their step and the next
  // run's step together, which leaves the next run's offset).
  int rs[3 * kMaxNpw];
  uint32_t ro[3 * kMaxNpw];
#pragma unroll
  for (int k = 0; k < kMaxNpw; ++k) {
    rs[k] = mst[k];
    ro[k] = (uint32_t)i0[k] * (uint32_t)a.mwpi - (uint32_t)rs[k];
    rs[kMaxNpw + k] = W + st[k];
    ro[kMaxNpw + k] = (uint32_t)i0[k] - (uint32_t)rs[kMaxNpw + k];
    rs[2 * kMaxNpw + k] = W + total + st[k] * HALF;
    ro[2 * kMaxNpw + k] = (uint32_t)i0[k] * HALF - (uint32_t)rs[2 * kMaxNpw + k];
  }
  uint32_t step[3 * kMaxNpw];
#pragma unroll
  for (int q = 1; q < 3 * kMaxNpw; ++q) step[q] = ro[q] - ro[q - 1];
  auto source = [&](int t) -> const uint32_t * {
    uint32_t off = ro[0] + (uint32_t)t;
#pragma unroll
    for (int q = 1; q < 3 * kMaxNpw; ++q) off += t >= rs[q] ? step[q] : 0u;
    const uint32_t *base = t < W ? a.mask : (t < W + total ? a.inc : slot32);
    return base + off;
  };
  // Loads are unconditional (the index is clamped to the last word, so every
  // address is valid) and only the LDS stores are masked: no branches between
  // the loads, so all kStage of them are in flight before the first wait.
  const int last = nwords > 0 ? nwords - 1 : 0;
  uint32_t buf[kStage];
  if (nwords > 0) {
#pragma unroll
    for (int r = 0; r < kStage; ++r) buf[r] = *source(min(r * kWave + lane, last));
  }
  // clear the accumulators while the staging loads are in flight
  for (int t = lane; t < apre; t += kWave) acc[t] = 0.0;
  if (nwords > 0) {
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
      const int t = r * kWave + lane;
      if (t < nwords) smask[t] = buf[r];
    }
  }
  for (int t0 = kStage * kWave; t0 < nwords; t0 += kStage * kWave) {  // large waves
#pragma unroll
    for (int r = 0; r < kStage; ++r) buf[r] = *source(min(t0 + r * kWave + lane, last));
#pragma unroll
    for (int r = 0; r < kStage; ++r) {
      const int t = t0 + r * kWave + lane;
      if (t < nwords) smask[t] = buf[r];
    }
  }
  __syncthreads();
  // Lanes: LPI per incidence, each computing the JB pair blocks (j, j + LPI,
  // ...) of its incidence, so the element record is loaded once for JB blocks.
  const int kk = lane / LPI, jl = lane - kk * LPI;
  // tables in the contraction's precision (the FP32 form reads FP32 records:
  // half the bytes per lane)
  const Real *tab_rec, *tab_gref, *tab_qw;
  const void *tab_R;
  if constexpr (sizeof(Real) == 4) {
    tab_rec = a.rec32; tab_gref = a.gref32; tab_qw = a.qw32; tab_R = a.R32;
  } else {
    tab_rec = a.rec; tab_gref = a.gref; tab_qw = a.qw; tab_R = a.R;
  }
  if constexpr (RK != 0) tab_R = a.Rq;  // integer reference tensor (both precisions)
  for (int c0 = 0; c0 < total; c0 += PER) {
    const int c = c0 + kk;
    if (kk < PER && c < total) {
      const int k = patch_of(c), nsk = pick4(ns, k);
      double *ak = acc + pick4(aoff, k);
      const uint32_t inc = sinc[c];
      const int64_t e = inc >> 5;
      const int i = inc & 31;
      const Real *rec = tab_rec + (int64_t)kRecStride * e;
      const double thr = (double)rec[11];
#pragma unroll 1
      for (int jb = 0; jb < JB; ++jb) {
        const int j = jl + jb * LPI;
        Real v[Q][Q];
        pair_block<Q, TERM, FORM, Real, RK>(rec, i, j, ND, a.nq, tab_gref, tab_qw, tab_R, v);
        const int s = sslot[c * ND + j];
        for (int x = 0; x < Q; ++x)
          for (int y = 0; y < Q; ++y) {
            const double val = (double)v[x][y];
            // a dropped entry adds +0.0 (the sum is unchanged: the
            // accumulator starts at +0.0 and is never -0.0), so the adds
            // need no branch
            atomicAdd(&ak[(x * nsk + s) * Q + y], fabs(val) > thr ? val : 0.0);
          }
      }
    }
  }
  __syncthreads();
  // The values are written once and not read again by this kernel: stream
  // them past the caches (non-temporal stores), which keeps L2 for the
  // element records and incidence data the neighbouring patches reuse.
  // A node's Q rows are consecutive CSR rows and consecutive runs of its
  // dense accumulator, so the node's values are one compaction of the
  // Q*ns*Q dense entries by the presence bits, starting at the node's first
  // CSR value.  The LDS reads of kOut chunks are issued together.
  constexpr int kOut = 4;
#pragma unroll
  for (int k = 0; k < kMaxNpw; ++k) {
    const double *ak = acc + aoff[k];
    const int n_d = Q * ns[k] * Q;  // 0 for an empty slot
    if (full[k]) {  // every entry present: one contiguous run, 16-byte stores
      double *dst = a.vals + rb[k];
      const int head = (int)(rb[k] & 1);  // one 8-byte value up to 16-byte alignment
      if (head && lane == 0 && n_d > 0) __builtin_nontemporal_store(ak[0], dst);
      const int npair = (n_d - head) >> 1;
      typedef double double2v __attribute__((ext_vector_type(2)));
      double2v *dst2 = reinterpret_cast<double2v *>(dst + head);
      for (int t0 = 0; t0 < npair; t0 += kOut * kWave) {
        double2v v2[kOut];
#pragma unroll
        for (int u = 0; u < kOut; ++u) {  // clamped: the reads are unconditional
          const int t = min(t0 + u * kWave + lane, max(npair - 1, 0));
          v2[u].x = ak[head + 2 * t];
          v2[u].y = ak[head + 2 * t + 1];
        }
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
          const int t = t0 + u * kWave + lane;
          if (t < npair) __builtin_nontemporal_store(v2[u], dst2 + t);
        }
      }
      if (((n_d - head) & 1) && lane == 0) __builtin_nontemporal_store(ak[n_d - 1], dst + n_d - 1);
      continue;
    }
    const uint32_t *sm = smask + mst[k];
    int64_t base = rb[k];
    for (int c0 = 0; c0 < n_d; c0 += kOut * kWave) {
      uint32_t word[kOut];
      double val[kOut];
#pragma unroll
      for (int u = 0; u < kOut; ++u) {  // clamped indices: the reads are unconditional
        const int dc = min(c0 + u * kWave + lane, n_d - 1);
        word[u] = sm[dc >> 5];
        val[u] = ak[dc];
      }
#pragma unroll
      for (int u = 0; u < kOut; ++u) {
        const int d = c0 + u * kWave + lane;
        const bool pr = (d < n_d) & (bool)((word[u] >> (d & 31)) & 1u);
        const uint64_t bal = __ballot(pr);
        if (pr) __builtin_nontemporal_store(val[u], a.vals + base + __popcll(bal & ((1ull << lane) - 1ull)));
        base += __popcll(bal);
      }
    }
  }
}

// ------------------------------------------------------------------------
// scans
// ------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanItems = 8, kScanBlock = kScanThreads * kScanItems;

template <typename T>
__device__ T block_exclusive_scan(T v, T *sh, T &total) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const T u = threadIdx.x >= off ? sh[threadIdx.x - off] : T(0);
    __syncthreads();
    sh[threadIdx.x] += u;
    __syncthreads();
  }
  total = sh[kScanThreads - 1];
  const T r = sh[threadIdx.x] - v;
  __syncthreads();
  return r;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const T *in, int64_t n, T *sums) {
  __shared__ T sh[kScanThreads];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  T s = 0;
  for (int t = 0; t < kScanItems; ++t)
    if (b0 + t < n) s += in[b0 + t];
  T total;
  block_exclusive_scan<T>(s, sh, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// out[i] = offset + exclusive prefix; out[n] = total when last block
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T *in, int64_t n, T *out,
                                                             const T *offs) {
  __shared__ T sh[kScanThreads];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  T loc[kScanItems];
  T s = 0;
  for (int t = 0; t < kScanItems; ++t) {
    loc[t] = (b0 + t < n) ? in[b0 + t] : T(0);
    s += loc[t];
  }
  T total;
  T run = block_exclusive_scan<T>(s, sh, total) + (offs ? offs[blockIdx.x] : T(0));
  for (int t = 0; t < kScanItems; ++t) {
    if (b0 + t < n) out[b0 + t] = run;
    run += loc[t];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[n] = run;
}

template <typename T>
hipError_t scan_exclusive(hipStream_t s, const T *in, T *out, int64_t n, T *tmp) {
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb <= 1) {
    k_scan_apply<T><<<1, kScanThreads, 0, s>>>(in, n, out, nullptr);
    return hipGetLastError();
  }
  T *sums = tmp, *soff = tmp + nb, *rest = tmp + 2 * nb + 1;
  k_scan_sums<T><<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, sums);
  hipError_t err = scan_exclusive<T>(s, sums, soff, nb, rest);
  if (err != hipSuccess) return err;
  k_scan_apply<T><<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, out, soff);
  return hipGetLastError();
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t t = gtid();
  if (t < n) p[t] = v;
}

}  // namespace

int64_t scan_tmp_elems(int64_t n) {
  int64_t tot = 0;
  while (true) {
    const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb <= 1) break;
    tot += 2 * nb + 1;
    n = nb;
  }
  return tot + 16;
}

hipError_t scan_exclusive_i64(hipStream_t s, const int64_t *in, int64_t *out, int64_t n,
                              int64_t *tmp) {
  return scan_exclusive<int64_t>(s, in, out, n, tmp);
}
hipError_t scan_exclusive_i32(hipStream_t s, const int32_t *in, int32_t *out, int64_t n,
                              int32_t *tmp) {
  return scan_exclusive<int32_t>(s, in, out, n, tmp);
}

static unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_records_tet(hipStream_t s, int64_t n_cv, const double *pts,
                              const int32_t *cvx, const double *ca, double ca_c,
                              const double *cb, double cb_c, const double *cc, double cc_c,
                              const double *thr, double r_scale, double *rec, float *rec32,
                              int *err) {
  if (n_cv == 0) return hipSuccess;
  k_records_tet<<<nblocks(n_cv, 256), 256, 0, s>>>(n_cv, pts, cvx, ca, ca_c, cb, cb_c, cc, cc_c,
                                                    thr, r_scale, rec, rec32, err);
  return hipGetLastError();
}

hipError_t launch_fill_desc(hipStream_t s, int64_t n, int qdim, const int32_t *order, const int32_t *inc_ptr,
                            const uint16_t *nslots, const int64_t *indptr, PatchDesc *desc) {
  if (n == 0) return hipSuccess;
  k_fill_desc<<<nblocks(n, 256), 256, 0, s>>>(n, qdim, order, inc_ptr, nslots, indptr, desc);
  return hipGetLastError();
}

hipError_t launch_incidences(hipStream_t s, int64_t n_cv, int nd, int qdim,
                             const int32_t *dofs, int64_t n_nodes, int32_t *inc_ptr,
                             uint32_t *inc, int32_t *tmp) {
  // tmp: n_nodes counts + n_nodes cursors + scan scratch
  int32_t *cnt = tmp, *cursor = tmp + n_nodes, *scr = tmp + 2 * n_nodes;
  const int64_t n = n_cv * nd;
  k_fill_i32<<<nblocks(2 * n_nodes, 256), 256, 0, s>>>(tmp, 2 * n_nodes, 0);
  k_inc_count<<<nblocks(n, 256), 256, 0, s>>>(n, nd, qdim, dofs, cnt);
  hipError_t err = scan_exclusive_i32(s, cnt, inc_ptr, n_nodes, scr);
  if (err != hipSuccess) return err;
  k_inc_fill<<<nblocks(n, 256), 256, 0, s>>>(n, nd, qdim, dofs, inc_ptr, cursor, inc);
  k_inc_sort<<<nblocks(n_nodes, 256), 256, 0, s>>>(n_nodes, inc_ptr, inc);
  return hipGetLastError();
}

hipError_t launch_keep(hipStream_t s, int term, const PatternArgs &a) {
  const unsigned nb = nblocks(a.n_cv, 4);
  const size_t lds = 4 * (size_t)a.kwords * sizeof(uint32_t);
  if (a.qdim == 1 && term == GFX_TERM_LAPLACE)
    k_keep<1, GFX_TERM_LAPLACE><<<nb, 256, lds, s>>>(a);
  else if (a.qdim == 3 && term == GFX_TERM_LAPLACE)
    k_keep<3, GFX_TERM_LAPLACE><<<nb, 256, lds, s>>>(a);
  else if (a.qdim == 3 && term == GFX_TERM_ISO_ELAST)
    k_keep<3, GFX_TERM_ISO_ELAST><<<nb, 256, lds, s>>>(a);
  else if (a.qdim == 3 && term == GFX_TERM_ISOTROPIC)
    k_keep<3, GFX_TERM_ISOTROPIC><<<nb, 256, lds, s>>>(a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_patch_build(hipStream_t s, const PatternArgs &a, int max_cand) {
  if (max_cand > kMaxCand) return hipErrorInvalidValue;
  if (a.n_nodes == 0) return hipSuccess;
  int pmax = 1;
  while (pmax < max_cand) pmax <<= 1;
  pmax = pmax < 64 ? 64 : pmax;
  // presence bits: Q*Q*ns <= Q*Q*pmax
  const size_t mwords = ((size_t)a.qdim * a.qdim * pmax + 31) / 32 + 1;
  const size_t lds = 8 * (size_t)pmax + 2 * (size_t)(pmax + (pmax & 1)) + 4 * mwords;
  // small lists: smaller workgroups, so more nodes run at once
  const int threads = pmax >= 1024 ? 256 : (pmax >= 256 ? 128 : 64);
  k_patch_build<<<(unsigned)a.n_nodes, threads, lds, s>>>(a, pmax);
  return hipGetLastError();
}

hipError_t launch_mask_fast(hipStream_t s, const PatternArgs &a) {
  if (a.n_nodes == 0) return hipSuccess;
  k_mask_fast<<<nblocks(a.n_nodes, 256), 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_mask_exact(hipStream_t s, const PatternArgs &a, int max_nslots) {
  if (a.n_nodes == 0) return hipSuccess;
  const size_t mwords = ((size_t)a.qdim * a.qdim * max_nslots + 31) / 32;
  const size_t lds = 4 * (mwords > 0 ? mwords : 1);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  k_mask_check<<<(unsigned)a.n_nodes, 128, lds, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_fill_columns(hipStream_t s, const PatternArgs &a) {
  if (a.n_nodes == 0) return hipSuccess;
  k_fill_cols<<<(unsigned)a.n_nodes, 64, 0, s>>>(a);
  return hipGetLastError();
}

template <int ND, int Q, int TERM>
static hipError_t launch_asm3(hipStream_t s, int form, int prec, const PatchArgs &a,
                              unsigned grid, size_t lds) {
  constexpr int F = GFX_FORM_FACTORIZED, P = GFX_FORM_PERQP;
  if (form == F) {
    const bool d = prec == GFX_PREC_FP64;
    if (a.rk == 1) {
      if (d) k_assemble_patch<ND, Q, TERM, F, double, 1><<<grid, kWave, lds, s>>>(a);
      else k_assemble_patch<ND, Q, TERM, F, float, 1><<<grid, kWave, lds, s>>>(a);
    } else if (a.rk == 2) {
      if (d) k_assemble_patch<ND, Q, TERM, F, double, 2><<<grid, kWave, lds, s>>>(a);
      else k_assemble_patch<ND, Q, TERM, F, float, 2><<<grid, kWave, lds, s>>>(a);
    } else {
      if (d) k_assemble_patch<ND, Q, TERM, F, double, 0><<<grid, kWave, lds, s>>>(a);
      else k_assemble_patch<ND, Q, TERM, F, float, 0><<<grid, kWave, lds, s>>>(a);
    }
  } else {
    if (prec == GFX_PREC_FP64)
      k_assemble_patch<ND, Q, TERM, P, double, 0><<<grid, kWave, lds, s>>>(a);
    else
      k_assemble_patch<ND, Q, TERM, P, float, 0><<<grid, kWave, lds, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_assemble_patch(hipStream_t s, int term, int form, int prec, const PatchArgs &a0) {
  // a0.n_waves wave descriptor groups; a0.lds_acc / a0.lds_flat: the largest
  // accumulator (doubles) and staging run (4-byte words) of any of its waves
  if (a0.n_waves == 0) return hipSuccess;
  PatchArgs a = a0;
  a.xcd_chunk = (int)((a.n_waves + 7) / 8);
  const unsigned grid = (unsigned)(8 * (int64_t)a.xcd_chunk);
  const size_t lds = 8 * (size_t)a.lds_acc + 4 * (size_t)a.lds_flat;
  if (a.lds_acc < 0 || a.lds_flat < 0 || lds > 64 * 1024) return hipErrorInvalidValue;
  const int nd = a.nd, Q = a.qdim;
  if (term == GFX_TERM_LAPLACE && Q == 1) {
    if (nd == 4) return launch_asm3<4, 1, GFX_TERM_LAPLACE>(s, form, prec, a, grid, lds);
    if (nd == 10) return launch_asm3<10, 1, GFX_TERM_LAPLACE>(s, form, prec, a, grid, lds);
    if (nd == 20) return launch_asm3<20, 1, GFX_TERM_LAPLACE>(s, form, prec, a, grid, lds);
  } else if (term == GFX_TERM_LAPLACE && Q == 3) {
    if (nd == 4) return launch_asm3<4, 3, GFX_TERM_LAPLACE>(s, form, prec, a, grid, lds);
    if (nd == 10) return launch_asm3<10, 3, GFX_TERM_LAPLACE>(s, form, prec, a, grid, lds);
  } else if (term == GFX_TERM_ISO_ELAST && Q == 3) {
    if (nd == 4) return launch_asm3<4, 3, GFX_TERM_ISO_ELAST>(s, form, prec, a, grid, lds);
    if (nd == 10) return launch_asm3<10, 3, GFX_TERM_ISO_ELAST>(s, form, prec, a, grid, lds);
    if (nd == 20) return launch_asm3<20, 3, GFX_TERM_ISO_ELAST>(s, form, prec, a, grid, lds);
  } else if (term == GFX_TERM_ISOTROPIC && Q == 3) {
    if (nd == 4) return launch_asm3<4, 3, GFX_TERM_ISOTROPIC>(s, form, prec, a, grid, lds);
    if (nd == 10) return launch_asm3<10, 3, GFX_TERM_ISOTROPIC>(s, form, prec, a, grid, lds);
    if (nd == 20) return launch_asm3<20, 3, GFX_TERM_ISOTROPIC>(s, form, prec, a, grid, lds);
  }
  return hipErrorInvalidValue;
}

}  // namespace gfx
