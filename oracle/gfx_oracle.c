/*
 * gfx_oracle.c -- CPU restatement of GetFEM's generic-assembly hot path
 * (ga_workspace::assembly executing compiled GWFL instructions) for the
 * supported subset.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, bench.py's cpu_baseline leg and
 * __graft_entry__.smoke() as the checker; never linked into getfem_amd.
 *
 * What pins each function (the checks live in tests/test_oracle_*.py):
 *  - orc_im_rule: point counts stated by the reference (1/15/43/27), weight
 *    sum = reference volume, exact integration of every monomial up to the
 *    rule degree (closed form a!b!c!/(a+b+c+3)! on the tetrahedron).
 *  - orc_fem_eval: Kronecker property at the nodes, partition of unity,
 *    gradients against central finite differences, node counts.
 *  - orc_regular_mesh / orc_enumerate_dof: point/dof counts ((n+1)^3,
 *    3*(2n+1)^3 ...), agreement with the independent vectorised generator
 *    of getfem_amd.meshes (exact lattice keys instead of coordinate
 *    matching), conformity (every interior face shared by two elements),
 *    total volume = 1.
 *  - orc_element_matrix: P1 Poisson closed form |T| grad(l_i).grad(l_j),
 *    zero row sums, symmetry, the six rigid-body modes in the kernel of the
 *    elasticity matrix, brute-force high-order quadrature on single
 *    elements, linearity in (lambda, mu).
 *  - orc_assemble_matrix: dense brute force (numpy) on small meshes, the
 *    stored-nnz counts GetFEM itself produced on the same inputs
 *    (tests/golden/reference_counts.json)
 *    (P1 Poisson 20^3x6: 128,113 stored of 128,581 topological;
 *    P2 elasticity 8^3x6: 1,096,305 of 1,141,065).
 *  - orc_assemble_vector (Neo-Hookean residual): finite differences of the
 *    strain energy; tangent = finite differences of the residual.
 *  - orc_assemble_blocks (matrices over several variables): one square
 *    block equals orc_assemble_matrix of the same element matrices; a
 *    rectangular block equals a dense brute force (per-element threshold,
 *    zero elements skipped, sorted rows per column).  The numpy element
 *    matrices that feed it (oracle.block_element_matrices) equal
 *    orc_element_matrix's Laplacian, the closed-form P1 mass
 *    |T|/20 (1 + delta_ij), and for the P1 x P2 coupling the same values
 *    under two exact rules and the exact integral of div u of a linear u.
 *  - orc_ilu / orc_ilu_apply: where ILU(0) drops no fill (tridiagonal,
 *    dense) the apply is the exact solve (numpy); on any symmetric pattern
 *    (L U)_ij = (A^T)_ij on the pattern; the apply equals the dense
 *    triangular solves with those factors; a zero first pivot is replaced.
 *  - orc_compliance: central finite differences on J.
 *  - orc_ildlt / orc_ildlt_apply: on a pattern without dropped fill
 *    (tridiagonal, dense) the factor is the exact LDL^T (numpy Cholesky);
 *    on any pattern (L D L^T)_ij = A_ij for every (i, j) of A's pattern,
 *    the defining property of a level-0 incomplete factorization; apply =
 *    the solve with L D L^T (dense numpy); pivot replacement on a matrix
 *    with a zero pivot.
 */
#include "gfx_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * Integration methods
 * ====================================================================== */

/* Orbit data of the simplex rules of getfem_im_list.h (IM_TETRAHEDRON(k),
 * data at getfem_im_list.h:3251-3340, node types at :4612-4616).  Each row
 * is (x, y, z, weight, type): type 0 = single point, type 1 = every
 * permutation of the barycentric coordinates (add_point_full_symmetric,
 * getfem_integration.cc:278-305). */
typedef struct { double x, y, z, w; int type; } orc_orbit;

static const orc_orbit tet1[] = {{0.25, 0.25, 0.25, 1.0 / 6.0, 0}};
static const orc_orbit tet2[] = {
    {0.13819660112501052, 0.13819660112501052, 0.13819660112501052,
     0.0416666666666666666666666666666666666666666666666666666666666666, 1}};
static const orc_orbit tet3[] = {
    {0.25, 0.25, 0.25, -0.1333333333333333333333333333333333333333333333333333, 0},
    {0.1666666666666666666666666666666666666666666666666666666666666666,
     0.1666666666666666666666666666666666666666666666666666666666666666,
     0.1666666666666666666666666666666666666666666666666666666666666666,
     0.075, 1}};
static const orc_orbit tet5[] = {
    {0.25, 0.25, 0.25, 0.019753086419753086, 0},
    {0.31979362782962991, 0.31979362782962991, 0.31979362782962991,
     0.011511367871045398, 1},
    {0.091971078052723033, 0.091971078052723033, 0.091971078052723033,
     0.01198951396316977, 1},
    {0.056350832689629156, 0.056350832689629156, 0.44364916731037084,
     0.008818342151675485, 1}};
static const orc_orbit tet6[] = {
    {0.214602871259152029288839219386284, 0.214602871259152029288839219386284,
     0.214602871259152029288839219386284, 0.00665379170969458201661510459291332, 1},
    {0.0406739585346113531155794489564100, 0.0406739585346113531155794489564100,
     0.0406739585346113531155794489564100, 0.00167953517588677382466887290765614, 1},
    {0.322337890142275510343994470762492, 0.322337890142275510343994470762492,
     0.322337890142275510343994470762492, 0.00922619692394245368252554630895433, 1},
    {0.0636610018750175252992355276057269, 0.0636610018750175252992355276057269,
     0.269672331458315808034097805727606, 0.00803571428571428571428571428571428, 1}};
static const orc_orbit tet8[] = {
    {0.25, 0.25, 0.25, -0.0205001886586399158405865177642941, 1},
    {0.206829931610673204083980900024961, 0.206829931610673204083980900024961,
     0.206829931610673204083980900024961, 0.0142503058228669012484397415358704, 1},
    {0.0821035883105467230906058078714215, 0.0821035883105467230906058078714215,
     0.0821035883105467230906058078714215, 0.00196703331313390098756280342445466, 1},
    {0.00578195050519799725317663886414270, 0.00578195050519799725317663886414270,
     0.00578195050519799725317663886414270, 0.000169834109092887379837744566704016, 1},
    {0.0505327400188942244256245285579071, 0.0505327400188942244256245285579071,
     0.449467259981105775574375471442092, 0.00457968382446728180074351446297276, 1},
    {0.229066536116811139600408854554753, 0.229066536116811139600408854554753,
     0.0356395827885340437169173969506114, 0.00570448580868191850680255862783040, 1},
    {0.0366077495531974236787738546327104, 0.0366077495531974236787738546327104,
     0.190486041934633455699433285315099, 0.00214051914116209259648335300092023, 1}};

static int add_point_norepeat(double *pts, double *w, int n, const double *p,
                              double wt) {
  for (int i = 0; i < n; ++i)
    if (fabs(pts[3 * i] - p[0]) + fabs(pts[3 * i + 1] - p[1]) +
            fabs(pts[3 * i + 2] - p[2]) < 1e-14)
      return n;
  pts[3 * n] = p[0]; pts[3 * n + 1] = p[1]; pts[3 * n + 2] = p[2];
  w[n] = wt;
  return n + 1;
}

/* expansion of a simplex orbit, same enumeration order as
 * approx_integration::add_point_full_symmetric (ind[0] fastest) */
static int expand_orbits(const orc_orbit *ob, int nob, double *pts, double *w) {
  int n = 0;
  for (int o = 0; o < nob; ++o) {
    double p3[4] = {ob[o].x, ob[o].y, ob[o].z, 0.0};
    p3[3] = 1.0 - ob[o].x - ob[o].y - ob[o].z;
    if (ob[o].type == 0) {
      n = add_point_norepeat(pts, w, n, p3, ob[o].w);
      continue;
    }
    for (int i2 = 0; i2 < 4; ++i2)
      for (int i1 = 0; i1 < 4; ++i1)
        for (int i0 = 0; i0 < 4; ++i0) {
          if (i0 == i1 || i0 == i2 || i1 == i2) continue;
          double p[3] = {p3[i0], p3[i1], p3[i2]};
          n = add_point_norepeat(pts, w, n, p, ob[o].w);
        }
  }
  return n;
}

/* Gauss-Legendre on [0,1] with m points, ascending, as the Gauss rule of
 * getfem_integration.cc:493-520 (weights (1-r^2)/(m P_{m-1}(r))^2). */
static void gauss01(int m, double *x, double *w) {
  for (int k = 0; k < m; ++k) {
    long double r = cosl(3.14159265358979323846264338327950L * (k + 0.75L) /
                         (m + 0.5L));
    long double p0 = 1, p1 = r, dp = 0;
    for (int it = 0; it < 100; ++it) {
      p0 = 1; p1 = r;
      for (int j = 2; j <= m; ++j) {
        long double p2 = ((2 * j - 1) * r * p1 - (j - 1) * p0) / j;
        p0 = p1; p1 = p2;
      }
      if (m == 1) { p0 = 1; p1 = r; }
      dp = m * (r * p1 - p0) / (r * r - 1);
      long double dr = p1 / dp;
      r -= dr;
      if (fabsl(dr) < 1e-30L) break;
    }
    /* recompute P_{m-1}(r) at the converged root */
    long double q0 = 1, q1 = r;
    if (m == 1) q1 = 1; /* P_0 */
    else {
      for (int j = 2; j <= m - 1; ++j) {
        long double q2 = ((2 * j - 1) * r * q1 - (j - 1) * q0) / j;
        q0 = q1; q1 = q2;
      }
    }
    if (m == 2) q1 = r;
    long double pm1 = (m == 1) ? 1.0L : q1;
    int idx = m - 1 - k; /* cos ordering is descending */
    x[idx] = (double)(0.5L + 0.5L * r);
    w[idx] = (double)((1.0L - r * r) / ((long double)m * pm1 * (long double)m * pm1));
  }
}

static int parse_int_args(const char *name, const char *prefix, int *a, int *b) {
  size_t lp = strlen(prefix);
  if (strncmp(name, prefix, lp) != 0) return 0;
  if (b) return sscanf(name + lp, "%d,%d)", a, b) == 2;
  return sscanf(name + lp, "%d)", a) == 1;
}

int orc_im_rule(const char *name, double *pts, double *w) {
  int k = 0, d = 0;
  if (parse_int_args(name, "IM_TETRAHEDRON(", &k, NULL)) {
    switch (k) {
      case 1: return expand_orbits(tet1, 1, pts, w);
      case 2: return expand_orbits(tet2, 1, pts, w);
      case 3: return expand_orbits(tet3, 2, pts, w);
      case 5: return expand_orbits(tet5, 4, pts, w);
      case 6: return expand_orbits(tet6, 4, pts, w);
      case 8: return expand_orbits(tet8, 7, pts, w);
      default: return -1;
    }
  }
  if (parse_int_args(name, "IM_GAUSS_PARALLELEPIPED(", &d, &k) && d == 3) {
    /* IM_PRODUCT(IM_PRODUCT(G1D,G1D),G1D), product index i1 + n1*i2
     * (getfem_integration.cc:700-720, :828-846) */
    int m = k / 2 + 1;
    double x[64], wx[64];
    if (m > 64) return -1;
    gauss01(m, x, wx);
    int n = 0;
    for (int iz = 0; iz < m; ++iz)
      for (int iy = 0; iy < m; ++iy)
        for (int ix = 0; ix < m; ++ix) {
          pts[3 * n] = x[ix]; pts[3 * n + 1] = x[iy]; pts[3 * n + 2] = x[iz];
          w[n] = (wx[ix] * wx[iy]) * wx[iz];
          ++n;
        }
    return n;
  }
  return -1;
}

int orc_im_npts(const char *name) {
  double p[3 * 4096], w[4096];
  return orc_im_rule(name, p, w);
}

/* ======================================================================
 * Lagrange finite elements
 * ====================================================================== */

/* PK lattice on the reference simplex, c[0] fastest with sum <= K
 * (K_simplex_of_ref_, bgeot_convex_ref.cc:311-321). */
static int pk_nodes(int K, int idx[][4]) {
  int n = 0;
  if (K == 0) { idx[0][0] = idx[0][1] = idx[0][2] = idx[0][3] = 0; return 1; }
  for (int c2 = 0; c2 <= K; ++c2)
    for (int c1 = 0; c1 + c2 <= K; ++c1)
      for (int c0 = 0; c0 + c1 + c2 <= K; ++c0) {
        idx[n][1] = c0; idx[n][2] = c1; idx[n][3] = c2;
        idx[n][0] = K - c0 - c1 - c2;
        ++n;
      }
  return n;
}

int orc_fem_nb_dof(int kind, int k) {
  if (kind == ORC_FEM_PK) return (k + 1) * (k + 2) * (k + 3) / 6;
  if (kind == ORC_FEM_QK) return (k + 1) * (k + 1) * (k + 1);
  if (kind == ORC_FEM_PK_DISC0) return 1;
  return -1;
}

int orc_fem_nodes(int kind, int k, double *nodes) {
  if (kind == ORC_FEM_PK) {
    int idx[1024][4];
    if (k < 1) return -1;
    int n = pk_nodes(k, idx);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) nodes[3 * i + c] = (double)idx[i][c + 1] / k;
    return n;
  }
  if (kind == ORC_FEM_QK) {
    int n = 0;
    for (int iz = 0; iz <= k; ++iz)
      for (int iy = 0; iy <= k; ++iy)
        for (int ix = 0; ix <= k; ++ix) {
          nodes[3 * n] = (double)ix / k; nodes[3 * n + 1] = (double)iy / k;
          nodes[3 * n + 2] = (double)iz / k; ++n;
        }
    return n;
  }
  if (kind == ORC_FEM_PK_DISC0) {
    nodes[0] = nodes[1] = nodes[2] = 0.25;
    return 1;
  }
  return -1;
}

/* prod_{j<m} (K*l - j)/(j+1) and its derivative in l */
static void lag_factor(int K, int m, double l, double *v, double *dv) {
  double val = 1.0, der = 0.0;
  for (int j = 0; j < m; ++j) {
    double f = (K * l - j) / (j + 1.0), df = K / (j + 1.0);
    der = der * f + val * df;
    val *= f;
  }
  *v = val; *dv = der;
}

int orc_fem_eval(int kind, int k, const double *x, double *val, double *grad) {
  if (kind == ORC_FEM_PK_DISC0) {
    val[0] = 1.0; if (grad) grad[0] = grad[1] = grad[2] = 0.0;
    return 1;
  }
  if (kind == ORC_FEM_PK) {
    /* barycentric product formula of PK_fem_::calc_base_func
     * (getfem_fem.cc:731-755) */
    int idx[1024][4];
    int n = pk_nodes(k, idx);
    double l[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    /* dl[b][c] = d l_b / d x_c */
    static const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < n; ++i) {
      double f[4], df[4];
      for (int b = 0; b < 4; ++b) lag_factor(k, idx[i][b], l[b], &f[b], &df[b]);
      val[i] = f[0] * f[1] * f[2] * f[3];
      if (grad) {
        for (int c = 0; c < 3; ++c) {
          double g = 0.0;
          for (int b = 0; b < 4; ++b) {
            double p = df[b] * dl[b][c];
            for (int bb = 0; bb < 4; ++bb) if (bb != b) p *= f[bb];
            g += p;
          }
          grad[3 * i + c] = g;
        }
      }
    }
    return n;
  }
  if (kind == ORC_FEM_QK) {
    /* FEM_PRODUCT(FEM_QK(2,k),FEM_PK(1,k)): first factor fastest
     * (tproduct_femi, getfem_fem.cc:795-826) */
    double v1[3][16], d1[3][16];
    for (int c = 0; c < 3; ++c)
      for (int m = 0; m <= k; ++m) {
        double a, da, b, db;
        lag_factor(k, k - m, 1.0 - x[c], &a, &da);
        lag_factor(k, m, x[c], &b, &db);
        v1[c][m] = a * b;
        d1[c][m] = -da * b + a * db;
      }
    int n = 0;
    for (int iz = 0; iz <= k; ++iz)
      for (int iy = 0; iy <= k; ++iy)
        for (int ix = 0; ix <= k; ++ix) {
          val[n] = v1[0][ix] * v1[1][iy] * v1[2][iz];
          if (grad) {
            grad[3 * n + 0] = d1[0][ix] * v1[1][iy] * v1[2][iz];
            grad[3 * n + 1] = v1[0][ix] * d1[1][iy] * v1[2][iz];
            grad[3 * n + 2] = v1[0][ix] * v1[1][iy] * d1[2][iz];
          }
          ++n;
        }
    return n;
  }
  return -1;
}

/* geometric transformation shape functions GT_PK(3,1) / GT_QK(3,1) */
static int gt_nv(int gt) { return gt == ORC_GT_PK31 ? 4 : 8; }
static void gt_eval(int gt, const double *x, double *N, double *dN) {
  if (gt == ORC_GT_PK31) {
    orc_fem_eval(ORC_FEM_PK, 1, x, N, dN);
  } else {
    orc_fem_eval(ORC_FEM_QK, 1, x, N, dN);
  }
}

/* ======================================================================
 * Regular meshes: two-stage construction of regular_unit_mesh
 * (getfem_regular_meshes.cc:237-285): first the structured mesh with
 * points deduplicated on insertion, then a copy through
 * add_convex_by_points, which renumbers points in first-encounter order.
 * ====================================================================== */

typedef struct {
  int64_t *keys; int64_t *vals; int64_t cap; int64_t n;
} hmap;

static void hm_init(hmap *h, int64_t n) {
  h->cap = 16; while (h->cap < 2 * n) h->cap <<= 1;
  h->keys = (int64_t *)malloc(sizeof(int64_t) * h->cap);
  h->vals = (int64_t *)malloc(sizeof(int64_t) * h->cap);
  for (int64_t i = 0; i < h->cap; ++i) h->keys[i] = INT64_MIN;
  h->n = 0;
}
static void hm_free(hmap *h) { free(h->keys); free(h->vals); }
static uint64_t hm_hash(int64_t k) {
  uint64_t x = (uint64_t)k;
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}
/* returns existing value or inserts v */
static int64_t hm_get_or_put(hmap *h, int64_t k, int64_t v) {
  uint64_t i = hm_hash(k) & (uint64_t)(h->cap - 1);
  for (;;) {
    if (h->keys[i] == INT64_MIN) { h->keys[i] = k; h->vals[i] = v; h->n++; return v; }
    if (h->keys[i] == k) return h->vals[i];
    i = (i + 1) & (uint64_t)(h->cap - 1);
  }
}
static int64_t hm_get(const hmap *h, int64_t k) {
  uint64_t i = hm_hash(k) & (uint64_t)(h->cap - 1);
  for (;;) {
    if (h->keys[i] == INT64_MIN) return -1;
    if (h->keys[i] == k) return h->vals[i];
    i = (i + 1) & (uint64_t)(h->cap - 1);
  }
}

/* coordinate key on a grid fine enough for tolerance matching */
#define ORC_QSCALE 4.0e7
static int64_t qkey3(int64_t a, int64_t b, int64_t c) {
  return ((a & 0x1FFFFF) << 42) | ((b & 0x1FFFFF) << 21) | (c & 0x1FFFFF);
}

/* point table with tolerance search (node_tab::search_node) */
typedef struct { hmap h; double *pts; int64_t n; } ptab;

static int64_t pt_find(const ptab *t, const double *p, double tol) {
  int64_t q[3];
  for (int c = 0; c < 3; ++c) q[c] = (int64_t)llround(p[c] * ORC_QSCALE);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int64_t v = hm_get(&t->h, qkey3(q[0] + dx, q[1] + dy, q[2] + dz));
        if (v >= 0) {
          const double *r = t->pts + 3 * v;
          if (fabs(r[0] - p[0]) <= tol && fabs(r[1] - p[1]) <= tol &&
              fabs(r[2] - p[2]) <= tol)
            return v;
        }
      }
  return -1;
}
static int64_t pt_add(ptab *t, const double *p, double tol) {
  int64_t v = pt_find(t, p, tol);
  if (v >= 0) return v;
  int64_t q[3];
  for (int c = 0; c < 3; ++c) q[c] = (int64_t)llround(p[c] * ORC_QSCALE);
  v = t->n++;
  memcpy(t->pts + 3 * v, p, 3 * sizeof(double));
  hm_get_or_put(&t->h, qkey3(q[0], q[1], q[2]), v);
  return v;
}

/* bgeot_convex_ref_simplexified.cc:34-39 */
static const int kuhn6[6][4] = {{3, 7, 0, 1}, {7, 0, 5, 4}, {7, 0, 1, 5},
                                {3, 7, 0, 2}, {6, 7, 0, 4}, {6, 7, 0, 2}};

int64_t orc_regular_mesh(int gt, int n, double *pts, int32_t *cvx) {
  int64_t npts_max = (int64_t)(n + 1) * (n + 1) * (n + 1);
  double h = 1.0 / (double)n;
  ptab stage1 = {{0}, (double *)malloc(sizeof(double) * 3 * npts_max), 0};
  hm_init(&stage1.h, npts_max);
  int nv = gt_nv(gt);
  int64_t ncv = (gt == ORC_GT_PK31) ? 6LL * n * n * n : (int64_t)n * n * n;
  int32_t *cv1 = (int32_t *)malloc(sizeof(int32_t) * nv * ncv);
  int64_t c = 0;
  double tol = 1e-9 * h;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < n; ++ix) {
        int64_t tab3[8];
        double a[3] = {ix * h, iy * h, iz * h};
        for (int k = 0; k < 8; ++k) {
          double p[3] = {a[0] + (k & 1) * h, a[1] + ((k >> 1) & 1) * h,
                         a[2] + ((k >> 2) & 1) * h};
          tab3[k] = pt_add(&stage1, p, tol);
        }
        if (gt == ORC_GT_PK31) {
          for (int s = 0; s < 6; ++s, ++c)
            for (int l = 0; l < 4; ++l) cv1[4 * c + l] = (int32_t)tab3[kuhn6[s][l]];
        } else {
          for (int k = 0; k < 8; ++k) cv1[8 * c + k] = (int32_t)tab3[k];
          ++c;
        }
      }
  /* stage 2: add_convex_by_points in convex order */
  ptab stage2 = {{0}, pts, 0};
  hm_init(&stage2.h, npts_max);
  for (int64_t e = 0; e < ncv; ++e)
    for (int l = 0; l < nv; ++l)
      cvx[nv * e + l] = (int32_t)pt_add(&stage2, stage1.pts + 3 * cv1[nv * e + l], tol);
  int64_t np = stage2.n;
  hm_free(&stage1.h); hm_free(&stage2.h);
  free(stage1.pts); free(cv1);
  return np;
}

/* ======================================================================
 * Dof enumeration: first-encounter numbering over elements in convex order
 * and local dofs in fem order, nodes identified by their real coordinates
 * (mesh_fem::enumerate_dof, getfem_mesh_fem.cc:320-440); a new node takes
 * Qdim consecutive dofs (nbdof += Qdim, :417-418).
 * ====================================================================== */
int64_t orc_enumerate_dof(int gt, int64_t n_pts, const double *pts,
                          int64_t n_cv, const int32_t *cvx, int fem_kind,
                          int fem_k, int qdim, int32_t *basic_dofs) {
  (void)n_pts;
  int nd = orc_fem_nb_dof(fem_kind, fem_k);
  int nv = gt_nv(gt);
  double nodes[3 * 1024];
  orc_fem_nodes(fem_kind, fem_k, nodes);
  double Nn[1024][8];
  for (int i = 0; i < nd; ++i) gt_eval(gt, nodes + 3 * i, Nn[i], NULL);
  int64_t nbdof = 0;
  if (fem_kind == ORC_FEM_PK_DISC0) { /* not linkable: one new dof per element */
    for (int64_t e = 0; e < n_cv; ++e) { basic_dofs[e] = (int32_t)nbdof; nbdof += qdim; }
    return nbdof;
  }
  ptab t = {{0}, (double *)malloc(sizeof(double) * 3 * (size_t)(n_cv * nd + 1)), 0};
  hm_init(&t.h, n_cv * nd / 2 + 16);
  int64_t *node_dof = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_cv * nd + 1));
  for (int64_t e = 0; e < n_cv; ++e) {
    /* element size for the tolerance */
    double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nv; ++v)
      for (int c = 0; c < 3; ++c) {
        double x = pts[3 * (int64_t)cvx[nv * e + v] + c];
        if (x < bmin[c]) bmin[c] = x;
        if (x > bmax[c]) bmax[c] = x;
      }
    double sz = fmax(bmax[0] - bmin[0], fmax(bmax[1] - bmin[1], bmax[2] - bmin[2]));
    double tol = 1e-6 * sz;
    for (int i = 0; i < nd; ++i) {
      double P[3] = {0, 0, 0};
      for (int v = 0; v < nv; ++v)
        for (int c = 0; c < 3; ++c) P[c] += Nn[i][v] * pts[3 * (int64_t)cvx[nv * e + v] + c];
      int64_t before = t.n;
      int64_t id = pt_add(&t, P, tol);
      if (id == before) node_dof[id] = nbdof, nbdof += qdim;
      basic_dofs[(int64_t)nd * e + i] = (int32_t)node_dof[id];
    }
  }
  hm_free(&t.h); free(t.pts); free(node_dof);
  return nbdof;
}

/* ======================================================================
 * Element computations (ga_exec loop, compile_and_exec.cc:8750-8869)
 * ====================================================================== */

/* K = G.grad(N) (bgeot_geometric_trans.cc:270-287), J by cofactors and
 * B = K^{-T} = cofactors / J (:321-355, :374-413). Kc is column-major. */
static double jacobian(const double *G /* nv x 3 rows=points */, int nv,
                       const double *dN /* nv x 3 */, double *B /* B[r*3+c] */,
                       double *Kout) {
  double K[9]; /* column-major K(r,c) = K[r + 3c] */
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      double a = G[r] * dN[c];
      for (int v = 1; v < nv; ++v) a += G[3 * v + r] * dN[3 * v + c];
      K[r + 3 * c] = a;
    }
  double a0 = K[4] * K[8] - K[5] * K[7];
  double a1 = K[5] * K[6] - K[3] * K[8];
  double a2 = K[3] * K[7] - K[4] * K[6];
  double J = K[0] * a0 + K[1] * a1 + K[2] * a2;
  /* B column-major B(r,c) = Bc[r + 3c]; returned row-major B[r*3+c] */
  double Bc[9];
  Bc[0] = a0 / J; Bc[1] = a1 / J; Bc[2] = a2 / J;
  Bc[3] = (K[2] * K[7] - K[1] * K[8]) / J;
  Bc[4] = (K[0] * K[8] - K[2] * K[6]) / J;
  Bc[5] = (K[1] * K[6] - K[0] * K[7]) / J;
  Bc[6] = (K[1] * K[5] - K[2] * K[4]) / J;
  Bc[7] = (K[2] * K[3] - K[0] * K[5]) / J;
  Bc[8] = (K[0] * K[4] - K[1] * K[3]) / J;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[3 * r + c] = Bc[r + 3 * c];
  if (Kout) for (int i = 0; i < 9; ++i) Kout[i] = K[i];
  return J;
}

typedef struct {
  int nv, nd, nq, qdim, nloc;
  double *val;   /* nq * nd */
  double *grad;  /* nq * nd * 3 reference gradients */
  double *gN;    /* nq * nv * 3 geotrans reference gradients */
} orc_tables;

static void make_tables(const orc_problem *p, orc_tables *t) {
  t->nv = gt_nv(p->gt);
  t->nd = orc_fem_nb_dof(p->fem_kind, p->fem_k);
  t->nq = p->nq;
  t->qdim = p->qdim;
  t->nloc = t->nd * t->qdim;
  t->val = (double *)malloc(sizeof(double) * t->nq * t->nd);
  t->grad = (double *)malloc(sizeof(double) * t->nq * t->nd * 3);
  t->gN = (double *)malloc(sizeof(double) * t->nq * t->nv * 3);
  double Nv[8];
  for (int q = 0; q < t->nq; ++q) {
    orc_fem_eval(p->fem_kind, p->fem_k, p->qp + 3 * q, t->val + q * t->nd,
                 t->grad + q * t->nd * 3);
    gt_eval(p->gt, p->qp + 3 * q, Nv, t->gN + q * t->nv * 3);
  }
}
static void free_tables(orc_tables *t) { free(t->val); free(t->grad); free(t->gN); }

/* Compressible Neo-Hookean (Bonet) law: W = mu/2 (I1 - 3 - ln I3)
 * + lambda/8 (ln I3)^2, S = mu (I - C^-1) + lambda/2 ln(I3) C^-1
 * (getfem_nonlinear_elasticity.cc:612-657), tangent of S w.r.t. C as in
 * grad_sigma (:659-696). Here in closed form on C^-1. */
static void nh_bonet(const double *F, double lam, double mu, double *S,
                     double *dSdE /* 81, d S_ij / d E_kl */) {
  double C[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += F[3 * k + i] * F[3 * k + j];
      C[3 * i + j] = s;
    }
  double c0 = C[4] * C[8] - C[5] * C[7];
  double c1 = C[5] * C[6] - C[3] * C[8];
  double c2 = C[3] * C[7] - C[4] * C[6];
  double I3 = C[0] * c0 + C[1] * c1 + C[2] * c2;
  double Ci[9];
  Ci[0] = c0 / I3; Ci[3] = c1 / I3; Ci[6] = c2 / I3;
  Ci[1] = (C[2] * C[7] - C[1] * C[8]) / I3;
  Ci[4] = (C[0] * C[8] - C[2] * C[6]) / I3;
  Ci[7] = (C[1] * C[6] - C[0] * C[7]) / I3;
  Ci[2] = (C[1] * C[5] - C[2] * C[4]) / I3;
  Ci[5] = (C[2] * C[3] - C[0] * C[5]) / I3;
  Ci[8] = (C[0] * C[4] - C[1] * C[3]) / I3;
  double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                F[2] * (F[3] * F[7] - F[4] * F[6]);
  double lnI3 = log(I3);
  for (int i = 0; i < 9; ++i)
    S[i] = mu * ((i % 4 == 0 ? 1.0 : 0.0) - Ci[i]) + 0.5 * lam * lnI3 * Ci[i];
  if (detF <= 0.0)
    for (int i = 0; i < 9; ++i) S[i] += 1e200 * C[i];
  /* dS/dE = 2 dS/dC;  d(C^-1)_ij/dC_kl = -1/2 (Ci_ik Ci_lj + Ci_il Ci_kj),
   * d ln I3 / dC_kl = Ci_kl */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double dCi = -0.5 * (Ci[3 * i + k] * Ci[3 * l + j] + Ci[3 * i + l] * Ci[3 * k + j]);
          double v = -mu * dCi + 0.5 * lam * (Ci[3 * k + l] * Ci[3 * i + j] + lnI3 * dCi);
          dSdE[27 * i + 9 * j + 3 * k + l] = 2.0 * v;
        }
}

/* real gradients at integration point q: g[i*3+k].  For the affine
 * transformation K, J and B are computed at the first point only and reused
 * (compile_and_exec.cc:8836-8837); per point for the trilinear one. */
typedef struct { int64_t cv; double B[9]; double J; } geo_cache;

static double element_geometry_c(const orc_problem *p, const orc_tables *t,
                                 int64_t cv, int q, double *g, geo_cache *gc) {
  double G[24], Bl[9];
  double *B = gc ? gc->B : Bl;
  double J;
  if (gc && p->gt == ORC_GT_PK31 && gc->cv == cv && q > 0) {
    J = gc->J;
  } else {
    for (int v = 0; v < t->nv; ++v)
      for (int c = 0; c < 3; ++c) G[3 * v + c] = p->pts[3 * (int64_t)p->cvx[(int64_t)t->nv * cv + v] + c];
    J = jacobian(G, t->nv, t->gN + q * t->nv * 3, B, NULL);
    if (gc) { gc->cv = cv; gc->J = J; }
  }
  const double *gr = t->grad + q * t->nd * 3;
  for (int i = 0; i < t->nd; ++i)
    for (int k = 0; k < 3; ++k)
      g[3 * i + k] = B[3 * k + 0] * gr[3 * i + 0] + B[3 * k + 1] * gr[3 * i + 1] +
                     B[3 * k + 2] * gr[3 * i + 2];
  return fabs(J);
}

static double element_geometry(const orc_problem *p, const orc_tables *t,
                               int64_t cv, int q, double *g) {
  return element_geometry_c(p, t, cv, q, g, NULL);
}

static double coef_of(const double *arr, double scal, int64_t cv) {
  return arr ? arr[cv] : scal;
}

static void nh_grad_u(const orc_problem *p, const orc_tables *t, int64_t cv,
                      const double *g, double *H) {
  for (int i = 0; i < 9; ++i) H[i] = 0.0;
  for (int n = 0; n < t->nd; ++n) {
    int64_t d = p->dofs[(int64_t)t->nd * cv + n];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) H[3 * a + b] += p->U[d + a] * g[3 * n + b];
  }
}

static int element_matrix_impl(const orc_problem *p, const orc_tables *t,
                               int64_t cv, double *elem, double *g, double *tq) {
  int nd = t->nd, Q = t->qdim, nl = t->nloc;
  double a = coef_of(p->coef_a, p->scal_a, cv);
  double b = coef_of(p->coef_b, p->scal_b, cv);
  double c = p->term == ORC_TERM_ISOTROPIC ? coef_of(p->coef_c, p->scal_c, cv) : b;
  geo_cache gc;
  gc.cv = -1;
  for (int q = 0; q < t->nq; ++q) {
    double J = element_geometry_c(p, t, cv, q, g, &gc);
    double coeff = J * p->qw[q];
    if (p->term == ORC_TERM_LAPLACE) {
      for (int j = 0; j < nd; ++j)
        for (int i = 0; i < nd; ++i) {
          double d = g[3 * i] * g[3 * j] + g[3 * i + 1] * g[3 * j + 1] + g[3 * i + 2] * g[3 * j + 2];
          for (int bb = 0; bb < Q; ++bb)
            for (int aa = 0; aa < Q; ++aa)
              tq[(i * Q + aa) + nl * (j * Q + bb)] = (aa == bb) ? a * d : 0.0;
        }
    } else if (p->term == ORC_TERM_ISO_ELAST || p->term == ORC_TERM_ISOTROPIC) {
      /* the merged tree's integrand: the div-div, transposed-gradient and
       * gradient:gradient contractions summed per point */
      for (int j = 0; j < nd; ++j)
        for (int i = 0; i < nd; ++i) {
          double d = g[3 * i] * g[3 * j] + g[3 * i + 1] * g[3 * j + 1] + g[3 * i + 2] * g[3 * j + 2];
          for (int bb = 0; bb < 3; ++bb)
            for (int aa = 0; aa < 3; ++aa)
              tq[(i * 3 + aa) + nl * (j * 3 + bb)] =
                  a * g[3 * i + aa] * g[3 * j + bb] + b * g[3 * i + bb] * g[3 * j + aa] +
                  (aa == bb ? c * d : 0.0);
        }
    } else if (p->term == ORC_TERM_NEOHOOKEAN) {
      /* derivative of ((Id+Grad_u)*S(Grad_u)):Grad_Test_u with respect to u:
       * Grad_Test2_u*S : Grad_Test_u + F*(dS/dE : dE)[Grad_Test2_u] : Grad_Test_u,
       * dE = sym(F^T Grad_Test2_u) */
      double H[9], F[9], S[9], A[81];
      nh_grad_u(p, t, cv, g, H);
      for (int i = 0; i < 9; ++i) F[i] = H[i] + (i % 4 == 0 ? 1.0 : 0.0);
      nh_bonet(F, a, b, S, A);
      for (int j = 0; j < nd; ++j)
        for (int bb = 0; bb < 3; ++bb) {
          /* trial function: Grad w = e_bb (x) g_j */
          double dE[9], dS[9];
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
              dE[3 * k + l] = 0.5 * (F[3 * bb + k] * g[3 * j + l] + F[3 * bb + l] * g[3 * j + k]);
          for (int m = 0; m < 9; ++m) {
            double s = 0.0;
            for (int kl = 0; kl < 9; ++kl) s += A[9 * m + kl] * dE[kl];
            dS[m] = s;
          }
          for (int i = 0; i < nd; ++i)
            for (int aa = 0; aa < 3; ++aa) {
              /* test function: Grad v = e_aa (x) g_i ;
               * (Grad w S):Grad v = delta(aa,bb) * (g_j . S g_i)
               * (F dS):Grad v = sum_l (F dS)_{aa l} g_i[l] */
              double v = 0.0;
              if (aa == bb) {
                for (int k = 0; k < 3; ++k)
                  for (int l = 0; l < 3; ++l) v += g[3 * j + k] * S[3 * k + l] * g[3 * i + l];
              }
              for (int l = 0; l < 3; ++l) {
                double fds = 0.0;
                for (int k = 0; k < 3; ++k) fds += F[3 * aa + k] * dS[3 * k + l];
                v += fds * g[3 * i + l];
              }
              tq[(i * 3 + aa) + nl * (j * 3 + bb)] = v;
            }
        }
    } else {
      return -1;
    }
    if (q == 0)
      for (int k = 0; k < nl * nl; ++k) elem[k] = coeff * tq[k];
    else
      for (int k = 0; k < nl * nl; ++k) elem[k] += coeff * tq[k];
  }
  return nl;
}

int orc_element_matrix(const orc_problem *p, int64_t cv, double *elem) {
  orc_tables t;
  make_tables(p, &t);
  double *g = (double *)malloc(sizeof(double) * 3 * t.nd);
  double *tq = (double *)malloc(sizeof(double) * t.nloc * t.nloc);
  int r = element_matrix_impl(p, &t, cv, elem, g, tq);
  free(g); free(tq); free_tables(&t);
  return r;
}

static int element_vector_impl(const orc_problem *p, const orc_tables *t,
                               int64_t cv, double *elem, double *g) {
  int nd = t->nd, nl = t->nloc;
  double a = coef_of(p->coef_a, p->scal_a, cv);
  double b = coef_of(p->coef_b, p->scal_b, cv);
  if (p->term != ORC_TERM_NEOHOOKEAN) return -1;
  for (int q = 0; q < t->nq; ++q) {
    double J = element_geometry(p, t, cv, q, g);
    double coeff = J * p->qw[q];
    double H[9], F[9], S[9], A[81], P[9];
    nh_grad_u(p, t, cv, g, H);
    for (int i = 0; i < 9; ++i) F[i] = H[i] + (i % 4 == 0 ? 1.0 : 0.0);
    nh_bonet(F, a, b, S, A);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += F[3 * i + k] * S[3 * k + j];
        P[3 * i + j] = s;
      }
    for (int i = 0; i < nd; ++i)
      for (int aa = 0; aa < 3; ++aa) {
        double v = P[3 * aa] * g[3 * i] + P[3 * aa + 1] * g[3 * i + 1] + P[3 * aa + 2] * g[3 * i + 2];
        if (q == 0) elem[3 * i + aa] = coeff * v; else elem[3 * i + aa] += coeff * v;
      }
  }
  return nl;
}

int orc_element_vector(const orc_problem *p, int64_t cv, double *elem) {
  orc_tables t;
  make_tables(p, &t);
  double *g = (double *)malloc(sizeof(double) * 3 * t.nd);
  int r = element_vector_impl(p, &t, cv, elem, g);
  free(g); free_tables(&t);
  return r;
}

/* ======================================================================
 * Global matrix: gmm::col_matrix<gmm::rsvector<double>> and add_elem_matrix
 * ====================================================================== */

typedef struct { int32_t *r; double *v; int32_t n, cap; } orc_col;
typedef struct { int64_t ncols; orc_col *cols; } orc_mat;

static orc_mat *mat_new(int64_t ncols) {
  orc_mat *m = (orc_mat *)malloc(sizeof(orc_mat));
  m->ncols = ncols;
  m->cols = (orc_col *)calloc((size_t)ncols, sizeof(orc_col));
  return m;
}

static void col_reserve(orc_col *c, int32_t need) {
  if (need <= c->cap) return;
  int32_t nc = c->cap ? c->cap : 8;
  while (nc < need) nc *= 2;
  c->r = (int32_t *)realloc(c->r, sizeof(int32_t) * nc);
  c->v = (double *)realloc(c->v, sizeof(double) * nc);
  c->cap = nc;
}

/* add_elem_matrix (compile_and_exec.cc:4853-4936): rows (dofs1, s1 of them)
 * insertion-sorted, then for each column dofs2[jc] a merge by binary search
 * into the sorted column; elem is column-major s1 x s2 */
static void add_elem_matrix_rect(orc_mat *K, const int32_t *dofs1, int s1, const int32_t *dofs2, int s2,
                                 const double *elem, double thr, int *sortv) {
  for (int i = 0; i < s1; ++i) {
    int j = i;
    while (j > 0 && dofs1[i] < dofs1[sortv[j - 1]]) { sortv[j] = sortv[j - 1]; --j; }
    sortv[j] = i;
  }
  for (int jc = 0; jc < s2; ++jc) {
    const double *it = elem + (size_t)s1 * jc;
    orc_col *col = &K->cols[dofs2[jc]];
    int32_t nb = col->n, ind = 0;
    col_reserve(col, nb + s1);
    for (int kk = 0; kk < s1; ++kk) {
      int k = sortv[kk];
      double e = it[k];
      if (!(fabs(e) > thr)) continue;
      int32_t c = dofs1[k];
      int32_t count = nb - ind;
      while (count > 0) {
        int32_t step = count / 2, l = ind + step;
        if (col->r[l] < c) { ind = l + 1; count -= step + 1; }
        else count = step;
      }
      if (ind != nb && col->r[ind] == c) {
        col->v[ind] += e;
      } else {
        memmove(col->r + ind + 1, col->r + ind, sizeof(int32_t) * (nb - ind));
        memmove(col->v + ind + 1, col->v + ind, sizeof(double) * (nb - ind));
        col->r[ind] = c; col->v[ind] = e;
        ++nb;
      }
      ++ind;
    }
    col->n = nb;
  }
}

static void add_elem_matrix(orc_mat *K, const int32_t *dofs, int s1, const double *elem, double thr,
                            int *sortv) {
  add_elem_matrix_rect(K, dofs, s1, dofs, s1, elem, thr, sortv);
}

/* ga_instruction_matrix_assembly_mf_mf::exec (compile_and_exec.cc:5074-5189)
 * for given element matrices: element by element, every block's matrix is
 * skipped when it is zero, else added with the threshold 1e-14 max|elem|. */
void *orc_assemble_blocks(int64_t n_cv, int64_t ndof, int nblk, const orc_block *blk) {
  orc_mat *K = mat_new(ndof);
  int smax = 1;
  for (int b = 0; b < nblk; ++b) if (blk[b].s1 > smax) smax = blk[b].s1;
  int *sortv = (int *)malloc(sizeof(int) * smax);
  for (int64_t cv = 0; cv < n_cv; ++cv)
    for (int b = 0; b < nblk; ++b) {
      const orc_block *B = &blk[b];
      const size_t ne = (size_t)B->s1 * B->s2;
      const double *elem = B->elem + ne * cv;
      double ninf = 0.0;
      for (size_t k = 0; k < ne; ++k) if (fabs(elem[k]) > ninf) ninf = fabs(elem[k]);
      if (ninf == 0.0) continue;
      add_elem_matrix_rect(K, B->rows + (int64_t)B->s1 * cv, B->s1, B->cols + (int64_t)B->s2 * cv, B->s2,
                           elem, ninf * 1e-14, sortv);
    }
  free(sortv);
  return K;
}

static void assemble_range(const orc_problem *p, orc_mat *K, int64_t b, int64_t e) {
  orc_tables t;
  make_tables(p, &t);
  int nl = t.nloc, Q = t.qdim;
  double *elem = (double *)malloc(sizeof(double) * nl * nl);
  double *tq = (double *)malloc(sizeof(double) * nl * nl);
  double *g = (double *)malloc(sizeof(double) * 3 * t.nd);
  int32_t *dofs = (int32_t *)malloc(sizeof(int32_t) * nl);
  int *sortv = (int *)malloc(sizeof(int) * nl);
  for (int64_t cv = b; cv < e; ++cv) {
    element_matrix_impl(p, &t, cv, elem, g, tq);
    double ninf = 0.0;
    for (int k = 0; k < nl * nl; ++k) if (fabs(elem[k]) > ninf) ninf = fabs(elem[k]);
    if (ninf == 0.0) continue;
    /* populate_dofs_vector (compile_and_exec.cc:5008-5020) */
    for (int i = 0; i < t.nd; ++i)
      for (int q = 0; q < Q; ++q) dofs[i * Q + q] = p->dofs[(int64_t)t.nd * cv + i] + q;
    add_elem_matrix(K, dofs, nl, elem, ninf * 1e-14, sortv);
  }
  free(elem); free(tq); free(g); free(dofs); free(sortv); free_tables(&t);
}

/* gmm::add of one col_matrix into another (sum of the thread copies) */
static void mat_add_into(orc_mat *dst, const orc_mat *src) {
  for (int64_t j = 0; j < dst->ncols; ++j) {
    const orc_col *s = &src->cols[j];
    if (!s->n) continue;
    orc_col *d = &dst->cols[j];
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * (d->n + s->n));
    double *v = (double *)malloc(sizeof(double) * (d->n + s->n));
    int32_t a = 0, bb = 0, n = 0;
    while (a < d->n || bb < s->n) {
      if (bb >= s->n || (a < d->n && d->r[a] < s->r[bb])) { r[n] = d->r[a]; v[n++] = d->v[a++]; }
      else if (a >= d->n || s->r[bb] < d->r[a]) { r[n] = s->r[bb]; v[n++] = s->v[bb++]; }
      else { r[n] = d->r[a]; v[n++] = d->v[a++] + s->v[bb++]; }
    }
    free(d->r); free(d->v);
    d->r = r; d->v = v; d->n = n; d->cap = d->n + s->n;
  }
}

void *orc_assemble_matrix(const orc_problem *p, int64_t cv_begin,
                          int64_t cv_end, int nthreads) {
  if (cv_end > p->n_cv) cv_end = p->n_cv;
  if (nthreads <= 1) {
    orc_mat *K = mat_new(p->ndof);
    assemble_range(p, K, cv_begin, cv_end);
    return K;
  }
  orc_mat **parts = (orc_mat **)malloc(sizeof(orc_mat *) * nthreads);
  int64_t len = cv_end - cv_begin;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
  for (int t = 0; t < nthreads; ++t) {
    parts[t] = mat_new(p->ndof);
    int64_t b = cv_begin + len * t / nthreads, e = cv_begin + len * (t + 1) / nthreads;
    assemble_range(p, parts[t], b, e);
  }
  for (int t = 1; t < nthreads; ++t) {
    mat_add_into(parts[0], parts[t]);
    orc_mat_free(parts[t]);
  }
  orc_mat *K = parts[0];
  free(parts);
  return K;
}

int64_t orc_mat_nnz(const void *mm) {
  const orc_mat *m = (const orc_mat *)mm;
  int64_t s = 0;
  for (int64_t j = 0; j < m->ncols; ++j) s += m->cols[j].n;
  return s;
}
int64_t orc_mat_ncols(const void *mm) { return ((const orc_mat *)mm)->ncols; }

void orc_mat_export_csc(const void *mm, int64_t *colptr, int32_t *rows, double *vals) {
  const orc_mat *m = (const orc_mat *)mm;
  int64_t s = 0;
  for (int64_t j = 0; j < m->ncols; ++j) {
    colptr[j] = s;
    const orc_col *c = &m->cols[j];
    if (rows) memcpy(rows + s, c->r, sizeof(int32_t) * c->n);
    if (vals) memcpy(vals + s, c->v, sizeof(double) * c->n);
    s += c->n;
  }
  colptr[m->ncols] = s;
}

void orc_mat_free(void *mm) {
  orc_mat *m = (orc_mat *)mm;
  if (!m) return;
  for (int64_t j = 0; j < m->ncols; ++j) { free(m->cols[j].r); free(m->cols[j].v); }
  free(m->cols);
  free(m);
}

int orc_assemble_vector(const orc_problem *p, double *V) {
  orc_tables t;
  make_tables(p, &t);
  int nl = t.nloc, Q = t.qdim;
  double *elem = (double *)malloc(sizeof(double) * nl);
  double *g = (double *)malloc(sizeof(double) * 3 * t.nd);
  int rc = 0;
  for (int64_t cv = 0; cv < p->n_cv; ++cv) {
    if (element_vector_impl(p, &t, cv, elem, g) < 0) { rc = -1; break; }
    for (int i = 0; i < t.nd; ++i)
      for (int q = 0; q < Q; ++q) V[p->dofs[(int64_t)t.nd * cv + i] + q] += elem[i * Q + q];
  }
  free(elem); free(g); free_tables(&t);
  return rc;
}

/* ======================================================================
 * SIMP compliance on P1 vector fields (IM_TETRAHEDRON(1)):
 *   J = sum_e E(rho_e) int_T lambda (div u)^2 + 2 mu eps(u):eps(u)
 *     = u^T K(rho) u,   E = emin + rho^penal (1 - emin)
 * dJ/drho_e = E'(rho_e) u_e^T Khat_e u_e;  dJ/dx_v = sum_e E w |J|
 * (W I - H^T P) g_v with H = grad u, P = dW/dH (shape derivative of the
 * pulled-back integrand).
 * ====================================================================== */
double orc_compliance(const orc_problem *p, const double *rho, double emin,
                      double penal, double *dJ_drho, double *dJ_dX) {
  orc_tables t;
  make_tables(p, &t);
  double lam = p->scal_a, mu = p->scal_b, Jtot = 0.0;
  double g[3 * 64];
  if (dJ_dX) memset(dJ_dX, 0, sizeof(double) * 3 * p->n_pts);
  for (int64_t cv = 0; cv < p->n_cv; ++cv) {
    double r = rho[cv];
    double E = emin + pow(r, penal) * (1.0 - emin);
    double dE = penal * pow(r, penal - 1.0) * (1.0 - emin);
    double Je = 0.0, Gx[4][3] = {{0}};
    for (int q = 0; q < t.nq; ++q) {
      double absJ = element_geometry(p, &t, cv, q, g);
      double H[9] = {0};
      for (int n = 0; n < t.nd; ++n) {
        int64_t d = p->dofs[(int64_t)t.nd * cv + n];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) H[3 * a + b] += p->U[d + a] * g[3 * n + b];
      }
      double tr = H[0] + H[4] + H[8];
      double eps[9], P[9], W = lam * tr * tr;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) eps[3 * a + b] = 0.5 * (H[3 * a + b] + H[3 * b + a]);
      for (int k = 0; k < 9; ++k) W += 2.0 * mu * eps[k] * eps[k];
      for (int k = 0; k < 9; ++k) P[k] = 4.0 * mu * eps[k] + (k % 4 == 0 ? 2.0 * lam * tr : 0.0);
      double wq = absJ * p->qw[q];
      Je += wq * W;
      if (dJ_dX && p->gt == ORC_GT_PK31 && p->fem_k == 1) {
        /* M = W I - H^T P ; dJ/dx_v += E wq M g_v */
        double M[9];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += H[3 * k + a] * P[3 * k + b];
            M[3 * a + b] = (a == b ? W : 0.0) - s;
          }
        for (int v = 0; v < 4; ++v)
          for (int a = 0; a < 3; ++a)
            Gx[v][a] += wq * (M[3 * a] * g[3 * v] + M[3 * a + 1] * g[3 * v + 1] + M[3 * a + 2] * g[3 * v + 2]);
      }
    }
    Jtot += E * Je;
    if (dJ_drho) dJ_drho[cv] = dE * Je;
    if (dJ_dX && p->gt == ORC_GT_PK31 && p->fem_k == 1)
      for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) dJ_dX[3 * (int64_t)p->cvx[4 * cv + v] + a] += E * Gx[v][a];
  }
  free_tables(&t);
  return Jtot;
}

/* ======================================================================
 * Incomplete LDL^T preconditioner (gmm::ildlt_precond)
 * ====================================================================== */

/* U's slots: the diagonal of row i first, then the columns j > i of A's row
 * i (gmm_precond_ildlt.h:146-166 builds the same layout in two passes). */
int64_t orc_ildlt_pattern(int64_t n, const int64_t *ip, const int32_t *ix, int64_t *uptr, int32_t *uind) {
  int64_t s = 0;
  for (int64_t i = 0; i < n; ++i) {
    uptr[i] = s;
    if (uind) uind[s] = (int32_t)i;
    ++s;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k)
      if (ix[k] > i) {
        if (uind) uind[s] = ix[k];
        ++s;
      }
  }
  uptr[n] = s;
  return s;
}

/* Right-looking level-0 LDL^T in U's slots (:168-196): for k = 0, 1, ...
 * the pivot z = U_kk (replaced by 1 when |z| <= maxp, with maxp starting at
 * eps |A_00| and raised to min(eps |z|, 1) after each pivot), row k scaled
 * by 1/z, then for every slot (k, h) of row k the entries (h, j) that row k
 * also holds receive -= (U_kh z) U_kj, in column order.  No fill: entries
 * outside U's slots are dropped.  Products and differences are rounded one
 * at a time (no fused multiply-add), as the reference's x86-64 build does. */
int64_t orc_ildlt(int64_t n, const int64_t *ip, const int32_t *ix, const double *av,
                  const int64_t *uptr, const int32_t *uind, double *uval) {
  if (n <= 0) return 0;
  const double eps = 2.220446049250313e-16; /* numeric_limits<double>::epsilon() */
  double a00 = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t s = uptr[i];
    uval[s++] = 0.0;
    for (int64_t k = ip[i]; k < ip[i + 1]; ++k) {
      if (ix[k] == i) uval[uptr[i]] = av[k];
      else if (ix[k] > i) uval[s++] = av[k];
    }
    if (i == 0) a00 = uval[uptr[0]];
  }
  if (a00 == 0.0) uval[uptr[0]] = 1.0;
  double maxp = fabs(a00) * eps;
  int64_t replaced = 0;
  for (int64_t k = 0; k < n; ++k) {
    const int64_t d = uptr[k], e = uptr[k + 1];
    double z = uval[d];
    if (fabs(z) <= maxp) {
      uval[d] = z = 1.0;
      ++replaced;
    }
    const double cap = fabs(z) * eps < 1.0 ? fabs(z) * eps : 1.0;
    if (cap > maxp) maxp = cap;
    for (int64_t t = d + 1; t < e; ++t) uval[t] /= z;
    for (int64_t t = d + 1; t < e; ++t) {
      volatile double zz = uval[t] * z; /* rounded before use */
      const int32_t h = uind[t];
      int64_t g = t; /* row k's slots from column h on */
      for (int64_t j = uptr[h]; j < uptr[h + 1] && g < e; ++j) {
        while (g < e && uind[g] < uind[j]) ++g;
        if (g < e && uind[g] == uind[j]) {
          volatile double prod = zz * uval[g];
          uval[j] = uval[j] - prod;
        }
      }
    }
  }
  return replaced;
}

/* z = (L D L^T)^{-1} r with L = U^T unit lower (:200-206): the column-wise
 * forward substitution of gmm::lower_tri_solve on conjugated(U)
 * (gmm_tri_solve.h:75-89), the division by D, then the row-wise backward
 * substitution of gmm::upper_tri_solve (:108-123), unit diagonals. */
void orc_ildlt_apply(int64_t n, const int64_t *uptr, const int32_t *uind, const double *uval,
                     const double *r, double *z) {
  for (int64_t i = 0; i < n; ++i) z[i] = r[i];
  for (int64_t j = 0; j < n; ++j) {
    const double xj = z[j];
    for (int64_t t = uptr[j] + 1; t < uptr[j + 1]; ++t) {
      volatile double prod = xj * uval[t];
      z[uind[t]] = z[uind[t]] - prod;
    }
  }
  for (int64_t i = 0; i < n; ++i) z[i] /= uval[uptr[i]];
  for (int64_t i = n - 1; i >= 0; --i) {
    double t = z[i];
    for (int64_t s = uptr[i] + 1; s < uptr[i + 1]; ++s) {
      volatile double prod = uval[s] * z[uind[s]];
      t = t - prod;
    }
    z[i] = t;
  }
}

/* ======================================================================
 * gmm::ilu_precond of GetFEM's column-major matrix (gmm_precond_ilu.h:
 * 104-219 on transposed(A), invert = true) and gmm::mult (:221-231)
 * ====================================================================== */

/* B = A^T by rows (CSR of B = CSC of A), sorted columns */
static void csr_transpose(int64_t n, const int64_t *ip, const int32_t *ix, const double *av,
                          int64_t *bp, int32_t *bx, double *bv) {
  for (int64_t i = 0; i <= n; ++i) bp[i] = 0;
  for (int64_t p = 0; p < ip[n]; ++p) bp[ix[p] + 1]++;
  for (int64_t i = 0; i < n; ++i) bp[i + 1] += bp[i];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  memcpy(fill, bp, sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = ip[i]; p < ip[i + 1]; ++p) {
      const int64_t d = fill[ix[p]]++;
      bx[d] = (int32_t)i;
      bv[d] = av[p];
    }
  free(fill);
}

int64_t orc_ilu(int64_t n, const int64_t *ip, const int32_t *ix, const double *av, double *F) {
  const int64_t nnz = ip[n];
  int64_t *bp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  int32_t *bx = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  double *bv = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
  csr_transpose(n, ip, ix, av, bp, bx, bv);
  /* L (strictly lower) and U (diagonal first, then the upper entries) of B */
  int64_t *Lp = (int64_t *)calloc(n + 1, sizeof(int64_t)), *Up = (int64_t *)calloc(n + 1, sizeof(int64_t));
  int32_t *Li = (int32_t *)malloc(sizeof(int32_t) * (nnz + 1)), *Ui = (int32_t *)malloc(sizeof(int32_t) * (nnz + n + 1));
  double *Lv = (double *)malloc(sizeof(double) * (nnz + 1)), *Uv = (double *)malloc(sizeof(double) * (nnz + n + 1));
  int64_t lo = 0, uo = 0;
  for (int64_t i = 0; i < n; ++i) {
    Ui[uo] = (int32_t)i; Uv[uo] = 0.0; ++uo;  /* the diagonal slot */
    for (int64_t p = bp[i]; p < bp[i + 1] && p - bp[i] < 1000; ++p) {
      const int32_t j = bx[p];
      if (j < i) { Li[lo] = j; Lv[lo] = bv[p]; ++lo; }
      else if (j == i) Uv[Up[i]] = bv[p];
      else { Ui[uo] = j; Uv[uo] = bv[p]; ++uo; }
    }
    Lp[i + 1] = lo;
    Up[i + 1] = uo;
  }
  const double prec = 2.220446049250313e-16;
  double max_pivot = fabs(Uv[0]) * prec;
  int64_t replaced = 0;
  if (n > 0 && Uv[0] == 0.0) { Uv[0] = 1.0; ++replaced; }
  for (int64_t i = 1; i < n; ++i) {
    int64_t pn = Up[i];
    if (fabs(Uv[pn]) <= max_pivot) { Uv[pn] = 1.0; ++replaced; }
    { double t = fabs(Uv[pn]) * prec; if (t > 1.0) t = 1.0; if (t > max_pivot) max_pivot = t; }
    for (int64_t j = Lp[i]; j < Lp[i + 1]; ++j) {
      pn = Up[Li[j]];
      const double multiplier = (Lv[j] /= Uv[pn]);
      int64_t qn = j + 1, rn = Up[i];
      for (pn++; pn < Up[Li[j] + 1] && Ui[pn] < i; pn++) {
        while (qn < Lp[i + 1] && Li[qn] < Ui[pn]) qn++;
        if (qn < Lp[i + 1] && Ui[pn] == Li[qn]) Lv[qn] -= multiplier * Uv[pn];
      }
      for (; pn < Up[Li[j] + 1]; pn++) {
        while (rn < Up[i + 1] && Ui[rn] < Ui[pn]) rn++;
        if (rn < Up[i + 1] && Ui[pn] == Ui[rn]) Uv[rn] -= multiplier * Uv[pn];
      }
    }
  }
  /* F aligned with A's CSR: the entry (i, j) of row i of B has A's pattern */
  for (int64_t i = 0; i < n; ++i) {
    int64_t l = Lp[i], u = Up[i] + 1;
    for (int64_t p = ip[i]; p < ip[i + 1]; ++p) {
      const int32_t j = ix[p];
      if (j < i) { F[p] = (l < Lp[i + 1] && Li[l] == j) ? Lv[l] : 0.0; ++l; }
      else if (j == i) F[p] = Uv[Up[i]];
      else { F[p] = (u < Up[i + 1] && Ui[u] == j) ? Uv[u] : 0.0; ++u; }
    }
  }
  free(bp); free(bx); free(bv); free(Lp); free(Up); free(Li); free(Ui); free(Lv); free(Uv);
  return replaced;
}

/* z = M^-1 r: lower_tri_solve(transposed(U), z, false) then
 * upper_tri_solve(transposed(L), z, true) (gmm_tri_solve.h:75-89, :45-57),
 * column-oriented on the transposes, i.e. over the rows of U and L; F as
 * orc_ilu returns it (A's pattern, which must be symmetric) */
void orc_ilu_apply(int64_t n, const int64_t *ip, const int32_t *ix, const double *F, const double *r, double *z) {
  for (int64_t i = 0; i < n; ++i) z[i] = r[i];
  for (int64_t j = 0; j < n; ++j) {  /* column j of U^T = row j of U */
    double dj = 0.0;
    for (int64_t p = ip[j]; p < ip[j + 1]; ++p) if (ix[p] == j) dj = F[p];
    z[j] /= dj;
    const double xj = z[j];
    for (int64_t p = ip[j]; p < ip[j + 1]; ++p) if (ix[p] > j) z[ix[p]] -= xj * F[p];
  }
  for (int64_t j = n - 1; j >= 0; --j) {  /* column j of L^T = row j of L, unit diagonal */
    const double xj = z[j];
    for (int64_t p = ip[j]; p < ip[j + 1]; ++p) if (ix[p] < j) z[ix[p]] -= xj * F[p];
  }
}

