// Reference-element data: integration rules, Lagrange elements, geometric
// transformations and first-encounter dof numbering.
#include "fem_ref.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>

namespace gfx {

namespace {

// Symmetric orbits of the simplex rules stored in getfem_im_list.h
// (IM_TETRAHEDRON(k): data :3251-3340, node types :4612-4616).  A row is a
// point in reference coordinates, its weight, and whether all permutations of
// its barycentric coordinates belong to the rule (add_point_full_symmetric,
// getfem_integration.cc:278-305).
struct Orbit { double x, y, z, w; bool sym; };

const Orbit kTet1[] = {{0.25, 0.25, 0.25, 1.0 / 6.0, false}};
const Orbit kTet2[] = {{0.13819660112501052, 0.13819660112501052, 0.13819660112501052,
                        1.0 / 24.0, true}};
const Orbit kTet3[] = {{0.25, 0.25, 0.25, -2.0 / 15.0, false},
                       {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.075, true}};
const Orbit kTet5[] = {
    {0.25, 0.25, 0.25, 0.019753086419753086, false},
    {0.31979362782962991, 0.31979362782962991, 0.31979362782962991, 0.011511367871045398, true},
    {0.091971078052723033, 0.091971078052723033, 0.091971078052723033, 0.01198951396316977, true},
    {0.056350832689629156, 0.056350832689629156, 0.44364916731037084, 0.008818342151675485, true}};
const Orbit kTet6[] = {
    {0.214602871259152029288839219386284, 0.214602871259152029288839219386284,
     0.214602871259152029288839219386284, 0.00665379170969458201661510459291332, true},
    {0.0406739585346113531155794489564100, 0.0406739585346113531155794489564100,
     0.0406739585346113531155794489564100, 0.00167953517588677382466887290765614, true},
    {0.322337890142275510343994470762492, 0.322337890142275510343994470762492,
     0.322337890142275510343994470762492, 0.00922619692394245368252554630895433, true},
    {0.0636610018750175252992355276057269, 0.0636610018750175252992355276057269,
     0.269672331458315808034097805727606, 0.00803571428571428571428571428571428, true}};
const Orbit kTet8[] = {
    {0.25, 0.25, 0.25, -0.0205001886586399158405865177642941, true},
    {0.206829931610673204083980900024961, 0.206829931610673204083980900024961,
     0.206829931610673204083980900024961, 0.0142503058228669012484397415358704, true},
    {0.0821035883105467230906058078714215, 0.0821035883105467230906058078714215,
     0.0821035883105467230906058078714215, 0.00196703331313390098756280342445466, true},
    {0.00578195050519799725317663886414270, 0.00578195050519799725317663886414270,
     0.00578195050519799725317663886414270, 0.000169834109092887379837744566704016, true},
    {0.0505327400188942244256245285579071, 0.0505327400188942244256245285579071,
     0.449467259981105775574375471442092, 0.00457968382446728180074351446297276, true},
    {0.229066536116811139600408854554753, 0.229066536116811139600408854554753,
     0.0356395827885340437169173969506114, 0.00570448580868191850680255862783040, true},
    {0.0366077495531974236787738546327104, 0.0366077495531974236787738546327104,
     0.190486041934633455699433285315099, 0.00214051914116209259648335300092023, true}};

void push_unique(IntegrationRule &r, const double p[3], double w) {
  for (int i = 0; i < r.n(); ++i) {
    const double *q = &r.pts[3 * i];
    if (std::fabs(q[0] - p[0]) + std::fabs(q[1] - p[1]) + std::fabs(q[2] - p[2]) < 1e-14) return;
  }
  r.pts.insert(r.pts.end(), p, p + 3);
  r.w.push_back(w);
}

void expand(IntegrationRule &r, const Orbit *ob, int n) {
  for (int o = 0; o < n; ++o) {
    const double b[4] = {ob[o].x, ob[o].y, ob[o].z, 1.0 - ob[o].x - ob[o].y - ob[o].z};
    if (!ob[o].sym) { push_unique(r, b, ob[o].w); continue; }
    // distinct index triples (i0, i1, i2) of {0..3}, i0 fastest
    for (int t = 0; t < 64; ++t) {
      int i0 = t & 3, i1 = (t >> 2) & 3, i2 = (t >> 4) & 3;
      if (i0 == i1 || i0 == i2 || i1 == i2) continue;
      const double p[3] = {b[i0], b[i1], b[i2]};
      push_unique(r, p, ob[o].w);
    }
  }
}

// Gauss-Legendre nodes/weights on [0,1], ascending.
void gauss_legendre01(int m, std::vector<double> &x, std::vector<double> &w) {
  x.assign(m, 0.0); w.assign(m, 0.0);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < (m + 1) / 2; ++k) {
    long double r = std::cos(pi * (k + 0.75L) / (m + 0.5L));
    long double pm = 0, pm1 = 0;
    for (int it = 0; it < 60; ++it) {
      long double a = 1, b = r;  // P0, P1
      if (m == 1) { pm = r; pm1 = 1; }
      else {
        for (int j = 2; j <= m; ++j) { long double c = ((2 * j - 1) * r * b - (j - 1) * a) / j; a = b; b = c; }
        pm = b; pm1 = a;
      }
      long double d = m * (r * pm - pm1) / (r * r - 1);
      long double dr = pm / d;
      r -= dr;
      if (std::fabs((double)dr) < 1e-22) break;
    }
    {
      long double a = 1, b = r;
      if (m == 1) { pm1 = 1; }
      else {
        for (int j = 2; j <= m; ++j) { long double c = ((2 * j - 1) * r * b - (j - 1) * a) / j; a = b; b = c; }
        pm1 = a;
      }
    }
    long double wt = (1 - r * r) / ((long double)m * pm1 * (long double)m * pm1);
    // r > 0 here (descending cosines): mirror pair
    x[m - 1 - k] = (double)(0.5L + 0.5L * r);
    w[m - 1 - k] = (double)wt;
    x[k] = (double)(0.5L - 0.5L * r);
    w[k] = (double)wt;
  }
  if (m & 1) x[m / 2] = 0.5;
}

bool parse2(const std::string &s, const char *prefix, int &a, int &b) {
  size_t lp = std::strlen(prefix);
  if (s.compare(0, lp, prefix) != 0) return false;
  char tail;
  return std::sscanf(s.c_str() + lp, "%d,%d%c", &a, &b, &tail) == 3 && tail == ')';
}
bool parse1(const std::string &s, const char *prefix, int &a) {
  size_t lp = std::strlen(prefix);
  if (s.compare(0, lp, prefix) != 0) return false;
  char tail;
  return std::sscanf(s.c_str() + lp, "%d%c", &a, &tail) == 2 && tail == ')';
}

std::string strip(const std::string &s) {
  std::string o;
  for (char c : s) if (c != ' ' && c != '\t') o += c;
  return o;
}

// prod_{j<m} (K l - j)/(j+1) and derivative in l
inline void lagrange_factor(int K, int m, double l, double &v, double &dv) {
  double val = 1.0, der = 0.0;
  for (int j = 0; j < m; ++j) {
    const double f = (K * l - j) / (j + 1.0), df = K / (j + 1.0);
    der = der * f + val * df;
    val *= f;
  }
  v = val; dv = der;
}

}  // namespace

bool integration_rule(const std::string &name_in, IntegrationRule &r) {
  const std::string name = strip(name_in);
  r = IntegrationRule();
  r.name = name;
  int k = 0, d = 0;
  if (parse1(name, "IM_TETRAHEDRON(", k)) {
    r.gt = GT_TET_P1;
    r.degree = k;
    switch (k) {
      case 1: expand(r, kTet1, 1); break;
      case 2: expand(r, kTet2, 1); break;
      case 3: expand(r, kTet3, 2); break;
      case 5: expand(r, kTet5, 4); break;
      case 6: expand(r, kTet6, 4); break;
      case 8: expand(r, kTet8, 7); break;
      default: return false;
    }
    return true;
  }
  if (parse2(name, "IM_GAUSS_PARALLELEPIPED(", d, k) && d == 3 && k >= 0 && k < 128) {
    // IM_PRODUCT of 1D Gauss rules with k/2+1 points, first factor fastest
    r.gt = GT_HEX_Q1;
    r.degree = k;
    const int m = k / 2 + 1;
    std::vector<double> x, w;
    gauss_legendre01(m, x, w);
    for (int iz = 0; iz < m; ++iz)
      for (int iy = 0; iy < m; ++iy)
        for (int ix = 0; ix < m; ++ix) {
          r.pts.push_back(x[ix]); r.pts.push_back(x[iy]); r.pts.push_back(x[iz]);
          r.w.push_back((w[ix] * w[iy]) * w[iz]);
        }
    return true;
  }
  return false;
}

bool fem_descriptor(const std::string &name_in, FemDesc &f) {
  const std::string name = strip(name_in);
  f = FemDesc();
  f.name = name;
  int d = 0, k = 0;
  if (parse2(name, "FEM_PK_DISCONTINUOUS(", d, k) && d == 3 && k == 0) {
    f.kind = FEM_PK_DISC0; f.k = 0; f.nd = 1; f.gt = GT_TET_P1;
    f.lattice = {0, 0, 0, 0};
    f.nodes = {0.25, 0.25, 0.25};
    return true;
  }
  if (parse2(name, "FEM_PK(", d, k) && d == 3 && k >= 1 && k <= 6) {
    f.kind = FEM_PK; f.k = k; f.gt = GT_TET_P1;
    // lattice of K_simplex_of_ref_ (bgeot_convex_ref.cc:311-321): first
    // reference coordinate fastest, coordinates summing to at most k
    for (int c2 = 0; c2 <= k; ++c2)
      for (int c1 = 0; c1 + c2 <= k; ++c1)
        for (int c0 = 0; c0 + c1 + c2 <= k; ++c0) {
          f.lattice.insert(f.lattice.end(), {k - c0 - c1 - c2, c0, c1, c2});
          f.nodes.insert(f.nodes.end(), {double(c0) / k, double(c1) / k, double(c2) / k});
        }
    f.nd = (int)f.nodes.size() / 3;
    return true;
  }
  if (parse2(name, "FEM_QK(", d, k) && d == 3 && k >= 1 && k <= 2) {
    // FEM_PRODUCT(FEM_QK(2,k), FEM_PK(1,k)): first factor fastest
    f.kind = FEM_QK; f.k = k; f.gt = GT_HEX_Q1;
    for (int iz = 0; iz <= k; ++iz)
      for (int iy = 0; iy <= k; ++iy)
        for (int ix = 0; ix <= k; ++ix) {
          f.lattice.insert(f.lattice.end(), {ix, iy, iz, 0});
          f.nodes.insert(f.nodes.end(), {double(ix) / k, double(iy) / k, double(iz) / k});
        }
    f.nd = (int)f.nodes.size() / 3;
    return true;
  }
  return false;
}

int geotrans_descriptor(const std::string &name_in) {
  const std::string name = strip(name_in);
  if (name == "GT_PK(3,1)") return GT_TET_P1;
  if (name == "GT_QK(3,1)") return GT_HEX_Q1;
  return -1;
}

int geotrans_nb_vertices(int gt) { return gt == GT_TET_P1 ? 4 : 8; }

void fem_eval(const FemDesc &f, const double *x, double *val, double *grad) {
  if (f.kind == FEM_PK_DISC0) {
    val[0] = 1.0;
    if (grad) grad[0] = grad[1] = grad[2] = 0.0;
    return;
  }
  if (f.kind == FEM_PK) {
    // product of barycentric Lagrange factors (getfem_fem.cc:731-755)
    const double lam[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    const double dlam[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < f.nd; ++i) {
      double v[4], dv[4];
      for (int b = 0; b < 4; ++b) lagrange_factor(f.k, f.lattice[4 * i + b], lam[b], v[b], dv[b]);
      val[i] = (v[0] * v[1]) * (v[2] * v[3]);
      if (!grad) continue;
      for (int c = 0; c < 3; ++c) {
        double g = 0.0;
        for (int b = 0; b < 4; ++b) {
          if (dlam[b][c] == 0.0) continue;
          double p = dv[b] * dlam[b][c];
          for (int o = 0; o < 4; ++o) if (o != b) p *= v[o];
          g += p;
        }
        grad[3 * i + c] = g;
      }
    }
    return;
  }
  // FEM_QK: tensor product of 1D Lagrange polynomials on 0, 1/k, ..., 1
  double v1[3][8], d1[3][8];
  for (int c = 0; c < 3; ++c)
    for (int m = 0; m <= f.k; ++m) {
      double a, da, b, db;
      lagrange_factor(f.k, f.k - m, 1.0 - x[c], a, da);
      lagrange_factor(f.k, m, x[c], b, db);
      v1[c][m] = a * b;
      d1[c][m] = b * (-da) + a * db;
    }
  for (int i = 0; i < f.nd; ++i) {
    const int ix = f.lattice[4 * i], iy = f.lattice[4 * i + 1], iz = f.lattice[4 * i + 2];
    val[i] = v1[0][ix] * v1[1][iy] * v1[2][iz];
    if (!grad) continue;
    grad[3 * i + 0] = d1[0][ix] * v1[1][iy] * v1[2][iz];
    grad[3 * i + 1] = v1[0][ix] * d1[1][iy] * v1[2][iz];
    grad[3 * i + 2] = v1[0][ix] * v1[1][iy] * d1[2][iz];
  }
}

void geotrans_eval(int gt, const double *x, double *N, double *dN) {
  FemDesc f;
  fem_descriptor(gt == GT_TET_P1 ? "FEM_PK(3,1)" : "FEM_QK(3,1)", f);
  fem_eval(f, x, N, dN);
}

namespace {

// key of a lattice node: the sorted multiset of the global vertices it is
// built from (tet) or of the sub-entity it lies on (hex), up to 8 ids
struct NodeKey {
  int32_t v[8];
  bool operator==(const NodeKey &o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};
struct NodeKeyHash {
  size_t operator()(const NodeKey &k) const {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < 8; ++i) { h ^= (uint32_t)k.v[i]; h *= 1099511628211ULL; h ^= h >> 29; }
    return (size_t)h;
  }
};

}  // namespace

int64_t enumerate_dofs(int gt, int64_t n_pts, const double *pts, int64_t n_cv,
                       const int32_t *cvx, const FemDesc &f, int qdim,
                       std::vector<int32_t> &dofs, std::string &err) {
  (void)pts;
  const int nv = geotrans_nb_vertices(gt);
  dofs.assign((size_t)n_cv * f.nd, 0);
  if (f.kind == FEM_PK_DISC0) {  // not linkable: a new dof per element
    for (int64_t e = 0; e < n_cv; ++e) dofs[e] = (int32_t)(e * qdim);
    return n_cv * qdim;
  }
  if (f.gt != gt) { err = "finite element does not match the geometric transformation"; return -1; }
  if (f.kind == FEM_QK && f.k > 2) { err = "FEM_QK(3,k) supported for k <= 2"; return -1; }
  std::unordered_map<NodeKey, int32_t, NodeKeyHash> seen;
  seen.reserve((size_t)(n_cv * f.nd / 3 + 16));
  int64_t nbdof = 0;
  for (int64_t e = 0; e < n_cv; ++e) {
    const int32_t *cv = cvx + e * nv;
    for (int i = 0; i < f.nd; ++i) {
      NodeKey key;
      for (int t = 0; t < 8; ++t) key.v[t] = -1;
      int n = 0;
      if (f.kind == FEM_PK) {
        for (int b = 0; b < 4; ++b)
          for (int r = 0; r < f.lattice[4 * i + b]; ++r) key.v[n++] = cv[b];
      } else {
        // hex vertex (dx,dy,dz) index dx + 2dy + 4dz; an interior coordinate
        // spans both sides
        const int *L = &f.lattice[4 * i];
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              const int d[3] = {dx, dy, dz};
              bool on = true;
              for (int c = 0; c < 3; ++c) {
                if (L[c] == 0 && d[c] != 0) on = false;
                if (L[c] == f.k && d[c] != 1) on = false;
              }
              if (on) key.v[n++] = cv[dx + 2 * dy + 4 * dz];
            }
      }
      if (n > 8 || (f.kind == FEM_PK && n > 8)) { err = "element degree too high"; return -1; }
      std::sort(key.v, key.v + n);
      auto it = seen.find(key);
      if (it == seen.end()) {
        if (nbdof + qdim > INT32_MAX) { err = "too many dofs for int32 indices"; return -1; }
        it = seen.emplace(key, (int32_t)nbdof).first;
        nbdof += qdim;
      }
      dofs[(size_t)e * f.nd + i] = it->second;
    }
  }
  (void)n_pts;
  return nbdof;
}

}  // namespace gfx
