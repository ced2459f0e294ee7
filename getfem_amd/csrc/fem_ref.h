// Reference-element data of the supported subset: integration rules,
// Lagrange elements and geometric transformations on the reference
// tetrahedron / cube.  Host-side; these tables are uploaded to HBM once per
// plan and read by the assembly kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gfx {

enum GeoTrans { GT_TET_P1 = 0, GT_HEX_Q1 = 1 };
enum FemKind { FEM_PK = 0, FEM_QK = 1, FEM_PK_DISC0 = 2 };

struct IntegrationRule {
  std::string name;
  int gt = GT_TET_P1;
  int degree = 0;          // polynomial degree integrated exactly
  std::vector<double> pts; // n * 3
  std::vector<double> w;   // n
  int n() const { return (int)w.size(); }
};

struct FemDesc {
  std::string name;
  int kind = FEM_PK;
  int k = 1;
  int nd = 0;    // scalar dofs per element
  int gt = GT_TET_P1;
  // lattice multi-index of each node: tet = barycentric counts (4 ints summing
  // to k), hex = (ix, iy, iz) in 0..k
  std::vector<int> lattice;
  std::vector<double> nodes;  // nd * 3 reference coordinates
};

// Parse "IM_TETRAHEDRON(k)" / "IM_GAUSS_PARALLELEPIPED(3,k)"; false if unknown.
bool integration_rule(const std::string &name, IntegrationRule &out);
// Parse "FEM_PK(3,k)" / "FEM_QK(3,k)" / "FEM_PK_DISCONTINUOUS(3,0)".
bool fem_descriptor(const std::string &name, FemDesc &out);
// Parse "GT_PK(3,1)" / "GT_QK(3,1)": returns GeoTrans or -1.
int geotrans_descriptor(const std::string &name);
int geotrans_nb_vertices(int gt);

// Values (nd) and reference gradients (nd*3) of the element base functions.
void fem_eval(const FemDesc &f, const double *x, double *val, double *grad);
// Shape functions of the geometric transformation (nv) and gradients (nv*3).
void geotrans_eval(int gt, const double *x, double *N, double *dN);

// Dof table (basic dof of every element node) by first-encounter numbering.
// Returns ndof, or -1 on unsupported input (message in err).
int64_t enumerate_dofs(int gt, int64_t n_pts, const double *pts, int64_t n_cv,
                       const int32_t *cvx, const FemDesc &f, int qdim,
                       std::vector<int32_t> &dofs, std::string &err);

}  // namespace gfx
