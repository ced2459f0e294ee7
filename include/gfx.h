/*
 * gfx.h -- C ABI of libgfx, the MI355X assembly engine behind GetFEM's
 * generic assembly (ga_workspace::assembly) for a subset of GWFL
 * expressions.
 *
 * Conventions
 *  - Every entry point returns an int status (GFX_OK = 0); nothing throws
 *    across the ABI.  gfx_ctx_last_error() gives the message of the last
 *    failure on that context.  GetFEM itself reports the same conditions by
 *    throwing gmm::gmm_error (src/gmm/gmm_except.h:55, :163).
 *  - Pointers named h_* are host memory, d_* device memory of the context's
 *    device; host arrays are only read during the call (GetFEM's workspace
 *    likewise borrows its inputs, src/getfem/getfem_generic_assembly.h:271-298).
 *  - Matrices are CSR with rows = test functions (Test_u) and sorted column
 *    indices = trial functions (Test2_u); the same arrays are GetFEM's
 *    col_matrix<rsvector> (CSC over trial dofs, src/gmm/gmm_vector.h:913-995)
 *    of the transposed matrix, and gfx_plan_export_csc gives GetFEM's own
 *    orientation.
 *  - One gfx_ctx per host thread / stream; objects created from a context
 *    must be used from one host thread at a time (a ga_workspace is not
 *    thread safe either, src/getfem_models.cc:2686-2706).
 */
#ifndef GFX_H
#define GFX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define GFX_OK 0
#define GFX_E_INVALID 1            /* bad argument / size                      */
#define GFX_E_UNSUPPORTED 2        /* expression, element or rule not supported */
#define GFX_E_DEGENERATE_ELEMENT 3 /* det K == 0 ("Non invertible matrix",
                                      bgeot_geometric_trans.cc:380)           */
#define GFX_E_OOM 4                /* device allocation failed                 */
#define GFX_E_HIP 5                /* HIP runtime error                        */
#define GFX_E_STATE 6              /* call out of order (e.g. no pattern yet)  */
#define GFX_E_PATTERN 7            /* a mesh patch exceeds the kernel limits   */

/* ---- terms of the supported GWFL subset (after derivation, order 2) ---- */
#define GFX_TERM_LAPLACE 0    /* a*Grad_u:Grad_Test_u (a scalar or per element)      */
#define GFX_TERM_ISO_ELAST 1  /* lambda*Div_u*Div_Test_u + 2*mu*Sym(Grad_u):Grad_Test_u */
#define GFX_TERM_NEOHOOKEAN 2 /* ((Id(meshdim)+Grad_u)*(Compressible_Neo_Hookean_Bonet_PK2(
                                 Grad_u,[lambda;mu]))):Grad_Test_u (tangent + residual)  */
#define GFX_TERM_ISOTROPIC 3  /* alpha*Div_u*Div_Test_u + beta*(Grad_u)':Grad_Test_u
                                 + gamma*Grad_u:Grad_Test_u on a vector field: every sum
                                 of the two terms above on one mesh_im, which GetFEM
                                 merges into one tree (src/getfem_generic_assembly_
                                 workspace.cc:474-495); coefficients alpha, beta, gamma */

/* ---- contraction forms / precisions ---- */
#define GFX_FORM_FACTORIZED 0 /* affine element: B R_ij B^T with R_ij = sum_q w_q
                                 grad(phi_i) grad(phi_j)^T on the reference element */
#define GFX_FORM_PERQP 1      /* per integration point B^T D B, as ga_exec does      */
#define GFX_PREC_FP64 0       /* element contraction in FP64                         */
#define GFX_PREC_FP32 1       /* element contraction in FP32 from FP32 element records
                                 (B, scales, threshold rounded once per element);
                                 geometry and global accumulation in FP64           */
#define GFX_PREC_BF16X2 2     /* isotropic elasticity, affine P2/P3 (vector) plans:
                                 element matrices B^T D B on the matrix cores (bf16
                                 hi + lo split, three products, FP32 accumulation),
                                 buffered in FP32 and gathered into the CSR in FP64;
                                 relative error about 1e-5 of the element norm; the
                                 form argument is ignored; needs mu > 0 and
                                 3 lambda + 2 mu > 0 (checked at assembly)          */
#define GFX_PREC_FP64_MFMA 3  /* the same element-batched B^T D B on the FP64 matrix
                                 cores (v_mfma_f64_16x16x4_f64), element matrices
                                 buffered in FP64 and gathered: the FP64 variant of
                                 the dense element-batched form (same plans and
                                 material conditions as GFX_PREC_BF16X2)           */
#define GFX_PREC_FP32_MFMA 4  /* the same on the FP32 matrix cores
                                 (v_mfma_f32_16x16x4_f32), FP32 accumulation,
                                 element matrices buffered in FP32                 */

/* ---- coefficient sources ---- */
#define GFX_COEF_CONST 0       /* one value                                 */
#define GFX_COEF_PER_ELEMENT 1 /* one value per element (P0 field / im_data) */

typedef struct {
  int kind;            /* GFX_COEF_* */
  double value;        /* GFX_COEF_CONST */
  const double *h_values; /* GFX_COEF_PER_ELEMENT: n_cv values (host) */
} gfx_coef;

typedef struct gfx_ctx gfx_ctx;
typedef struct gfx_mesh gfx_mesh;
typedef struct gfx_mesh_fem gfx_mesh_fem;
typedef struct gfx_mesh_im gfx_mesh_im;
typedef struct gfx_plan gfx_plan;

/* ======================= host-only utilities =============
 * No device needed.  Reference-element data and dof numbering exactly as the
 * plans use them. */
/* integration rule: *npts receives the count; pts (npts*3) / w may be NULL */
int gfx_integration_rule(const char *im_name, int *npts, double *h_pts, double *h_w);
/* base values (nd) and reference gradients (nd*3) at x; *nd receives nd */
int gfx_fem_eval(const char *fem_name, const double *x, int *nd, double *h_val,
                 double *h_grad);
/* mesh_fem::enumerate_dof (src/getfem_mesh_fem.cc:320-440) without a device:
 * h_out receives n_cv*nd basic dofs, *ndof the number of dofs */
int gfx_enumerate_dofs(const char *geotrans, const char *fem_name, int qdim,
                       int64_t n_pts, const double *h_pts, int64_t n_cv,
                       const int32_t *h_cvx, int32_t *h_out, int64_t *ndof);

/* the GWFL compile stage alone: decls is "name:kind:n;..." with kind
 * var (fem variable, n components), const (fixed-size constant, n values) or
 * p0 (per-element fem constant); exprs are separated by '\n' and summed;
 * out receives the normal form of the order-2 term (as gfx_ws_describe) */
int gfx_gwfl_compile(const char *exprs, const char *decls, char *out, int64_t len);
/* test-function order of one expression (0: potential, 1: with Test_u, 2:
 * with Test_u and Test2_u), as ga_workspace::add_expression computes it
 * before deriving the higher orders (src/getfem_generic_assembly_workspace.cc:
 * 545-600); the workspace runs an expression in assembly(k) for order <= k <=
 * its derivation order */
int gfx_gwfl_order(const char *expr, const char *decls, int *order);

/* The polynomial route's compile stage alone (order 0: the potential;
 * order 1: the derivative with respect to the fem variable `var`; order 2:
 * the block of test variable and trial variable named by `var` as "u,p",
 * output slot_u * 4 qdim_p + slot_p with slot 0..qdim-1 the value and
 * qdim + 3a + b the gradient of a test function), as text: a first line
 * "nout=<outputs> zero=<variables held at zero>", a second "syms=<data
 * symbols>", then one line per monomial "<output> <coefficient> <symbol
 * indices>".  decls as gfx_gwfl_compile, plus the kinds "p0var" (a fem
 * variable on a P0 mesh_fem) and "data" (a fem constant of another fem). */
int gfx_gwfl_compile_scalar(const char *exprs, const char *decls, int order, const char *var, char *out,
                            int64_t len);
/* ======================= context ======================= */
int gfx_ctx_create(int device, gfx_ctx **out);
int gfx_ctx_destroy(gfx_ctx *ctx);
/* use an external HIP stream (e.g. torch's current stream); NULL restores the
 * context's own stream */
int gfx_ctx_set_stream(gfx_ctx *ctx, void *hip_stream);
void *gfx_ctx_stream(gfx_ctx *ctx);
int gfx_ctx_synchronize(gfx_ctx *ctx);
const char *gfx_ctx_last_error(gfx_ctx *ctx);
const char *gfx_version(void);

/* ======================= mesh =========================
 * Replaces the arrays behind getfem::mesh (points, convex -> point lists),
 * src/getfem/getfem_mesh.h; geotrans "GT_PK(3,1)" or "GT_QK(3,1)"
 * (bgeot_geometric_trans.cc). */
int gfx_mesh_create(gfx_ctx *ctx, const char *geotrans, int64_t n_pts,
                    const double *h_pts, int64_t n_cv, const int32_t *h_cvx,
                    gfx_mesh **out);
/* move the nodes (same connectivity); h_pts is n_pts*3 host doubles */
int gfx_mesh_set_points(gfx_mesh *m, const double *h_pts);
int gfx_mesh_destroy(gfx_mesh *m);

/* ======================= mesh_fem =====================
 * getfem::mesh_fem with a uniform Lagrange element and Qdim
 * (src/getfem_mesh_fem.cc:320-440, enumerate_dof): dofs are numbered in
 * first-encounter order over elements, Qdim consecutive dofs per node. */
int gfx_mesh_fem_create(gfx_mesh *m, const char *fem_name, int qdim,
                        gfx_mesh_fem **out);
/* same, with the caller's dof table: basic (component 0) dof of every element
 * node, i.e. mesh_fem::ind_scalar_basic_dof_of_element
 * (src/getfem/getfem_mesh_fem.h:461) for every element. */
int gfx_mesh_fem_create_from_dofs(gfx_mesh *m, const char *fem_name, int qdim,
                                  const int32_t *h_basic_dofs, int64_t ndof,
                                  gfx_mesh_fem **out);
int64_t gfx_mesh_fem_nb_dof(gfx_mesh_fem *mf);
int gfx_mesh_fem_nb_dof_per_element(gfx_mesh_fem *mf);
int gfx_mesh_fem_basic_dofs(gfx_mesh_fem *mf, int32_t *h_out);
/* coordinates of the node of every basic dof (ndof/qdim x 3) */
int gfx_mesh_fem_dof_nodes(gfx_mesh_fem *mf, double *h_out);
int gfx_mesh_fem_destroy(gfx_mesh_fem *mf);

/* ======================= mesh_im ======================
 * getfem::mesh_im with a uniform method: "IM_TETRAHEDRON(k)",
 * "IM_GAUSS_PARALLELEPIPED(3,k)" (src/getfem_integration.cc,
 * src/getfem/getfem_im_list.h). */
int gfx_mesh_im_create(gfx_mesh *m, const char *im_name, gfx_mesh_im **out);
int gfx_mesh_im_nb_points(gfx_mesh_im *mim);
int gfx_mesh_im_points(gfx_mesh_im *mim, double *h_pts, double *h_w);
int gfx_mesh_im_destroy(gfx_mesh_im *mim);

/* ======================= assembly plans ===============
 * A plan is one compiled order-2 term: it replaces, for one (mesh_im,
 * region = all elements, Test_u, Test2_u) tree, what ga_compile emits and
 * ga_exec runs (src/getfem_generic_assembly_compile_and_exec.cc:7910-8643,
 * :8750-9047). */
int gfx_plan_create(gfx_mesh_fem *mf, gfx_mesh_im *mim, int term,
                    const gfx_coef *coefs, int n_coefs, gfx_plan **out);
int gfx_plan_set_coef(gfx_plan *p, int slot, const gfx_coef *c);
int gfx_plan_set_form(gfx_plan *p, int form, int precision);
/* Pattern: GetFEM's stored pattern, i.e. entries with
 * |e| > 1e-14 * max|elem| in at least one element matrix
 * (compile_and_exec.cc:4853-4936, :5438-5467), built on the GPU. */
int gfx_plan_build_pattern(gfx_plan *p, int64_t *nnz);
/* assembly(2) into the pattern's values (device resident); for a
 * GFX_TERM_NEOHOOKEAN plan also the residual (order-1 assembly of the same
 * expression, ga_instruction_vector_assembly_mf, compile_and_exec.cc:4669-4735) */
int gfx_plan_assemble(gfx_plan *p);
/* nonlinear plans: the current displacement (ndof host doubles) and the
 * assembled residual vector */
int gfx_plan_set_state(gfx_plan *p, const double *h_U);
/* the same from a device vector (nb_dof doubles, stream-ordered, no host
 * synchronisation): a Newton loop keeps its state on the device */
int gfx_plan_set_state_device(gfx_plan *p, const double *d_U);
/* Re-thresholding.  GetFEM thresholds every element matrix at every assembly
 * (compile_and_exec.cc:5441-5453), so its stored pattern follows the values.
 * After gfx_plan_set_coef, gfx_mesh_set_points (on the plan's mesh) or a new
 * state, the next gfx_plan_assemble re-derives every element's threshold and
 * keep bits at the new values and compares the pattern they give with the
 * stored one: the same pattern is kept, a different one is rebuilt (new CSR
 * arrays; gfx_plan_pattern_id changes).  A pinned pattern is never rebuilt:
 * the assembly returns GFX_E_STATE instead, so that device pointers a caller
 * holds (gfx_plan_device_csr) stay valid. */
int gfx_plan_pin_pattern(gfx_plan *p, int pin);
int gfx_plan_pattern_id(gfx_plan *p, int64_t *id);
/* order-1 assembly alone (the residual; ga_workspace::assembly(1),
 * src/getfem_generic_assembly_workspace.cc:832-847 with
 * ga_instruction_vector_assembly_mf, compile_and_exec.cc:4669-4735), after
 * the pattern is built; the matrix values are left as they are */
int gfx_plan_assemble_residual(gfx_plan *p);
int gfx_plan_residual(gfx_plan *p, double *h_V);
int gfx_plan_copy_residual_device(gfx_plan *p, double *d_V);
int gfx_plan_dims(gfx_plan *p, int64_t *nrows, int64_t *nnz);
int gfx_plan_export_csr(gfx_plan *p, int64_t *h_indptr, int32_t *h_indices,
                        double *h_vals);
/* GetFEM orientation: CSC over trial dofs with sorted test rows, uint32 like
 * the Python interface's gsparse export (interface/src/getfemint.cc:712-737) */
int gfx_plan_export_csc(gfx_plan *p, int64_t *h_colptr, uint32_t *h_rows,
                        double *h_vals);
int gfx_plan_device_csr(gfx_plan *p, const int64_t **d_indptr,
                        const int32_t **d_indices, double **d_vals);
/* copy the CSR arrays into caller-owned device buffers of the same device
 * (stream-ordered on the context's stream; any pointer may be NULL) */
int gfx_plan_copy_csr_device(gfx_plan *p, int64_t *d_indptr, int32_t *d_indices,
                             double *d_vals);
/* d_y = K d_x with the assembled values (device vectors of nb_dof doubles,
 * stream-ordered): the product of a linear brick's matrix with the current
 * variable value that model::assembly adds to the residual
 * (gmm::mult_add(matj, vec_linear2, ...), src/getfem_models.cc:2560-2575) */
int gfx_plan_spmv(gfx_plan *p, const double *d_x, double *d_y);
/* d_y += K d_x */
int gfx_plan_spmv_add(gfx_plan *p, const double *d_x, double *d_y);
/* Add the assembled values into a caller-owned device matrix, as GetFEM's
 * assembly adds into the matrix given by ga_workspace::set_assembled_matrix
 * (src/getfem/getfem_generic_assembly.h:378-381; the workspace clears K only
 * when it owns it, src/getfem_generic_assembly_workspace.cc:804-807).  The
 * target's pattern must hold the plan's (GFX_E_PATTERN otherwise; entries it
 * lacks are dropped).  GFX_CSR: rows = test dofs, sorted columns; GFX_CSC:
 * GetFEM's col_matrix orientation, columns = trial dofs, sorted rows.  n is
 * the number of dofs; synchronises the context's stream (error check). */
#define GFX_CSR 0
#define GFX_CSC 1
int gfx_plan_add_to_matrix(gfx_plan *p, int orientation, int64_t n, const int64_t *d_ptr,
                           const int32_t *d_ind, double *d_vals);
/* Solve K x = b on the device by the preconditioned conjugate gradient
 * method, for the coercive (symmetric positive definite) systems GetFEM sends
 * to gmm::cg (default_linear_solver picks linear_solver_cg_preconditioned_ildlt,
 * src/getfem/getfem_model_solvers.h:90-98, :652-655; the preconditioner here is
 * block Jacobi, not ILDLT: the inverse 3x3 diagonal block of each node for a
 * qdim-3 field, the inverse diagonal otherwise).
 * Stopping test as gmm::cg (src/gmm/gmm_solver_cg.h:85-120): ||b - K x|| <=
 * rtol ||b||; b = 0 gives x = 0.  d_x: initial guess in, solution out (device,
 * nb_dof doubles).  iters and rel_res receive the iteration count and the
 * final relative residual; reaching maxit is not an error (compare rel_res
 * with rtol, as gmm::iteration::converged). */
int gfx_plan_solve_cg(gfx_plan *p, const double *d_b, double *d_x, double rtol, int maxit,
                      int *iters, double *rel_res);
/* The same solve with a choice of preconditioner: GFX_PRECOND_BLOCK_JACOBI
 * (gfx_plan_solve_cg) or GFX_PRECOND_ILDLT, GetFEM's default for coercive
 * problems: gmm::ildlt_precond (src/gmm/gmm_precond_ildlt.h:136-206), the
 * level-0 incomplete LDL^T on the pattern of K's upper triangle, rebuilt from
 * the current values at every solve as linear_solver_cg_preconditioned_ildlt
 * does (src/getfem/getfem_model_solvers.h:90-98), inside gmm::cg
 * (src/gmm/gmm_solver_cg.h:84-118).  The factor and its triangular solves
 * equal the reference's sequential algorithm bit for bit (same operations,
 * same order, no fused multiply-add); rows are scheduled in level sets of
 * node-sized units.  Needs a symmetric pattern with every diagonal entry
 * (GFX_E_UNSUPPORTED otherwise) and rows of at most 1489 entries; a pivot the
 * reference would replace by 1 ("pivot too small") returns GFX_E_UNSUPPORTED. */
#define GFX_PRECOND_BLOCK_JACOBI 0
#define GFX_PRECOND_ILDLT 1
#define GFX_PRECOND_ILU 2        /* gmm::ilu_precond (GMRES) */
#define GFX_PRECOND_DIAGONAL 3   /* the inverse diagonal (GMRES) */
int gfx_plan_solve_pcg(gfx_plan *p, int precond, const double *d_b, double *d_x, double rtol,
                       int maxit, int *iters, double *rel_res);
/* ILDLT pieces (for inspection and tests): factor the current values (level
 * counts out, either may be NULL); the factor aligned with the CSR (nnz
 * doubles: D_i on the diagonal, U_ij / D_i at (i, j) and at (j, i)); and
 * z = (L D L^T)^{-1} r on the device (stream-ordered). */
int gfx_plan_ildlt_factor(gfx_plan *p, int64_t *n_levels_forward, int64_t *n_levels_backward);
int gfx_plan_ildlt_values(gfx_plan *p, double *h_F);
int gfx_plan_ildlt_apply(gfx_plan *p, const double *d_r, double *d_z);
/* Dof constraints as model::assembly applies them after the bricks
 * ("Post simplification for dof constraints", src/getfem_models.cc:
 * 2806-2870; added by add_Dirichlet_condition_with_simplification, :4864-4940
 * and :5019-5030).  h_dofs[k] (k < n_fixed) gets the prescribed value h_go[k].
 * GFX_BUILD_RHS with d_rhs: linear model -> if GFX_MODEL_SYMMETRIC, d_rhs -=
 * K(:, fixed) go with the assembled K, then d_rhs[fixed] = go; nonlinear
 * model -> d_rhs[fixed] = go - d_pr[fixed] (d_pr: the current values, device,
 * NULL = 0).  GFX_BUILD_MATRIX: rows of the fixed dofs cleared (and their
 * columns when GFX_MODEL_SYMMETRIC), diagonal set to 1.  Cleared entries stay
 * in the pattern as explicit zeros (gmm's rsvector erases them). */
#define GFX_BUILD_MATRIX 1
#define GFX_BUILD_RHS 2
#define GFX_MODEL_LINEAR 4
#define GFX_MODEL_SYMMETRIC 8
int gfx_plan_dof_constraints(gfx_plan *p, int64_t n_fixed, const int32_t *h_dofs,
                             const double *h_go, const double *d_pr, double *d_rhs, int flags);
/* record HIP events around the phases of every gfx_plan_assemble call (on
 * the context's stream, no host synchronisation); enabling resets the record.
 * gfx_plan_last_timings: durations (ms) of the last call, geometry kernel and
 * contraction + scatter kernel; gfx_plan_timing_summary: number of recorded
 * calls and summed durations, then resets the record. */
int gfx_plan_set_timing(gfx_plan *p, int enable);
int gfx_plan_last_timings(gfx_plan *p, float *ms_geometry, float *ms_assemble);
int gfx_plan_timing_summary(gfx_plan *p, int *n_calls, double *ms_geometry,
                            double *ms_assemble);
/* the recorded calls one by one (at most max_calls): durations of each
 * call's geometry and contraction + scatter phases; resets the record */
int gfx_plan_timing_calls(gfx_plan *p, int max_calls, int *n_calls, float *ms_geometry,
                          float *ms_assemble);
int gfx_plan_destroy(gfx_plan *p);

/* ======================= workspace (ga_workspace) =======
 * The drop-in for getfem::ga_workspace (src/getfem/getfem_generic_assembly.h:
 * 262-597) on the supported GWFL subset.  Names follow GWFL rules; a fem
 * variable is differentiated, constants are not.  Expressions of derivation
 * order 2 are summed (ga_workspace::add_tree merges trees with the same test
 * functions, src/getfem_generic_assembly_workspace.cc:474-495), compiled by
 * gfx_ws_compile into kernel parameters and assembled by gfx_ws_assembly(2)
 * into the plan returned by gfx_ws_plan (CSR on the device; export with the
 * gfx_plan_* functions).  Values of variables/constants are copied. */
typedef struct gfx_workspace gfx_workspace;
int gfx_ws_create(gfx_ctx *ctx, gfx_workspace **out);
int gfx_ws_destroy(gfx_workspace *w);
/* ga_workspace::add_fem_variable(name, mf, I, V) (.h:453-455): V has
 * mf's nb_dof values (NULL = zeros); first = I.first() */
int gfx_ws_add_fem_variable(gfx_workspace *w, const char *name, gfx_mesh_fem *mf,
                            int64_t first, const double *h_V);
/* ga_workspace::add_fem_constant(name, mf, V) (.h:465-466) */
int gfx_ws_add_fem_constant(gfx_workspace *w, const char *name, gfx_mesh_fem *mf,
                            const double *h_V);
/* ga_workspace::add_fixed_size_constant(name, V) (.h:467-468) */
int gfx_ws_add_fixed_size_constant(gfx_workspace *w, const char *name, const double *h_V,
                                   int64_t n);
/* new values for a registered name (same size).  The plans keep their
 * patterns unless the new values change GetFEM's threshold-filtered pattern
 * (re-derived at the next assembly, see gfx_plan_pin_pattern). */
int gfx_ws_set_values(gfx_workspace *w, const char *name, const double *h_V);
/* new values of the fem variable from a device vector (stream-ordered, no
 * host synchronisation; compiles the workspace first if needed) */
int gfx_ws_set_values_device(gfx_workspace *w, const char *name, const double *d_V);
/* ga_workspace::add_expression(expr, mim, region = all, der_order) (.h:425-428) */
int gfx_ws_add_expression(gfx_workspace *w, const char *expr, gfx_mesh_im *mim, int der_order);
int gfx_ws_set_form(gfx_workspace *w, int form, int precision);
/* the compile stage alone (GFX_E_UNSUPPORTED with a message outside the subset) */
int gfx_ws_compile(gfx_workspace *w);
/* ga_workspace::assembly(order) (.h:564).  order 0: the potential, the sum
 * of the expressions of test-function order 0 (read it with
 * gfx_ws_assembled_potential).  Orders 0, 1 and 2 outside the matrix kernels'
 * subset take the polynomial route: every integrand that is a polynomial in
 * the values and gradients of fem fields (variables and fem constants on
 * affine tetrahedra or Q1 hexahedra, P0 included), fixed-size constants, pow(x, k), Det and
 * Inv(Id + Grad_w) at w = 0 (the pulled-back coordinates), compiled by
 * symbolic derivation into per-point programs run by one kernel; the order-1
 * vector then holds the derivative with respect to every fem variable at its
 * interval, and the order-2 matrix (nvec x nvec, nvec = the end of the last
 * variable's interval) every block test variable x trial variable at their
 * intervals, each element matrix thresholded at 1e-14 max|elem| and the
 * entries summed in element order (ga_instruction_matrix_assembly_mf_mf and
 * add_elem_matrix, src/getfem_generic_assembly_compile_and_exec.cc:
 * 5074-5189, :4853-4936).  order 2: the matrix (for the
 * Neo-Hookean term the pattern is rebuilt when the state or the coefficients
 * changed, since GetFEM's stored pattern depends on the values).  order 1: the
 * vector of the expression's first derivative at the unknown's current value
 * (Neo-Hookean: the residual; linear terms: K u); read it with
 * gfx_ws_assembled_vector.  Expressions must have been added with a
 * derivation order >= order. */
int gfx_ws_assembly(gfx_workspace *w, int order);
/* ga_workspace::assembled_vector() (.h:389-390): gfx_ws_vector_size doubles
 * to the host */
int gfx_ws_assembled_vector(gfx_workspace *w, double *h_V);
/* size of the last assembled vector: the unknown's nb_dof on the matrix
 * kernels' route, max(first + nb_dof) over the fem variables on the order-0/1
 * route (GetFEM's nb_prim_dof) */
int gfx_ws_vector_size(gfx_workspace *w, int64_t *n);
/* ga_workspace::assembled_potential() (.h:394-401): the value of the last
 * assembly(0) */
int gfx_ws_assembled_potential(gfx_workspace *w, double *J);
/* the same vector into a caller-owned device buffer (stream-ordered) */
int gfx_ws_copy_assembled_vector_device(gfx_workspace *w, double *d_V);
/* the plan of a workspace with one tree (GFX_E_STATE with several) */
int gfx_ws_plan(gfx_workspace *w, gfx_plan **out);
/* number of order-2 trees: one per mesh_im with order-2 expressions (GetFEM
 * merges the expressions of one (mesh_im, region, test functions) into one
 * tree, src/getfem_generic_assembly_workspace.cc:474-495) */
int gfx_ws_nb_trees(gfx_workspace *w);
/* The workspace's assembled matrix (CSR, rows = Test_u): the single tree's
 * plan CSR, or with several trees the workspace's own matrix, whose pattern
 * is the union of the trees' patterns and whose values are their sum (ga_exec
 * adds every tree into K).  Device pointers stay valid until the next
 * assembly that changes a pattern. */
int gfx_ws_device_csr(gfx_workspace *w, int64_t *nrows, int64_t *nnz, const int64_t **d_indptr,
                      const int32_t **d_indices, double **d_vals);
int gfx_ws_matrix_dims(gfx_workspace *w, int64_t *nrows, int64_t *nnz);
int gfx_ws_export_csr(gfx_workspace *w, int64_t *h_indptr, int32_t *h_indices, double *h_vals);
/* GetFEM orientation (CSC over trial dofs, uint32 rows) */
int gfx_ws_export_csc(gfx_workspace *w, int64_t *h_colptr, uint32_t *h_rows, double *h_vals);
/* ga_workspace::set_assembled_matrix (src/getfem/getfem_generic_assembly.h:
 * 378-381): assembly(2) then adds every tree into this caller-owned device
 * matrix (GFX_CSR or GFX_CSC, its pattern holding the trees' patterns) and
 * does not clear it, as GetFEM leaves a matrix it does not own uncleared
 * (src/getfem_generic_assembly_workspace.cc:804-807).  d_ptr NULL: back to
 * the workspace's own matrix. */
int gfx_ws_set_assembled_matrix(gfx_workspace *w, int orientation, int64_t n, const int64_t *d_ptr,
                                const int32_t *d_ind, double *d_vals);
/* normal form of the compiled term, e.g. "isotropic elasticity unknown=u ..." */
int gfx_ws_describe(gfx_workspace *w, char *buf, int64_t len);
/* CG on the workspace's assembled matrix (the single tree's plan CSR, the
 * union of several trees, or the polynomial route's matrix over all fem
 * variables), diagonal preconditioner, stopping test and arguments as
 * gfx_plan_solve_cg; what model::solve's default_linear_solver hands a
 * coercive system of several variables (src/getfem/getfem_model_solvers.h:
 * 652-655) */
int gfx_ws_solve_cg(gfx_workspace *w, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                    double *rel_res);
/* Restarted GMRES on the workspace's assembled matrix, after gmm::gmres
 * (src/gmm/gmm_solver_gmres.h:85-157), the solver default_linear_solver gives
 * a model that is not coercive (linear_solver_gmres_preconditioned_ilu,
 * src/getfem/getfem_model_solvers.h:101-109, :652-660; restart 500 there).
 * Left preconditioned by GFX_PRECOND_ILU, GetFEM's gmm::ilu_precond of its
 * column-major matrix (src/gmm/gmm_precond_ilu.h: the rows of K^T
 * factored, bit for bit, pivots GetFEM replaces by 1 replaced alike; the
 * pattern must be symmetric with every diagonal entry and rows of at most
 * 1000 entries), or GFX_PRECOND_DIAGONAL, the inverse diagonal.  The
 * stopping test is ||M (b - K x)|| <= rtol ||M b||; ten restarts without
 * progress end it; maxit bounds the total iterations.  The Krylov basis holds
 * (restart + 1) nb_dof doubles, the restart lowered to keep it within 4 GiB.
 * d_x: initial guess in, solution out; iters and rel_res as
 * gfx_plan_solve_cg. */
int gfx_ws_solve_gmres(gfx_workspace *w, int precond, const double *d_b, double *d_x, double rtol, int maxit,
                       int restart, int *iters, double *rel_res);
/* The ILU of gfx_ws_solve_gmres on its own: factor the workspace's current
 * matrix (replaced: pivots GetFEM replaces by 1, may be NULL), then z =
 * M^-1 r on the device (stream-ordered); an order-2 assembly invalidates the
 * factor */
int gfx_ws_ilu_factor(gfx_workspace *w, int64_t *replaced);
int gfx_ws_ilu_apply(gfx_workspace *w, const double *d_r, double *d_z);
/* gfx_plan_dof_constraints on the workspace's assembled matrix (dofs are
 * global: the variables' intervals) */
int gfx_ws_dof_constraints(gfx_workspace *w, int64_t n_fixed, const int32_t *h_dofs, const double *h_go,
                           const double *d_pr, double *d_rhs, int flags);

/* ======================= adjoint (SIMP compliance) =====
 * J = u^T K(rho, X) u with K the isotropic elasticity matrix of a P1 vector
 * mesh_fem, E_e = emin + rho_e^penal (1 - emin) scaling (lambda, mu).  The
 * gradients are what GetFEM's order-1 assembly of the order-0 potential
 * gives when rho and the (pulled-back) coordinates are variables
 * (src/getfem_generic_assembly_workspace.cc:520-541, :832-847).
 * All arrays are device pointers on the context's device, stream-ordered on
 * its stream; d_X (n_pts*3) may be NULL for the mesh's own points; any
 * output but d_J may be NULL. */
typedef struct gfx_compliance gfx_compliance;
int gfx_compliance_create(gfx_mesh_fem *mf, gfx_mesh_im *mim, double lambda, double mu,
                          double emin, double penal, gfx_compliance **out);
/* kernels only on the context's stream (no allocation, copy, memset or host
 * synchronisation): a call may be captured into a HIP graph and replayed */
int gfx_compliance_eval(gfx_compliance *c, const double *d_X, const double *d_rho,
                        const double *d_u, double *d_J, double *d_dJ_drho,
                        double *d_dJ_dX, double *d_dJ_du);
/* gfx_compliance_eval without the object's scratch buffers: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the J partials and the per-element node contributions live in one static
 * device array of 64 MiB per GPU, allocated when the library's code is
 * loaded on that device (its address is looked up by gfx_ctx_create).  The
 * call allocates nothing; inputs and outputs are the caller's, as for
 * gfx_compliance_eval, with the same results.  Limit: 8 * (2 * ceil(n_cv /
 * 256) rounded up to 32) + 96 * n_cv bytes for each requested node gradient
 * (d_dJ_dX, d_dJ_du) must fit in 64 MiB (about 349,000 elements with both);
 * larger meshes return GFX_E_INVALID.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (two streams, two threads) race
 * and must be ordered by the caller. */
int gfx_compliance_eval_noalloc(gfx_compliance *c, const double *d_X, const double *d_rho,
                                const double *d_u, double *d_J, double *d_dJ_drho,
                                double *d_dJ_dX, double *d_dJ_du);
/* GFX_E_DEGENERATE_ELEMENT if an evaluation since the last call met an
 * element with det K = 0 (GetFEM: "Non invertible matrix",
 * bgeot_geometric_trans.cc:380); synchronises the context's stream and clears
 * the flag.  gfx_compliance_eval itself never synchronises. */
int gfx_compliance_status(gfx_compliance *c);
int gfx_compliance_destroy(gfx_compliance *c);

#ifdef __cplusplus
}
#endif
#endif
