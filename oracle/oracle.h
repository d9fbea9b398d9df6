/*
 * oracle.h - CPU restatement of DPGSolver's explicit DG residual path.  TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the GPU path is compared against.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg may load it; the product library never links or calls it.
 *
 * It restates the reference algorithm as the reference executes it, with dense per-element operators
 * (src/simulation/solvers/solve.c:133-149 -> dg/solve_dg.c:130-136, 182-201, 265-282):
 *   volume:   s_vc = cv0_vs_vc u; f = Euler flux; fr = metrics . f; rhs += sum_j tw1_vt_vc[j] fr_j
 *             (compute_volume_rlhs_T.c:145-177, 277-285; compute_rlhs_T.c:71-110; flux_euler_T.c:71-143)
 *   face:     s_l = cv0_vs_fc[lf_l] u_l, s_r = cv0_vs_fc[lf_r] u_r; nnf = LF | Roe-Pike; nnf *= |J_f|;
 *             rhs_l -= tw0_vt_fc[lf_l] nnf; rhs_r -= tw0_vt_fc[lf_r] (-nnf)
 *             (compute_face_rlhs_T.c:51-146, 297-359; boundary_T.c:42-92,134-145;
 *              numerical_flux_euler_T.c:60-327, 470-778)
 *   inverse mass: rhs = m_inv rhs, m = V^T diag(w J) V, m_inv by LU (volume_solver_T.c:115-148)
 *   SSP-RK3 + positivity (solve_explicit.c:252-301; solve.c:187-233, 356-374)
 * Operators are built like the reference's: orthonormal Legendre basis, a Vandermonde solve for the
 * Lagrange/Bezier bases, Kronecker products for QUAD/HEX (element_operators_tp.c:133-241).
 * Geometry is recomputed here from the geometry nodes the same way (geometry_T.c:208-325, 412-445,
 * 484-590, 653-659, 734-804; 3-D curl form deprecated/src/setup_geom_factors.c:182-250).
 *
 * Every function of this file is pinned by a check in tests/ (see each test's docstring): golden
 * vectors from the reference's unit-test data, closed forms, invariants and complex-step identities.
 */
#ifndef DGX_ORACLE_H
#define DGX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* 1-D nodes by Golub-Welsch (eigenvalues of the Jacobi matrix): family 0 = GL, 1 = GLL. */
int orc_nodes_1d(int family, int p, double* x, double* w);

/* Orthonormal-basis route to the Lagrange operators: out = Phi^(der)(y) Phi(xs)^-1 (ny x ns, row-major). */
int orc_lagrange_1d(int ns, const double* xs, int ny, const double* y, int der, double* out);

/* Dense Kronecker product of d row-major factors: op[(i0 + m0 (i1 + m1 i2)), (j0 + n0 (j1 + n1 j2))]. */
int orc_kron(int d, const int* m, const int* n, const double* const* A, double* op);

/* General dense matrix product C = A B (row-major), the reference's mm_d (matrix_math_T.c:178-201). */
void orc_mm(int m, int k, int n, const double* A, const double* B, double* C);

/* Context: mesh as in dgx_mesh, solution order p, flux 0 = LF / 1 = Roe-Pike, positivity 0/1.
 * xyz_g may be NULL: the geometry is then left unset (zero) until orc_set_geometry provides it. */
orc_ctx* orc_create(int d, int p, int p_g, int64_t n_elem, int64_t n_face, const double* xyz_g,
                    const int32_t* face_elem, const int8_t* face_lf, int num_flux, int positivity);
void orc_destroy(orc_ctx* c);

/* Geometry factors computed by the oracle, in the dgx_geometry layouts. */
void orc_geometry(const orc_ctx* c, double* metrics_vc, double* jdet_vc, double* normals_fc, double* jdet_fc);

/* Switch the residual to Navier-Stokes: constant viscosity mu, Prandtl number Pr, BR2-stable viscous
 * flux with the central numerical solution and flux (penalty 1.01 x faces per element). */
void orc_set_viscous(orc_ctx* c, double mu, double Pr);

/* Test hook: BR2 penalty used instead of 1.01 x NFMAX (NFMAX = 4 for QUAD, 6 for HEX). */
void orc_set_penalty(orc_ctx* c, double eta);
/* Threads of orc_residual / orc_rk_stage (OpenMP; default 1).  With more than one thread the face terms
 * are accumulated per volume in local-face order instead of the reference's face order (round-off
 * differences only); used to time the baseline on all host cores. */
void orc_set_threads(orc_ctx* c, int nthreads);
/* Sum-factorized application of the element operators (on != 0): every operator the residual applies is
 * a Kronecker product of 1-D factors and is applied by one mode product per direction, M^-1 as the
 * diagonal 1/(w J) of the Gauss-Legendre collocation, the positivity's Bezier maps likewise; the same
 * operators as the dense path up to summation order.  A CPU baseline stronger than the reference's dense
 * per-element operators.  Returns -1 (and leaves the dense path) when the basis is not collocated. */
int orc_set_sum_factorized(orc_ctx* c, int on);

/* Weak gradients of the Navier-Stokes path at state u, each [E][d][nv][Ns]: grad_v (volume term) and
 * grad (volume term plus the liftings of all faces of the element). */
void orc_gradient(orc_ctx* c, const double* u, double* grad_v, double* grad);

/* Replace the oracle's geometry by given factors (same layouts) and rebuild the mass matrices: the
 * residual path is then compared on exactly the geometry the GPU path consumes. */
void orc_set_geometry(orc_ctx* c, const double* metrics_vc, const double* jdet_vc, const double* normals_fc,
                      const double* jdet_fc);

/* Dense element operators (for fixture checks): name = "cv0_vs_vc", "tw1_vt_vc0".., "cv0_vs_fc0"..,
 * "tw0_vt_fc0"..,"ccSB0", "ccBS0", "w_vc"; returns the number of values written (or -1). */
int orc_operator(const orc_ctx* c, const char* name, double* out);

/* rhs = [M^-1] (volume + face) of u (layout u[e][v][n]); returns max_e ||rhs_e||_inf. */
double orc_residual(orc_ctx* c, const double* u, double* rhs, int apply_minv);

/* Complex-step directional derivative of the residual: out = Im R(u + i h w) / h. */
void orc_jvp_cs(orc_ctx* c, const double* u, const double* w, double h, double* out, int apply_minv);

/* Apply enforce_positivity_highorder to every element of u in place; returns -1 on success or the
 * first element whose average density/pressure is below EPS_PHYS. */
int64_t orc_positivity(orc_ctx* c, double* u);

/* Face node correspondence nc_fc[ord] (src/element/nodes_correspondence.c): d_face = 1 (LINE, n1 nodes,
 * ord 0..1) or 2 (QUAD, n1 x n1 nodes, ord 0..7); out[q] = the right-volume face node seen at left face
 * node q.  Returns 0, or 1 for an unsupported case. */
int orc_face_corr(int d_face, int n1, int ord, int32_t* out);
/* neigh_info[1].ind_ord of every face (NULL: all aligned): the right trace and the flux lifted into the
 * right volume are gathered through nc_fc[ind_ord] as the reference does.  Euler only (returns 2 for a
 * Navier-Stokes context); 1 for an out-of-range order. */
int orc_set_face_ord(orc_ctx* c, const int8_t* face_ord);

/* Explicit time stepper of orc_rk_stage (test_case->solver_type_e): 0 SSP-RK3 (default), 1 forward
 * Euler, 2 low-storage RK 5-stage 4th order (solve_explicit.c:226-350).  Returns 0, or 1 for a bad value. */
int orc_set_time_stepper(orc_ctx* c, int time_stepper);
/* Stages per step of the current time stepper (3, 1 or 5). */
int orc_n_stages(const orc_ctx* c);

/* One stage of the time stepper on (u, u_p) in place (u_p: sol_coef_p, the RK register); returns
 * max_rhs of the stage's residual.  *bad_elem is set to the first positivity failure (or -1). */
double orc_rk_stage(orc_ctx* c, int stage, double dt, double* u, double* u_p, int64_t* bad_elem);

/* Physical flux / numerical flux at one point (d = ctx d): used by the flux unit tests.
 * f[dim + d*eq] = F_dim(eq); nf[eq]. */
void orc_flux_point(int d, const double* u, double* f);
void orc_numflux_point(int d, int num_flux, const double* ul, const double* ur, const double* n, double* nf);
/* batched forms: u [n][nv] -> f [n][nv][d]; ul, ur [n][nv], nrm [n][d] -> nf [n][nv] */
void orc_flux_batch(int d, int64_t n, const double* u, double* f);
void orc_numflux_batch(int d, int num_flux, int64_t n, const double* ul, const double* ur, const double* nrm,
                       double* nf);

#ifdef __cplusplus
}
#endif
#endif
