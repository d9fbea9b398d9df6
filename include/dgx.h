/*
 * dgx.h - C-ABI of the MI355X explicit DG residual path (volume + face residual, inverse mass,
 * SSP-RK3 stage with positivity enforcement) for the Euler equations on QUAD/HEX meshes.
 *
 * Each entry point replaces one reference entry point on the explicit path; the citation is given at
 * each declaration.  Conventions that differ from the reference on purpose:
 *   - every function returns a dgx_status and never aborts (the reference calls abort() via EXIT_*);
 *   - there is no global state: all device memory and the HIP stream belong to a dgx_ctx
 *     (one context per host thread at a time);
 *   - host arrays are caller-owned and copied; nothing here keeps a pointer to caller memory.
 *
 * Host state layout DGX_LAYOUT_REF is the reference's per-volume sol_coef: for element e the
 * (Ns, nv) column-major block, i.e. u[e][v][n] (src/simulation/solvers/volume_solver_T.h:31-104).
 * Node n = i + (p+1)*(j + (p+1)*k) (r fastest), variables (rho, rho u, rho v, [rho w], E).
 */
#ifndef DGX_H
#define DGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dgx_status {
    DGX_OK = 0,
    DGX_ERR_INVALID_ARG = 1,      /* bad pointer, size or unsupported option */
    DGX_ERR_HIP = 2,              /* a HIP runtime call failed (message in dgx_last_error) */
    DGX_ERR_POSITIVITY = 3,       /* element-average density or pressure < EPS_PHYS (solve.c:204-208) */
    DGX_ERR_NONFINITE = 4,        /* NaN/Inf in the residual */
    DGX_ERR_GEOMETRY = 5,         /* det J <= 0 (geometry_T.c:537-543) */
    DGX_ERR_NO_DEVICE = 6         /* no GPU or the device code is missing */
} dgx_status;

enum { DGX_PDE_EULER = 0, DGX_PDE_NAVIER_STOKES = 1 };   /* test_case.c pde_index: euler, navier_stokes */
enum { DGX_FLUX_LAX_FRIEDRICHS = 0, DGX_FLUX_ROE_PIKE = 1 };   /* numerical_flux_euler_T.c:60-327, 470-778 */
/* DGX_PREC_FP64: the plain form.  DGX_PREC_MIXED: the residual's operator contractions (face traces, weak
 * derivative, and for Navier-Stokes the volume gradient) contract, along every line, the FP64 differences
 * from the line's first node as FP32 hi/lo parts with FP32 accumulation (the line's constant part in
 * FP64); the face lift and all else stay FP64 (Euler HEX p=3 on aligned meshes and Navier-Stokes HEX p=4;
 * dgx_create and dgx_rhs return DGX_ERR_INVALID_ARG elsewhere).  No reference counterpart: the reference
 * is FP64. */
enum { DGX_PREC_FP64 = 0, DGX_PREC_MIXED = 1 };
enum { DGX_LAYOUT_REF = 0, DGX_LAYOUT_SOA = 1 };               /* SOA = device layout u[v][e][n] */
/* Explicit time stepper, the reference's test_case->solver_type_e (solve_explicit.c:150-170):
 * SSP-RK3 (3 stages, time_step_ssp_rk_33 :252-301), forward Euler (1 stage, time_step_euler :226-250),
 * low-storage RK 5-stage 4th order (5 stages, time_step_ls_rk_54 :303-350). */
enum { DGX_SOLVER_E_SSP_RK_33 = 0, DGX_SOLVER_E_EULER = 1, DGX_SOLVER_E_LS_RK_54 = 2 };
/* Node families of the 1-D reference line (src/element/nodes.c:70-130: GL = "GLe", GLL = "GLoL"). */
enum { DGX_NODES_GL = 0, DGX_NODES_GLL = 1 };
/* Ordering of tensor-product nodes: node n = i + (p+1)*(j + (p+1)*k), r fastest (element_operators_tp.c:90-241). */
enum { DGX_KRON_R_FASTEST = 0 };
enum { DGX_PTR_HOST = 0, DGX_PTR_DEVICE = 1 };

typedef struct dgx_config {
    int32_t d;            /* 2 (QUAD) or 3 (HEX) */
    int32_t p;            /* solution order; GL solution nodes = GL cubature, p+1 per direction */
    int32_t pde;          /* DGX_PDE_EULER or DGX_PDE_NAVIER_STOKES */
    int32_t num_flux;     /* DGX_FLUX_* */
    int32_t positivity;   /* 1: enforce_positivity_highorder after every RK stage (solve.c:187-233) */
    int32_t precision;    /* DGX_PREC_* used by dgx_step / dgx_rk_stage */
    /* Navier-Stokes only (ignored for Euler): constant dynamic viscosity "mu" and Prandtl number "Pr"
     * of the reference's input file (solution_navier_stokes_T.c:221-240, flux_navier_stokes_T.c:461-483);
     * the viscous flux is BR2-stable with the central numerical solution and flux, penalty 1.01 x faces
     * per element (compute_grad_coef_dg_T.c, compute_face_rlhs_dg_T.c:276-296) */
    double mu;
    double Pr;
    int32_t solver_type_e;   /* DGX_SOLVER_E_*: stages of dgx_rk_stage / dgx_step (0 = SSP-RK3) */
    int32_t reserved;        /* 0 */
} dgx_config;

typedef struct dgx_mesh {
    int32_t d;
    int32_t p_g;                 /* geometry order; GLL geometry nodes (p_g+1)^d per element */
    int64_t n_elem;
    int64_t n_face;
    const double*  xyz_g;        /* [n_elem][d][(p_g+1)^d] (reference geom_coef, Lagrange at GLL) */
    const int32_t* face_elem;    /* [n_face][2]  neigh_info[0/1].volume (face.h:47-95) */
    const int8_t*  face_lf;      /* [n_face][2]  neigh_info[0/1].ind_lf */
    const int8_t*  face_ord;     /* [n_face]     neigh_info[1].ind_ord; only 0 (aligned) is supported */
} dgx_mesh;

typedef struct dgx_geometry {
    const double* metrics_vc;    /* [n_elem][d][d][Nc]  m(n,i,j) = J dxi_j/dx_i  (metrics_vc) */
    const double* jdet_vc;       /* [n_elem][Nc]        (jacobian_det_vc) */
    const double* normals_fc;    /* [n_face][Nf][d]     (normals_fc, from the left volume) */
    const double* jdet_fc;       /* [n_face][Nf]        (jacobian_det_fc) */
} dgx_geometry;

/* The 1-D operator set of a GL solution/cubature order p (collocated GL/GL):
 *   x[n], w[n]                       GL nodes and weights, n = p+1
 *   D[n*n]   D[i*n + a] = l_a'(x_i)  derivative of the solution basis at the cubature nodes
 *   lm[n], lp[n]                     l_a(-1), l_a(+1)  (face trace rows; cv0_vs_fc for a LINE)
 *   sb[n*n]  ccSB0 (nodal -> Bezier coefficients), row-major
 *   bs[n*n]  ccBS0 (Bezier -> nodal), row-major
 * Reference: the LINE sub-operators that element_operators_tp.c:90-241 Kronecker-multiplies into the
 * tensor-product operators cv0_vs_vc, cv1_vt_vc, cv0_vs_fc, w_vc, ccSB0_vs_vs, ccBS0_vs_vs. */
typedef struct dgx_ops1d {
    int32_t p;
    int32_t n;
    double x[8], w[8];
    double D[64];
    double lm[8], lp[8];
    double sb[64], bs[64];
} dgx_ops1d;

/* The operator set the kernels apply by sum factorization: the 1-D factors and the tag of the
 * Kronecker ordering that builds the element operators from them (the reference's Solver_Element
 * operators, element_solver.h:31-74, as their LINE factors). */
typedef struct dgx_operators {
    int32_t node_type;    /* DGX_NODES_GL (solution nodes = cubature nodes; GLL is not supported) */
    int32_t kron_order;   /* DGX_KRON_R_FASTEST */
    dgx_ops1d line;       /* 1-D factors of order cfg->p */
} dgx_operators;

typedef struct dgx_ctx dgx_ctx;

/* Create a context on the current HIP device: validates the inputs, copies mesh connectivity and
 * geometry to device memory in the SoA layouts the kernels read, and loads the operator set.
 * ops: the 1-D factors (nullable: then built for cfg->p by dgx_ops1d_build).  They must be the GL/GL
 * set of order cfg->p in r-fastest order, entry by entry within 1e-13 (relative to max(1, |entry|)) of
 * dgx_ops1d_build's, else DGX_ERR_INVALID_ARG: the kernels' diagonal inverse mass and collocated
 * cubature are exact for that set only.  The values passed are the ones applied.
 * Replaces the explicit-path part of constructor_derived_computational_elements(IL_SOLVER_DG)
 * (solve_explicit.c:88-89 -> dg/volume_solver_dg_T.c:49-78: sol_coef_p, m_inv), the operator set of
 * constructor_derived_Solver_Element (element_solver.c:179-300) and set_up_solver_geometry's consumers. */
dgx_status dgx_create(const dgx_config* cfg, const dgx_mesh* mesh, const dgx_geometry* geom,
                      const dgx_operators* ops, dgx_ctx** out);
dgx_status dgx_destroy(dgx_ctx* ctx);
const char* dgx_last_error(const dgx_ctx* ctx);
const char* dgx_status_string(dgx_status s);

/* Returns the hipStream_t every call of this context is ordered on. */
void* dgx_get_stream(dgx_ctx* ctx);

/* Sizes: n_elem * nv * Ns doubles for a state. */
int64_t dgx_state_size(const dgx_ctx* ctx);

/* Upload / download the state (sol_coef of every volume).  ptr_kind: DGX_PTR_HOST or DGX_PTR_DEVICE
 * (a device pointer must hold the same layout).  Uploading also refreshes the face traces. */
dgx_status dgx_set_state(dgx_ctx* ctx, int layout, int ptr_kind, const double* u);
dgx_status dgx_get_state(dgx_ctx* ctx, int layout, int ptr_kind, double* u);

/* Residual of the current state.  apply_minv = 1: rhs = M^-1 (volume + face) as compute_rhs
 * (solve.c:133-149 -> dg/solve_dg.c:130-136, 191-201); apply_minv = 0: the pre-M^-1 residual
 * (volume + face), the operator test_support_solve_dg.c:164-184 linearizes.
 * max_rhs (nullable) = max over volumes of ||rhs||_inf (compute_rlhs.c:59-71).  rhs (nullable) in
 * `layout` on the host or device.  precision: DGX_PREC_*. */
dgx_status dgx_rhs(dgx_ctx* ctx, int precision, int apply_minv, int layout, int ptr_kind, double* rhs,
                   double* max_rhs);

/* One stage (0 .. n-1; n = 3, 1 or 5 for cfg->solver_type_e) of time_step_ssp_rk_33 / time_step_euler /
 * time_step_ls_rk_54 (solve_explicit.c:226-350): residual, stage combination, then
 * enforce_positivity_highorder (solve.c:187-233) when enabled.  max_rhs (nullable) as for dgx_rhs, of
 * this stage's residual.  Stages run in the order 0, 1, .., n-1, as the reference's step functions run
 * them (another order returns DGX_ERR_INVALID_ARG; dgx_set_state restarts).
 * Positivity failure: the failing stage completes on every element (the failing element keeps its
 * un-limited stage value) and DGX_ERR_POSITIVITY is returned; every later stage is a no-op that leaves
 * the state as the failing stage wrote it, until dgx_set_state (the reference aborts, solve.c:204-208). */
dgx_status dgx_rk_stage(dgx_ctx* ctx, int stage, double dt, double* max_rhs);

/* dgx_stages_per_step: n of dgx_rk_stage (3, 1 or 5). */
int dgx_stages_per_step(const dgx_ctx* ctx);

/* nsteps full steps of the time stepper with fixed dt, launched as a captured hipGraph.  max_rhs
 * (nullable) of the last stage.  Replaces the stepping loop of solve_explicit (solve_explicit.c:78-118).
 * A positivity failure stops the computation at the failing stage (the stages after it do not change
 * the state; see dgx_rk_stage) and returns DGX_ERR_POSITIVITY. */
dgx_status dgx_step(dgx_ctx* ctx, double dt, int nsteps, double* max_rhs);

/* Wait for all work on the context's stream and report a sticky device-side error
 * (positivity failure with the first failing element in *elem, or a non-finite residual). */
dgx_status dgx_sync(dgx_ctx* ctx, int64_t* elem);

/* Residual linearization at the current state, on R = pre-M^-1 residual (apply_minv = 0) or M^-1 R
 * (apply_minv = 1).  in/out in `layout`, host or device pointers.
 *   dgx_vjp: out = J^T v   (hand-written adjoint kernels; cubature states are recomputed)
 *   dgx_jvp: out = J w     (forward-mode kernels)
 * Positivity enforcement is not part of R (test_support_solve_dg.c:164-184). */
dgx_status dgx_vjp(dgx_ctx* ctx, int apply_minv, int layout, int ptr_kind, const double* v, double* out);
dgx_status dgx_jvp(dgx_ctx* ctx, int apply_minv, int layout, int ptr_kind, const double* w, double* out);

/* Volume-diagonal blocks of the linearization at the current state (Euler): for every volume e and
 * unit direction j = v*Ns + n of that volume (neighbours' directions zero), blocks_t[e][j][i] =
 * d(R or M^-1 R)_i / d u_j with i = v'*Ns + n' (the transposed block, row-major; device pointer of
 * n_elem * (nv*Ns)^2 doubles, volumes in list order).  Computed on the device without probing; equals
 * the JVPs of the unit directions restricted to their volume.  Replaces the assembly of the volume-diagonal entries of the
 * reference's implicit matrix (compute_rlhs_dg with the lhs, dg/solve_dg.c:138-145) for a block
 * preconditioner. */
dgx_status dgx_jacobian_blocks(dgx_ctx* ctx, int apply_minv, double* blocks_t);
/* The same blocks rounded to FP32 (half the memory: 107 GB instead of 215 GB for 64^3 HEX p=3). */
dgx_status dgx_jacobian_blocks_f32(dgx_ctx* ctx, int apply_minv, float* blocks_t);

/* Batched block inversion for a block-Jacobi preconditioner: inv[b] = (shift[b] I - blocks[b])^-1 for
 * n_blocks dense n x n blocks (n <= 320), blocks row-major FP64, inv row-major FP32, all device
 * pointers; singular[b] = 0, k + 1 when column k had no non-zero pivot, or n + 1 when the block or its
 * inverse holds a non-finite value (a NaN input, or a denormal pivot whose reciprocal overflows): the
 * block's output is then meaningless.  Gauss-Jordan with partial pivoting, one workgroup per block, launched on `stream`
 * (a hipStream_t; NULL = the null stream) without synchronising.  Replaces the preconditioner set-up of
 * the reference's KSP (constructor_petsc_ksp, solve_implicit.c:562-590: ILU on the assembled matrix),
 * which a matrix-free solver cannot factor; the blocks come from dgx_jacobian_blocks. */
dgx_status dgx_block_inverse(int64_t n_blocks, int n, const double* blocks, const double* shift, float* inv,
                             int32_t* singular, void* stream);
/* The same from FP32 blocks; inv may equal blocks (inversion in place). */
dgx_status dgx_block_inverse_f32(int64_t n_blocks, int n, const float* blocks, const double* shift, float* inv,
                                 int32_t* singular, void* stream);

/* Block-diagonal product y[b] = A[b] x[b] for n_blocks FP32 n x n blocks (n <= 320) stored column-major
 * (the row-major inverse of dgx_block_inverse read as A^T, or its transpose copied row-major read as
 * A), FP64 x and y, device pointers, launched on `stream` without synchronising.  The preconditioner
 * application of the matrix-free Krylov solver; no reference counterpart (PETSc applies its ILU
 * factors inside KSPSolve, solve_implicit.c:404). */
dgx_status dgx_block_apply(int64_t n_blocks, int n, const float* blocks_colmajor, const double* x, double* y,
                           void* stream);
/* The same for the blocks sel[0..n_sel) only (block indices into blocks_colmajor, x and y; the other
 * entries of y are left untouched): one colour of a multicolour block Gauss-Seidel sweep. */
dgx_status dgx_block_apply_sel(int64_t n_sel, const int32_t* sel, int n, const float* blocks_colmajor,
                               const double* x, double* y, void* stream);

/* Time of the last dgx_step / dgx_rk_stage kernel work in milliseconds, from HIP events recorded on
 * the context stream around the launches (0 when unavailable). */
double dgx_last_elapsed_ms(const dgx_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
