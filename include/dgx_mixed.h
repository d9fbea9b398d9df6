/*
 * dgx_mixed.h - C-ABI of the batching form: the explicit DG residual and SSP-RK3 stage on a mixed
 * HEX / WEDGE / TET / PYR mesh with per-element orders, on one GPU.
 *
 * The reference evaluates the same DG residual for every element type through per-type operators
 * (compute_volume_rlhs_T.c:145-177, compute_face_rlhs_T.c:166-170,297-359, inverse mass per volume
 * solve_dg.c:265-272) and a face order max(p_L, p_R) (compute_face_reference_order, face_solver.c:
 * 123-132).  Here the elements are grouped by (type, p) and every group's operators are dense tables
 * uploaded once; a stage is two kernels (face numerical fluxes, then volume + lift + M^-1 + RK update)
 * launched either as one persistent grid that takes (type, p) tiles from a work counter, or as one
 * launch per group.
 *
 * Conventions (as dgx.h): every function returns a dgx_status and never aborts; all device memory and
 * the stream belong to the context; host arrays are copied, never kept.
 *
 * Element basis: an orthonormal basis of the element's polynomial space on the reference element, so
 * that the mass matrix of an affine element is det(J) times the identity; the state of element e is
 * its modal coefficients [var][mode] (the reference's sol_coef (Ns, nv) column-major block,
 * volume_solver_T.h:31-104), elements in group order, variables (rho, rho u, rho v, rho w, E).
 */
#ifndef DGX_MIXED_H
#define DGX_MIXED_H

#include <stdint.h>

#include "dgx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { DGX_ELEM_HEX = 0, DGX_ELEM_WEDGE = 1, DGX_ELEM_TET = 2, DGX_ELEM_PYR = 3 };
/* Stage launch form: one persistent grid over all (type, p) tiles (work counter), one launch per
 * (type, p) group, or one grouped launch with a workgroup per tile of every group. */
enum { DGX_MIXED_LAUNCH_PERSISTENT = 0, DGX_MIXED_LAUNCH_PER_GROUP = 1, DGX_MIXED_LAUNCH_GROUPED = 2 };

/* Mesh, geometry and operator tables of a straight-sided (affine) mixed mesh.  Elements must be sorted
 * by group: group g holds elements group_start[g] .. group_start[g+1]-1. */
typedef struct dgx_mixed_mesh {
    int64_t n_elem;
    int64_t n_face;
    int32_t n_group;
    int32_t reserved;
    const int32_t* group_type;   /* [n_group] DGX_ELEM_* */
    const int32_t* group_p;      /* [n_group] order p (1..5) */
    const int32_t* group_ns;     /* [n_group] modes Ns */
    const int32_t* group_nc;     /* [n_group] volume cubature nodes Nc */
    const int64_t* group_start;  /* [n_group+1] */
    /* volume operators of group g at vol_off[g]: B [Nc][Ns] (basis at the cubature nodes), then
     * G [3][Nc][Ns] (reference-gradient of the basis times the cubature weight: w_q d phi_j/d xi_d) -
     * the reference's cv0_vs_vc and tw1_vt_vc (element_operators.c) */
    const double* vol_ops;
    const int64_t* vol_off;      /* [n_group] */
    /* nullable; offset in vol_ops of a group's factor tables, -1 for the dense path.  The kernels then
     * apply the volume operators by sum factorization (element_operators_tp.c:90-241 builds the
     * reference's tensor-product operators from the same LINE / TRI factors).
     * HEX: B1 [n][n] (orthonormal Legendre at the n = p+1 Gauss-Legendre nodes), D1 [n][n] (their
     * derivatives), w1 [n] (weights), tm, tp [n] (L_i(-1), L_i(+1)); B = B1 (x) B1 (x) B1 (r fastest), G_d
     * the same with D1 in direction d.
     * WEDGE: B1, D1, w1, then Bt, Gr, Gs [n^2][Nt] (orthonormal TRI basis at the collapsed TRI nodes and its
     * r, s derivatives, Nt = n(n+1)/2), wt [n^2], E [3][n][Nt] (the TRI modes on the edge under quad face e
     * in the edge's Legendre basis), tm, tp [n]; B = B1 (x) Bt (TRI fastest).
     * TET, PYR (collapsed coordinates a, b, c): A, Ad [n][n]; B, Bd (TET [n][n][n] as [i][q][j], PYR
     * [n][n]); C, Cd (TET [Ns][n]: row = mode (i, j, k) in the mode order, i slowest, k fastest, then r -
     * only the (i, j, k) entries of its modes; PYR [n][n][n] as [max(i,j)][r][k]); the nodes
     * xa, xb, xc [n]; the node weights [n^3]; the mode scales s [Ns]: phi = s A B C. */
    const int64_t* vol_tp_off;
    const double* jdet;          /* [n_elem] det of the affine map */
    const double* metrics;       /* [n_elem][3][3] m(i,j) = J d xi_j / d x_i (metrics_vc, constant) */
    /* faces: left volume face_elem[f][0], right face_elem[f][1]; outward unit normal of the left volume;
     * face Jacobian = physical face area / reference face area */
    const int32_t* face_elem;    /* [n_face][2] */
    const int32_t* face_lf;      /* [n_face][2] local face of each side (neigh_info[s].ind_lf); nullable
                                  * unless HEX or WEDGE groups are sum-factorized */
    const int32_t* face_op;      /* [n_face][2] face operator of each side */
    const int32_t* face_wt;      /* [n_face]    face weight table */
    const double* face_normal;   /* [n_face][3] */
    const double* face_jac;      /* [n_face] */
    /* face operators: basis of the side's volume at the face cubature nodes, as listed by the left
     * volume (the right side's node correspondence folded in): B_f [Nq][Ns] at fop_off[o] (cv0_vs_fc) */
    int32_t n_face_op;
    int32_t n_face_wt;
    const int32_t* fop_nq;       /* [n_face_op] */
    const int32_t* fop_ns;       /* [n_face_op] */
    const int64_t* fop_off;      /* [n_face_op] */
    /* nullable; offset in fop_data of a side's face-modal rows BF, -1 if none (required for the sides of
     * sum-factorized HEX and WEDGE groups: their lifts and traces contract the face-modal form).
     * HEX: BF [Nq][n^2] with B_f[q][(i, a, b)] = tau_i BF[q][a + n b], tau = L_i(-1) or L_i(+1) along the
     * face's normal direction and (a, b) the other two directions, lower one fastest.  WEDGE quad face
     * lf (0..2, over TRI edge lf): BF [Nq][n^2] with B_f[q][(k, m)] = sum_a BF[q][a + n k] E_lf[a][m];
     * WEDGE TRI face (lf 3, 4: t = -1, +1): BF [Nq][Nt] with B_f[q][(k, m)] = tau_k BF[q][m]. */
    const int64_t* fop_red_off;
    const int32_t* fop_red_n;    /* [n_face_op] columns of the face-modal rows (nullable with fop_red_off) */
    const double* fop_data;
    const int32_t* fw_n;         /* [n_face_wt] face cubature nodes */
    const int64_t* fw_off;       /* [n_face_wt] */
    const double* fw_data;       /* face cubature weights */
} dgx_mixed_mesh;

typedef struct dgx_mixed_ctx dgx_mixed_ctx;

/* Create a context on the current HIP device: validates the tables (group order, sizes, face sides,
 * every element's faces matching its type), builds the element -> face lists and the tile lists, and
 * copies everything to device memory.  num_flux: DGX_FLUX_*.  Replaces, for the batching form, the
 * explicit-path set-up of constructor_derived_computational_elements (solve_explicit.c:88-89). */
dgx_status dgx_mixed_create(const dgx_mixed_mesh* mesh, int32_t num_flux, dgx_mixed_ctx** out);
dgx_status dgx_mixed_destroy(dgx_mixed_ctx* ctx);
/* The context's last error message; with ctx NULL, the message of this thread's last failed create. */
const char* dgx_mixed_last_error(const dgx_mixed_ctx* ctx);
int64_t dgx_mixed_n_dof(const dgx_mixed_ctx* ctx);

/* State in group order ([var][mode] per element), ptr_kind DGX_PTR_HOST or DGX_PTR_DEVICE. */
dgx_status dgx_mixed_set_state(dgx_mixed_ctx* ctx, int ptr_kind, const double* u);
dgx_status dgx_mixed_get_state(dgx_mixed_ctx* ctx, int ptr_kind, double* u);

/* The residual of the current state, with or without M^-1 (compute_rhs, solve.c:133-149). */
dgx_status dgx_mixed_rhs(dgx_mixed_ctx* ctx, int apply_minv, int ptr_kind, double* rhs);

/* nsteps SSP-RK3 steps (time_step_ssp_rk_33, solve_explicit.c:252-301), the 3 stages of a step
 * replayed as one captured hipGraph.  No positivity limiter on this path. */
dgx_status dgx_mixed_step(dgx_mixed_ctx* ctx, double dt, int nsteps);

/* Launch form and tile shapes.  tile_elems[g]: elements per workgroup tile of group g (a power of two,
 * 1..64, with the group's lanes per element = 256 / tile_elems[g] >= 4); each launch form keeps its own
 * tile shapes, and NULL selects the ones it has (the autotuner's, or the last given for it).
 * blocks_per_cu: workgroups per CU of the persistent grid (1..8).  Drops the captured graph. */
dgx_status dgx_mixed_set_launch(dgx_mixed_ctx* ctx, int mode, const int32_t* tile_elems, int32_t blocks_per_cu);
dgx_status dgx_mixed_get_launch(const dgx_mixed_ctx* ctx, int* mode, int32_t* tile_elems, int32_t* blocks_per_cu);

/* Form of the quad face maps of sum-factorized HEX and WEDGE sides (the new state's traces at the face
 * nodes, and the face flux's projection onto the side's face-modal rows for the lift): 1 (default) applies
 * them as a tensor product, BF[q][a + n b] = E[ia][a] E[ib][b] with the 1-D factor E at the face's m points
 * per direction and (ia, ib) the node's place in the side's grid under a symmetry of the square, wherever
 * dgx_mixed_create found the rows to factor that way (checked to 1e-12 of the table); 0 applies every
 * row as given.  No counterpart in the reference (its face operators are applied as dense matrices,
 * compute_face_rlhs_T.c); both forms give the same residual to round-off.  Drops the captured graph. */
dgx_status dgx_mixed_set_face_maps(dgx_mixed_ctx* ctx, int tensor);
/* Face operators that have the tensor form (-1 for a NULL context). */
int32_t dgx_mixed_tensor_face_ops(const dgx_mixed_ctx* ctx);

/* The autotuner (both launch forms; the current form is kept): for every group, times the volume kernel
 * of that group alone at each legal tile_elems (reps launches each, HIP events on the context's stream)
 * and keeps the fastest for the per-group form; for the persistent form, whose one launch has one LDS
 * size, caps those tiles to an LDS budget (24..60 KB, halving a tile until it fits) and times the whole
 * persistent volume launch for every budget, every blocks_per_cu the budget lets fit on a CU and
 * every occupancy variant of the kernel (dgx_mixed_persistent_waves), keeping the fastest; the grouped
 * form takes the budget whose one launch is fastest.  Runs the residual form of the kernels: the state is untouched.  timings
 * (nullable) receives per group the best microseconds per launch. */
dgx_status dgx_mixed_autotune(dgx_mixed_ctx* ctx, int reps, double* timings);
/* The occupancy variant of the persistent volume kernel the autotuner chose: 2, 3 or 4 waves per SIMD
 * (the kernel holds every (type, p) path; the 3- and 4-wave variants cap its registers). */
int dgx_mixed_persistent_waves(const dgx_mixed_ctx* ctx);

/* Mean microseconds per stage of nstages SSP-RK3 stages (graph replays, events on the context's
 * stream), and of its kernels: [0] stage, [1] face kernel, [2] volume kernel (each timed alone). */
dgx_status dgx_mixed_time_stage(dgx_mixed_ctx* ctx, double dt, int nstages, double* us);

#ifdef __cplusplus
}
#endif

#endif
