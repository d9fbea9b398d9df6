/*
 * dgx_setup.h - host-side (C99) set-up producers for the explicit DG residual path.
 *
 * These functions produce the inputs that the solver entry points in dgx.h consume: reference-element
 * nodes and 1-D operators, a structured periodic QUAD/HEX mesh with optional isoparametric curving,
 * the geometry factors (cofactor metrics, det J, face normals and face Jacobians) and initial states.
 * They play the role of the reference's set-up chain
 *   constructor_Simulation -> constructor_derived_Solver_Element -> set_up_solver_geometry -> set_initial_solution
 * (src/simulation/simulation.c:128-144, src/simulation/solvers/element_solver.c:179-300,
 *  src/simulation/geometry/geometry_T.c:165-325,734-804, src/simulation/test_case/solution/solution_T.c:127-156).
 *
 * All arrays are caller-allocated; sizes follow from the integer parameters as documented. No function
 * here allocates memory that outlives the call, touches a GPU, or aborts: errors come back as dgx_status.
 */
#ifndef DGX_SETUP_H
#define DGX_SETUP_H

#include <stdint.h>
#include "dgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 1-D nodes and weights on [-1,1], n = p+1 points (GLL requires p >= 1).  x[n], w[n].
 * family: DGX_NODES_GL or DGX_NODES_GLL (dgx.h). */
dgx_status dgx_nodes_1d(int family, int p, double* x, double* w);

/* 1-D Lagrange basis on nodes xs[ns] evaluated (der = 0) or differentiated (der = 1) at points y[ny]:
 * out[i*ns + j] = l_j^(der)(y_i)  (row-major ny x ns). */
dgx_status dgx_lagrange_1d(int ns, const double* xs, int ny, const double* y, int der, double* out);

/* The 1-D operator set (dgx_ops1d, dgx.h) of the GL solution/cubature order p. */
dgx_status dgx_ops1d_build(int p, dgx_ops1d* ops);

/* Number of elements / faces / geometry nodes of a structured periodic mesh with n[0..d-1] cells. */
int64_t dgx_mesh_structured_n_elem(int d, const int* n);
int64_t dgx_mesh_structured_n_face(int d, const int* n);

/* Structured periodic QUAD (d=2) / HEX (d=3) mesh on the box [lo,hi].
 * Element e = ix + n0*(iy + n1*iz); face f = dir*n_elem + e is the +dir face of e.  Its left volume is
 * the one met first in volume order, as the reference's face construction takes it (face.c:109-146):
 * e (local face 2*dir+1) with its +dir neighbour on the right (local face 2*dir), except on the periodic
 * wrap-around, where the neighbour at index 0 of that direction is left (local face 2*dir) and e right.  Geometry nodes are GLL of order p_g per direction,
 * ordered r fastest.  When amp != 0 every node is displaced by
 *     x_k += amp * h_k * prod_{m != k} sin(x_m + phi[k][m]),   phi[k][m] = 2*pi*U(seed)
 * evaluated at the undisplaced position (integer wavenumber, so the curving is periodic on [0,2pi]^d).
 * Outputs: xyz_g[n_elem][d][(p_g+1)^d], face_elem[n_face][2], face_lf[n_face][2], face_ord[n_face]. */
dgx_status dgx_mesh_structured(int d, const int* n, const double* lo, const double* hi, int p_g,
                               double amp, uint64_t seed,
                               double* xyz_g, int32_t* face_elem, int8_t* face_lf, int8_t* face_ord);

/* Mesh ingest in the reference's on-disk format: a Gmsh 2.2 ASCII file of QUAD (d=2) or HEX (d=3)
 * volumes with its $Periodic section (mesh_readers.c:56-440, reorder_nodes_gmsh :132-162;
 * mesh_connectivity.c:93-127,285-393; mesh_periodic.c:102-140).  Volumes are numbered in file order,
 * their vertices put in tensor-product order (Gmsh QUAD 0 1 2 3 -> 0 1 3 2, HEX -> 0 1 3 2 4 5 7 6) and
 * used as p_g = 1 geometry nodes.  Local faces lf = 2 dn + side are matched by their sorted vertices, the
 * vertices of a face on a periodic slave entity replaced by the master's ($Periodic node pairs).
 * Faces are numbered in the order the volumes (then their local faces) meet them; the first volume is
 * the face's left volume.  face_ord = 0 when both volumes list the face's vertices in the same order.
 * Call with null array pointers to get n_elem / n_face, then with arrays of
 *   xyz_g[n_elem][d][2^d], face_elem[n_face][2], face_lf[n_face][2], face_ord[n_face]
 * (all four or none) and, nullable, the reference's connectivity for checking:
 *   v_to_v[n_elem][2d]   neighbour volume of each local face, -1 on a boundary
 *   v_to_lf[n_elem][2d]  the neighbour's local face, or the boundary face's physical tag
 *   node_nums[n_elem][2^d] vertex (node index in file order) of each volume, tensor-product order.
 * DGX_ERR_INVALID_ARG for unreadable files, other volume types, or a face shared by three volumes. */
dgx_status dgx_mesh_read_gmsh(const char* path, int d, int64_t* n_elem, int64_t* n_face, double* xyz_g,
                              int32_t* face_elem, int8_t* face_lf, int8_t* face_ord, int32_t* v_to_v,
                              int32_t* v_to_lf, int32_t* node_nums);

/* Geometry factors for solution order p on GL cubature (collocated GL/GL), following
 * geometry_T.c:734-804 (metric terms at the GLL metric nodes, interpolated to the cubature nodes),
 * geometry_T.c:484-590 (det J; 2-D cofactors) and geometry_T.c:412-445,653-659 (normals, face det J);
 * the 3-D cofactors use the conservative curl form of deprecated/src/setup_geom_factors.c:182-250.
 * Outputs (reference layouts):
 *   metrics_vc[n_elem][d][d][Nc]   per element (Nc,d,d) column-major: m(n,i,j) = J dxi_j/dx_i
 *   jdet_vc[n_elem][Nc]
 *   normals_fc[n_face][Nf][d]      unit normals, outward from the face's left element
 *   jdet_fc[n_face][Nf]            |C n_ref|
 * Returns DGX_ERR_GEOMETRY when det J <= 0 at any cubature node (geometry_T.c:537-543). */
dgx_status dgx_geometry_build(const dgx_mesh* mesh, int p,
                              double* metrics_vc, double* jdet_vc, double* normals_fc, double* jdet_fc);

/* dgx_geometry_build on the GPU (one workgroup per volume; same definitions, outputs and layouts),
 * for p <= 4 and p_g <= 4.  The mesh arrays are host pointers; the outputs are host or device pointers
 * (ptr_kind DGX_PTR_HOST / DGX_PTR_DEVICE).  Synchronous; kernel_ms (nullable) receives the kernel's
 * time from HIP events around its launch.  DGX_ERR_GEOMETRY when det J <= 0 at any
 * cubature node, DGX_ERR_NO_DEVICE without a GPU.  Replaces the geometry set-up of the reference
 * (setup_geometry, geometry_T.c:208-325,734-804) for large meshes, where the host build dominates the
 * set-up time. */
dgx_status dgx_geometry_build_gpu(const dgx_mesh* mesh, int p, int ptr_kind, double* metrics_vc,
                                  double* jdet_vc, double* normals_fc, double* jdet_fc, double* kernel_ms);

/* dgx_geometry_build_gpu that allocates nothing: no reference counterpart.  For callers that cannot
 * allocate device memory or create streams (HIP graph capture, e.g. a moving-mesh time loop that
 * rebuilds the geometry every step, or a loop that forbids allocation).  Same kernels and results (bit
 * for bit) as dgx_geometry_build_gpu.  Every pointer in mesh (xyz_g, face_elem, face_lf) and every
 * output is a DEVICE pointer; the work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without synchronising.  *flags (device int32) is written with 0, or with bit 1
 * when det J <= 0 at some cubature node and bit 2 when a face record is out of range or two faces claim
 * one (volume, local face).  Scratch: a 32 MiB device array per GPU held by the library until the
 * process ends (the face weight table and a face table of 2 d int32 per volume): at most 1,391,701
 * volumes in 3-D (2,087,552 in 2-D); larger meshes get DGX_ERR_INVALID_ARG, with the limit printed to
 * stderr.  Calls on one GPU share that scratch: calls that overlap in time on one GPU race. */
dgx_status dgx_geometry_build_gpu_noalloc(const dgx_mesh* mesh, int p, double* metrics_vc, double* jdet_vc,
                                          double* normals_fc, double* jdet_fc, int32_t* flags, void* stream);

/* Restart file of the state sol[n_elem][d+2][Ns] (reference layout, host) in the reference's format
 * (output_restart_gmsh, test_case/solution/restart/restart_writers.c:149-420): a Gmsh 2.2 mesh whose
 * nodes are the volumes' vertices sorted (last coordinate first) and made unique within a relative
 * 1e-10, its volumes (type 3 QUAD / 5 HEX, 2 tags), and $SolutionCoefficients: per volume
 * "index p_ref n_dof n_var" and the Bezier coefficients of the solution, (n_dof, n_var) column-major,
 * "% .15e".  Periodic meshes have no boundary faces, so no face elements are written. */
dgx_status dgx_restart_write(const char* path, const dgx_mesh* mesh, int p, const double* sol);

/* The state of a restart file written for the same mesh (same volume order and order p): Bezier
 * coefficients back to nodal values.  The reference's reader also handles non-nested meshes by
 * locating nodes in the restart mesh (restart_T.c); this one requires the same mesh and returns
 * DGX_ERR_INVALID_ARG for another volume count, order or variable count. */
dgx_status dgx_restart_read(const char* path, int d, int p, int64_t n_elem, double* sol);

/* VTK XML unstructured-grid file of the Euler solution (output_visualization_vtk_sol,
 * visualization.c:302-349,675-865): one Piece per volume on the equally spaced plotting nodes of
 * order p (nodes.c:113-120) split into p^d linear cells (nodes_plotting.c:255-276), with the point data
 * index, p_ref, ml, density, velocity, pressure, entropy log(p rho^-gamma) and Mach number. */
dgx_status dgx_vtu_write(const char* path, const dgx_mesh* mesh, int p, const double* sol);

/* Physical coordinates of the solution nodes: xyz_s[n_elem][d][Ns] (interpolated from the geometry nodes). */
dgx_status dgx_solution_nodes_xyz(const dgx_mesh* mesh, int p, double* xyz_s);

/* Initial states in the reference per-element layout u[n_elem][nv][Ns] (conservative variables).
 *  - isentropic vortex (Shu): free stream rho=u=v=p=1 (2-D), strength beta, centred at (xc,yc);
 *  - Taylor-Green vortex: rho=1, u=sin x cos y cos z, v=-cos x sin y cos z, w=0,
 *    p = 1/(gamma M^2) + (cos 2x + cos 2y)(cos 2z + 2)/16 (3-D). */
dgx_status dgx_ic_isentropic_vortex(const dgx_mesh* mesh, int p, double beta, double xc, double yc, double* u);
dgx_status dgx_ic_taylor_green(const dgx_mesh* mesh, int p, double mach, double* u);

/* Fixed time step from the initial state: cfl * h_min / ((2p+1) * d * max(|V| + c)). */
double dgx_dt_cfl(int d, int p, int64_t n_elem, const double* u, double h_min, double cfl);

/* Counter-based normal deviates: out[i] = N(0,1) from splitmix64(seed, i) via Box-Muller. */
void dgx_randn(uint64_t seed, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
