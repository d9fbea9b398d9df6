/* Restart and visualisation output of the state (host C): the reference's restart file (Gmsh 2.2 mesh
 * + Bezier solution coefficients; test_case/solution/restart/restart_writers.c:149-420) with its reader
 * for the same mesh, and the VTK XML solution file of visualization.c:302-349,675-865. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dgx_setup.h"

#define OUT_EPS 1e-15          /* values below it are printed as it (visualization.c:1176-1240) */
#define NODE_TOL 1e-10         /* vertex identification tolerance (restart_writers.c:186) */
#define GAMMA_OUT 1.4

static int ipow_o(int b, int e)
{
    int r = 1;
    while (e-- > 0)
        r *= b;
    return r;
}

/* out = (M x M [x M]) in on the n_in^d tensor grid; M is n_out x n_in, row-major */
static void tp_apply_o(int d, int n_out, int n_in, const double* M, const double* in, double* out)
{
    double buf[2][4096];
    int ext[3] = { n_in, d > 1 ? n_in : 1, d > 2 ? n_in : 1 };
    const double* src = in;
    for (int k = 0; k < d; ++k) {
        int eo[3] = { ext[0], ext[1], ext[2] };
        eo[k] = n_out;
        double* dst = (k == d - 1) ? out : buf[k & 1];
        for (int c = 0; c < eo[2]; ++c)
            for (int b = 0; b < eo[1]; ++b)
                for (int a = 0; a < eo[0]; ++a) {
                    const int io[3] = { a, b, c };
                    double s = 0.0;
                    for (int q = 0; q < n_in; ++q) {
                        int ii[3] = { a, b, c };
                        ii[k] = q;
                        s += M[io[k] * n_in + q] * src[ii[0] + ext[0] * (ii[1] + ext[1] * ii[2])];
                    }
                    dst[a + eo[0] * (b + eo[1] * c)] = s;
                }
        ext[0] = eo[0];
        ext[1] = eo[1];
        ext[2] = eo[2];
        src = dst;
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* restart file                                                                                    */
/* ---------------------------------------------------------------------------------------------- */

typedef struct vnode {
    int idx;
    int d;          /* coordinates compared (the mesh dimension); kept per node so the sort is reentrant */
    double x[3];
} vnode;

/* equal_spec_rel_T with den = x0 (math_functions_T.c:34-48) */
static int equal_rel(double x0, double x1, double tol)
{
    return (fabs(x0) < tol && fabs(x1 - x0) < tol) || fabs((x0 - x1) / x0) < tol;
}

/* cmp_DIM_d (containers/matrix.c:123-136): last coordinate first, equal within the tolerance */
static int cmp_vnode(const void* pa, const void* pb)
{
    const vnode* a = (const vnode*)pa;
    const vnode* b = (const vnode*)pb;
    for (int k = a->d - 1; k >= 0; --k) {
        if (equal_rel(a->x[k], b->x[k], NODE_TOL))
            continue;
        return a->x[k] < b->x[k] ? -1 : 1;
    }
    return 0;
}

dgx_status dgx_restart_write(const char* path, const dgx_mesh* mesh, int p, const double* sol)
{
    if (!path || !mesh || !mesh->xyz_g || !sol || (mesh->d != 2 && mesh->d != 3) || p < 1 || p > 7 ||
        mesh->p_g < 1 || mesh->n_elem < 1)
        return DGX_ERR_INVALID_ARG;
    const int d = mesh->d, ng = mesh->p_g + 1, Ng = ipow_o(ng, d), nve = 1 << d, nv = d + 2;
    const int n1 = p + 1, Ns = ipow_o(n1, d);
    const int64_t E = mesh->n_elem, NV = E * nve;
    dgx_ops1d ops;
    if (dgx_ops1d_build(p, &ops) != DGX_OK)
        return DGX_ERR_INVALID_ARG;

    /* vertices of every volume in volume order, tensor-product vertex order (the volume's xyz_ve) */
    vnode* nodes = (vnode*)malloc(sizeof(vnode) * (size_t)NV);
    int* sorted_pos = (int*)malloc(sizeof(int) * (size_t)NV);
    int* uniq = (int*)malloc(sizeof(int) * (size_t)NV);
    int* first_u = (int*)malloc(sizeof(int) * (size_t)NV);
    FILE* f = NULL;
    dgx_status rc = DGX_OK;
    if (!nodes || !sorted_pos || !uniq || !first_u) {
        rc = DGX_ERR_INVALID_ARG;
        goto done;
    }
    for (int64_t e = 0; e < E; ++e)
        for (int v = 0; v < nve; ++v) {
            const int a = (ng - 1) * ((v & 1) + ng * (((v >> 1) & 1) + ng * ((v >> 2) & 1)));
            vnode* nd = &nodes[e * nve + v];
            nd->idx = (int)(e * nve + v);
            nd->d = d;
            for (int k = 0; k < 3; ++k)
                nd->x[k] = k < d ? mesh->xyz_g[(e * d + k) * Ng + a] : 0.0;
        }
    /* row_sort_DIM_Matrix_d + make_unique_row_Multiarray_d (restart_writers.c:186-196) */
    qsort(nodes, (size_t)NV, sizeof(vnode), cmp_vnode);
    for (int64_t i = 0; i < NV; ++i)
        sorted_pos[nodes[i].idx] = (int)i;
    int n_u = 1;
    first_u[0] = 0;
    uniq[0] = 0;
    for (int64_t i = 1; i < NV; ++i) {
        const vnode* u = &nodes[first_u[n_u - 1]];
        int unique = 0;
        for (int k = 0; k < d; ++k)
            if (!equal_rel(u->x[k], nodes[i].x[k], NODE_TOL)) {
                unique = 1;
                break;
            }
        if (unique)
            first_u[n_u++] = (int)i;
        uniq[i] = n_u - 1;
    }

    f = fopen(path, "w");
    if (!f) {
        rc = DGX_ERR_INVALID_ARG;
        goto done;
    }
    fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n");
    fprintf(f, "$Nodes\n%d\n", n_u);
    for (int n = 0; n < n_u; ++n) {
        fprintf(f, "%d", n + 1);
        for (int k = 0; k < 3; ++k)
            fprintf(f, " %g", k < d ? nodes[first_u[n]].x[k] : 0.0);
        fprintf(f, "\n");
    }
    fprintf(f, "$EndNodes\n");
    /* periodic meshes have no boundary faces, so only volumes are listed (type 3 QUAD, 5 HEX; 2 tags) */
    fprintf(f, "$Elements\n%lld\n", (long long)E);
    {
        static const int tp_to_gmsh[8] = { 0, 1, 3, 2, 4, 5, 7, 6 };
        for (int64_t e = 0; e < E; ++e) {
            fprintf(f, "%lld %d %d %d %d", (long long)(e + 1), d == 2 ? 3 : 5, 2, 0, 0);
            for (int v = 0; v < nve; ++v)
                fprintf(f, " %d", uniq[sorted_pos[e * nve + tp_to_gmsh[v]]] + 1);
            fprintf(f, "\n");
        }
    }
    fprintf(f, "$EndElements\n");
    /* Bezier coefficients, (Ns, nv) column-major per volume (fprintf_solution_bezier_gmsh) */
    fprintf(f, "$SolutionCoefficients\n%lld\n", (long long)E);
    {
        double* b = (double*)malloc(sizeof(double) * (size_t)Ns);
        if (!b) {
            rc = DGX_ERR_INVALID_ARG;
            goto done;
        }
        for (int64_t e = 0; e < E; ++e) {
            fprintf(f, "%lld %d %d %d", (long long)(e + 1), p, Ns, nv);
            for (int v = 0; v < nv; ++v) {
                tp_apply_o(d, n1, n1, ops.sb, sol + (e * nv + v) * Ns, b);
                for (int n = 0; n < Ns; ++n)
                    fprintf(f, " % .15e", b[n]);
            }
            fprintf(f, "\n");
        }
        free(b);
    }
    fprintf(f, "$EndSolutionCoefficients\n");
    if (ferror(f))
        rc = DGX_ERR_INVALID_ARG;
done:
    if (f && fclose(f) != 0)
        rc = DGX_ERR_INVALID_ARG;
    free(nodes);
    free(sorted_pos);
    free(uniq);
    free(first_u);
    return rc;
}

dgx_status dgx_restart_read(const char* path, int d, int p, int64_t n_elem, double* sol)
{
    if (!path || !sol || (d != 2 && d != 3) || p < 1 || p > 7 || n_elem < 1)
        return DGX_ERR_INVALID_ARG;
    const int n1 = p + 1, Ns = ipow_o(n1, d), nv = d + 2;
    dgx_ops1d ops;
    if (dgx_ops1d_build(p, &ops) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "r");
    if (!f)
        return DGX_ERR_INVALID_ARG;
    char line[256];
    int found = 0;
    while (fgets(line, sizeof line, f))
        if (strncmp(line, "$SolutionCoefficients", 21) == 0) {
            found = 1;
            break;
        }
    dgx_status rc = DGX_OK;
    long long nvol = -1;
    double* b = (double*)malloc(sizeof(double) * (size_t)Ns * nv);
    if (!found || !b || fscanf(f, "%lld", &nvol) != 1 || nvol != n_elem) {
        rc = DGX_ERR_INVALID_ARG;
        goto done;
    }
    for (int64_t e = 0; e < n_elem; ++e) {
        long long id;
        int pr, nd, nvar;
        if (fscanf(f, "%lld %d %d %d", &id, &pr, &nd, &nvar) != 4 || id != e + 1 || pr != p || nd != Ns ||
            nvar != nv) {
            rc = DGX_ERR_INVALID_ARG;
            goto done;
        }
        for (int i = 0; i < Ns * nv; ++i)
            if (fscanf(f, "%lf", &b[i]) != 1) {
                rc = DGX_ERR_INVALID_ARG;
                goto done;
            }
        for (int v = 0; v < nv; ++v)
            tp_apply_o(d, n1, n1, ops.bs, b + v * Ns, sol + (e * nv + v) * Ns);
    }
done:
    free(b);
    fclose(f);
    return rc;
}

/* ---------------------------------------------------------------------------------------------- */
/* VTK XML solution file                                                                           */
/* ---------------------------------------------------------------------------------------------- */

static void vtk_scalar(FILE* f, const char* name, const double* v, int n)
{
    fprintf(f, "\t\t<DataArray type=\"Float32\" Name=\"%s\" format=\"ascii\">\n", name);
    for (int i = 0; i < n; ++i) {
        if (i % 8 == 0)
            fprintf(f, "\t\t");
        fprintf(f, " % .8e", fabs(v[i]) < OUT_EPS ? OUT_EPS : v[i]);
        if ((i + 1) % 8 == 0 || i == n - 1)
            fprintf(f, "\n");
    }
    fprintf(f, "\t\t</DataArray>\n");
}

static void vtk_vector(FILE* f, const char* name, const double* v, int n, int d)
{
    if (name)
        fprintf(f, "\t\t<DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                name);
    else
        fprintf(f, "\t\t<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int i = 0; i < n; ++i) {
        fprintf(f, "\t\t");
        for (int k = 0; k < 3; ++k) {
            if (k < d) {
                const double x = v[k * n + i];
                fprintf(f, " % .8e", fabs(x) < OUT_EPS ? OUT_EPS : x);
            } else {
                fprintf(f, " %d", 0);
            }
        }
        fprintf(f, "\n");
    }
    fprintf(f, "\t\t</DataArray>\n");
}

static void vtk_int(FILE* f, const char* name, int value, int n)
{
    fprintf(f, "\t\t<DataArray type=\"Int32\" Name=\"%s\" format=\"ascii\">\n", name);
    for (int i = 0; i < n; ++i) {
        if (i % 8 == 0)
            fprintf(f, "\t\t");
        fprintf(f, " %d", value);
        if ((i + 1) % 8 == 0 || i == n - 1)
            fprintf(f, "\n");
    }
    fprintf(f, "\t\t</DataArray>\n");
}

dgx_status dgx_vtu_write(const char* path, const dgx_mesh* mesh, int p, const double* sol)
{
    if (!path || !mesh || !mesh->xyz_g || !sol || (mesh->d != 2 && mesh->d != 3) || p < 1 || p > 7 ||
        mesh->p_g < 1 || mesh->n_elem < 1)
        return DGX_ERR_INVALID_ARG;
    const int d = mesh->d, ng = mesh->p_g + 1, Ng = ipow_o(ng, d), nv = d + 2;
    const int n1 = p + 1, Ns = ipow_o(n1, d), NP = Ns, NC = ipow_o(p, d);
    const int64_t E = mesh->n_elem;
    double xg[8], wg[8], xs[8], ws[8], rp[8], Lg[64], Ls[64];
    if (dgx_nodes_1d(DGX_NODES_GLL, mesh->p_g, xg, wg) != DGX_OK || dgx_nodes_1d(DGX_NODES_GL, p, xs, ws) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    for (int i = 0; i < n1; ++i)
        rp[i] = -1.0 + 2.0 * i / p;                  /* NODES_PLOT (nodes.c:113-120) */
    dgx_lagrange_1d(ng, xg, n1, rp, 0, Lg);           /* geometry basis at the plotting nodes */
    dgx_lagrange_1d(n1, xs, n1, rp, 0, Ls);           /* solution basis at the plotting nodes */
    (void)Ng;

    FILE* f = fopen(path, "w");
    if (!f)
        return DGX_ERR_INVALID_ARG;
    double* xyz = (double*)malloc(sizeof(double) * 3 * NP);
    double* w = (double*)malloc(sizeof(double) * nv * NP);
    double* prim = (double*)malloc(sizeof(double) * (nv + 2) * NP);
    if (!xyz || !w || !prim) {
        free(xyz);
        free(w);
        free(prim);
        fclose(f);
        return DGX_ERR_INVALID_ARG;
    }
    fprintf(f, "<?xml version=\"1.0\"?>\n");
    fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    fprintf(f, "<UnstructuredGrid>\n");
    const int N = n1, N2 = n1 * n1;
    for (int64_t e = 0; e < E; ++e) {
        for (int k = 0; k < d; ++k)
            tp_apply_o(d, n1, ng, Lg, mesh->xyz_g + (e * d + k) * Ng, xyz + k * NP);
        for (int v = 0; v < nv; ++v)
            tp_apply_o(d, n1, n1, Ls, sol + (e * nv + v) * Ns, w + v * NP);
        /* primitive variables, entropy log(p rho^-gamma), Mach (solution_euler.c:72-118) */
        double* rho = prim;
        double* vel = prim + NP;
        double* pr = prim + (1 + d) * NP;
        double* ent = prim + (2 + d) * NP;
        double* mach = prim + (3 + d) * NP;
        for (int i = 0; i < NP; ++i) {
            const double r = w[i];
            double V2 = 0.0;
            for (int k = 0; k < d; ++k) {
                vel[k * NP + i] = w[(1 + k) * NP + i] / r;
                V2 += vel[k * NP + i] * vel[k * NP + i];
            }
            rho[i] = r;
            pr[i] = (GAMMA_OUT - 1.0) * (w[(d + 1) * NP + i] - 0.5 * r * V2);
            ent[i] = log(pr[i] * pow(r, -GAMMA_OUT));
            mach[i] = sqrt(V2 / (GAMMA_OUT * pr[i] / r));
        }
        fprintf(f, "\n<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n", NP, NC);
        fprintf(f, "\t<Points>\n");
        vtk_vector(f, NULL, xyz, NP, d);
        fprintf(f, "\t</Points>\n");
        fprintf(f, "\t<Cells>\n");
        fprintf(f, "\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
        /* sub-cells of the plotting grid (nodes_plotting.c:255-276) */
        for (int k = 0; k < (d == 3 ? p : 1); ++k)
            for (int j = 0; j < p; ++j)
                for (int i = 0; i < p; ++i) {
                    const int q[4] = { i + N * j, i + 1 + N * j, i + 1 + N * (j + 1), i + N * (j + 1) };
                    fprintf(f, "\t\t");
                    for (int l = 0; l < 4; ++l)
                        fprintf(f, " %d", q[l] + N2 * k);
                    if (d == 3)
                        for (int l = 0; l < 4; ++l)
                            fprintf(f, " %d", q[l] + N2 * (k + 1));
                    fprintf(f, "\n");
                }
        fprintf(f, "\t\t</DataArray>\n");
        fprintf(f, "\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n\t\t");
        for (int c = 0; c < NC; ++c)
            fprintf(f, " %d", (c + 1) * (1 << d));
        fprintf(f, "\n\t\t</DataArray>\n");
        fprintf(f, "\t\t<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n\t\t");
        for (int c = 0; c < NC; ++c)
            fprintf(f, " %d", d == 2 ? 9 : 12);
        fprintf(f, "\n\t\t</DataArray>\n");
        fprintf(f, "\t</Cells>\n");
        fprintf(f, "\t<PointData Scalars=\"Scalars\" Vectors=\"Vectors\">\n");
        vtk_int(f, "index", (int)e, NP);
        vtk_int(f, "p_ref", p, NP);
        vtk_int(f, "ml", 0, NP);
        vtk_scalar(f, "$\\rho$", rho, NP);
        vtk_vector(f, "$\\mathbf{v}$", vel, NP, d);
        vtk_scalar(f, "$\\mathit{p}$", pr, NP);
        vtk_scalar(f, "$\\mathit{s}$", ent, NP);
        vtk_scalar(f, "mach", mach, NP);
        fprintf(f, "\t</PointData>\n");
        fprintf(f, "</Piece>\n\n");
    }
    fprintf(f, "</UnstructuredGrid>\n</VTKFile>\n");
    const int bad = ferror(f);
    free(xyz);
    free(w);
    free(prim);
    if (fclose(f) != 0 || bad)
        return DGX_ERR_INVALID_ARG;
    return DGX_OK;
}
