/*
 * Host-side set-up producers: reference-element nodes and 1-D operators, structured periodic meshes,
 * geometry factors and initial states.  C99, FP64, CPU only.  See include/dgx_setup.h for the contract.
 *
 * Operators are built directly from the Lagrange product formula on the nodes (no Vandermonde
 * inversion); element operators are never formed densely: every multi-dimensional operator is applied
 * as a sequence of 1-D contractions (sum factorization), the application format the reference stores
 * in Operator::ops_tp (src/simulation/operators/operator.h:40-44) but leaves disabled on its hot path.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dgx_setup.h"

#define DGX_PI 3.14159265358979323846
#define GAMMA 1.4
#define GM1 0.4

/* ------------------------------------------------------------------------------------------------ */
/* Legendre polynomials and 1-D node families                                                       */
/* ------------------------------------------------------------------------------------------------ */

/* P_n(x) and P_n'(x) by the three-term recurrence. */
static void legendre(int n, double x, double* pn, double* dpn)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    *pn = p1;
    /* P_n' = n (x P_n - P_{n-1}) / (x^2 - 1), away from the end points */
    if (fabs(fabs(x) - 1.0) < 1e-300)
        *dpn = 0.5 * n * (n + 1.0) * (x > 0 ? 1.0 : ((n % 2) ? 1.0 : -1.0));
    else
        *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

static void sort_asc(int n, double* x, double* w)
{
    for (int i = 1; i < n; ++i) {
        double xi = x[i], wi = w[i];
        int j = i - 1;
        while (j >= 0 && x[j] > xi) { x[j + 1] = x[j]; w[j + 1] = w[j]; --j; }
        x[j + 1] = xi;
        w[j + 1] = wi;
    }
}

dgx_status dgx_nodes_1d(int family, int p, double* x, double* w)
{
    if (!x || !w || p < 0 || p > 15)
        return DGX_ERR_INVALID_ARG;
    const int n = p + 1;
    if (family == DGX_NODES_GL) {
        /* roots of P_n by Newton from the Chebyshev-like guess cos(pi (i+3/4)/(n+1/2)) */
        for (int i = 0; i < n; ++i) {
            double xi = cos(DGX_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                double pn, dpn;
                legendre(n, xi, &pn, &dpn);
                const double dx = pn / dpn;
                xi -= dx;
                if (fabs(dx) < 1e-16)
                    break;
            }
            double pn, dpn;
            legendre(n, xi, &pn, &dpn);
            x[i] = xi;
            w[i] = 2.0 / ((1.0 - xi * xi) * dpn * dpn);
        }
        sort_asc(n, x, w);
        return DGX_OK;
    }
    if (family == DGX_NODES_GLL) {
        if (p < 1)
            return DGX_ERR_INVALID_ARG;
        /* end points and the roots of P_p'; Newton on P_p' with P_p'' from the Legendre ODE
         * (1-x^2) P'' = 2x P' - p(p+1) P, started at the Chebyshev-Gauss-Lobatto points */
        x[0] = -1.0;
        x[p] = 1.0;
        for (int i = 1; i < p; ++i) {
            double xi = -cos(DGX_PI * i / p);
            for (int it = 0; it < 100; ++it) {
                double pp, dpp;
                legendre(p, xi, &pp, &dpp);
                const double d2 = (2.0 * xi * dpp - p * (p + 1.0) * pp) / (1.0 - xi * xi);
                const double dx = dpp / d2;
                xi -= dx;
                if (fabs(dx) < 1e-16)
                    break;
            }
            x[i] = xi;
        }
        for (int i = 0; i <= p; ++i) {
            double pp, dpp;
            legendre(p, x[i], &pp, &dpp);
            w[i] = 2.0 / (p * (p + 1.0) * pp * pp);
        }
        sort_asc(n, x, w);
        return DGX_OK;
    }
    return DGX_ERR_INVALID_ARG;
}

/* ------------------------------------------------------------------------------------------------ */
/* 1-D Lagrange operators                                                                           */
/* ------------------------------------------------------------------------------------------------ */

static double lagr(int ns, const double* xs, int j, double y)
{
    double v = 1.0;
    for (int k = 0; k < ns; ++k)
        if (k != j)
            v *= (y - xs[k]) / (xs[j] - xs[k]);
    return v;
}

static double dlagr(int ns, const double* xs, int j, double y)
{
    /* l_j'(y) = sum_{m != j} 1/(x_j - x_m) prod_{k != j,m} (y - x_k)/(x_j - x_k); valid at the nodes too */
    double s = 0.0;
    for (int m = 0; m < ns; ++m) {
        if (m == j)
            continue;
        double t = 1.0 / (xs[j] - xs[m]);
        for (int k = 0; k < ns; ++k)
            if (k != j && k != m)
                t *= (y - xs[k]) / (xs[j] - xs[k]);
        s += t;
    }
    return s;
}

dgx_status dgx_lagrange_1d(int ns, const double* xs, int ny, const double* y, int der, double* out)
{
    if (ns < 1 || ny < 1 || !xs || !y || !out || (der != 0 && der != 1))
        return DGX_ERR_INVALID_ARG;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < ns; ++j)
            out[i * ns + j] = der ? dlagr(ns, xs, j, y[i]) : lagr(ns, xs, j, y[i]);
    return DGX_OK;
}

/* In-place inverse of a small row-major matrix by Gauss-Jordan elimination with partial pivoting. */
static int invert_small(int n, double* a)
{
    double aug[16 * 32];
    if (n > 16)
        return -1;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j)
            aug[i * 2 * n + j] = (j < n) ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(aug[r * 2 * n + c]) > fabs(aug[piv * 2 * n + c]))
                piv = r;
        if (aug[piv * 2 * n + c] == 0.0)
            return -1;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) {
                double t = aug[c * 2 * n + j];
                aug[c * 2 * n + j] = aug[piv * 2 * n + j];
                aug[piv * 2 * n + j] = t;
            }
        const double inv = 1.0 / aug[c * 2 * n + c];
        for (int j = 0; j < 2 * n; ++j)
            aug[c * 2 * n + j] *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == c)
                continue;
            const double f = aug[r * 2 * n + c];
            if (f != 0.0)
                for (int j = 0; j < 2 * n; ++j)
                    aug[r * 2 * n + j] -= f * aug[c * 2 * n + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            a[i * n + j] = aug[i * 2 * n + n + j];
    return 0;
}

static double binom(int n, int k)
{
    double c = 1.0;
    for (int i = 1; i <= k; ++i)
        c = c * (n - k + i) / i;
    return c;
}

dgx_status dgx_ops1d_build(int p, dgx_ops1d* ops)
{
    if (!ops || p < 1 || p > 7)
        return DGX_ERR_INVALID_ARG;
    memset(ops, 0, sizeof(*ops));
    const int n = p + 1;
    ops->p = p;
    ops->n = n;
    dgx_status st = dgx_nodes_1d(DGX_NODES_GL, p, ops->x, ops->w);
    if (st != DGX_OK)
        return st;
    dgx_lagrange_1d(n, ops->x, n, ops->x, 1, ops->D);
    const double m1 = -1.0, p1 = 1.0;
    dgx_lagrange_1d(n, ops->x, 1, &m1, 0, ops->lm);
    dgx_lagrange_1d(n, ops->x, 1, &p1, 0, ops->lp);
    /* Bernstein basis B_k(x_i) = C(p,k) t^k (1-t)^(p-k), t = (1+x)/2 (src/element/bases.c:1013-1019):
     * nodal values = bs . bezier coefficients, so ccBS0 = B and ccSB0 = B^-1. */
    for (int i = 0; i < n; ++i) {
        const double t = 0.5 * (1.0 + ops->x[i]);
        for (int k = 0; k < n; ++k)
            ops->bs[i * n + k] = binom(p, k) * pow(t, k) * pow(1.0 - t, p - k);
    }
    memcpy(ops->sb, ops->bs, sizeof(double) * n * n);
    if (invert_small(n, ops->sb) != 0)
        return DGX_ERR_INVALID_ARG;
    return DGX_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* Sum-factorized tensor-product application                                                         */
/* ------------------------------------------------------------------------------------------------ */

/* out = (A[d-1] x ... x A[0]) in, with in indexed i0 + n0*(i1 + n1*i2) and A[k] row-major m[k] x n[k];
 * A[k] == NULL stands for the identity (m[k] == n[k]). */
static void tp_apply(int d, const int* m, const int* n, const double* const* A, const double* in, double* out)
{
    double buf0[4096], buf1[4096];
    int ext[3] = { n[0], d > 1 ? n[1] : 1, d > 2 ? n[2] : 1 };
    int last = -1;
    for (int k = 0; k < d; ++k)
        if (A[k])
            last = k;
    if (last < 0) {
        memcpy(out, in, sizeof(double) * ext[0] * ext[1] * ext[2]);
        return;
    }
    const double* src = in;
    double* dst = buf0;
    for (int k = 0; k < d; ++k) {
        if (!A[k])
            continue;
        int ext_o[3] = { ext[0], ext[1], ext[2] };
        ext_o[k] = m[k];
        double* target = (k == last) ? out : dst;
        for (int c = 0; c < ext_o[2]; ++c)
            for (int b = 0; b < ext_o[1]; ++b)
                for (int a = 0; a < ext_o[0]; ++a) {
                    const int io[3] = { a, b, c };
                    double s = 0.0;
                    for (int q = 0; q < n[k]; ++q) {
                        int ii[3] = { a, b, c };
                        ii[k] = q;
                        s += A[k][io[k] * n[k] + q] * src[ii[0] + ext[0] * (ii[1] + ext[1] * ii[2])];
                    }
                    target[a + ext_o[0] * (b + ext_o[1] * c)] = s;
                }
        ext[0] = ext_o[0];
        ext[1] = ext_o[1];
        ext[2] = ext_o[2];
        src = target;
        dst = (dst == buf0) ? buf1 : buf0;
    }
}

static int ipow(int b, int e)
{
    int r = 1;
    for (int i = 0; i < e; ++i)
        r *= b;
    return r;
}

/* ------------------------------------------------------------------------------------------------ */
/* Counter-based random numbers                                                                      */
/* ------------------------------------------------------------------------------------------------ */

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* i-th output of the splitmix64 sequence started at seed */
static uint64_t splitmix_at(uint64_t seed, uint64_t i)
{
    return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

static double unif_at(uint64_t seed, uint64_t i)
{
    return (double)(splitmix_at(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

void dgx_randn(uint64_t seed, int64_t n, double* out)
{
    for (int64_t i = 0; i < n; i += 2) {
        double u1 = unif_at(seed, (uint64_t)i), u2 = unif_at(seed, (uint64_t)i + 1);
        if (u1 < 1e-300)
            u1 = 1e-300;
        const double r = sqrt(-2.0 * log(u1)), th = 2.0 * DGX_PI * u2;
        out[i] = r * cos(th);
        if (i + 1 < n)
            out[i + 1] = r * sin(th);
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* Structured periodic mesh                                                                           */
/* ------------------------------------------------------------------------------------------------ */

int64_t dgx_mesh_structured_n_elem(int d, const int* n)
{
    int64_t e = 1;
    for (int k = 0; k < d; ++k)
        e *= n[k];
    return e;
}

int64_t dgx_mesh_structured_n_face(int d, const int* n)
{
    return d * dgx_mesh_structured_n_elem(d, n);
}

dgx_status dgx_mesh_structured(int d, const int* n, const double* lo, const double* hi, int p_g,
                               double amp, uint64_t seed,
                               double* xyz_g, int32_t* face_elem, int8_t* face_lf, int8_t* face_ord)
{
    if ((d != 2 && d != 3) || !n || !lo || !hi || p_g < 1 || p_g > 7 || !xyz_g || !face_elem || !face_lf ||
        !face_ord)
        return DGX_ERR_INVALID_ARG;
    for (int k = 0; k < d; ++k)
        if (n[k] < 1 || !(hi[k] > lo[k]))
            return DGX_ERR_INVALID_ARG;
    const int64_t E = dgx_mesh_structured_n_elem(d, n);
    if (E > (1 << 28))
        return DGX_ERR_INVALID_ARG;
    const int ng = p_g + 1, Ng = ipow(ng, d);
    double xg[8], wg[8];
    dgx_nodes_1d(DGX_NODES_GLL, p_g, xg, wg);

    double h[3] = { 0, 0, 0 }, phi[3][3];
    for (int k = 0; k < d; ++k)
        h[k] = (hi[k] - lo[k]) / n[k];
    for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 3; ++m)
            phi[k][m] = 2.0 * DGX_PI * unif_at(seed, (uint64_t)(3 * k + m));

    const int n1 = (d > 1) ? n[1] : 1;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        const int ix = (int)(e % n[0]), iy = (int)((e / n[0]) % n1), iz = (int)(e / ((int64_t)n[0] * n1));
        const int icell[3] = { ix, iy, iz };
        double* xe = xyz_g + e * d * Ng;
        for (int a = 0; a < Ng; ++a) {
            const int ia[3] = { a % ng, (a / ng) % ng, a / (ng * ng) };
            double x0[3] = { 0, 0, 0 };
            for (int k = 0; k < d; ++k)
                x0[k] = lo[k] + h[k] * (icell[k] + 0.5 * (xg[ia[k]] + 1.0));
            for (int k = 0; k < d; ++k) {
                double xk = x0[k];
                if (amp != 0.0) {
                    double prod = 1.0;
                    for (int m = 0; m < d; ++m)
                        if (m != k)
                            prod *= sin(x0[m] + phi[k][m]);
                    xk += amp * h[k] * prod;
                }
                xe[k * Ng + a] = xk;
            }
        }
        /* the face on the +dir side of this element.  Its left volume is the one the reference's face
         * construction meets first in volume order (constructor_Faces, face.c:109-146): this element for
         * an interior face, the neighbour across the periodic boundary (through its -dir face) for the
         * wrap-around face.  The Roe-Pike entropy fix is not symmetric in left and right, so the choice
         * is part of the result. */
        for (int dir = 0; dir < d; ++dir) {
            int jc[3] = { ix, iy, iz };
            const int wrap = (jc[dir] + 1 == n[dir]);
            jc[dir] = (jc[dir] + 1) % n[dir];
            const int64_t er = jc[0] + (int64_t)n[0] * (jc[1] + (int64_t)n1 * jc[2]);
            const int64_t f = dir * E + e;
            face_elem[2 * f + 0] = (int32_t)(wrap ? er : e);
            face_elem[2 * f + 1] = (int32_t)(wrap ? e : er);
            face_lf[2 * f + 0] = (int8_t)(wrap ? 2 * dir : 2 * dir + 1);
            face_lf[2 * f + 1] = (int8_t)(wrap ? 2 * dir + 1 : 2 * dir);
            face_ord[f] = 0;
        }
    }
    return DGX_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* Geometry                                                                                           */
/* ------------------------------------------------------------------------------------------------ */

typedef struct geo_ops {
    int d, p, p_g, ng, nc, Ng, Nc, Nf;
    double xg[8], wg[8], xc[8], wc[8];
    double Dg[64];        /* GLL derivative at GLL nodes, ng x ng */
    double Igc[64];       /* GLL Lagrange at GL nodes, nc x ng */
    double Dgc[64];       /* GLL Lagrange derivative at GL nodes, nc x ng */
    double Igm[8], Igp[8];/* GLL Lagrange at -1 / +1, 1 x ng */
    double Ig[64];        /* identity ng x ng */
} geo_ops;

static dgx_status geo_ops_build(int d, int p, int p_g, geo_ops* g)
{
    memset(g, 0, sizeof(*g));
    g->d = d; g->p = p; g->p_g = p_g;
    g->ng = p_g + 1; g->nc = p + 1;
    g->Ng = ipow(g->ng, d); g->Nc = ipow(g->nc, d); g->Nf = ipow(g->nc, d - 1);
    if (dgx_nodes_1d(DGX_NODES_GLL, p_g, g->xg, g->wg) != DGX_OK ||
        dgx_nodes_1d(DGX_NODES_GL, p, g->xc, g->wc) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    const double m1 = -1.0, p1 = 1.0;
    dgx_lagrange_1d(g->ng, g->xg, g->ng, g->xg, 1, g->Dg);
    dgx_lagrange_1d(g->ng, g->xg, g->nc, g->xc, 0, g->Igc);
    dgx_lagrange_1d(g->ng, g->xg, g->nc, g->xc, 1, g->Dgc);
    dgx_lagrange_1d(g->ng, g->xg, 1, &m1, 0, g->Igm);
    dgx_lagrange_1d(g->ng, g->xg, 1, &p1, 0, g->Igp);
    for (int i = 0; i < g->ng; ++i)
        for (int j = 0; j < g->ng; ++j)
            g->Ig[i * g->ng + j] = (i == j) ? 1.0 : 0.0;
    return DGX_OK;
}

/* d_j X at the geometry (= metric) nodes */
static void geo_grad_vm(const geo_ops* g, const double* f, int j, double* out)
{
    const double* A[3] = { NULL, NULL, NULL };
    const int m[3] = { g->ng, g->ng, g->ng };
    A[j] = g->Dg;
    tp_apply(g->d, m, m, A, f, out);
}

/* metric terms m(n,i,j) = J dxi_j/dx_i at the metric nodes, mvm[(i + d*j)*Ng + a] */
static void geo_metrics_vm(const geo_ops* g, const double* X, double* mvm)
{
    const int d = g->d, Ng = g->Ng;
    double J[9][512];
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j)
            geo_grad_vm(g, X + i * Ng, j, J[i * 3 + j]);
    if (d == 2) {
        for (int a = 0; a < Ng; ++a) {
            mvm[(0 + 2 * 0) * Ng + a] = J[1 * 3 + 1][a];   /*  y_s */
            mvm[(0 + 2 * 1) * Ng + a] = -J[1 * 3 + 0][a];  /* -y_r */
            mvm[(1 + 2 * 0) * Ng + a] = -J[0 * 3 + 1][a];  /* -x_s */
            mvm[(1 + 2 * 1) * Ng + a] = J[0 * 3 + 0][a];   /*  x_r */
        }
        return;
    }
    /* conservative curl form: C_ic = 1/2 [curl_xi(X_a grad_xi X_b) - curl_xi(X_b grad_xi X_a)]_c,
     * (a,b) = (i+1, i+2) mod 3, computed with the discrete derivative on the metric nodes */
    double V[3][512], dV[512];
    for (int i = 0; i < 3; ++i) {
        const int ia = (i + 1) % 3, ib = (i + 2) % 3;
        double curl[2][3][512];
        for (int s = 0; s < 2; ++s) {
            const int a = s ? ib : ia, b = s ? ia : ib;
            for (int k = 0; k < 3; ++k)
                for (int q = 0; q < Ng; ++q)
                    V[k][q] = X[a * Ng + q] * J[b * 3 + k][q];
            for (int c = 0; c < 3; ++c) {
                const int k1 = (c + 1) % 3, k2 = (c + 2) % 3;
                /* curl_c = d_k1 V_k2 - d_k2 V_k1 */
                geo_grad_vm(g, V[k2], k1, curl[s][c]);
                geo_grad_vm(g, V[k1], k2, dV);
                for (int q = 0; q < Ng; ++q)
                    curl[s][c][q] -= dV[q];
            }
        }
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < Ng; ++q)
                mvm[(i + 3 * c) * Ng + q] = 0.5 * (curl[0][c][q] - curl[1][c][q]);
    }
}

/* Face geometry of one face from the metric terms of its left element (geometry_T.c:274-325, 412-445). */
static void geo_face(const geo_ops* g, const double* mvm, int lf, double* normals, double* jdet)
{
    const int d = g->d, Ng = g->Ng, Nf = g->Nf;
    const int mg[3] = { g->ng, g->ng, g->ng };
    const int dn = lf / 2;
    const double sgn = (lf % 2) ? 1.0 : -1.0;
    const double* A[3] = { g->Igc, g->Igc, g->Igc };
    A[dn] = (lf % 2) ? g->Igp : g->Igm;
    int mf[3] = { g->nc, g->nc, g->nc };
    mf[dn] = 1;
    double nrm[3][512];
    for (int i = 0; i < d; ++i) {
        tp_apply(d, mf, mg, A, mvm + (i + d * dn) * Ng, nrm[i]);
        for (int q = 0; q < Nf; ++q)
            nrm[i][q] *= sgn;
    }
    for (int q = 0; q < Nf; ++q) {
        double s = 0.0;
        for (int i = 0; i < d; ++i)
            s += nrm[i][q] * nrm[i][q];
        s = sqrt(s);
        jdet[q] = s;
        for (int i = 0; i < d; ++i)
            normals[q * d + i] = nrm[i][q] / s;
    }
}

dgx_status dgx_geometry_build(const dgx_mesh* mesh, int p,
                              double* metrics_vc, double* jdet_vc, double* normals_fc, double* jdet_fc)
{
    if (!mesh || !mesh->xyz_g || !mesh->face_elem || !mesh->face_lf || !metrics_vc || !jdet_vc ||
        !normals_fc || !jdet_fc || p < 1 || p > 7 || (mesh->d != 2 && mesh->d != 3))
        return DGX_ERR_INVALID_ARG;
    geo_ops g;
    if (geo_ops_build(mesh->d, p, mesh->p_g, &g) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    const int d = g.d, Ng = g.Ng, Nc = g.Nc, Nf = g.Nf;
    const int64_t E = mesh->n_elem, F = mesh->n_face;

    /* faces grouped by their left element, so each element's metric terms are computed once */
    int64_t* first = (int64_t*)calloc((size_t)E + 1, sizeof(int64_t));
    int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)(F > 0 ? F : 1));
    if (!first || !list) {
        free(first);
        free(list);
        return DGX_ERR_INVALID_ARG;
    }
    for (int64_t f = 0; f < F; ++f) {
        const int64_t eL = mesh->face_elem[2 * f];
        const int lf = mesh->face_lf[2 * f];
        if (eL < 0 || eL >= E || lf < 0 || lf >= 2 * d) {
            free(first);
            free(list);
            return DGX_ERR_INVALID_ARG;
        }
        first[eL + 1]++;
    }
    for (int64_t e = 0; e < E; ++e)
        first[e + 1] += first[e];
    {
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)E);
        memcpy(fill, first, sizeof(int64_t) * (size_t)E);
        for (int64_t f = 0; f < F; ++f)
            list[fill[mesh->face_elem[2 * f]]++] = f;
        free(fill);
    }

    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t e = 0; e < E; ++e) {
        const int mc[3] = { g.nc, g.nc, g.nc }, mg[3] = { g.ng, g.ng, g.ng };
        double mvm[9 * 512], Jc[9][512], X[3 * 512];
        /* metric terms are translation invariant; element-local coordinates keep the curl form's
         * products X_a dX_b at the element's own scale (less cancellation than absolute coordinates) */
        const double* Xa = mesh->xyz_g + e * d * Ng;
        for (int i = 0; i < d; ++i)
            for (int a = 0; a < Ng; ++a)
                X[i * Ng + a] = Xa[i * Ng + a] - Xa[i * Ng];
        geo_metrics_vm(&g, X, mvm);
        const double* Ai[3] = { g.Igc, g.Igc, g.Igc };
        for (int c = 0; c < d * d; ++c)
            tp_apply(d, mc, mg, Ai, mvm + c * Ng, metrics_vc + (e * d * d + c) * Nc);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                const double* A[3] = { g.Igc, g.Igc, g.Igc };
                A[j] = g.Dgc;
                tp_apply(d, mc, mg, A, X + i * Ng, Jc[i * 3 + j]);
            }
        double* jd = jdet_vc + e * Nc;
        for (int n = 0; n < Nc; ++n) {
            double det;
            if (d == 2) {
                det = Jc[0][n] * Jc[4][n] - Jc[1][n] * Jc[3][n];
            } else {
                det = Jc[0][n] * (Jc[4][n] * Jc[8][n] - Jc[5][n] * Jc[7][n]) -
                      Jc[1][n] * (Jc[3][n] * Jc[8][n] - Jc[5][n] * Jc[6][n]) +
                      Jc[2][n] * (Jc[3][n] * Jc[7][n] - Jc[4][n] * Jc[6][n]);
            }
            jd[n] = det;
            if (!(det > 0.0))
                bad |= 1;
        }
        for (int64_t k = first[e]; k < first[e + 1]; ++k) {
            const int64_t f = list[k];
            geo_face(&g, mvm, mesh->face_lf[2 * f], normals_fc + f * Nf * d, jdet_fc + f * Nf);
        }
    }
    free(first);
    free(list);
    return bad ? DGX_ERR_GEOMETRY : DGX_OK;
}

dgx_status dgx_solution_nodes_xyz(const dgx_mesh* mesh, int p, double* xyz_s)
{
    if (!mesh || !mesh->xyz_g || !xyz_s || p < 1 || p > 7)
        return DGX_ERR_INVALID_ARG;
    geo_ops g;
    if (geo_ops_build(mesh->d, p, mesh->p_g, &g) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    const int d = g.d;
    const int mc[3] = { g.nc, g.nc, g.nc }, mg[3] = { g.ng, g.ng, g.ng };
    const double* A[3] = { g.Igc, g.Igc, g.Igc };
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < mesh->n_elem; ++e)
        for (int i = 0; i < d; ++i)
            tp_apply(d, mc, mg, A, mesh->xyz_g + (e * d + i) * g.Ng, xyz_s + (e * d + i) * g.Nc);
    return DGX_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* Initial states and time step                                                                      */
/* ------------------------------------------------------------------------------------------------ */

dgx_status dgx_ic_isentropic_vortex(const dgx_mesh* mesh, int p, double beta, double xc, double yc, double* u)
{
    if (!mesh || mesh->d != 2 || !u)
        return DGX_ERR_INVALID_ARG;
    const int Ns = (p + 1) * (p + 1), nv = 4;
    double* xyz = (double*)malloc(sizeof(double) * mesh->n_elem * 2 * Ns);
    if (!xyz)
        return DGX_ERR_INVALID_ARG;
    dgx_status st = dgx_solution_nodes_xyz(mesh, p, xyz);
    if (st != DGX_OK) {
        free(xyz);
        return st;
    }
    for (int64_t e = 0; e < mesh->n_elem; ++e)
        for (int n = 0; n < Ns; ++n) {
            const double x = xyz[(e * 2 + 0) * Ns + n] - xc, y = xyz[(e * 2 + 1) * Ns + n] - yc;
            const double r2 = x * x + y * y;
            const double ex = exp(0.5 * (1.0 - r2));
            const double du = -beta / (2.0 * DGX_PI) * ex * y, dv = beta / (2.0 * DGX_PI) * ex * x;
            const double T = 1.0 - GM1 * beta * beta / (8.0 * GAMMA * DGX_PI * DGX_PI) * exp(1.0 - r2);
            const double rho = pow(T, 1.0 / GM1), pr = pow(rho, GAMMA);
            const double uu = 1.0 + du, vv = 1.0 + dv;
            double* ue = u + e * nv * Ns;
            ue[0 * Ns + n] = rho;
            ue[1 * Ns + n] = rho * uu;
            ue[2 * Ns + n] = rho * vv;
            ue[3 * Ns + n] = pr / GM1 + 0.5 * rho * (uu * uu + vv * vv);
        }
    free(xyz);
    return DGX_OK;
}

dgx_status dgx_ic_taylor_green(const dgx_mesh* mesh, int p, double mach, double* u)
{
    if (!mesh || mesh->d != 3 || !u || !(mach > 0.0))
        return DGX_ERR_INVALID_ARG;
    const int Ns = (p + 1) * (p + 1) * (p + 1), nv = 5;
    double* xyz = (double*)malloc(sizeof(double) * mesh->n_elem * 3 * Ns);
    if (!xyz)
        return DGX_ERR_INVALID_ARG;
    dgx_status st = dgx_solution_nodes_xyz(mesh, p, xyz);
    if (st != DGX_OK) {
        free(xyz);
        return st;
    }
    const double p0 = 1.0 / (GAMMA * mach * mach);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < mesh->n_elem; ++e)
        for (int n = 0; n < Ns; ++n) {
            const double x = xyz[(e * 3 + 0) * Ns + n], y = xyz[(e * 3 + 1) * Ns + n],
                         z = xyz[(e * 3 + 2) * Ns + n];
            const double rho = 1.0;
            const double uu = sin(x) * cos(y) * cos(z), vv = -cos(x) * sin(y) * cos(z), ww = 0.0;
            const double pr = p0 + (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0) / 16.0;
            double* ue = u + e * nv * Ns;
            ue[0 * Ns + n] = rho;
            ue[1 * Ns + n] = rho * uu;
            ue[2 * Ns + n] = rho * vv;
            ue[3 * Ns + n] = rho * ww;
            ue[4 * Ns + n] = pr / GM1 + 0.5 * rho * (uu * uu + vv * vv + ww * ww);
        }
    free(xyz);
    return DGX_OK;
}

double dgx_dt_cfl(int d, int p, int64_t n_elem, const double* u, double h_min, double cfl)
{
    const int Ns = ipow(p + 1, d), nv = d + 2;
    double smax = 0.0;
    for (int64_t e = 0; e < n_elem; ++e)
        for (int n = 0; n < Ns; ++n) {
            const double* ue = u + e * nv * Ns;
            const double rho = ue[n];
            double v2 = 0.0;
            for (int k = 0; k < d; ++k) {
                const double vk = ue[(1 + k) * Ns + n] / rho;
                v2 += vk * vk;
            }
            const double pr = GM1 * (ue[(d + 1) * Ns + n] - 0.5 * rho * v2);
            const double s = sqrt(v2) + sqrt(GAMMA * pr / rho);
            if (s > smax)
                smax = s;
        }
    /* the d in the denominator sums the directional CFL numbers; without it SSP-RK3 at p=3 in 3-D is
     * unstable at cfl = 0.5 (the oracle blows up after ~15 steps on a curved 8^3 mesh) */
    return cfl * h_min / ((2.0 * p + 1.0) * d * smax);
}

/* ------------------------------------------------------------------------------------------------ */
/* Gmsh 2.2 ASCII mesh ingest                                                                        */
/* ------------------------------------------------------------------------------------------------ */

/* Nodes of a Gmsh element type (the types a QUAD/HEX mesh file holds; 0 = unsupported). */
static int gmsh_type_nodes(int type)
{
    switch (type) {
    case 15: return 1;   /* point */
    case 1: return 2;    /* line */
    case 2: return 3;    /* triangle */
    case 3: return 4;    /* quadrangle */
    case 4: return 4;    /* tetrahedron */
    case 5: return 8;    /* hexahedron */
    case 6: return 6;    /* prism */
    case 7: return 5;    /* pyramid */
    default: return 0;
    }
}

typedef struct {
    int type, ntag, tag0, tag1;
    int nodes[8];   /* node indices (0-based, file order) */
} gm_elem;

typedef struct {
    int slave, master;    /* entity tags */
    int n;
    int* s;               /* slave node indices */
    int* m;               /* master node indices */
} gm_per;

typedef struct {
    int key[4];           /* sorted (substituted) vertices of a volume's local face */
    int64_t v;
    int lf;
} gm_fkey;

static int cmp_key(const int* x, const int* y);

static int cmp_fkey(const void* a, const void* b)
{
    const gm_fkey* x = (const gm_fkey*)a;
    const gm_fkey* y = (const gm_fkey*)b;
    const int c = cmp_key(x->key, y->key);
    if (c != 0)
        return c;
    if (x->v != y->v)
        return x->v < y->v ? -1 : 1;
    return x->lf - y->lf;
}

static int cmp_key(const int* x, const int* y)
{
    for (int k = 0; k < 4; ++k)
        if (x[k] != y[k])
            return x[k] < y[k] ? -1 : 1;
    return 0;
}

static void sort_small(int* a, int n)
{
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && a[j - 1] > a[j]; --j) {
            const int t = a[j];
            a[j] = a[j - 1];
            a[j - 1] = t;
        }
}

/* Vertices of local face lf = 2*dn + side of a tensor-product element (vertex i0 + 2 i1 + 4 i2), in
 * tensor-product order over the remaining directions. */
/* neigh_info[1].ind_ord of a face from its vertices as listed by the left volume (a) and the right volume
 * (b), both in the reference's face-vertex order (element.c f_ve): matches_R_to_L[i] = the position in b of
 * a[i], looked up in the table of possible matches (set_ind_ord, face.c:496-552: LINE 2 cases, QUAD 8).
 * Returns the case, or -1 when the lists do not hold the same vertices. */
static int face_vertex_ord(int nfv, const int* a, const int* b)
{
    static const int line[2][2] = { { 0, 1 }, { 1, 0 } };
    static const int quad[8][4] = { { 0, 1, 2, 3 }, { 1, 0, 3, 2 }, { 2, 3, 0, 1 }, { 3, 2, 1, 0 },
                                    { 0, 2, 1, 3 }, { 2, 0, 3, 1 }, { 1, 3, 0, 2 }, { 3, 1, 2, 0 } };
    int match[4] = { -1, -1, -1, -1 };
    for (int i = 0; i < nfv; ++i)
        for (int j = 0; j < nfv; ++j)
            if (b[j] == a[i])
                match[i] = j;
    const int n_case = (nfv == 2) ? 2 : 8;
    for (int k = 0; k < n_case; ++k) {
        const int* row = (nfv == 2) ? line[k] : quad[k];
        int same = 1;
        for (int i = 0; i < nfv; ++i)
            same &= (match[i] == row[i]);
        if (same)
            return k;
    }
    return -1;
}

static int tp_face_vertices(int d, int lf, int* out)
{
    const int dn = lf >> 1, side = lf & 1;
    int k = 0;
    for (int v = 0; v < (1 << d); ++v)
        if (((v >> dn) & 1) == side)
            out[k++] = v;
    return k;
}

typedef struct {
    int d;
    int64_t n_nodes, n_vol;
    double* xyz;          /* [n_nodes][3] */
    int* node_of_id;      /* Gmsh node id -> index */
    int max_id;
    gm_elem* el;
    int64_t n_el;
    gm_per* per;
    int n_per;
} gm_file;

static void gm_free(gm_file* g)
{
    free(g->xyz);
    free(g->node_of_id);
    free(g->el);
    for (int i = 0; i < g->n_per; ++i) {
        free(g->per[i].s);
        free(g->per[i].m);
    }
    free(g->per);
    memset(g, 0, sizeof(*g));
}

static int gm_node(const gm_file* g, long id)
{
    return (id >= 0 && id <= g->max_id) ? g->node_of_id[id] : -1;
}

/* Reads $MeshFormat (2.2 ASCII), $Nodes, $Elements and $Periodic (mesh_readers.c:215-440; periodic
 * entities mesh_readers.c, mesh_periodic.c:102-140). */
static dgx_status gm_read(const char* path, int d, gm_file* g)
{
    memset(g, 0, sizeof(*g));
    g->d = d;
    FILE* fp = fopen(path, "r");
    if (!fp)
        return DGX_ERR_INVALID_ARG;
    char tok[256];
    dgx_status st = DGX_OK;
    int have_nodes = 0, have_elems = 0;
    while (st == DGX_OK && fscanf(fp, "%255s", tok) == 1) {
        if (strcmp(tok, "$MeshFormat") == 0) {
            double ver;
            int ftype, dsize;
            if (fscanf(fp, "%lf %d %d", &ver, &ftype, &dsize) != 3 || ver < 2.0 || ver >= 3.0 || ftype != 0)
                st = DGX_ERR_INVALID_ARG;   /* only the ASCII 2.x format the reference reads */
        } else if (strcmp(tok, "$Nodes") == 0) {
            long n;
            if (fscanf(fp, "%ld", &n) != 1 || n <= 0) {
                st = DGX_ERR_INVALID_ARG;
                break;
            }
            g->n_nodes = n;
            g->xyz = (double*)malloc(sizeof(double) * 3 * (size_t)n);
            long* ids = (long*)malloc(sizeof(long) * (size_t)n);
            long max_id = 0;
            for (long i = 0; i < n && st == DGX_OK; ++i) {
                if (fscanf(fp, "%ld %lf %lf %lf", &ids[i], &g->xyz[3 * i], &g->xyz[3 * i + 1], &g->xyz[3 * i + 2]) != 4 ||
                    ids[i] < 0 || ids[i] > 100000000L)
                    st = DGX_ERR_INVALID_ARG;
                else if (ids[i] > max_id)
                    max_id = ids[i];
            }
            if (st == DGX_OK) {
                g->max_id = (int)max_id;
                g->node_of_id = (int*)malloc(sizeof(int) * (size_t)(max_id + 1));
                for (long i = 0; i <= max_id; ++i)
                    g->node_of_id[i] = -1;
                for (long i = 0; i < n; ++i)
                    g->node_of_id[ids[i]] = (int)i;
                have_nodes = 1;
            }
            free(ids);
        } else if (strcmp(tok, "$Elements") == 0) {
            long n;
            if (!have_nodes || fscanf(fp, "%ld", &n) != 1 || n <= 0) {
                st = DGX_ERR_INVALID_ARG;
                break;
            }
            g->n_el = n;
            g->el = (gm_elem*)calloc((size_t)n, sizeof(gm_elem));
            for (long i = 0; i < n && st == DGX_OK; ++i) {
                long id;
                int type, ntag;
                gm_elem* e = &g->el[i];
                if (fscanf(fp, "%ld %d %d", &id, &type, &ntag) != 3 || ntag < 0 || gmsh_type_nodes(type) == 0) {
                    st = DGX_ERR_INVALID_ARG;
                    break;
                }
                e->type = type;
                e->ntag = ntag;
                for (int t = 0; t < ntag; ++t) {
                    int tv;
                    if (fscanf(fp, "%d", &tv) != 1) {
                        st = DGX_ERR_INVALID_ARG;
                        break;
                    }
                    if (t == 0) e->tag0 = tv;
                    if (t == 1) e->tag1 = tv;
                }
                for (int k = 0; k < gmsh_type_nodes(type) && st == DGX_OK; ++k) {
                    long nid;
                    if (fscanf(fp, "%ld", &nid) != 1 || (e->nodes[k] = gm_node(g, nid)) < 0)
                        st = DGX_ERR_INVALID_ARG;
                }
            }
            have_elems = (st == DGX_OK);
        } else if (strcmp(tok, "$Periodic") == 0) {
            int n;
            if (!have_nodes || fscanf(fp, "%d", &n) != 1 || n < 0) {
                st = DGX_ERR_INVALID_ARG;
                break;
            }
            g->per = (gm_per*)calloc((size_t)(n > 0 ? n : 1), sizeof(gm_per));
            for (int i = 0; i < n && st == DGX_OK; ++i) {
                int dim, np;
                gm_per* p = &g->per[g->n_per++];
                if (fscanf(fp, "%d %d %d %d", &dim, &p->slave, &p->master, &np) != 4 || np < 0) {
                    st = DGX_ERR_INVALID_ARG;
                    break;
                }
                p->s = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
                p->m = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
                p->n = np;
                for (int k = 0; k < np && st == DGX_OK; ++k) {
                    long a, b;
                    if (fscanf(fp, "%ld %ld", &a, &b) != 2 || (p->s[k] = gm_node(g, a)) < 0 ||
                        (p->m[k] = gm_node(g, b)) < 0)
                        st = DGX_ERR_INVALID_ARG;
                }
                if (dim != d - 1)
                    p->n = 0;   /* only face entities carry face node correspondences */
            }
        }
    }
    fclose(fp);
    if (st == DGX_OK && !have_elems)
        st = DGX_ERR_INVALID_ARG;
    if (st != DGX_OK)
        gm_free(g);
    return st;
}

dgx_status dgx_mesh_read_gmsh(const char* path, int d, int64_t* n_elem, int64_t* n_face, double* xyz_g,
                              int32_t* face_elem, int8_t* face_lf, int8_t* face_ord, int32_t* v_to_v,
                              int32_t* v_to_lf, int32_t* node_nums)
{
    if (!path || (d != 2 && d != 3) || !n_elem || !n_face)
        return DGX_ERR_INVALID_ARG;
    gm_file g;
    dgx_status st = gm_read(path, d, &g);
    if (st != DGX_OK)
        return st;
    const int vol_type = (d == 2) ? 3 : 5, face_type = (d == 2) ? 1 : 3;
    const int nvert = 1 << d, nfv = 1 << (d - 1), nlf = 2 * d;
    /* Gmsh -> tensor-product vertex order (mesh_readers.c:132-162) */
    static const int ord2[4] = { 0, 1, 3, 2 };
    static const int ord3[8] = { 0, 1, 3, 2, 4, 5, 7, 6 };
    const int* ord = (d == 2) ? ord2 : ord3;

    int64_t E = 0;
    for (int64_t i = 0; i < g.n_el; ++i) {
        const int t = g.el[i].type;
        const int dim = (t == 15) ? 0 : (t == 1) ? 1 : (t == 2 || t == 3) ? 2 : 3;
        if (dim == d && t != vol_type)
            st = DGX_ERR_INVALID_ARG;   /* mixed element types: not built */
        if (t == vol_type)
            ++E;
    }
    if (st != DGX_OK || E == 0) {
        gm_free(&g);
        return DGX_ERR_INVALID_ARG;
    }
    int* vn = (int*)malloc(sizeof(int) * (size_t)E * nvert);   /* volume vertices, TP order */
    {
        int64_t e = 0;
        for (int64_t i = 0; i < g.n_el; ++i)
            if (g.el[i].type == vol_type) {
                for (int k = 0; k < nvert; ++k)
                    vn[e * nvert + k] = g.el[i].nodes[ord[k]];
                ++e;
            }
    }
    /* boundary entities of the face elements: sorted vertices -> (physical tag, entity tag) */
    int64_t nb = 0;
    for (int64_t i = 0; i < g.n_el; ++i)
        nb += (g.el[i].type == face_type);
    gm_fkey* bkeys = (gm_fkey*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(gm_fkey));
    {
        int64_t b = 0;
        for (int64_t i = 0; i < g.n_el; ++i)
            if (g.el[i].type == face_type) {
                for (int k = 0; k < 4; ++k)
                    bkeys[b].key[k] = (k < nfv) ? g.el[i].nodes[k] : -1;
                sort_small(bkeys[b].key, nfv);
                bkeys[b].v = i;   /* element index in the file */
                ++b;
            }
        qsort(bkeys, (size_t)nb, sizeof(gm_fkey), cmp_fkey);
    }
    /* local faces with slave-entity vertices replaced by their masters (correct_f_ve_for_periodic,
     * mesh_periodic.c:102-140) */
    gm_fkey* fk = (gm_fkey*)malloc(sizeof(gm_fkey) * (size_t)E * nlf);
    int* fv_sub = (int*)malloc(sizeof(int) * (size_t)E * nlf * nfv);   /* substituted, TP face order */
    int* bc = (int*)malloc(sizeof(int) * (size_t)E * nlf);              /* physical tag of a boundary face */
    for (int64_t e = 0; e < E; ++e)
        for (int lf = 0; lf < nlf; ++lf) {
            int loc[4], ids[4] = { -1, -1, -1, -1 };
            tp_face_vertices(d, lf, loc);
            for (int k = 0; k < nfv; ++k)
                ids[k] = vn[e * nvert + loc[k]];
            gm_fkey probe;
            memcpy(probe.key, ids, sizeof(ids));
            sort_small(probe.key, nfv);
            /* the boundary element on this face, if any */
            int64_t lo = 0, hi = nb;
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if (cmp_key(bkeys[mid].key, probe.key) < 0)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            int phys = 0, ent = 0;
            if (lo < nb) {
                int same = 1;
                for (int k = 0; k < nfv; ++k)
                    same &= (bkeys[lo].key[k] == probe.key[k]);
                if (same) {
                    phys = g.el[bkeys[lo].v].tag0;
                    ent = g.el[bkeys[lo].v].tag1;
                }
            }
            bc[e * nlf + lf] = phys;
            for (int p = 0; p < g.n_per && ent != 0; ++p) {
                if (g.per[p].slave != ent || g.per[p].n == 0)
                    continue;
                for (int k = 0; k < nfv; ++k)
                    for (int q = 0; q < g.per[p].n; ++q)
                        if (g.per[p].s[q] == ids[k]) {
                            ids[k] = g.per[p].m[q];
                            break;
                        }
                break;
            }
            gm_fkey* f = &fk[e * nlf + lf];
            memcpy(f->key, ids, sizeof(ids));
            memcpy(&fv_sub[(e * nlf + lf) * nfv], ids, sizeof(int) * nfv);
            sort_small(f->key, nfv);
            f->v = e;
            f->lf = lf;
        }
    qsort(fk, (size_t)E * nlf, sizeof(gm_fkey), cmp_fkey);
    int32_t* vv = (int32_t*)malloc(sizeof(int32_t) * (size_t)E * nlf);
    int32_t* vl = (int32_t*)malloc(sizeof(int32_t) * (size_t)E * nlf);
    for (int64_t i = 0; i < E * nlf; ++i)
        vv[i] = -1;
    for (int64_t i = 0; i < E * nlf && st == DGX_OK;) {
        int64_t j = i + 1;
        while (j < E * nlf && cmp_key(fk[j].key, fk[i].key) == 0)
            ++j;
        if (j - i == 2) {
            const gm_fkey *a = &fk[i], *b = &fk[i + 1];
            vv[a->v * nlf + a->lf] = (int32_t)b->v;
            vl[a->v * nlf + a->lf] = b->lf;
            vv[b->v * nlf + b->lf] = (int32_t)a->v;
            vl[b->v * nlf + b->lf] = a->lf;
        } else if (j - i == 1) {
            vl[fk[i].v * nlf + fk[i].lf] = bc[fk[i].v * nlf + fk[i].lf];   /* boundary: its condition */
        } else {
            st = DGX_ERR_INVALID_ARG;   /* a face shared by more than two volumes */
        }
        i = j;
    }
    /* faces in the order the volumes meet them: the first volume is the face's left volume */
    int64_t F = 0;
    for (int64_t e = 0; e < E && st == DGX_OK; ++e)
        for (int lf = 0; lf < nlf; ++lf) {
            const int64_t nb_e = vv[e * nlf + lf];
            const int nb_lf = vl[e * nlf + lf];
            if (nb_e < 0 || nb_e < e || (nb_e == e && nb_lf < lf))
                continue;
            if (face_elem) {
                face_elem[2 * F] = (int32_t)e;
                face_elem[2 * F + 1] = (int32_t)nb_e;
                face_lf[2 * F] = (int8_t)lf;
                face_lf[2 * F + 1] = (int8_t)nb_lf;
                /* neigh_info[1].ind_ord: the case of the vertex correspondence seen from the left */
                face_ord[F] = (int8_t)face_vertex_ord(nfv, &fv_sub[(e * nlf + lf) * nfv],
                                                      &fv_sub[(nb_e * nlf + nb_lf) * nfv]);
            }
            ++F;
        }
    if (st == DGX_OK) {
        *n_elem = E;
        *n_face = F;
        for (int64_t e = 0; e < E; ++e) {
            if (xyz_g)
                for (int k = 0; k < d; ++k)
                    for (int v = 0; v < nvert; ++v)
                        xyz_g[(e * d + k) * nvert + v] = g.xyz[3 * vn[e * nvert + v] + k];
            if (node_nums)
                for (int v = 0; v < nvert; ++v)
                    node_nums[e * nvert + v] = vn[e * nvert + v];
            for (int lf = 0; lf < nlf; ++lf) {
                if (v_to_v)
                    v_to_v[e * nlf + lf] = vv[e * nlf + lf];
                if (v_to_lf)
                    v_to_lf[e * nlf + lf] = vl[e * nlf + lf];
            }
        }
    }
    free(vn); free(bkeys); free(fk); free(fv_sub); free(bc); free(vv); free(vl);
    gm_free(&g);
    return st;
}
