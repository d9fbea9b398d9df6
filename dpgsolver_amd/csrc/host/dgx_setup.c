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
        /* faces on the +dir side of this element */
        for (int dir = 0; dir < d; ++dir) {
            int jc[3] = { ix, iy, iz };
            jc[dir] = (jc[dir] + 1) % n[dir];
            const int64_t er = jc[0] + (int64_t)n[0] * (jc[1] + (int64_t)n1 * jc[2]);
            const int64_t f = dir * E + e;
            face_elem[2 * f + 0] = (int32_t)e;
            face_elem[2 * f + 1] = (int32_t)er;
            face_lf[2 * f + 0] = (int8_t)(2 * dir + 1);
            face_lf[2 * f + 1] = (int8_t)(2 * dir);
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
