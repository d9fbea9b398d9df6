/*
 * Oracle (test infrastructure only): CPU restatement of the reference's explicit DG residual path with
 * dense per-element operators.  See oracle.h for the scope and the reference files it follows.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GAMMA 1.4                    /* definitions_physics.h:23 */
#define GM1 0.4                      /* definitions_physics.h:24 */
#define EPS_PHYS 1.0e-13             /* definitions_physics.h:30 */
#define EPS 1.0e-15                  /* definitions_tol.h:23 */
#define ORC_PI 3.1415926535897932384626433832795029

struct orc_ctx {
    int d, p, p_g, num_flux, positivity;
    int n1, Ns, Nc, Nf, ng, Ng;
    int64_t n_elem, n_face;
    int32_t* face_elem;
    int8_t* face_lf;
    /* dense element operators (row-major) */
    double* cv0_vs_vc;          /* Nc x Ns */
    double* tw1_vt_vc[3];       /* Ns x Nc */
    double* cv1_vs_vc[3];       /* Nc x Ns: derivative along xi_j at the volume cubature nodes */
    double* tw0_vt_vc;          /* Ns x Nc: cv0_vs_vc^T diag(w_vc) */
    double* cv0_vs_fc[6];       /* Nf x Ns */
    double* tw0_vt_fc[6];       /* Ns x Nf */
    double* ccSB0;              /* Ns x Ns */
    double* ccBS0;              /* Ns x Ns */
    double* w_vc;               /* Nc */
    /* geometry */
    double* metrics_vc;         /* [E][d][d][Nc] */
    double* jdet_vc;            /* [E][Nc] */
    double* normals_fc;         /* [F][Nf][d] */
    double* jdet_fc;            /* [F][Nf] */
    double* m_inv;              /* [E][Ns*Ns], built on first use after the geometry is set */
    int minv_ready;
    /* Navier-Stokes (constant viscosity) with the BR2-stable second-order flux when ns != 0 */
    int ns;
    double mu, Pr;
    double eta;                 /* BR2 penalty override (> 0), else 1.01 x NFMAX */
    /* threads of the residual and stage loops (1: the reference's sequential order) */
    int nthreads;
    /* explicit time stepper of orc_rk_stage: 0 SSP-RK3, 1 forward Euler, 2 low-storage RK 5-stage 4th order */
    int time_stepper;
    int8_t* face_ord;           /* [F] neigh_info[1].ind_ord (NULL: every face aligned) */
    int64_t* ef_face;           /* [E][2d] face of each local face, built on first threaded use */
    int8_t* ef_side;            /* [E][2d] 0: the volume is the face's left volume, 1: its right */
    /* sum-factorized application (orc_set_sum_factorized): every element operator above is a Kronecker
     * product of these 1-D factors, applied by one mode product per direction instead of densely; M^-1 is
     * the diagonal 1/(w J) of the GL collocation (cv0 = I) */
    int sf;
    double f_cv0[256], f_cv1[256], f_tw0[256], f_tw1[256], f_fm[16], f_fp[16], f_B[256], f_Binv[256];
    double* minv_diag;          /* [E][Ns], with the geometry */
};

/* operators of the sum-factorized application */
enum { OP_CV0_VC, OP_TW1_VC, OP_CV1_VC, OP_TW0_VC, OP_CV0_FC, OP_TW0_FC };

/* 1-D factors F[k] (row-major m[k] x n[k]) of operator op (direction or local face j): the factors
 * orc_create builds the dense operator from */
static void kron_of(const orc_ctx* c, int op, int j, const double** F, int* m, int* n)
{
    const int nn = c->n1;
    for (int k = 0; k < 3; ++k) {
        m[k] = nn;
        n[k] = nn;
        F[k] = (op == OP_TW1_VC || op == OP_TW0_VC || op == OP_TW0_FC) ? c->f_tw0 : c->f_cv0;
    }
    if (op == OP_TW1_VC)
        F[j] = c->f_tw1;
    else if (op == OP_CV1_VC)
        F[j] = c->f_cv1;
    else if (op == OP_CV0_FC) {
        F[j / 2] = (j & 1) ? c->f_fp : c->f_fm;
        m[j / 2] = 1;
    } else if (op == OP_TW0_FC) {
        F[j / 2] = (j & 1) ? c->f_fp : c->f_fm;   /* the trace row as a column: tw0_vt_fc = cv0_vs_fc^T diag(w_fc) */
        n[j / 2] = 1;
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* small dense linear algebra                                                                        */
/* ------------------------------------------------------------------------------------------------ */

void orc_mm(int m, int k, int n, const double* A, const double* B, double* C)
{
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int q = 0; q < k; ++q)
                s += A[i * k + q] * B[q * n + j];
            C[i * n + j] = s;
        }
}

/* LU with partial pivoting; solves A X = B for X in place of B (n x nrhs, row-major). */
static int lu_solve(int n, double* A, int nrhs, double* B)
{
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r * n + c]) > fabs(A[piv * n + c]))
                piv = r;
        if (A[piv * n + c] == 0.0)
            return -1;
        if (piv != c) {
            for (int j = 0; j < n; ++j) {
                double t = A[c * n + j]; A[c * n + j] = A[piv * n + j]; A[piv * n + j] = t;
            }
            for (int j = 0; j < nrhs; ++j) {
                double t = B[c * nrhs + j]; B[c * nrhs + j] = B[piv * nrhs + j]; B[piv * nrhs + j] = t;
            }
        }
        for (int r = c + 1; r < n; ++r) {
            const double f = A[r * n + c] / A[c * n + c];
            if (f == 0.0)
                continue;
            for (int j = c; j < n; ++j)
                A[r * n + j] -= f * A[c * n + j];
            for (int j = 0; j < nrhs; ++j)
                B[r * nrhs + j] -= f * B[c * nrhs + j];
        }
    }
    for (int c = n - 1; c >= 0; --c)
        for (int j = 0; j < nrhs; ++j) {
            double s = B[c * nrhs + j];
            for (int q = c + 1; q < n; ++q)
                s -= A[c * n + q] * B[q * nrhs + j];
            B[c * nrhs + j] = s / A[c * n + c];
        }
    return 0;
}

static int invert(int n, const double* A, double* Ainv)
{
    double* a = (double*)malloc(sizeof(double) * n * n);
    memcpy(a, A, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            Ainv[i * n + j] = (i == j) ? 1.0 : 0.0;
    const int r = lu_solve(n, a, n, Ainv);
    free(a);
    return r;
}

/* Cyclic Jacobi eigen-decomposition of a small symmetric matrix: eigenvalues in ev, vectors in columns of V. */
static void jacobi_eig(int n, double* A, double* ev, double* V)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
                off += A[i * n + j] * A[i * n + j];
        if (off < 1e-34)
            break;
        for (int pp = 0; pp < n; ++pp)
            for (int q = pp + 1; q < n; ++q) {
                const double apq = A[pp * n + q];
                if (fabs(apq) < 1e-300)
                    continue;
                const double theta = (A[q * n + q] - A[pp * n + pp]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + pp], akq = A[k * n + q];
                    A[k * n + pp] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[pp * n + k], aqk = A[q * n + k];
                    A[pp * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + pp], vkq = V[k * n + q];
                    V[k * n + pp] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i)
        ev[i] = A[i * n + i];
}

/* ------------------------------------------------------------------------------------------------ */
/* nodes and 1-D bases                                                                               */
/* ------------------------------------------------------------------------------------------------ */

static double legendre_p(int n, double x)
{
    double p0 = 1.0, p1 = x;
    if (n == 0)
        return 1.0;
    for (int k = 2; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    return p1;
}

/* orthonormal Legendre phi_k = sqrt((2k+1)/2) P_k and its derivative (via P_k' recurrence) */
static void ortho_basis(int nb, double x, double* phi, double* dphi)
{
    double P[16], dP[16];
    P[0] = 1.0; dP[0] = 0.0;
    if (nb > 1) { P[1] = x; dP[1] = 1.0; }
    for (int k = 2; k < nb; ++k) {
        P[k] = ((2.0 * k - 1.0) * x * P[k - 1] - (k - 1.0) * P[k - 2]) / k;
        dP[k] = dP[k - 2] + (2.0 * k - 1.0) * P[k - 1];
    }
    for (int k = 0; k < nb; ++k) {
        const double s = sqrt(0.5 * (2.0 * k + 1.0));
        phi[k] = s * P[k];
        dphi[k] = s * dP[k];
    }
}

static void sort2(int n, double* x, double* w)
{
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && x[j - 1] > x[j]; --j) {
            double t = x[j]; x[j] = x[j - 1]; x[j - 1] = t;
            t = w[j]; w[j] = w[j - 1]; w[j - 1] = t;
        }
}

int orc_nodes_1d(int family, int p, double* x, double* w)
{
    const int n = p + 1;
    double J[256], V[256];
    if (p < 0 || p > 15)
        return -1;
    if (family == 0) {
        /* Golub-Welsch: Jacobi matrix of the Legendre recurrence, b_k = k / sqrt(4k^2 - 1) */
        memset(J, 0, sizeof(J));
        for (int k = 1; k < n; ++k) {
            const double b = k / sqrt(4.0 * k * k - 1.0);
            J[(k - 1) * n + k] = b;
            J[k * n + k - 1] = b;
        }
        jacobi_eig(n, J, x, V);
        for (int i = 0; i < n; ++i)
            w[i] = 2.0 * V[0 * n + i] * V[0 * n + i];
        sort2(n, x, w);
        return 0;
    }
    if (family == 1) {
        if (p < 1)
            return -1;
        /* interior GLL nodes = zeros of P_p' (Jacobi (1,1)): b_k^2 = k(k+2)/((2k+1)(2k+3)) */
        const int m = p - 1;
        memset(J, 0, sizeof(J));
        for (int k = 1; k < m; ++k) {
            const double b = sqrt(k * (k + 2.0) / ((2.0 * k + 1.0) * (2.0 * k + 3.0)));
            J[(k - 1) * m + k] = b;
            J[k * m + k - 1] = b;
        }
        double xi[16];
        if (m > 0)
            jacobi_eig(m, J, xi, V);
        x[0] = -1.0;
        for (int i = 0; i < m; ++i)
            x[1 + i] = xi[i];
        x[p] = 1.0;
        for (int i = 0; i <= p; ++i) {
            const double P = legendre_p(p, x[i]);
            w[i] = 2.0 / (p * (p + 1.0) * P * P);
        }
        sort2(n, x, w);
        return 0;
    }
    return -1;
}

int orc_lagrange_1d(int ns, const double* xs, int ny, const double* y, int der, double* out)
{
    double Phi[256], Phi_inv[256], dummy[16];
    if (ns > 16)
        return -1;
    for (int i = 0; i < ns; ++i)
        ortho_basis(ns, xs[i], &Phi[i * ns], dummy);
    if (invert(ns, Phi, Phi_inv) != 0)
        return -1;
    for (int i = 0; i < ny; ++i) {
        double phi[16], dphi[16];
        ortho_basis(ns, y[i], phi, dphi);
        for (int j = 0; j < ns; ++j) {
            double s = 0.0;
            for (int k = 0; k < ns; ++k)
                s += (der ? dphi[k] : phi[k]) * Phi_inv[k * ns + j];
            out[i * ns + j] = s;
        }
    }
    return 0;
}

int orc_kron(int d, const int* m, const int* n, const double* const* A, double* op)
{
    const int M0 = m[0], M1 = d > 1 ? m[1] : 1, M2 = d > 2 ? m[2] : 1;
    const int N0 = n[0], N1 = d > 1 ? n[1] : 1, N2 = d > 2 ? n[2] : 1;
    const int NC = N0 * N1 * N2;
    for (int i2 = 0; i2 < M2; ++i2)
        for (int i1 = 0; i1 < M1; ++i1)
            for (int i0 = 0; i0 < M0; ++i0)
                for (int j2 = 0; j2 < N2; ++j2)
                    for (int j1 = 0; j1 < N1; ++j1)
                        for (int j0 = 0; j0 < N0; ++j0) {
                            double v = A[0][i0 * N0 + j0];
                            if (d > 1) v *= A[1][i1 * N1 + j1];
                            if (d > 2) v *= A[2][i2 * N2 + j2];
                            op[(i0 + M0 * (i1 + M1 * i2)) * NC + (j0 + N0 * (j1 + N1 * j2))] = v;
                        }
    return 0;
}

static double binom(int n, int k)
{
    double c = 1.0;
    for (int i = 1; i <= k; ++i)
        c = c * (n - k + i) / i;
    return c;
}

static int ipow(int b, int e)
{
    int r = 1;
    for (int i = 0; i < e; ++i)
        r *= b;
    return r;
}

/* Local face -> (face, side) of every volume, for the threaded residual. */
static void build_elem_faces(orc_ctx* c)
{
    if (c->ef_face)
        return;
    const int nfe = 2 * c->d;
    c->ef_face = (int64_t*)malloc(sizeof(int64_t) * (size_t)c->n_elem * nfe);
    c->ef_side = (int8_t*)malloc((size_t)c->n_elem * nfe);
    for (int64_t fi = 0; fi < c->n_face; ++fi)
        for (int side = 0; side < 2; ++side) {
            const int64_t e = c->face_elem[2 * fi + side];
            const int lf = c->face_lf[2 * fi + side];
            c->ef_face[e * nfe + lf] = fi;
            c->ef_side[e * nfe + lf] = (int8_t)side;
        }
}

/* ------------------------------------------------------------------------------------------------ */
/* templated residual (double and double complex)                                                    */
/* ------------------------------------------------------------------------------------------------ */
static void build_minv(orc_ctx* c);


#define T double
#define SFX(name) name##_d
#define RE(x) (x)
#define ABSV(x) fabs(x)
#define SQRTT(x) sqrt(x)
#include "oracle_residual.inc"
#undef T
#undef SFX
#undef RE
#undef ABSV
#undef SQRTT

#define T double complex
#define SFX(name) name##_c
#define RE(x) creal(x)
#define ABSV(x) cabs(x)
#define SQRTT(x) csqrt(x)
#include "oracle_residual.inc"
#undef T
#undef SFX
#undef RE
#undef ABSV
#undef SQRTT

void orc_flux_point(int d, const double* u, double* f)
{
    flux_euler_d(d, u, f);
}

void orc_numflux_point(int d, int num_flux, const double* ul, const double* ur, const double* n, double* nf)
{
    if (num_flux == 0)
        numflux_lf_d(d, ul, ur, n, nf);
    else
        numflux_roe_d(d, ul, ur, n, nf);
}

/* Batched forms over n points (the mixed-mesh restatement evaluates every cubature point of a group at
 * once): u [n][nv], f [n][nv][d] (the pointwise layout), ul/ur [n][nv], nrm [n][d], nf [n][nv]. */
void orc_flux_batch(int d, int64_t n, const double* u, double* f)
{
    const int nv = d + 2;
    for (int64_t i = 0; i < n; ++i)
        flux_euler_d(d, u + i * nv, f + i * nv * d);
}

void orc_numflux_batch(int d, int num_flux, int64_t n, const double* ul, const double* ur, const double* nrm,
                       double* nf)
{
    const int nv = d + 2;
    for (int64_t i = 0; i < n; ++i)
        orc_numflux_point(d, num_flux, ul + i * nv, ur + i * nv, nrm + i * d, nf + i * nv);
}


/* ------------------------------------------------------------------------------------------------ */
/* context: operators, geometry, mass matrices                                                      */
/* ------------------------------------------------------------------------------------------------ */

/* reference normals of the QUAD/HEX faces (src/element/element.c:708-779) */
static void reference_normal(int d, int lf, double* nr)
{
    static const double te_q[] = { 0.0, 0.0, 0.0, 0.0 };
    static const double tz_q[] = { ORC_PI, 0.0, 1.5 * ORC_PI, 0.5 * ORC_PI };
    static const double te_h[] = { 0.0, 0.0, 0.0, 0.0, 0.5 * ORC_PI, 1.5 * ORC_PI };
    static const double tz_h[] = { ORC_PI, 0.0, 1.5 * ORC_PI, 0.5 * ORC_PI, 0.0, ORC_PI };
    const double te = (d == 2) ? te_q[lf] : te_h[lf], tz = (d == 2) ? tz_q[lf] : tz_h[lf];
    nr[0] = cos(te) * cos(tz);
    nr[1] = cos(te) * sin(tz);
    if (d > 2)
        nr[2] = -sin(te);
}

static void dense_apply(int m, int k, const double* A, const double* x, double* y)
{
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int q = 0; q < k; ++q)
            s += A[i * k + q] * x[q];
        y[i] = s;
    }
}

static void build_geometry(orc_ctx* c, const double* xyz_g)
{
    const int d = c->d, ng = c->ng, nc = c->n1, Ng = c->Ng, Nc = c->Nc, Nf = c->Nf;
    double xg[16], wg[16], xc[16], wc[16];
    orc_nodes_1d(1, c->p_g, xg, wg);
    orc_nodes_1d(0, c->p, xc, wc);
    double I_gg[256], D_gg[256], I_gc[256], D_gc[256], I_gm[16], I_gp[16];
    const double m1 = -1.0, p1 = 1.0;
    orc_lagrange_1d(ng, xg, ng, xg, 0, I_gg);
    orc_lagrange_1d(ng, xg, ng, xg, 1, D_gg);
    orc_lagrange_1d(ng, xg, nc, xc, 0, I_gc);
    orc_lagrange_1d(ng, xg, nc, xc, 1, D_gc);
    orc_lagrange_1d(ng, xg, 1, &m1, 0, I_gm);
    orc_lagrange_1d(ng, xg, 1, &p1, 0, I_gp);

    /* dense geometry operators: cv1_vg_vm[j], cv1_vg_vc[j], vv0_vm_vc, vv0_vm_fc[lf] */
    double* cv1_vg_vm[3];
    double* cv1_vg_vc[3];
    double* vv0_vm_fc[6];
    const int mg[3] = { ng, ng, ng }, mc[3] = { nc, nc, nc };
    for (int j = 0; j < d; ++j) {
        const double* A[3] = { I_gg, I_gg, I_gg };
        A[j] = D_gg;
        cv1_vg_vm[j] = (double*)malloc(sizeof(double) * Ng * Ng);
        orc_kron(d, mg, mg, A, cv1_vg_vm[j]);
        const double* B[3] = { I_gc, I_gc, I_gc };
        B[j] = D_gc;
        cv1_vg_vc[j] = (double*)malloc(sizeof(double) * Nc * Ng);
        orc_kron(d, mc, mg, B, cv1_vg_vc[j]);
    }
    double* vv0_vm_vc = (double*)malloc(sizeof(double) * Nc * Ng);
    {
        const double* A[3] = { I_gc, I_gc, I_gc };
        orc_kron(d, mc, mg, A, vv0_vm_vc);
    }
    double* vv0_vg_vm = (double*)malloc(sizeof(double) * Ng * Ng);
    {
        const double* A[3] = { I_gg, I_gg, I_gg };
        orc_kron(d, mg, mg, A, vv0_vg_vm);
    }
    for (int lf = 0; lf < 2 * d; ++lf) {
        const int dn = lf / 2;
        const double* A[3] = { I_gc, I_gc, I_gc };
        A[dn] = (lf % 2) ? I_gp : I_gm;
        int mf[3] = { nc, nc, nc };
        mf[dn] = 1;
        vv0_vm_fc[lf] = (double*)malloc(sizeof(double) * Nf * Ng);
        orc_kron(d, mf, mg, A, vv0_vm_fc[lf]);
    }

    double* metrics_vm = (double*)malloc(sizeof(double) * (size_t)c->n_elem * d * d * Ng);
    double J[9][512], Jc[9][512];
    for (int64_t e = 0; e < c->n_elem; ++e) {
        const double* X = xyz_g + e * d * Ng;
        double* mvm = metrics_vm + e * d * d * Ng;
        /* jacobian_vm(n, i, j) = cv1_vg_vm[j] . X_i (geometry_T.c:781-782) */
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j)
                dense_apply(Ng, Ng, cv1_vg_vm[j], X + i * Ng, J[i * 3 + j]);
        if (d == 2) {
            /* compute_cofactors_T (geometry_T.c:566-580) */
            for (int a = 0; a < Ng; ++a) {
                mvm[(0 + 2 * 0) * Ng + a] = J[4][a];
                mvm[(0 + 2 * 1) * Ng + a] = -J[3][a];
                mvm[(1 + 2 * 0) * Ng + a] = -J[1][a];
                mvm[(1 + 2 * 1) * Ng + a] = J[0][a];
            }
        } else {
            /* curl form (deprecated/src/setup_geom_factors.c:182-250): XYZ at the metric nodes, the
             * products X_a dX_b/dxi_k, their discrete curl and C = (curl(y grad z) - curl(z grad y))/2 ... */
            static const int order_curl[12] = { 1, 2, 2, 1, 2, 0, 0, 2, 0, 1, 1, 0 };
            double Xm[3][512], XJ[6][3][512], CJ[6][3][512], tmp[512];
            for (int i = 0; i < 3; ++i)
                dense_apply(Ng, Ng, vv0_vg_vm, X + i * Ng, Xm[i]);
            for (int i = 0; i < 6; ++i)
                for (int dim = 0; dim < 3; ++dim)
                    for (int q = 0; q < Ng; ++q)
                        XJ[i][dim][q] = Xm[order_curl[2 * i]][q] * J[order_curl[2 * i + 1] * 3 + dim][q];
            for (int i = 0; i < 6; ++i)
                for (int dim = 0; dim < 3; ++dim) {
                    const int* oc = &order_curl[4 * dim];
                    dense_apply(Ng, Ng, cv1_vg_vm[oc[0]], XJ[i][oc[2]], CJ[i][dim]);
                    dense_apply(Ng, Ng, cv1_vg_vm[oc[1]], XJ[i][oc[3]], tmp);
                    for (int q = 0; q < Ng; ++q)
                        CJ[i][dim][q] -= tmp[q];
                }
            for (int row = 0; row < 3; ++row)
                for (int col = 0; col < 3; ++col)
                    for (int q = 0; q < Ng; ++q)
                        mvm[(row + 3 * col) * Ng + q] = 0.5 * (CJ[2 * row][col][q] - CJ[2 * row + 1][col][q]);
        }
        /* metrics_vc = vv0_vm_vc . metrics_vm ; det J from cv1_vg_vc (geometry_T.c:784-801) */
        for (int cc = 0; cc < d * d; ++cc)
            dense_apply(Nc, Ng, vv0_vm_vc, mvm + cc * Ng, c->metrics_vc + (e * d * d + cc) * Nc);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j)
                dense_apply(Nc, Ng, cv1_vg_vc[j], X + i * Ng, Jc[i * 3 + j]);
        for (int q = 0; q < Nc; ++q) {
            double det;
            if (d == 2)
                det = Jc[0][q] * Jc[4][q] - Jc[1][q] * Jc[3][q];
            else
                det = Jc[0][q] * (Jc[4][q] * Jc[8][q] - Jc[5][q] * Jc[7][q]) -
                      Jc[1][q] * (Jc[3][q] * Jc[8][q] - Jc[5][q] * Jc[6][q]) +
                      Jc[2][q] * (Jc[3][q] * Jc[7][q] - Jc[4][q] * Jc[6][q]);
            c->jdet_vc[e * Nc + q] = det;
        }
    }
    /* faces: metrics_fc from the left volume, normals = normalize(C n_ref) (geometry_T.c:274-325, 412-445) */
    for (int64_t f = 0; f < c->n_face; ++f) {
        const int64_t eL = c->face_elem[2 * f];
        const int lf = c->face_lf[2 * f];
        double nr[3];
        reference_normal(d, lf, nr);
        double mfc[9][512];
        for (int cc = 0; cc < d * d; ++cc)
            dense_apply(Nf, Ng, vv0_vm_fc[lf], metrics_vm + (eL * d * d + cc) * Ng, mfc[cc]);
        for (int q = 0; q < Nf; ++q) {
            double nv[3] = { 0, 0, 0 }, s = 0.0;
            for (int i = 0; i < d; ++i) {
                for (int j = 0; j < d; ++j)
                    nv[i] += nr[j] * mfc[i + d * j][q];
                s += nv[i] * nv[i];
            }
            s = sqrt(s);
            c->jdet_fc[f * Nf + q] = s;
            for (int i = 0; i < d; ++i)
                c->normals_fc[(f * Nf + q) * d + i] = nv[i] / s;
        }
    }
    for (int j = 0; j < d; ++j) {
        free(cv1_vg_vm[j]);
        free(cv1_vg_vc[j]);
    }
    for (int lf = 0; lf < 2 * d; ++lf)
        free(vv0_vm_fc[lf]);
    free(vv0_vm_vc);
    free(vv0_vg_vm);
    free(metrics_vm);
}

orc_ctx* orc_create(int d, int p, int p_g, int64_t n_elem, int64_t n_face, const double* xyz_g,
                    const int32_t* face_elem, const int8_t* face_lf, int num_flux, int positivity)
{
    if ((d != 2 && d != 3) || p < 1 || p > 7 || p_g < 1 || p_g > 7 || n_elem < 1 || !face_elem || !face_lf)
        return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->d = d; c->p = p; c->p_g = p_g; c->num_flux = num_flux; c->positivity = positivity;
    c->n1 = p + 1; c->ng = p_g + 1;
    c->Ns = ipow(c->n1, d); c->Nc = c->Ns; c->Nf = ipow(c->n1, d - 1); c->Ng = ipow(c->ng, d);
    c->n_elem = n_elem; c->n_face = n_face;
    c->face_elem = (int32_t*)malloc(sizeof(int32_t) * 2 * n_face);
    c->face_lf = (int8_t*)malloc(2 * n_face);
    memcpy(c->face_elem, face_elem, sizeof(int32_t) * 2 * n_face);
    memcpy(c->face_lf, face_lf, 2 * n_face);
    const int n = c->n1, Ns = c->Ns, Nc = c->Nc, Nf = c->Nf;

    /* LINE operators for the solution basis (Lagrange on GL), GL cubature of p+1 points
     * (nodes_operators.c:329-347, 587-605 with p_c_x = 2, p_c_p = 0) */
    double xs[16], ws[16];
    orc_nodes_1d(0, p, xs, ws);
    double cv0[256], cv1[256], tw0[256], tw1[256], fm[16], fp[16], B[256], Binv[256];
    orc_lagrange_1d(n, xs, n, xs, 0, cv0);
    orc_lagrange_1d(n, xs, n, xs, 1, cv1);
    for (int a = 0; a < n; ++a)
        for (int q = 0; q < n; ++q) {
            tw0[a * n + q] = cv0[q * n + a] * ws[q];       /* tw0 = cv0^T diag(w) */
            tw1[a * n + q] = cv1[q * n + a] * ws[q];       /* tw1 = cv1^T diag(w) */
        }
    const double m1 = -1.0, p1 = 1.0;
    orc_lagrange_1d(n, xs, 1, &m1, 0, fm);
    orc_lagrange_1d(n, xs, 1, &p1, 0, fp);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double t = 0.5 * (1.0 + xs[i]);
            B[i * n + k] = binom(p, k) * pow(t, k) * pow(1.0 - t, p - k);
        }
    invert(n, B, Binv);

    memcpy(c->f_cv0, cv0, sizeof(double) * n * n);
    memcpy(c->f_cv1, cv1, sizeof(double) * n * n);
    memcpy(c->f_tw0, tw0, sizeof(double) * n * n);
    memcpy(c->f_tw1, tw1, sizeof(double) * n * n);
    memcpy(c->f_fm, fm, sizeof(double) * n);
    memcpy(c->f_fp, fp, sizeof(double) * n);
    memcpy(c->f_B, B, sizeof(double) * n * n);
    memcpy(c->f_Binv, Binv, sizeof(double) * n * n);
    const int ms[3] = { n, n, n };
    c->cv0_vs_vc = (double*)malloc(sizeof(double) * Nc * Ns);
    {
        const double* A[3] = { cv0, cv0, cv0 };
        orc_kron(d, ms, ms, A, c->cv0_vs_vc);
    }
    for (int j = 0; j < d; ++j) {
        const double* A[3] = { tw0, tw0, tw0 };
        A[j] = tw1;
        c->tw1_vt_vc[j] = (double*)malloc(sizeof(double) * Ns * Nc);
        orc_kron(d, ms, ms, A, c->tw1_vt_vc[j]);
        const double* B[3] = { cv0, cv0, cv0 };
        B[j] = cv1;
        c->cv1_vs_vc[j] = (double*)malloc(sizeof(double) * Nc * Ns);
        orc_kron(d, ms, ms, B, c->cv1_vs_vc[j]);
    }
    {
        const double* A[3] = { tw0, tw0, tw0 };
        c->tw0_vt_vc = (double*)malloc(sizeof(double) * Ns * Nc);
        orc_kron(d, ms, ms, A, c->tw0_vt_vc);
    }
    for (int lf = 0; lf < 2 * d; ++lf) {
        const int dn = lf / 2;
        const double* A[3] = { cv0, cv0, cv0 };
        A[dn] = (lf % 2) ? fp : fm;
        int mf[3] = { n, n, n };
        mf[dn] = 1;
        c->cv0_vs_fc[lf] = (double*)malloc(sizeof(double) * Nf * Ns);
        orc_kron(d, mf, ms, A, c->cv0_vs_fc[lf]);
        /* tw0_vt_fc = cv0_vt_fc^T diag(w_fc), w_fc = Kronecker of the face-tangential weights */
        c->tw0_vt_fc[lf] = (double*)malloc(sizeof(double) * Ns * Nf);
        for (int a = 0; a < Ns; ++a)
            for (int q = 0; q < Nf; ++q) {
                double wq = 1.0;
                int rem = q, t = 0;
                for (int k = 0; k < d; ++k) {
                    if (k == dn)
                        continue;
                    const int iq = rem % n;
                    rem /= n;
                    wq *= ws[iq];
                    ++t;
                }
                c->tw0_vt_fc[lf][a * Nf + q] = c->cv0_vs_fc[lf][q * Ns + a] * wq;
            }
    }
    c->ccSB0 = (double*)malloc(sizeof(double) * Ns * Ns);
    c->ccBS0 = (double*)malloc(sizeof(double) * Ns * Ns);
    {
        const double* A[3] = { Binv, Binv, Binv };
        orc_kron(d, ms, ms, A, c->ccSB0);
        const double* Bk[3] = { B, B, B };
        orc_kron(d, ms, ms, Bk, c->ccBS0);
    }
    c->w_vc = (double*)malloc(sizeof(double) * Nc);
    for (int q = 0; q < Nc; ++q) {
        double wq = 1.0;
        int rem = q;
        for (int k = 0; k < d; ++k) {
            wq *= ws[rem % n];
            rem /= n;
        }
        c->w_vc[q] = wq;
    }

    c->metrics_vc = (double*)calloc((size_t)n_elem * d * d * Nc, sizeof(double));
    c->jdet_vc = (double*)calloc((size_t)n_elem * Nc, sizeof(double));
    c->normals_fc = (double*)calloc((size_t)n_face * Nf * d, sizeof(double));
    c->jdet_fc = (double*)calloc((size_t)n_face * Nf, sizeof(double));
    c->m_inv = (double*)calloc((size_t)n_elem * Ns * Ns, sizeof(double));
    if (xyz_g)
        build_geometry(c, xyz_g);
    c->minv_ready = 0;   /* m = V^T diag(w J) V ; m_inv = m^-1 (volume_solver_T.c:115-148), built lazily */
    return c;
}

void orc_destroy(orc_ctx* c)
{
    if (!c)
        return;
    free(c->face_elem); free(c->face_lf); free(c->face_ord);
    free(c->minv_diag);
    free(c->cv0_vs_vc);
    for (int j = 0; j < 3; ++j) {
        free(c->tw1_vt_vc[j]);
        free(c->cv1_vs_vc[j]);
    }
    for (int lf = 0; lf < 6; ++lf) {
        free(c->cv0_vs_fc[lf]);
        free(c->tw0_vt_fc[lf]);
    }
    free(c->tw0_vt_vc);
    free(c->ccSB0); free(c->ccBS0); free(c->w_vc);
    free(c->metrics_vc); free(c->jdet_vc); free(c->normals_fc); free(c->jdet_fc); free(c->m_inv);
    free(c->ef_face); free(c->ef_side);
    free(c);
}

void orc_geometry(const orc_ctx* c, double* metrics_vc, double* jdet_vc, double* normals_fc, double* jdet_fc)
{
    const int d = c->d;
    if (metrics_vc)
        memcpy(metrics_vc, c->metrics_vc, sizeof(double) * (size_t)c->n_elem * d * d * c->Nc);
    if (jdet_vc)
        memcpy(jdet_vc, c->jdet_vc, sizeof(double) * (size_t)c->n_elem * c->Nc);
    if (normals_fc)
        memcpy(normals_fc, c->normals_fc, sizeof(double) * (size_t)c->n_face * c->Nf * d);
    if (jdet_fc)
        memcpy(jdet_fc, c->jdet_fc, sizeof(double) * (size_t)c->n_face * c->Nf);
}

static void build_minv(orc_ctx* c)
{
    const int Ns = c->Ns, Nc = c->Nc;
    if (c->sf) {
        if (!c->minv_diag)
            c->minv_diag = (double*)malloc(sizeof(double) * (size_t)c->n_elem * Ns);
        for (int64_t e = 0; e < c->n_elem; ++e)
            for (int i = 0; i < Ns; ++i)
                c->minv_diag[e * Ns + i] = 1.0 / (c->w_vc[i] * c->jdet_vc[e * Nc + i]);
        c->minv_ready = 1;
        return;
    }
    double* m = (double*)malloc(sizeof(double) * Ns * Ns);
    for (int64_t e = 0; e < c->n_elem; ++e) {
        for (int i = 0; i < Ns; ++i)
            for (int j = 0; j < Ns; ++j) {
                double s = 0.0;
                for (int q = 0; q < Nc; ++q)
                    s += c->cv0_vs_vc[q * Ns + i] * c->w_vc[q] * c->jdet_vc[e * Nc + q] * c->cv0_vs_vc[q * Ns + j];
                m[i * Ns + j] = s;
            }
        invert(Ns, m, c->m_inv + (size_t)e * Ns * Ns);
    }
    free(m);
    c->minv_ready = 1;
}

void orc_set_geometry(orc_ctx* c, const double* metrics_vc, const double* jdet_vc, const double* normals_fc,
                      const double* jdet_fc)
{
    const int d = c->d;
    memcpy(c->metrics_vc, metrics_vc, sizeof(double) * (size_t)c->n_elem * d * d * c->Nc);
    memcpy(c->jdet_vc, jdet_vc, sizeof(double) * (size_t)c->n_elem * c->Nc);
    memcpy(c->normals_fc, normals_fc, sizeof(double) * (size_t)c->n_face * c->Nf * d);
    memcpy(c->jdet_fc, jdet_fc, sizeof(double) * (size_t)c->n_face * c->Nf);
    c->minv_ready = 0;
}

void orc_gradient(orc_ctx* c, const double* u, double* grad_v, double* grad)
{
    double* grad_f = (double*)malloc(sizeof(double) * (size_t)c->Ns * (c->d + 2) * c->d * 2 * c->n_face);
    gradients_d(c, u, grad_v, grad_f, grad);
    free(grad_f);
}

void orc_set_viscous(orc_ctx* c, double mu, double Pr)
{
    c->ns = 1;
    c->mu = mu;
    c->Pr = Pr;
}

void orc_set_penalty(orc_ctx* c, double eta)
{
    c->eta = eta;
}

int orc_operator(const orc_ctx* c, const char* name, double* out)
{
    const int Ns = c->Ns, Nc = c->Nc, Nf = c->Nf;
    const double* src = NULL;
    int cnt = 0;
    if (!strcmp(name, "cv0_vs_vc")) { src = c->cv0_vs_vc; cnt = Nc * Ns; }
    else if (!strcmp(name, "ccSB0")) { src = c->ccSB0; cnt = Ns * Ns; }
    else if (!strcmp(name, "ccBS0")) { src = c->ccBS0; cnt = Ns * Ns; }
    else if (!strcmp(name, "w_vc")) { src = c->w_vc; cnt = Nc; }
    else if (!strncmp(name, "tw1_vt_vc", 9)) {
        const int j = name[9] - '0';
        if (j < 0 || j >= c->d) return -1;
        src = c->tw1_vt_vc[j]; cnt = Ns * Nc;
    } else if (!strncmp(name, "cv0_vs_fc", 9)) {
        const int lf = name[9] - '0';
        if (lf < 0 || lf >= 2 * c->d) return -1;
        src = c->cv0_vs_fc[lf]; cnt = Nf * Ns;
    } else if (!strncmp(name, "tw0_vt_fc", 9)) {
        const int lf = name[9] - '0';
        if (lf < 0 || lf >= 2 * c->d) return -1;
        src = c->tw0_vt_fc[lf]; cnt = Ns * Nf;
    } else
        return -1;
    memcpy(out, src, sizeof(double) * cnt);
    return cnt;
}

/* ------------------------------------------------------------------------------------------------ */
/* residual entry points                                                                             */
/* ------------------------------------------------------------------------------------------------ */

double orc_residual(orc_ctx* c, const double* u, double* rhs, int apply_minv)
{
    return residual_d(c, u, rhs, apply_minv);
}

void orc_jvp_cs(orc_ctx* c, const double* u, const double* w, double h, double* out, int apply_minv)
{
    const int64_t n = c->n_elem * (c->d + 2) * c->Ns;
    double complex* uc = (double complex*)calloc((size_t)n, sizeof(double complex));
    double complex* rc = (double complex*)malloc(sizeof(double complex) * n);
    for (int64_t i = 0; i < n; ++i)
        uc[i] = u[i] + I * h * w[i];
    residual_c(c, uc, rc, apply_minv);
    for (int64_t i = 0; i < n; ++i)
        out[i] = cimag(rc[i]) / h;
    free(uc);
    free(rc);
}

/* ------------------------------------------------------------------------------------------------ */
/* positivity and SSP-RK3                                                                            */
/* ------------------------------------------------------------------------------------------------ */

/* convert_variables_T (solution_euler_T.c:136-200) on an (n, nv) column-major block */
static void cons_to_prim(int d, int n, double* s)
{
    const int nv = d + 2;
    for (int i = 0; i < n; ++i) {
        const double rho = s[i], rho_inv = 1.0 / rho;
        double rho2V2 = 0.0;
        for (int k = 0; k < d; ++k) {
            rho2V2 += s[(1 + k) * n + i] * s[(1 + k) * n + i];
            s[(1 + k) * n + i] *= rho_inv;
        }
        s[(nv - 1) * n + i] = GM1 * (s[(nv - 1) * n + i] - 0.5 * rho2V2 * rho_inv);
    }
}

static void prim_to_cons(int d, int n, double* s)
{
    const int nv = d + 2;
    for (int i = 0; i < n; ++i) {
        const double rho = s[i];
        double V2 = 0.0;
        for (int k = 0; k < d; ++k) {
            const double uk = s[(1 + k) * n + i];
            V2 += uk * uk;
            s[(1 + k) * n + i] = rho * uk;
        }
        s[(nv - 1) * n + i] = s[(nv - 1) * n + i] / GM1 + 0.5 * rho * V2;
    }
}

/* enforce_positivity_highorder for one element (solve.c:187-233; correct_coef solve.c:356-374) */
static int positivity_elem(const orc_ctx* c, double* ue, double* sb)
{
    const int d = c->d, nv = d + 2, Ns = c->Ns;
    int ok = 1;
    if (c->sf) {
        const double* F[3] = { c->f_Binv, c->f_Binv, c->f_Binv };
        const int m3[3] = { c->n1, c->n1, c->n1 };
        kron_apply_d(d, m3, m3, F, nv, 1.0, ue, Ns, sb, Ns, 0);
    } else {
        for (int v = 0; v < nv; ++v)
            dense_apply(Ns, Ns, c->ccSB0, ue + v * Ns, sb + v * Ns);
    }
    cons_to_prim(d, Ns, sb);
    for (int vr = 0; vr < nv; vr += nv - 1) {
        const double* vd = sb + vr * Ns;
        double avg = 0.0, mn = vd[0];
        for (int q = 0; q < Ns; ++q) {
            avg += vd[q];
            if (vd[q] < mn)
                mn = vd[q];
        }
        avg /= Ns;
        if (avg < EPS_PHYS) {
            ok = 0;
            break;
        }
        if (mn < EPS_PHYS) {
            double p_ho = (avg - EPS_PHYS) / (avg - mn);
            p_ho = p_ho < 1.0 ? p_ho : 1.0;
            p_ho = p_ho > 0.0 ? p_ho : 0.0;
            for (int v = 0; v < nv; ++v) {
                double* dv = sb + v * Ns;
                double a = 0.0;
                for (int q = 0; q < Ns; ++q)
                    a += dv[q];
                a /= Ns;
                if ((v == 0 || v == nv - 1) && a < EPS_PHYS)
                    a = EPS_PHYS;
                for (int q = 0; q < Ns; ++q)
                    dv[q] = p_ho * dv[q] + (1 - p_ho) * a;
            }
        }
    }
    if (!ok)
        return 0;
    prim_to_cons(d, Ns, sb);
    if (c->sf) {
        const double* F[3] = { c->f_B, c->f_B, c->f_B };
        const int m3[3] = { c->n1, c->n1, c->n1 };
        kron_apply_d(d, m3, m3, F, nv, 1.0, sb, Ns, ue, Ns, 0);
    } else {
        for (int v = 0; v < nv; ++v)
            dense_apply(Ns, Ns, c->ccBS0, sb + v * Ns, ue + v * Ns);
    }
    return 1;
}

int64_t orc_positivity(orc_ctx* c, double* u)
{
    const int nv = c->d + 2, Ns = c->Ns;
    double* sb = (double*)malloc(sizeof(double) * Ns * nv);
    int64_t bad = -1;
    for (int64_t e = 0; e < c->n_elem; ++e)
        if (!positivity_elem(c, u + e * nv * Ns, sb) && bad < 0)
            bad = e;
    free(sb);
    return bad;
}

/* SSP-RK3 combination (time_step_ssp_rk_33, solve_explicit.c:271-293) and positivity of volume e; 0 when
 * the volume fails the positivity check. */
/* Low-storage RK 5-stage 4th order: Carpenter & Kennedy 1994, solution 3 of table 1, rational form
 * (the coefficient table of time_step_ls_rk_54, solve_explicit.c:310-321). */
static const double k_ls54_a[5] = { 0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                    -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0 };
static const double k_ls54_b[5] = { 1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                    17201463215490.0 / 20902069494980.0, 3134564353537.0 / 4481467310338.0,
                                    2277821191437.0 / 14882151754819.0 };

/* Stage update of one volume, then enforce_positivity_highorder:
 *   SSP-RK3  time_step_ssp_rk_33 (solve_explicit.c:252-301, Gottlieb 2001 eq. 4.2)
 *   Euler    time_step_euler (solve_explicit.c:226-250)
 *   LS-RK54  time_step_ls_rk_54 (solve_explicit.c:303-350): sp = a sp + dt r, s += b sp */
static int rk_elem(const orc_ctx* c, int stage, double dt, int64_t e, double* u, double* u_p, const double* rhs,
                   double* sb)
{
    const int nv = c->d + 2, Ns = c->Ns;
    double* s = u + e * nv * Ns;
    double* sp = u_p + e * nv * Ns;
    const double* r = rhs + e * nv * Ns;
    for (int i = 0; i < nv * Ns; ++i) {
        if (c->time_stepper == 1) {
            s[i] += dt * r[i];
        } else if (c->time_stepper == 2) {
            sp[i] *= k_ls54_a[stage];
            sp[i] += dt * r[i];
            s[i] += k_ls54_b[stage] * sp[i];
        } else if (stage == 0) {
            sp[i] = s[i];
            s[i] += dt * r[i];
        } else if (stage == 1) {
            s[i] = (1.0 / 4.0) * (3.0 * sp[i] + s[i] + dt * r[i]);
        } else {
            s[i] = (1.0 / 3.0) * (sp[i] + 2.0 * s[i] + 2.0 * dt * r[i]);
        }
    }
    return !c->positivity || positivity_elem(c, s, sb);
}

/* Face node correspondence nc_fc[ord] of a LINE (d_face = 1, n1 nodes) or QUAD (d_face = 2, n1 x n1 nodes)
 * face, as constructor_face_corr builds it (src/element/nodes_correspondence.c:87-160): QUAD cases 4-7 start
 * from the transposed ordering, cases 1, 3, 6, 7 reverse every 1-D block, cases 2, 3, 5, 7 swap the blocks
 * end for end; a LINE reverses for case 1. */
int orc_face_corr(int d_face, int n1, int ord, int32_t* out)
{
    if (d_face == 1) {
        if (ord < 0 || ord > 1)
            return 1;
        for (int i = 0; i < n1; ++i)
            out[i] = ord ? n1 - 1 - i : i;
        return 0;
    }
    if (d_face != 2 || ord < 0 || ord > 7)
        return 1;
    const int n = n1 * n1;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
            out[i * n1 + j] = ord >= 4 ? i + j * n1 : i * n1 + j;
    if (ord == 1 || ord == 3 || ord == 6 || ord == 7)
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1 / 2; ++j) {
                const int t = out[i * n1 + j];
                out[i * n1 + j] = out[i * n1 + n1 - 1 - j];
                out[i * n1 + n1 - 1 - j] = t;
            }
    if (ord == 2 || ord == 3 || ord == 5 || ord == 7)
        for (int i = 0; i < n1 / 2; ++i)
            for (int j = 0; j < n1; ++j) {
                const int t = out[i * n1 + j];
                out[i * n1 + j] = out[(n1 - 1 - i) * n1 + j];
                out[(n1 - 1 - i) * n1 + j] = t;
            }
    (void)n;
    return 0;
}

int orc_set_face_ord(orc_ctx* c, const int8_t* face_ord)
{
    if (c->ns && face_ord)
        return 2;   /* the Navier-Stokes gradient path is restated for aligned faces only */
    free(c->face_ord);
    c->face_ord = NULL;
    if (!face_ord)
        return 0;
    const int maxo = (c->d == 2) ? 1 : 7;
    for (int64_t f = 0; f < c->n_face; ++f)
        if (face_ord[f] < 0 || face_ord[f] > maxo)
            return 1;
    c->face_ord = (int8_t*)malloc((size_t)c->n_face);
    memcpy(c->face_ord, face_ord, (size_t)c->n_face);
    return 0;
}

int orc_set_time_stepper(orc_ctx* c, int time_stepper)
{
    if (time_stepper < 0 || time_stepper > 2)
        return 1;
    c->time_stepper = time_stepper;
    return 0;
}

int orc_n_stages(const orc_ctx* c)
{
    return c->time_stepper == 1 ? 1 : (c->time_stepper == 2 ? 5 : 3);
}

int orc_set_sum_factorized(orc_ctx* c, int on)
{
    if (on) {
        /* the diagonal M^-1 needs the GL collocation: the Lagrange basis at its own nodes (cv0 = I) */
        for (int i = 0; i < c->n1; ++i)
            for (int k = 0; k < c->n1; ++k)
                if (fabs(c->f_cv0[i * c->n1 + k] - (i == k ? 1.0 : 0.0)) > 1e-14)
                    return -1;
    }
    c->sf = on ? 1 : 0;
    c->minv_ready = 0;
    return 0;
}

void orc_set_threads(orc_ctx* c, int nthreads)
{
    c->nthreads = nthreads > 1 ? nthreads : 1;
}

double orc_rk_stage(orc_ctx* c, int stage, double dt, double* u, double* u_p, int64_t* bad_elem)
{
    const int nv = c->d + 2, Ns = c->Ns;
    const int64_t n = c->n_elem * nv * Ns;
    double* rhs = (double*)malloc(sizeof(double) * n);
    double* sb = (double*)malloc(sizeof(double) * Ns * nv);
    const double max_rhs = residual_d(c, u, rhs, 1);
    int64_t bad = -1;
    if (c->nthreads <= 1) {
        for (int64_t e = 0; e < c->n_elem; ++e)
            if (!rk_elem(c, stage, dt, e, u, u_p, rhs, sb) && bad < 0)
                bad = e;
    } else {
        /* threaded (timing baseline only): the first failing volume is the smallest failing index */
        int64_t first = c->n_elem;
#pragma omp parallel num_threads(c->nthreads)
        {
            double* sbt = (double*)malloc(sizeof(double) * Ns * nv);
            int64_t mine = c->n_elem;
#pragma omp for schedule(static)
            for (int64_t e = 0; e < c->n_elem; ++e)
                if (!rk_elem(c, stage, dt, e, u, u_p, rhs, sbt) && e < mine)
                    mine = e;
#pragma omp critical
            {
                if (mine < first)
                    first = mine;
            }
            free(sbt);
        }
        bad = first < c->n_elem ? first : -1;
    }
    if (bad_elem)
        *bad_elem = bad;
    free(rhs);
    free(sb);
    return max_rhs;
}
