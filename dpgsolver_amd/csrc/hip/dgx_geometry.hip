// Geometry set-up on the GPU: the metric terms, det J, face normals and face det J that
// dgx_geometry_build computes on the host (geometry_T.c:208-325,734-804; 3-D cofactors in the
// conservative curl form of deprecated/src/setup_geom_factors.c:182-250), one workgroup per volume.
//
// The volume's geometry nodes (element-local coordinates, as on the host) sit in LDS; each thread owns
// one geometry node for the derivative and curl phases and one cubature node for the interpolation
// phase.  Faces are computed by their left volume from its metric terms (the host's grouping): the table
// vface[e][lf] holds the face whose left volume is e on its local face lf (-1 otherwise), scattered on
// the device from the face records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dgx.h"
#include "dgx_setup.h"

namespace dgx {
namespace {

constexpr int NMAX = 5;                 // geometry and solution orders up to 4
constexpr int VMAX = NMAX * NMAX * NMAX;
constexpr int BLOCK = 128;

struct GeoOps {
    int ng, nc, Ng, Nc, Nf;
    double Dg[NMAX * NMAX];             // d l_m / dx at the GLL geometry nodes, [a][m]
    double Igc[NMAX * NMAX];            // l_m at the GL cubature nodes, [n][m]
    double Dgc[NMAX * NMAX];            // d l_m / dx at the GL cubature nodes, [n][m]
    double Igm[NMAX], Igp[NMAX];        // l_m at -1 / +1
};

// d_j f at a geometry node: the line through it along j starts at `base` with stride `st`, the node is
// number `i` on that line
__device__ inline double grad_line(const GeoOps& g, const double* f, int i, int base, int st)
{
    double s = 0.0;
    for (int m = 0; m < g.ng; ++m)
        s = fma(g.Dg[i * NMAX + m], f[base + m * st], s);
    return s;
}

// one 1-D pass of the sum-factorised interpolation: out[p] = sum_m A[i][m] in[base + m * st], A = n_out x ng
__device__ inline double pass_line(const double* A, const double* in, int i, int base, int st, int nin)
{
    double s = 0.0;
    for (int m = 0; m < nin; ++m)
        s = fma(A[i * NMAX + m], in[base + m * st], s);
    return s;
}

// Wf[(lf * Ng + a) * Nf + q]: interpolation weight of geometry node a at face node q of local face lf
// (tangential directions in order), products of the 1-D operators (face_weight_kernel, once per build), a-major so
// that consecutive threads (face nodes) read consecutive weights.
template <int D, int VM>
__global__ __launch_bounds__(BLOCK) void geometry_kernel(GeoOps gk, const double* __restrict__ xyz_g, int64_t E,
                                                         const double* __restrict__ Wf,
                                                         const int32_t* __restrict__ vface,
                                                         double* __restrict__ met, double* __restrict__ jdet,
                                                         double* __restrict__ nrm, double* __restrict__ jdf,
                                                         int* __restrict__ bad, int64_t base)
{
    // geometry-node fields, then the two intermediate grids of the sum-factorised interpolation: the
    // pass-s output (18 fields) reuses sM and sX, which are no longer read once pass r is complete
    __shared__ double sbuf[33][VM];     // VM: the largest of the geometry and cubature grids
    double(*sX)[VM] = sbuf;             // [3]  element-local coordinates
    double(*sJ)[VM] = sbuf + 3;         // [9]  d_j X_i at the geometry nodes; pass-r output (with sV)
    double(*sV)[VM] = sbuf + 12;        // [6]  curl products; pass-r output
    double(*sM)[VM] = sbuf + 18;        // [9]  metric terms at the geometry nodes
    // pass-s output slot k: sM for k < 9, sX for 9 <= k < 12, the last 6 rows for 12 <= k < 18
#define SW(k) sbuf[(k) < 9 ? 18 + (k) : ((k) < 12 ? (k) - 9 : 15 + (k))]
    __shared__ GeoOps g;                // the 1-D operators, read with per-thread indices
    const int64_t e = base + blockIdx.x;
    const int t = threadIdx.x;
    if (t == 0)
        g = gk;
    __syncthreads();
    const int ng = gk.ng, nc = gk.nc, Ng = gk.Ng, Nc = gk.Nc, Nf = gk.Nf;
    const double* Xe = xyz_g + e * D * Ng;

    // thread t owns geometry node t (Ng <= BLOCK): its position on the three lines through it
    const bool gnode = t < Ng;
    int ia[3] = { t % ng, (t / ng) % ng, D == 3 ? t / (ng * ng) : 0 };
    const int gst[3] = { 1, ng, ng * ng };
    int gbase[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        gbase[j] = t - ia[j] * gst[j];

    // element-local coordinates: the curl form's products X_a dX_b at the element's own scale
    if (gnode)
        for (int i = 0; i < D; ++i)
            sX[i][t] = Xe[i * Ng + t] - Xe[i * Ng];
    __syncthreads();
    if (gnode)
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j)
                sJ[i * 3 + j][t] = grad_line(g, sX[i], ia[j], gbase[j], gst[j]);
    __syncthreads();

    if constexpr (D == 2) {
        if (gnode) {
            sM[0 + 2 * 0][t] = sJ[1 * 3 + 1][t];        //  y_s
            sM[0 + 2 * 1][t] = -sJ[1 * 3 + 0][t];       // -y_r
            sM[1 + 2 * 0][t] = -sJ[0 * 3 + 1][t];       // -x_s
            sM[1 + 2 * 1][t] = sJ[0 * 3 + 0][t];        //  x_r
        }
    } else {
        // C_ic = 1/2 [curl(X_a grad X_b) - curl(X_b grad X_a)]_c, (a, b) = (i+1, i+2) mod 3
        if (gnode)
            for (int c = 0; c < 9; ++c)
                sM[c][t] = 0.0;
        for (int i = 0; i < 3; ++i) {
            const int ia_ = (i + 1) % 3, ib_ = (i + 2) % 3;
            for (int s = 0; s < 2; ++s) {
                const int a = s ? ib_ : ia_, b = s ? ia_ : ib_;
                __syncthreads();                        // sV free (previous curl done)
                if (gnode)
                    for (int k = 0; k < 3; ++k)
                        sV[k][t] = sX[a][t] * sJ[b * 3 + k][t];
                __syncthreads();
                const double sg = s ? -0.5 : 0.5;
                if (gnode)
                    for (int c = 0; c < 3; ++c) {
                        const int k1 = (c + 1) % 3, k2 = (c + 2) % 3;
                        const double curl = grad_line(g, sV[k2], ia[k1], gbase[k1], gst[k1]) -
                                            grad_line(g, sV[k1], ia[k2], gbase[k2], gst[k2]);
                        sM[i + 3 * c][t] += sg * curl;
                    }
            }
        }
    }
    __syncthreads();

    // faces whose left volume is e, from the metric terms at the geometry nodes: thread t serves local
    // face t / Nf, node t % Nf, when e is that face's left volume
    for (int u = t; u < 2 * D * Nf; u += BLOCK) {
        const int lf = u / Nf, q = u - lf * Nf;
        const int64_t f = vface[e * 2 * D + lf];
        if (f < 0)
            continue;
        const int dn = lf / 2;
        const double sgn = (lf & 1) ? 1.0 : -1.0;
        double v[D];
#pragma unroll
        for (int i = 0; i < D; ++i)
            v[i] = 0.0;
        for (int a = 0; a < Ng; ++a) {
            const double w = Wf[((int64_t)lf * Ng + a) * Nf + q];
#pragma unroll
            for (int i = 0; i < D; ++i)
                v[i] = fma(w, sM[i + D * dn][a], v[i]);
        }
        double s2 = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            v[i] *= sgn;
            s2 += v[i] * v[i];
        }
        const double sn = sqrt(s2);
        jdf[f * Nf + q] = sn;
#pragma unroll
        for (int i = 0; i < D; ++i)
            nrm[(f * Nf + q) * D + i] = v[i] / sn;
    }

    // cubature values by sum factorisation (r, then s, then t, as the host build): the metric terms
    // interpolated, d_j X_i with the derivative factor in direction j
    double J[3][3];
    if constexpr (D == 2) {
        // pass r on the (nc, ng) grid: sJ[0..3] = I_r M_c, sV[0..1] = I_r X_i, sV[2..3] = D_r X_i
        const bool p1 = t < nc * ng;
        const int ir = t % nc, as = t / nc;
        if (p1) {
            for (int c = 0; c < 4; ++c)
                sJ[c][t] = pass_line(g.Igc, sM[c], ir, ng * as, 1, ng);
            for (int i = 0; i < 2; ++i) {
                sV[i][t] = pass_line(g.Igc, sX[i], ir, ng * as, 1, ng);
                sV[2 + i][t] = pass_line(g.Dgc, sX[i], ir, ng * as, 1, ng);
            }
        }
        __syncthreads();
        if (t < Nc) {
            const int jr = t % nc, js = t / nc;
            for (int c = 0; c < 4; ++c)
                met[(e * 4 + c) * Nc + t] = pass_line(g.Igc, sJ[c], js, jr, nc, ng);
            for (int i = 0; i < 2; ++i) {
                J[i][0] = pass_line(g.Igc, sV[2 + i], js, jr, nc, ng);
                J[i][1] = pass_line(g.Dgc, sV[i], js, jr, nc, ng);
            }
        }
    } else {
        // pass r on the (nc, ng, ng) grid: sJ[c] = I_r M_c, sV[i] = I_r X_i, sV[3 + i] = D_r X_i
        const bool p1 = t < nc * ng * ng;
        const int ir = t % nc, rest1 = t / nc;              // rest1 = as + ng * at
        if (p1) {
            for (int c = 0; c < 9; ++c)
                sJ[c][t] = pass_line(g.Igc, sM[c], ir, ng * rest1, 1, ng);
            for (int i = 0; i < 3; ++i) {
                sV[i][t] = pass_line(g.Igc, sX[i], ir, ng * rest1, 1, ng);
                sV[3 + i][t] = pass_line(g.Dgc, sX[i], ir, ng * rest1, 1, ng);
            }
        }
        __syncthreads();
        // pass s on the (nc, nc, ng) grid: SW(c) = I_s (I_r M_c); SW(9+i) = I_s I_r X_i, SW(12+i) = D_s I_r X_i,
        // SW(15+i) = I_s D_r X_i
        const bool p2 = t < nc * nc * ng;
        const int jr = t % nc, js = (t / nc) % nc, at = t / (nc * nc);
        if (p2) {
            const int base = jr + nc * ng * at;             // (jr, m, at) on the pass-r grid: + nc * m
            for (int c = 0; c < 9; ++c)
                SW(c)[t] = pass_line(g.Igc, sJ[c], js, base, nc, ng);
            for (int i = 0; i < 3; ++i) {
                SW(9 + i)[t] = pass_line(g.Igc, sV[i], js, base, nc, ng);
                SW(12 + i)[t] = pass_line(g.Dgc, sV[i], js, base, nc, ng);
                SW(15 + i)[t] = pass_line(g.Igc, sV[3 + i], js, base, nc, ng);
            }
        }
        __syncthreads();
        if (t < Nc) {
            const int kt = t / (nc * nc), base = t - kt * nc * nc;   // (jr, js, m) on the pass-s grid: + nc^2 m
            for (int c = 0; c < 9; ++c)
                met[(e * 9 + c) * Nc + t] = pass_line(g.Igc, SW(c), kt, base, nc * nc, ng);
            for (int i = 0; i < 3; ++i) {
                J[i][0] = pass_line(g.Igc, SW(15 + i), kt, base, nc * nc, ng);
                J[i][1] = pass_line(g.Igc, SW(12 + i), kt, base, nc * nc, ng);
                J[i][2] = pass_line(g.Dgc, SW(9 + i), kt, base, nc * nc, ng);
            }
        }
    }
    if (t < Nc) {
        double det;
        if constexpr (D == 2)
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else
            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        jdet[e * Nc + t] = det;
        if (!(det > 0.0))
            atomicOr(bad, 1);
    }
#undef SW
}

// Wf[(lf * Ng + a) * Nf + q] (see geometry_kernel), one thread per entry, the factors multiplied in
// direction order
__global__ void face_weight_kernel(GeoOps g, int D, double* __restrict__ wf)
{
    const int n = 2 * D * g.Ng * g.Nf;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n)
        return;
    const int ng = g.ng, nc = g.nc;
    const int q = idx % g.Nf, a = (idx / g.Nf) % g.Ng, lf = idx / (g.Nf * g.Ng);
    const int ia[3] = { a % ng, (a / ng) % ng, a / (ng * ng) };
    int iq[3] = { 0, 0, 0 }, r = q;
    for (int k = 0; k < D; ++k)
        if (k != lf / 2) {
            iq[k] = r % nc;
            r /= nc;
        }
    double v = 1.0;
    for (int k = 0; k < D; ++k)
        v *= (k == lf / 2) ? ((lf & 1) ? g.Igp : g.Igm)[ia[k]] : g.Igc[iq[k] * NMAX + ia[k]];
    wf[idx] = v;
}

// grid-stride fill (launched with at most 65536 workgroups)
__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// vface[e][lf] = f for the left side (face_elem[2f], face_lf[2f]) of every face; a record out of range or
// a (volume, local face) claimed twice sets bit 2 of *bad.  vface must hold -1 everywhere before.
__global__ void face_table_kernel(const int32_t* __restrict__ face_elem, const int8_t* __restrict__ face_lf,
                                  int64_t F, int64_t E, int D, int32_t* __restrict__ vface, int* __restrict__ bad)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F)
        return;
    const int64_t eL = face_elem[2 * f];
    const int lf = face_lf[2 * f];
    if (eL < 0 || eL >= E || lf < 0 || lf >= 2 * D) {
        atomicOr(bad, 2);
        return;
    }
    if (atomicCAS(&vface[eL * 2 * D + lf], -1, (int32_t)f) != -1)
        atomicOr(bad, 2);
}

// The whole device pipeline of one geometry build on `st` (no allocation, no synchronisation): the weight
// table into wf, the face table into vface, *bad cleared and then set (bit 1: det J <= 0, bit 2: bad
// face record), the metric terms, det J, normals and face det J into om, oj, on, of.
hipError_t run_geometry(int D, const GeoOps& g, const double* x, int64_t E, const int32_t* face_elem,
                        const int8_t* face_lf, int64_t F, double* wf, int32_t* vface, int* bad, double* om, double* oj,
                        double* on, double* of, hipStream_t st);

template <int D>
void launch_geometry(bool small, const GeoOps& g, const double* x, int64_t E, const double* wf, const int32_t* vface,
                     double* om, double* oj, double* on, double* of, int* bad, hipStream_t st)
{
    // at most 2^20 workgroups per dispatch (a dispatch holds at most 2^32 - 1 work-items)
    const int64_t kMaxGroups = int64_t(1) << 20;
    for (int64_t base = 0; base < E; base += kMaxGroups) {
        const unsigned cnt = (unsigned)((E - base < kMaxGroups) ? E - base : kMaxGroups);
        if (small)
            hipLaunchKernelGGL((geometry_kernel<D, 64>), dim3(cnt), dim3(BLOCK), 0, st, g, x, E, wf, vface, om, oj,
                               on, of, bad, base);
        else
            hipLaunchKernelGGL((geometry_kernel<D, VMAX>), dim3(cnt), dim3(BLOCK), 0, st, g, x, E, wf, vface, om, oj,
                               on, of, bad, base);
    }
}

hipError_t run_geometry(int D, const GeoOps& g, const double* x, int64_t E, const int32_t* face_elem,
                        const int8_t* face_lf, int64_t F, double* wf, int32_t* vface, int* bad, double* om, double* oj,
                        double* on, double* of, hipStream_t st)
{
    const int nw = 2 * D * g.Ng * g.Nf;
    hipLaunchKernelGGL(face_weight_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, g, D, wf);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, st, (int32_t*)bad, (int64_t)1, 0);
    const int64_t nv = E * 2 * D;
    const int64_t nb = (nv + 255) / 256;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, st, vface, nv, -1);
    if (F > 0)
        hipLaunchKernelGGL(face_table_kernel, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, st, face_elem, face_lf,
                           F, E, D, vface, bad);
    // grids of at most 64 points (orders up to 3 in 3-D, 7 in 2-D) use the smaller LDS footprint
    const int vm = std::max(g.Ng, std::max(g.Nc, g.nc * g.ng * (D == 3 ? g.ng : 1)));
    const bool small = vm <= 64;
    if (D == 2)
        launch_geometry<2>(small, g, x, E, wf, vface, om, oj, on, of, bad, st);
    else
        launch_geometry<3>(small, g, x, E, wf, vface, om, oj, on, of, bad, st);
    return hipGetLastError();
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// The 1-D operators of a build: GLL geometry nodes of order p_g, GL cubature of order p (host, stack only).
dgx_status make_geo_ops(int d, int p, int p_g, GeoOps* out)
{
    GeoOps& g = *out;
    std::memset(&g, 0, sizeof(g));
    g.ng = p_g + 1;
    g.nc = p + 1;
    g.Ng = g.ng * g.ng * (d == 3 ? g.ng : 1);
    g.Nc = g.nc * g.nc * (d == 3 ? g.nc : 1);
    g.Nf = d == 3 ? g.nc * g.nc : g.nc;
    double xg[8], wg[8], xc[8], wc[8], tmp[64];
    const double m1 = -1.0, p1 = 1.0;
    if (dgx_nodes_1d(DGX_NODES_GLL, p_g, xg, wg) != DGX_OK || dgx_nodes_1d(DGX_NODES_GL, p, xc, wc) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    const int ng = g.ng, nc = g.nc;
    dgx_lagrange_1d(ng, xg, ng, xg, 1, tmp);
    for (int a = 0; a < ng; ++a)
        for (int m = 0; m < ng; ++m)
            g.Dg[a * NMAX + m] = tmp[a * ng + m];
    dgx_lagrange_1d(ng, xg, nc, xc, 0, tmp);
    for (int n = 0; n < nc; ++n)
        for (int m = 0; m < ng; ++m)
            g.Igc[n * NMAX + m] = tmp[n * ng + m];
    dgx_lagrange_1d(ng, xg, nc, xc, 1, tmp);
    for (int n = 0; n < nc; ++n)
        for (int m = 0; m < ng; ++m)
            g.Dgc[n * NMAX + m] = tmp[n * ng + m];
    dgx_lagrange_1d(ng, xg, 1, &m1, 0, g.Igm);
    dgx_lagrange_1d(ng, xg, 1, &p1, 0, g.Igp);
    return DGX_OK;
}

}  // namespace
}  // namespace dgx

extern "C" dgx_status dgx_geometry_build_gpu(const dgx_mesh* mesh, int p, int ptr_kind, double* metrics_vc,
                                             double* jdet_vc, double* normals_fc, double* jdet_fc, double* kernel_ms)
{
    using namespace dgx;
    if (!mesh || !mesh->xyz_g || !mesh->face_elem || !mesh->face_lf || !metrics_vc || !jdet_vc ||
        !normals_fc || !jdet_fc || (mesh->d != 2 && mesh->d != 3) || p < 1 || p > NMAX - 1 ||
        mesh->p_g < 1 || mesh->p_g > NMAX - 1 || mesh->n_elem < 1 || mesh->n_elem > 0x7fffffff ||
        mesh->n_face < 0 || mesh->n_face > 0x7fffffff || (ptr_kind != DGX_PTR_HOST && ptr_kind != DGX_PTR_DEVICE))
        return DGX_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return DGX_ERR_NO_DEVICE;
    const int d = mesh->d;
    const int64_t E = mesh->n_elem, F = mesh->n_face;

    GeoOps g;
    if (make_geo_ops(d, p, mesh->p_g, &g) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    // face records: the left side of every face in range (the device scatter checks the rest)
    for (int64_t f = 0; f < F; ++f) {
        const int64_t eL = mesh->face_elem[2 * f];
        const int lf = mesh->face_lf[2 * f];
        if (eL < 0 || eL >= E || lf < 0 || lf >= 2 * d)
            return DGX_ERR_INVALID_ARG;
    }
    const size_t b_xyz = sizeof(double) * (size_t)E * d * g.Ng;
    const size_t b_met = sizeof(double) * (size_t)E * d * d * g.Nc, b_jd = sizeof(double) * (size_t)E * g.Nc;
    const size_t b_nrm = sizeof(double) * (size_t)(F > 0 ? F : 1) * g.Nf * d;
    const size_t b_jdf = sizeof(double) * (size_t)(F > 0 ? F : 1) * g.Nf;
    const size_t b_wf = sizeof(double) * (size_t)2 * d * g.Ng * g.Nf;
    DevBuf dx, dwf, dvface, dfe, dlf, dbad, dmet, djd, dnrm, djdf;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    dgx_status rc = DGX_OK;
    int bad = 0;
#define GEO_TRY(x)                        \
    do {                                  \
        if ((x) != hipSuccess) {          \
            rc = DGX_ERR_HIP;             \
            goto done;                    \
        }                                 \
    } while (0)
    GEO_TRY(hipStreamCreate(&st));
    GEO_TRY(hipEventCreate(&ev0));
    GEO_TRY(hipEventCreate(&ev1));
    GEO_TRY(hipMalloc(&dx.p, b_xyz));
    GEO_TRY(hipMalloc(&dwf.p, b_wf));
    GEO_TRY(hipMalloc(&dvface.p, sizeof(int32_t) * (size_t)E * 2 * d));
    GEO_TRY(hipMalloc(&dfe.p, sizeof(int32_t) * 2 * (size_t)(F > 0 ? F : 1)));
    GEO_TRY(hipMalloc(&dlf.p, 2 * (size_t)(F > 0 ? F : 1)));
    GEO_TRY(hipMalloc(&dbad.p, sizeof(int)));
    if (ptr_kind == DGX_PTR_HOST) {
        GEO_TRY(hipMalloc(&dmet.p, b_met));
        GEO_TRY(hipMalloc(&djd.p, b_jd));
        GEO_TRY(hipMalloc(&dnrm.p, b_nrm));
        GEO_TRY(hipMalloc(&djdf.p, b_jdf));
    }
    {
        double* om = ptr_kind == DGX_PTR_HOST ? (double*)dmet.p : metrics_vc;
        double* oj = ptr_kind == DGX_PTR_HOST ? (double*)djd.p : jdet_vc;
        double* on = ptr_kind == DGX_PTR_HOST ? (double*)dnrm.p : normals_fc;
        double* of = ptr_kind == DGX_PTR_HOST ? (double*)djdf.p : jdet_fc;
        GEO_TRY(hipMemcpyAsync(dx.p, mesh->xyz_g, b_xyz, hipMemcpyHostToDevice, st));
        if (F > 0) {
            GEO_TRY(hipMemcpyAsync(dfe.p, mesh->face_elem, sizeof(int32_t) * 2 * F, hipMemcpyHostToDevice, st));
            GEO_TRY(hipMemcpyAsync(dlf.p, mesh->face_lf, 2 * F, hipMemcpyHostToDevice, st));
        }
        GEO_TRY(hipEventRecord(ev0, st));
        GEO_TRY(run_geometry(d, g, (const double*)dx.p, E, (const int32_t*)dfe.p, (const int8_t*)dlf.p, F,
                             (double*)dwf.p, (int32_t*)dvface.p, (int*)dbad.p, om, oj, on, of, st));
        GEO_TRY(hipEventRecord(ev1, st));
        if (ptr_kind == DGX_PTR_HOST) {
            GEO_TRY(hipMemcpyAsync(metrics_vc, om, b_met, hipMemcpyDeviceToHost, st));
            GEO_TRY(hipMemcpyAsync(jdet_vc, oj, b_jd, hipMemcpyDeviceToHost, st));
            if (F > 0) {
                GEO_TRY(hipMemcpyAsync(normals_fc, on, b_nrm, hipMemcpyDeviceToHost, st));
                GEO_TRY(hipMemcpyAsync(jdet_fc, of, b_jdf, hipMemcpyDeviceToHost, st));
            }
        }
        GEO_TRY(hipMemcpyAsync(&bad, dbad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        GEO_TRY(hipStreamSynchronize(st));
        if (kernel_ms) {
            float ms = 0.f;
            GEO_TRY(hipEventElapsedTime(&ms, ev0, ev1));
            *kernel_ms = ms;
        }
    }
    if (bad & 2)
        rc = DGX_ERR_INVALID_ARG;
    else if (bad)
        rc = DGX_ERR_GEOMETRY;
done:
#undef GEO_TRY
    if (st)
        (void)hipStreamSynchronize(st), (void)hipStreamDestroy(st);
    if (ev0)
        (void)hipEventDestroy(ev0);
    if (ev1)
        (void)hipEventDestroy(ev1);
    return rc;
}

// ---------------------------------------------------------------------------------------------------
// Allocation-free variant.  Its scratch memory (the face weight table and the per-volume face table) is
// this file-scope array: the loader gives every GPU its own copy, held until the process ends.  In host
// code `scratch` names a host-side stand-in, so its device address is looked up once per device
// (hipGetSymbolAddress) and only that address is passed to kernels.  Calls on one GPU share the array:
// calls that overlap in time on one GPU race.
// ---------------------------------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per GPU
constexpr size_t kWfBytes = 153600;                   // >= 2 * 3 * 125 * 25 doubles (p, p_g <= 4), 256-aligned
constexpr int kMaxDevices = 64;
std::atomic<unsigned char*> g_scratch_dev[kMaxDevices];
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this device's scratch array (looked up on the first call on each device).
static unsigned char* scratch_address(int dev)
{
    unsigned char* ptr = g_scratch_dev[dev].load(std::memory_order_acquire);
    if (ptr)
        return ptr;
    void* addr = nullptr;
    if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr)
        return nullptr;
    g_scratch_dev[dev].store((unsigned char*)addr, std::memory_order_release);
    return (unsigned char*)addr;
}

extern "C" dgx_status dgx_geometry_build_gpu_noalloc(const dgx_mesh* mesh, int p, double* metrics_vc,
                                                     double* jdet_vc, double* normals_fc, double* jdet_fc,
                                                     int32_t* flags, void* stream)
{
    using namespace dgx;
    if (!mesh || !mesh->xyz_g || !mesh->face_elem || !mesh->face_lf || !metrics_vc || !jdet_vc || !normals_fc ||
        !jdet_fc || !flags || (mesh->d != 2 && mesh->d != 3) || p < 1 || p > NMAX - 1 || mesh->p_g < 1 ||
        mesh->p_g > NMAX - 1 || mesh->n_elem < 1 || mesh->n_elem > 0x7fffffff || mesh->n_face < 0 ||
        mesh->n_face > 0x7fffffff)
        return DGX_ERR_INVALID_ARG;
    const int d = mesh->d;
    const int64_t E = mesh->n_elem, F = mesh->n_face;
    const size_t need = sizeof(int32_t) * (size_t)E * 2 * d;
    if (need > kScratchBytes - kWfBytes) {
        fprintf(stderr,
                "dgx_geometry_build_gpu_noalloc: %lld volumes need %zu bytes of face table, more than the %zu "
                "bytes of scratch (at most %zu volumes in %d-D)\n",
                (long long)E, need, kScratchBytes - kWfBytes, (kScratchBytes - kWfBytes) / (sizeof(int32_t) * 2 * d),
                d);
        return DGX_ERR_INVALID_ARG;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0)
        return DGX_ERR_NO_DEVICE;
    if (dev >= kMaxDevices)
        return DGX_ERR_INVALID_ARG;
    unsigned char* base = scratch_address(dev);
    if (!base)
        return DGX_ERR_HIP;
    GeoOps g;
    if (make_geo_ops(d, p, mesh->p_g, &g) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    const hipError_t e = run_geometry(d, g, mesh->xyz_g, E, mesh->face_elem, mesh->face_lf, F, (double*)base,
                                      (int32_t*)(base + kWfBytes), flags, metrics_vc, jdet_vc, normals_fc, jdet_fc,
                                      (hipStream_t)stream);
    return e == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}
