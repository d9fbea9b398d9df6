// Batching form: explicit DG residual and SSP-RK3 stage on a mixed HEX/WEDGE/TET/PYR mesh with
// per-element orders, elements grouped by (type, p).
//
// A stage is two kernels:
//   face    one lane per face cubature node: the two sides' traces from the trace buffer, the numerical
//           flux, times the cubature weight and face Jacobian -> the face flux buffer
//           (compute_face_rlhs_T.c:166-170, numerical_flux_euler_T.c)
//   volume  one tile of elements of one group per workgroup: coefficients -> LDS, state at the volume
//           cubature nodes and the metric-contracted Euler flux -> LDS, weak divergence (one lane per
//           mode, compute_volume_rlhs_T.c:145-177), lift of the element's faces from the face flux
//           buffer (compute_face_rlhs_T.c:297-359; the right side sees -nnf), M^-1 = 1/det J
//           (orthonormal basis on an affine element, solve_dg.c:265-272), the SSP-RK3 combination
//           (solve_explicit.c:252-301), then the new state's traces at the face cubature nodes of each
//           of its faces (basis rows of the side at the nodes as the face's left volume lists them: the
//           right side's node correspondence is folded into its table) -> the trace buffer
// The state is read once per stage (by the volume kernel); the face kernel reads traces only.  A trace
// pass (the volume kernel's last phase alone) fills the trace buffer after a state upload.
// Launch forms: a persistent grid whose workgroups take tiles from a work counter (tiles of every
// group in one launch, largest first), one grouped launch with a workgroup per tile of every group, or
// one launch per (type, p) group / per face size class.  The
// volume kernel updates its elements in place (an element reads only its own coefficients; neighbour
// data come through the trace and face flux buffers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dgx_kernels.hpp"
#include "dgx_mixed.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxTileLds = 60 * 1024;   // bytes of LDS one volume tile may use

struct GroupDev {
    int32_t type, p, ns, nc;
    int32_t tile;          // elements per volume tile
    int32_t pad;
    int64_t start, count;
    int64_t op_off;        // BT [Ns][Nc] then G [3][Nc][Ns]
    int64_t tp_off;        // HEX: B1 [n][n], D1 [n][n], w1 [n] (sum-factorized path); -1 otherwise
};

constexpr int kTpLds = 96;   // doubles at the start of a volume tile's LDS: the 1-D factors
constexpr int kGroupedBpc = 4;   // workgroups of the volume kernels one CU holds at their register count

// Modes of a (type, p) element of the sum-factorized paths.
__host__ __device__ constexpr int sf_modes(int type, int p)
{
    const int n = p + 1;
    return type == DGX_ELEM_HEX ? n * n * n
           : type == DGX_ELEM_WEDGE ? n * n * (n + 1) / 2
           : type == DGX_ELEM_TET ? n * (n + 1) * (n + 2) / 6 : n * (n + 1) * (2 * n + 1) / 6;
}

// Doubles of a sum-factorized group's factor block: HEX B1, D1 [n][n], w1, L(-1), L(+1) [n]; WEDGE also
// Bt, Gr, Gs [n^2][Nt], wt [n^2], the TRI edge matrices E [3][n][Nt] and tm, tp [n]; TET / PYR: A, Ad
// [n][n]; B, Bd; C, Cd (TET: only the entries of its modes, [Ns][n]); xa, xb, xc [n]; node weights [n^3]; mode
// scales [Ns].
__host__ __device__ constexpr int sf_tab_len(int type, int p)
{
    const int n = p + 1, nt = n * (n + 1) / 2;
    if (type == DGX_ELEM_TET || type == DGX_ELEM_PYR) {
        const int sb = type == DGX_ELEM_TET ? n * n * n : n * n;
        const int sc = type == DGX_ELEM_TET ? sf_modes(type, p) * n : n * n * n;
        return 2 * n * n + 2 * sb + 2 * sc + 3 * n + n * n * n + sf_modes(type, p);
    }
    return type == DGX_ELEM_WEDGE ? 2 * n * n + n + 3 * n * n * nt + n * n + 3 * n * nt + 2 * n
                                  : 2 * n * n + 3 * n;
}

// LDS doubles ahead of a tile's element slots: the whole factor block (the TRI and collapsed-coordinate
// tables are read at lane-dependent rows, which from global memory cost one cache line per lane), at
// least kTpLds, in whole 128-byte lines.
__host__ __device__ constexpr int sf_tab_lds(int type, int p)
{
    const int r = (sf_tab_len(type, p) + 15) / 16 * 16;
    return r > kTpLds ? r : kTpLds;
}

// LDS doubles of one HEX element slot: 15 n^3 (nodal fluxes) and at least 5 n^3 + 30 n^2 (mode work +
// face-modal work of the 6 faces)
__host__ __device__ constexpr int hex_slot(int n)
{
    return 15 * n * n * n > 5 * n * n * n + 30 * n * n ? 15 * n * n * n : 5 * n * n * n + 30 * n * n;
}

struct VolTile {
    int64_t e0;
    int32_t group, count;
};

struct FaceTile {
    int64_t i0;            // first entry of face_order
    int32_t count, lg_tpf; // faces, log2 lanes per face
};

struct MixArgs {
    const double* u;       // stage input
    const double* reg;     // u^n (stages 1, 2)
    double* out;           // stage output or residual
    int stage;             // 0..2, -1: residual, -2: traces of u only
    int apply_minv;
    double dt;
    const GroupDev* groups;
    const double* ops;
    const int64_t* dof;    // [E+1]
    const double* jdet;
    const double* met;     // [E][9]
    const int32_t* ef;     // [E][6]: 2 f + side, -1 = none
    const int32_t* face_lf_of;  // [E][6]: the local face index of each entry of ef
    const int32_t* face_elem;
    const int32_t* face_op;
    const int32_t* face_wt;
    const double* face_normal;
    const double* face_jac;
    const int64_t* fb_off;
    double* fb;            // face flux buffer [sum Nq][5]
    const int64_t* fl_off; // [2F]: the side's face-modal lift coefficients in fl, -1 for sides lifted densely
    double* fl;            // face-modal lift coefficients of HEX / WEDGE sides, [5][width] each
    double* tr;            // trace buffer: face f at 2 fb_off[f]: [side][Nq][5]
    const int32_t* fop_nq;
    const int32_t* fop_ns;
    const int64_t* fop_off;    // Bf [Nq][Ns]
    const int64_t* fopt_off;   // BfT [Ns][Nq]
    const int64_t* fopr_off;   // HEX sides: face-modal rows BF [Nq][n^2] (-1: none)
    const int64_t* foprt_off;  // and their transpose BFT [n^2][Nq]
    const int32_t* fopr_n;     // columns of the face-modal rows (HEX, WEDGE quad: n^2; WEDGE TRI: Nt)
    const int32_t* ftc;        // quad face-modal rows in tensor form: m | symmetry << 8 | n << 12, 0: none
    const int64_t* fte_off;    // and their 1-D factor E [m][n] in fop
    int tensor_faces;          // 1: the quad face maps run in tensor form where they have one
    const double* fop;
    const int64_t* fw_off;
    const double* fw;
    const int32_t* face_order;
    const VolTile* vtiles;
    const int32_t* vorder;     // persistent order of the volume tiles
    int64_t n_vtiles;
    const FaceTile* ftiles;
    int64_t n_ftiles;
    unsigned int* counter;     // [2] work counters: face, volume
};

// --------------------------------------------------------------------------------------------------
// device code
// --------------------------------------------------------------------------------------------------

// The tensor form of a quad face's face-modal rows, BF[q][a + n b] = E[ia][a] E[ib][b]: face node
// q = qx + m qy (the face rule's order, as listed by the face's left volume) sits at (ia, ib) of the side's
// m x m grid under one of the 8 symmetries of the square, code = swap | flip_a << 1 | flip_b << 2:
// (u, v) = swap ? (qy, qx) : (qx, qy), ia = flip_a ? m-1-u : u, ib = flip_b ? m-1-v : v.
__device__ __forceinline__ void face_grid_of(int q, int m, int code, int& ia, int& ib)
{
    const int qx = q % m, qy = q / m;
    const int u = (code & 1) ? qy : qx, v = (code & 1) ? qx : qy;
    ia = (code & 2) ? m - 1 - u : u;
    ib = (code & 4) ? m - 1 - v : v;
}

// The face node at (ia, ib) of the side's grid (the inverse of face_grid_of).
__device__ __forceinline__ int face_node_of(int ia, int ib, int m, int code)
{
    const int u = (code & 2) ? m - 1 - ia : ia, v = (code & 4) ? m - 1 - ib : ib;
    return (code & 1) ? v + m * u : u + m * v;
}

// A face node's value of the side's face-modal expansion (S: [n^2] coefficients, a fastest) through the
// tensor form: sum_b E[ib][b] sum_a E[ia][a] S[a + n b], the 2n factor entries loaded once.
template <int N>
__device__ __forceinline__ void face_trace_tensor(const double* E, int ia, int ib, const double* S, int sv,
                                                  double w[5])
{
    double ea[N], eb[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        ea[i] = E[ia * N + i];
        eb[i] = E[ib * N + i];
    }
#pragma unroll
    for (int b = 0; b < N; ++b) {
        double t[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int v = 0; v < 5; ++v)
                t[v] = fma(ea[i], S[v * sv + i + N * b], t[v]);
#pragma unroll
        for (int v = 0; v < 5; ++v)
            w[v] = fma(eb[b], t[v], w[v]);
    }
}

template <int FLUX>
__device__ __attribute__((always_inline)) inline void face_tile(const MixArgs& a, const FaceTile& t)
{
    __shared__ double s_nf[kBlock * 5];   // the tile's scaled fluxes, [lane][v]
    const int lg = t.lg_tpf, tpf = 1 << lg;
    const int slot = threadIdx.x >> lg, l = threadIdx.x & (tpf - 1);
    const bool face = slot < t.count;
    const int f = face ? a.face_order[t.i0 + slot] : 0;
    const int nq = face ? a.fop_nq[a.face_op[2 * f]] : 0;
    if (face && l < nq) {
        double wl[5], wr[5];
        const double* tf = a.tr + 2 * a.fb_off[f] + 5 * l;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            wl[v] = tf[v];
            wr[v] = tf[5 * nq + v];
        }
        const double n[3] = { a.face_normal[3 * f], a.face_normal[3 * f + 1], a.face_normal[3 * f + 2] };
        double nf[5];
        dgx::numflux<3, FLUX>(wl, wr, n, nf);
        const double sc = a.fw[a.fw_off[a.face_wt[f]] + l] * a.face_jac[f];
        // the flux buffer is read only by the sides lifted from the face nodes (the dense lifts): a face
        // whose two sides both lift in face-modal form does not write it
        const bool to_fb = a.fl_off[2 * f] < 0 || a.fl_off[2 * f + 1] < 0;
        double* o = a.fb + a.fb_off[f] + 5 * l;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (to_fb)
                o[v] = nf[v] * sc;
            s_nf[5 * threadIdx.x + v] = nf[v] * sc;
        }
    }
    __syncthreads();
    // the lift of each HEX / WEDGE side in face-modal form, -+sum_q BF[q][x] nnf[q][v] (the left side takes
    // -nnf): computed here, where the face's fluxes sit in LDS, the volume kernel reads 5 x width values
    // per face instead of a vector load per face node and row.  Sides whose rows have the tensor form,
    // BF[q][a + n b] = E[ia][a] E[ib][b], contract one direction at a time through LDS:
    //   R[v][a][ib] = sum_ia E[ia][a] nnf(node at (ia, ib)),  c[v][a + n b] = sum_ib E[ib][b] R[v][a][ib]
    // (5 n m (n + m) multiply-adds per side instead of 5 n^2 m^2).
    __shared__ double s_R[kBlock * 10];    // per face slot 10 tpf doubles: [side][v][a][ib], 5 n m <= 5 tpf each
    int64_t offL = -1, offR = -1;
    int oL = 0, oR = 0, wL = 0, wR = 0, tcL = 0, tcR = 0;
    const double* nfs = s_nf + 5 * (slot << lg);
    double* Rs = s_R + 10 * (slot << lg);
    if (face) {
        offL = a.fl_off[2 * f];
        offR = a.fl_off[2 * f + 1];
        oL = a.face_op[2 * f];
        oR = a.face_op[2 * f + 1];
        wL = offL >= 0 ? a.fopr_n[oL] : 0;
        wR = offR >= 0 ? a.fopr_n[oR] : 0;
        tcL = (a.tensor_faces && offL >= 0) ? a.ftc[oL] : 0;
        tcR = (a.tensor_faces && offR >= 0) ? a.ftc[oR] : 0;
        // first direction: items (a, ib) of each tensor side
        const int r1L = tcL ? (tcL >> 12) * (tcL & 255) : 0, r1R = tcR ? (tcR >> 12) * (tcR & 255) : 0;
        for (int it = l; it < r1L + r1R; it += tpf) {
            const int side = it >= r1L ? 1 : 0;
            const int tc = side ? tcR : tcL;
            const int m = tc & 255, code = (tc >> 8) & 7, n = tc >> 12;
            const int y = it - side * r1L, xa = y % n, ib = y / n;
            const double* E = a.fop + a.fte_off[side ? oR : oL];
            double ea[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                ea[i] = i < m ? E[i * n + xa] : 0.0;
            double t[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
            for (int ia = 0; ia < 8; ++ia) {
                if (ia < m) {
                    const double* nv = nfs + 5 * face_node_of(ia, ib, m, code);
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        t[v] = fma(ea[ia], nv[v], t[v]);
                }
            }
            double* R = Rs + (side ? 5 * r1L : 0) + y;
#pragma unroll
            for (int v = 0; v < 5; ++v)
                R[v * n * m] = t[v];
        }
    }
    __syncthreads();
    if (face) {
        const int r1L = tcL ? (tcL >> 12) * (tcL & 255) : 0;
        for (int it = l; it < wL + wR; it += tpf) {
            const int side = it >= wL ? 1 : 0;
            const int x = it - side * wL, wdt = side ? wR : wL;
            const int o = side ? oR : oL;
            const int tc = side ? tcR : tcL;
            double acc[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
            if (tc) {
                // second direction: c[v][a + n b] = sum_ib E[ib][b] R[v][a][ib]
                const int m = tc & 255, n = tc >> 12;
                const int xa = x % n, xb = x / n;
                const double* E = a.fop + a.fte_off[o];
                const double* R = Rs + (side ? 5 * r1L : 0) + xa;
                double eb[8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    eb[i] = i < m ? E[i * n + xb] : 0.0;
#pragma unroll
                for (int ib = 0; ib < 8; ++ib) {
                    if (ib < m) {
#pragma unroll
                        for (int v = 0; v < 5; ++v)
                            acc[v] = fma(eb[ib], R[v * n * m + n * ib], acc[v]);
                    }
                }
            } else {
                const double* bf = a.fop + a.fopr_off[o];
#pragma unroll 6
                for (int q = 0; q < nq; ++q) {
                    const double b = bf[q * wdt + x];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        acc[v] = fma(b, nfs[5 * q + v], acc[v]);
                }
            }
            double* out = a.fl + (side ? offR : offL) + x;
#pragma unroll
            for (int v = 0; v < 5; ++v)
                out[v * wdt] = side ? acc[v] : -acc[v];
        }
    }
    __syncthreads();   // s_nf is reused by the next tile of a persistent workgroup
}

__device__ inline double rk_mix(int stage, double dt, double s, double r, const double* reg)
{
    if (stage == 0)
        return s + dt * r;
    if (stage == 1)
        return (1.0 / 4.0) * (3.0 * (*reg) + s + dt * r);
    return (1.0 / 3.0) * ((*reg) + 2.0 * s + 2.0 * dt * r);
}

// Per-mode end of the volume phase: face lifts, M^-1, the residual or the RK update (the new state also
// goes to U, the tile's LDS copy, for the trace phase).  s_old: the stage input value of each variable.
__device__ __attribute__((always_inline)) inline void finish_mode(const MixArgs& a, int64_t e, int ns, int j, double r[5], const double s_old[5],
                                   double* U, bool dense_lift = true, bool wave_elem = false)
{
    if (dense_lift && wave_elem) {
        // the wavefront serves one element: its face records and face fluxes are read as scalars (the
        // same values on every lane; through vector memory they cost five loads per face node), the
        // lift rows as vector loads over the modes
        using cip = const __attribute__((address_space(4))) int32_t*;
        using clp = const __attribute__((address_space(4))) int64_t*;
        using cdp = const __attribute__((address_space(4))) double*;
        const int eu = __builtin_amdgcn_readfirstlane((int)e);
#pragma unroll 1
        for (int k = 0; k < 6; ++k) {
            const int ent = ((cip)a.ef)[6 * eu + k];
            if (ent < 0)
                break;
            const int f = ent >> 1, side = ent & 1;
            const int o = ((cip)a.face_op)[2 * f + side];
            const int nq = ((cip)a.fop_nq)[o];
            const double* bf = a.fop + ((clp)a.fop_off)[o];
            const cdp nf = (cdp)a.fb + ((clp)a.fb_off)[f];
            double s[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 6
            for (int q = 0; q < nq; ++q) {
                const double b = bf[q * ns + j];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    s[v] = fma(b, nf[5 * q + v], s[v]);
            }
#pragma unroll
            for (int v = 0; v < 5; ++v)
                r[v] = side ? r[v] + s[v] : r[v] - s[v];
        }
        dense_lift = false;
    }
#pragma unroll 1
    for (int k = 0; k < 6 && dense_lift; ++k) {
        const int ent = a.ef[6 * e + k];
        if (ent < 0)
            break;
        const int f = ent >> 1, side = ent & 1;
        const int o = a.face_op[2 * f + side];
        const int nq = a.fop_nq[o];
        const double* bf = a.fop + a.fop_off[o];
        const double* nf = a.fb + a.fb_off[f];
        double s[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 6
        for (int q = 0; q < nq; ++q) {
            const double b = bf[q * ns + j];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                s[v] = fma(b, nf[5 * q + v], s[v]);
        }
#pragma unroll
        for (int v = 0; v < 5; ++v)
            r[v] = side ? r[v] + s[v] : r[v] - s[v];
    }
    double* oe = a.out + a.dof[e];
    const double ij = a.jdet[e];
    if (a.stage < 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
            oe[v * ns + j] = a.apply_minv ? r[v] / ij : r[v];
        return;
    }
    const double* re = a.reg ? a.reg + a.dof[e] : nullptr;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const int i = v * ns + j;
        const double un = rk_mix(a.stage, a.dt, s_old[v], r[v] / ij, re ? re + i : nullptr);
        oe[i] = un;
        U[i] = un;
    }
}

// Traces of U at every face's cubature nodes, [side][Nq][5] at 2 fb_off[f]; the element's lanes run
// over its face nodes one after another.
__device__ __attribute__((always_inline)) inline void write_traces(const MixArgs& a, int64_t e, int ns, const double* U, int l, int tpe)
{
    int nqk[6], total = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int ent = a.ef[6 * e + k];
        nqk[k] = ent < 0 ? 0 : a.fop_nq[a.face_op[2 * (ent >> 1) + (ent & 1)]];
        total += nqk[k];
    }
    for (int it = l; it < total; it += tpe) {
        int k = 0, q = it;
        while (q >= nqk[k]) {
            q -= nqk[k];
            ++k;
        }
        const int ent = a.ef[6 * e + k];
        const int nq = nqk[k];
        const int f = ent >> 1, side = ent & 1;
        const int o = a.face_op[2 * f + side];
        const double* bt = a.fop + a.fopt_off[o];
        double w[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
        // partly unrolled: ns is a compile-time constant on the sum-factorized paths
#pragma unroll 6
        for (int j = 0; j < ns; ++j) {
            const double b = bt[j * nq + q];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                w[v] = fma(b, U[v * ns + j], w[v]);
        }
        double* t = a.tr + 2 * a.fb_off[f] + (int64_t)side * 5 * nq + 5 * q;
#pragma unroll
        for (int v = 0; v < 5; ++v)
            t[v] = w[v];
    }
}

// One volume tile of a group with dense operators: `count` elements starting at e0; every lane of the
// block reaches every barrier.
__device__ __attribute__((always_inline)) inline void vol_tile(const MixArgs& a, const GroupDev& G, const VolTile& t, double* lds)
{
    const int ns = G.ns, nc = G.nc;
    const int tpe = kBlock / G.tile;
    const int slot = threadIdx.x / tpe, l = threadIdx.x - slot * tpe;
    const bool act = slot < t.count;
    const int64_t e = t.e0 + slot;
    double* U = lds + kTpLds + slot * (5 * ns + 15 * nc);
    double* Fr = U + 5 * ns;
    if (act) {
        const double* ue = a.u + a.dof[e];
        for (int i = l; i < 5 * ns; i += tpe)
            U[i] = ue[i];
    }
    __syncthreads();
    if (a.stage != -2 && act) {
        const double* bt = a.ops + G.op_off;
        double m[9];
#pragma unroll
        for (int i = 0; i < 9; ++i)
            m[i] = a.met[9 * e + i];
        for (int q = l; q < nc; q += tpe) {
            double w[5];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                w[v] = 0.0;
            for (int j = 0; j < ns; ++j) {
                const double b = bt[j * nc + q];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    w[v] = fma(b, U[v * ns + j], w[v]);
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                // reference flux along xi_d: sum_i m(i, d) F_i
                const double col[3] = { m[d], m[3 + d], m[6 + d] };
                double fd[5];
                dgx::euler_flux_n<3>(w, col, fd);
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    Fr[(q * 3 + d) * 5 + v] = fd[v];
            }
        }
    }
    __syncthreads();
    if (a.stage != -2 && act) {
        const double* gm = a.ops + G.op_off + (int64_t)ns * nc;
        for (int j = l; j < ns; j += tpe) {
            double r[5], s_old[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                r[v] = 0.0;
                s_old[v] = U[v * ns + j];
            }
            for (int q = 0; q < nc; ++q) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double g = gm[(d * nc + q) * ns + j];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        r[v] = fma(g, Fr[(q * 3 + d) * 5 + v], r[v]);
                }
            }
            // no lane reads U after the barrier that closed the interpolation phase but its own modes
            finish_mode(a, e, ns, j, r, s_old, U);
        }
    }
    __syncthreads();
    if (a.stage != -1 && act)
        write_traces(a, e, ns, U, l, tpe);
    __syncthreads();
}

// One volume tile of HEX p = P elements by sum factorization: the orthonormal Legendre basis at the
// Gauss-Legendre nodes is B1 (x) B1 (x) B1 (modes and nodes r fastest), its reference gradient D1 in one
// direction, the weights w1 (x) w1 (x) w1.  Each element works in place in 15 n^3 doubles of LDS:
//   interpolation   three line passes (x, y, z) over the 5 variables, one lane per line
//   fluxes          one lane per node: the 3 metric-contracted fluxes times the node weight
//   weak divergence x: D1^T on fr_0, B1^T on fr_1, fr_2; y: B1^T x_0 + D1^T x_1, B1^T x_2;
//                   z: B1^T y_01 + D1^T y_2  -> the volume residual in mode order
// then the per-mode lifts / M^-1 / update and the traces as the dense path.
template <int P>
__device__ __attribute__((always_inline)) inline void vol_tile_hex(const MixArgs& a, const GroupDev& G, const VolTile& t, double* lds)
{
    constexpr int n = P + 1, n2 = n * n, n3 = n2 * n;
    const int tpe = kBlock / G.tile;
    const int slot = threadIdx.x / tpe, l = threadIdx.x - slot * tpe;
    const bool act = slot < t.count;
    const int64_t e = t.e0 + slot;
    constexpr int per = hex_slot(n);
    // B1 and D1 are read at compile-time indices (the same entry on every lane) through the constant
    // address space: scalar loads and scalar FMA operands, where an LDS copy costs an LDS read per
    // coefficient in passes that are already bound by their LDS line traffic
    using cdp = const __attribute__((address_space(4))) double*;
    const cdp B1 = (cdp)(a.ops + G.tp_off);
    const cdp D1 = B1 + n2;
    double* w1 = lds + 2 * n2;
    double* tm1 = lds + 2 * n2 + n;        // L_i(-1), L_i(+1): the traces of the 1-D basis
    double* tp1 = lds + 2 * n2 + 2 * n;
    double* W = lds + sf_tab_lds(DGX_ELEM_HEX, P) + slot * per;
    double* S = W + 5 * n3;                // face-modal work [face][v][n^2]
    if (threadIdx.x < 2 * n2 + 3 * n)
        lds[threadIdx.x] = a.ops[G.tp_off + threadIdx.x];
    if (act) {
        const double* ue = a.u + a.dof[e];
        for (int i = l; i < 5 * n3; i += tpe)
            W[i] = ue[i];
    }
    __syncthreads();
    if (a.stage != -2) {
        // interpolation: x, y, z line passes (stride 1, n, n^2), lines (v, two other indices)
#pragma unroll
        for (int pass = 0; pass < 3; ++pass) {
            const int st = pass == 0 ? 1 : (pass == 1 ? n : n2);
            if (act) {
                for (int ln = l; ln < 5 * n2; ln += tpe) {
                    const int v = ln / n2, rr = ln - v * n2;
                    const int a0 = rr % n, a1 = rr / n;
                    // base of the line: the two fixed indices in the other directions
                    const int base = v * n3 + (pass == 0 ? a0 * n + a1 * n2 : (pass == 1 ? a0 + a1 * n2 : a0 + a1 * n));
                    double in[n];
#pragma unroll
                    for (int i = 0; i < n; ++i)
                        in[i] = W[base + i * st];
#pragma unroll
                    for (int q = 0; q < n; ++q) {
                        double acc = 0.0;
#pragma unroll
                        for (int i = 0; i < n; ++i)
                            acc = fma(B1[q * n + i], in[i], acc);
                        W[base + q * st] = acc;
                    }
                }
            }
            __syncthreads();
        }
        // fluxes at the nodes (a lane touches its own node only)
        if (act) {
            double m[9];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                m[i] = a.met[9 * e + i];
            for (int q = l; q < n3; q += tpe) {
                double w[5];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    w[v] = W[v * n3 + q];
                const double wq = w1[q % n] * w1[(q / n) % n] * w1[q / n2];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double col[3] = { m[d], m[3 + d], m[6 + d] };
                    double fd[5];
                    dgx::euler_flux_n<3>(w, col, fd);
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        W[(d * 5 + v) * n3 + q] = wq * fd[v];
                }
            }
        }
        __syncthreads();
        // x: lines (d, v, qy, qz): D1^T for d = 0, then B1^T for d = 1, 2 (one operator per loop, so that
        // every lane reads the same coefficient)
        if (act) {
            for (int ln = l; ln < 15 * n2; ln += tpe) {
                const int dv = ln / n2, rr = ln - dv * n2;
                const int base = dv * n3 + rr * n;
                if (ln >= 5 * n2)
                    break;
                double in[n];
#pragma unroll
                for (int q = 0; q < n; ++q)
                    in[q] = W[base + q];
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q)
                        acc = fma(D1[q * n + i], in[q], acc);
                    W[base + i] = acc;
                }
            }
            for (int ln = 5 * n2 + l; ln < 15 * n2; ln += tpe) {
                const int dv = ln / n2, rr = ln - dv * n2;
                const int base = dv * n3 + rr * n;
                double in[n];
#pragma unroll
                for (int q = 0; q < n; ++q)
                    in[q] = W[base + q];
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q)
                        acc = fma(B1[q * n + i], in[q], acc);
                    W[base + i] = acc;
                }
            }
        }
        __syncthreads();
        // y: lines (v, i, qz): d = 0 with d = 1 (B1^T, D1^T), and d = 2 (B1^T)
        if (act) {
            for (int ln = l; ln < 10 * n2; ln += tpe) {
                const bool two = ln >= 5 * n2;
                const int l2 = two ? ln - 5 * n2 : ln;
                const int v = l2 / n2, rr = l2 - v * n2;
                const int i = rr % n, kz = rr / n;
                const int b0 = ((two ? 10 : 0) + v) * n3 + i + kz * n2;
                const int b1 = (5 + v) * n3 + i + kz * n2;
                double in0[n], in1[n];
#pragma unroll
                for (int q = 0; q < n; ++q) {
                    in0[q] = W[b0 + q * n];
                    in1[q] = two ? 0.0 : W[b1 + q * n];
                }
#pragma unroll
                for (int j = 0; j < n; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q) {
                        acc = fma(B1[q * n + j], in0[q], acc);
                        if (!two)
                            acc = fma(D1[q * n + j], in1[q], acc);
                    }
                    W[b0 + j * n] = acc;
                }
            }
        }
        __syncthreads();
        // z: lines (v, i, j): B1^T y_01 + D1^T y_2 -> residual in mode order in W[0 .. 5 n^3)
        if (act) {
            for (int ln = l; ln < 5 * n2; ln += tpe) {
                const int v = ln / n2, rr = ln - v * n2;
                const int b0 = v * n3 + rr, b2 = (10 + v) * n3 + rr;
                double in0[n], in2[n];
#pragma unroll
                for (int q = 0; q < n; ++q) {
                    in0[q] = W[b0 + q * n2];
                    in2[q] = W[b2 + q * n2];
                }
#pragma unroll
                for (int k = 0; k < n; ++k) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q) {
                        acc = fma(B1[q * n + k], in0[q], acc);
                        acc = fma(D1[q * n + k], in2[q], acc);
                    }
                    W[b0 + k * n2] = acc;
                }
            }
        }
        __syncthreads();
        // face lifts in face-modal form, S_k[v][ab] = -+sum_q BF_k[q][ab] nnf_k[q][v], as the face kernel left
        // them (lanes over (k, ab))
        if (act) {
            for (int it = l; it < 6 * n2; it += tpe) {
                const int k = it / n2, ab = it - k * n2;
                const double* fl = a.fl + a.fl_off[a.ef[6 * e + k]] + ab;
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    S[(k * 5 + v) * n2 + ab] = fl[v * n2];
            }
        }
        __syncthreads();
        if (act) {
            const double* ue = a.u + a.dof[e];
            for (int j = l; j < n3; j += tpe) {
                const int i0 = j % n, i1 = (j / n) % n, i2 = j / n2;
                double r[5], s_old[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    r[v] = W[v * n3 + j];
                    s_old[v] = ue[v * n3 + j];
                }
                // r(i0, i1, i2) += tau_lf(i_dn) S_k(the other two indices), faces in local order lf
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int lf = a.face_lf_of[6 * e + k];
                    const int dn = lf >> 1;
                    const double* tau = (lf & 1) ? tp1 : tm1;
                    const int id = dn == 0 ? i0 : (dn == 1 ? i1 : i2);
                    const int ab = dn == 0 ? i1 + n * i2 : (dn == 1 ? i0 + n * i2 : i0 + n * i1);
                    const double tv = tau[id];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        r[v] = fma(tv, S[(k * 5 + v) * n2 + ab], r[v]);
                }
                // each lane reads and then overwrites its own modes only
                finish_mode(a, e, n3, j, r, s_old, W, false);
            }
        }
        __syncthreads();
    }
    if (a.stage != -1) {
        // traces in face-modal form: C_k[v][ab] = sum_i tau_lf(i) U(i along dn; ab), then
        // trace_k(q) = sum_ab BF_k[q][ab] C_k[v][ab]
        if (act) {
            for (int it = l; it < 30 * n2; it += tpe) {
                const int kv = it / n2, ab = it - kv * n2;
                const int k = kv / 5, v = kv - 5 * k;
                const int lf = a.face_lf_of[6 * e + k];
                const int dn = lf >> 1;
                const double* tau = (lf & 1) ? tp1 : tm1;
                const int aa = ab % n, bb = ab / n;
                const int base = v * n3 + (dn == 0 ? n * aa + n2 * bb : (dn == 1 ? aa + n2 * bb : aa + n * bb));
                const int st = dn == 0 ? 1 : (dn == 1 ? n : n2);
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < n; ++i)
                    acc = fma(tau[i], W[base + i * st], acc);
                S[kv * n2 + ab] = acc;
            }
        }
        __syncthreads();
        if (act) {
            int nqk[6], total = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int ent = a.ef[6 * e + k];
                nqk[k] = a.fop_nq[a.face_op[2 * (ent >> 1) + (ent & 1)]];
                total += nqk[k];
            }
            for (int it = l; it < total; it += tpe) {
                int k = 0, q = it;
                while (q >= nqk[k]) {
                    q -= nqk[k];
                    ++k;
                }
                const int ent = a.ef[6 * e + k];
                const int nq = nqk[k];
                const int f = ent >> 1, side = ent & 1;
                const int o = a.face_op[2 * f + side];
                const int tc = a.tensor_faces ? a.ftc[o] : 0;
                double w[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
                if (tc) {
                    int ia, ib;
                    face_grid_of(q, tc & 255, (tc >> 8) & 7, ia, ib);
                    face_trace_tensor<n>(a.fop + a.fte_off[o], ia, ib, S + k * 5 * n2, n2, w);
                } else {
                    const double* bt = a.fop + a.foprt_off[o];
#pragma unroll 6
                    for (int ab = 0; ab < n2; ++ab) {
                        const double b = bt[ab * nq + q];
#pragma unroll
                        for (int v = 0; v < 5; ++v)
                            w[v] = fma(b, S[(k * 5 + v) * n2 + ab], w[v]);
                    }
                }
                double* t = a.tr + 2 * a.fb_off[f] + (int64_t)side * 5 * nq + 5 * q;
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    t[v] = w[v];
            }
        }
    }
    __syncthreads();
}

// One volume tile of WEDGE p = P elements by sum factorization over TRI x LINE: the basis at the nodes
// is B1 (x) Bt (TRI index fastest, line slowest, as the reference's WEDGE operators), its r and s
// derivatives B1 (x) Gr, B1 (x) Gs, its t derivative D1 (x) Bt; weights w1 (x) wt.  LDS per element:
// 15 n^3 doubles (nodal work) + 5 n Nt (modal work).
//   interpolation   TRI (lines (v, k): Nt modes -> n^2 nodes), then LINE (lines (v, qt))
//   fluxes          one lane per node, times the node weight
//   weak divergence LINE^T (lines (d, v, qt): B1^T for r, s; D1^T for t), then TRI^T (lines (v, k):
//                   Gr^T y_0 + Gs^T y_1 + Bt^T y_2 -> Nt modes)
template <int P>
__device__ __attribute__((always_inline)) inline void vol_tile_wedge(const MixArgs& a, const GroupDev& G, const VolTile& t, double* lds)
{
    constexpr int n = P + 1, n2 = n * n, n3 = n2 * n, nt = n * (n + 1) / 2, ns = n * nt;
    const int tpe = kBlock / G.tile;
    const int slot = threadIdx.x / tpe, l = threadIdx.x - slot * tpe;
    const bool act = slot < t.count;
    const int64_t e = t.e0 + slot;
    constexpr int TAB = sf_tab_lds(DGX_ELEM_WEDGE, P);
    // the LINE factors are read at compile-time indices: scalar operands (as in the HEX path); the TRI
    // tables at lane-dependent rows from the LDS copy
    using cdp = const __attribute__((address_space(4))) double*;
    const cdp B1 = (cdp)(a.ops + G.tp_off);
    const cdp D1 = B1 + n2;
    const double* w1 = lds + 2 * n2;
    const double* tri = lds + 2 * n2 + n;                  // Bt, Gr, Gs [n2][nt], wt [n2]
    const double* Bt = tri;
    const double* Gr = tri + n2 * nt;
    const double* Gs = tri + 2 * n2 * nt;
    const double* wt = tri + 3 * n2 * nt;
    const double* Ew = wt + n2;                             // E [3][n][Nt]: the TRI edge matrices
    const double* tmw = Ew + 3 * n * nt;                    // L_k(-1), L_k(+1)
    const double* tpw = tmw + n;
    double* W = lds + TAB + slot * (15 * n3 + 5 * ns);
    double* Um = W + 15 * n3;                               // modal work [v][k][m]
    double* S = W;                                          // face-modal work [face][v][n^2] (25 n^2 <= 15 n^3)
    for (int i = threadIdx.x; i < sf_tab_len(DGX_ELEM_WEDGE, P); i += kBlock)
        lds[i] = a.ops[G.tp_off + i];
    if (act) {
        const double* ue = a.u + a.dof[e];
        for (int i = l; i < 5 * ns; i += tpe)
            Um[i] = ue[i];
    }
    __syncthreads();
    if (a.stage != -2) {
        // TRI: lines (k, all 5 variables): one operator read serves the 5 variables
        if (act) {
            for (int it = l; it < n * n2; it += tpe) {
                const int k = it / n2, qt = it - k * n2;
                const double* in = Um + k * nt;
                double acc[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 3
                for (int m = 0; m < nt; ++m) {
                    const double b = Bt[qt * nt + m];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        acc[v] = fma(b, in[v * ns + m], acc[v]);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    W[(v * n + k) * n2 + qt] = acc[v];   // [v][k][qt]
            }
        }
        __syncthreads();
        // LINE: lines (v, qt), stride n2, in place
        if (act) {
            for (int ln = l; ln < 5 * n2; ln += tpe) {
                const int v = ln / n2, qt = ln - v * n2;
                const int base = v * n3 + qt;
                double in[n];
#pragma unroll
                for (int k = 0; k < n; ++k)
                    in[k] = W[base + k * n2];
#pragma unroll
                for (int q = 0; q < n; ++q) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < n; ++k)
                        acc = fma(B1[q * n + k], in[k], acc);
                    W[base + q * n2] = acc;         // [v][ql][qt]
                }
            }
        }
        __syncthreads();
        if (act) {
            double m[9];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                m[i] = a.met[9 * e + i];
            for (int q = l; q < n3; q += tpe) {
                double w[5];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    w[v] = W[v * n3 + q];
                const double wq = wt[q % n2] * w1[q / n2];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double col[3] = { m[d], m[3 + d], m[6 + d] };
                    double fd[5];
                    dgx::euler_flux_n<3>(w, col, fd);
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        W[(d * 5 + v) * n3 + q] = wq * fd[v];
                }
            }
        }
        __syncthreads();
        // LINE^T: lines (d, v, qt), in place: B1^T for d = 0, 1, then D1^T for d = 2 (one operator per loop)
        if (act) {
            for (int ln = l; ln < 10 * n2; ln += tpe) {
                const int dv = ln / n2, qt = ln - dv * n2;
                const int base = dv * n3 + qt;
                double in[n];
#pragma unroll
                for (int q = 0; q < n; ++q)
                    in[q] = W[base + q * n2];
#pragma unroll
                for (int k = 0; k < n; ++k) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q)
                        acc = fma(B1[q * n + k], in[q], acc);
                    W[base + k * n2] = acc;         // [d][v][k][qt]
                }
            }
            for (int ln = 10 * n2 + l; ln < 15 * n2; ln += tpe) {
                const int dv = ln / n2, qt = ln - dv * n2;
                const int base = dv * n3 + qt;
                double in[n];
#pragma unroll
                for (int q = 0; q < n; ++q)
                    in[q] = W[base + q * n2];
#pragma unroll
                for (int k = 0; k < n; ++k) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q)
                        acc = fma(D1[q * n + k], in[q], acc);
                    W[base + k * n2] = acc;
                }
            }
        }
        __syncthreads();
        // TRI^T: outputs (k, m) of all 5 variables -> Um (the three table reads serve the 5 variables)
        if (act) {
            for (int it = l; it < ns; it += tpe) {
                const int k = it / nt, mm = it - k * nt;
                const double* y0 = W + k * n2;                  // d = 0: [v][k] blocks, v stride n3
                const double* y1 = W + 5 * n3 + k * n2;
                const double* y2 = W + 10 * n3 + k * n2;
                double acc[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 2
                for (int qt = 0; qt < n2; ++qt) {
                    const double gr = Gr[qt * nt + mm], gs = Gs[qt * nt + mm], bt = Bt[qt * nt + mm];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        acc[v] = fma(gr, y0[v * n3 + qt], acc[v]);
                        acc[v] = fma(gs, y1[v * n3 + qt], acc[v]);
                        acc[v] = fma(bt, y2[v * n3 + qt], acc[v]);
                    }
                }
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    Um[v * ns + it] = acc[v];
            }
        }
        __syncthreads();
        // face lifts in face-modal form (the nodal work area is free), S_k[v][x] = -+sum_q BF_k[q][x] nnf[q][v],
        // as the face kernel left them
        if (act) {
            for (int it = l; it < 5 * n2; it += tpe) {
                const int k = it / n2, x = it - k * n2;
                const int ent = a.ef[6 * e + k];
                const int wdt = a.fopr_n[a.face_op[ent]];
                if (x >= wdt)
                    continue;
                const double* fl = a.fl + a.fl_off[ent] + x;
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    S[(k * 5 + v) * n2 + x] = fl[v * wdt];
            }
        }
        __syncthreads();
        if (act) {
            const double* ue = a.u + a.dof[e];
            for (int j = l; j < ns; j += tpe) {
                const int kl = j / nt, mm = j - kl * nt;
                double r[5], s_old[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    r[v] = Um[v * ns + j];
                    s_old[v] = ue[v * ns + j];
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const int lf = a.face_lf_of[6 * e + k];
                    if (lf < 3) {
                        // quad face over TRI edge lf: sum_a E[a][m] S[a + n kl]
                        const double* E = Ew + lf * n * nt;
#pragma unroll
                        for (int aa = 0; aa < n; ++aa) {
                            const double ev = E[aa * nt + mm];
#pragma unroll
                            for (int v = 0; v < 5; ++v)
                                r[v] = fma(ev, S[(k * 5 + v) * n2 + aa + n * kl], r[v]);
                        }
                    } else {
                        const double tv = (lf == 3 ? tmw : tpw)[kl];
#pragma unroll
                        for (int v = 0; v < 5; ++v)
                            r[v] = fma(tv, S[(k * 5 + v) * n2 + mm], r[v]);
                    }
                }
                // each lane reads and then overwrites its own modes only
                finish_mode(a, e, ns, j, r, s_old, Um, false);
            }
        }
        __syncthreads();
    }
    if (a.stage != -1) {
        // traces in face-modal form: quad faces c[a + n k] = sum_m E[a][m] U[k][m], TRI faces
        // c[m] = sum_k tau_k U[k][m]; then trace(q) = sum_x BF[q][x] c[x]
        if (act) {
            for (int it = l; it < 25 * n2; it += tpe) {
                const int kv = it / n2, x = it - kv * n2;
                const int k = kv / 5, v = kv - 5 * k;
                const int lf = a.face_lf_of[6 * e + k];
                double acc = 0.0;
                if (lf < 3) {
                    const int aa = x % n, kl = x / n;
                    const double* E = Ew + lf * n * nt + aa * nt;
                    const double* u = Um + v * ns + kl * nt;
                    for (int mm = 0; mm < nt; ++mm)
                        acc = fma(E[mm], u[mm], acc);
                } else {
                    if (x >= nt)
                        continue;
                    const double* tau = lf == 3 ? tmw : tpw;
#pragma unroll
                    for (int kl = 0; kl < n; ++kl)
                        acc = fma(tau[kl], Um[v * ns + kl * nt + x], acc);
                }
                S[kv * n2 + x] = acc;
            }
        }
        __syncthreads();
        if (act) {
            int nqk[5], total = 0;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const int ent = a.ef[6 * e + k];
                nqk[k] = a.fop_nq[a.face_op[2 * (ent >> 1) + (ent & 1)]];
                total += nqk[k];
            }
            for (int it = l; it < total; it += tpe) {
                int k = 0, q = it;
                while (q >= nqk[k]) {
                    q -= nqk[k];
                    ++k;
                }
                const int ent = a.ef[6 * e + k];
                const int nq = nqk[k];
                const int f = ent >> 1, side = ent & 1;
                const int o = a.face_op[2 * f + side];
                const int wdt = a.fopr_n[o];
                const int tc = a.tensor_faces ? a.ftc[o] : 0;
                double w[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
                if (tc) {
                    // a quad face: its coefficients c[a + n k] in tensor form
                    int ia, ib;
                    face_grid_of(q, tc & 255, (tc >> 8) & 7, ia, ib);
                    face_trace_tensor<n>(a.fop + a.fte_off[o], ia, ib, S + k * 5 * n2, n2, w);
                } else {
                    const double* bt = a.fop + a.foprt_off[o];
#pragma unroll 6
                    for (int x = 0; x < wdt; ++x) {
                        const double b = bt[x * nq + q];
#pragma unroll
                        for (int v = 0; v < 5; ++v)
                            w[v] = fma(b, S[(k * 5 + v) * n2 + x], w[v]);
                    }
                }
                double* t = a.tr + 2 * a.fb_off[f] + (int64_t)side * 5 * nq + 5 * q;
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    t[v] = w[v];
            }
        }
    }
    __syncthreads();
}

template <int P, int TYPE>
struct Collapsed {
    static constexpr int n = P + 1;
    __host__ __device__ static constexpr int jmax(int i) { return TYPE == DGX_ELEM_TET ? P - i : P; }
    __host__ __device__ static constexpr int kmax(int i, int j)
    {
        return TYPE == DGX_ELEM_TET ? P - i - j : P - (i > j ? i : j);
    }
    // first mode of the (i, j) column
    __host__ __device__ static constexpr int off(int i, int j)
    {
        int o = 0;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b <= jmax(a); ++b) {
                if (a == i && b == j)
                    return o;
                o += kmax(a, b) + 1;
            }
        return o;
    }
    __host__ __device__ static constexpr int n_pairs()
    {
        int c = 0;
        for (int a = 0; a < n; ++a)
            c += jmax(a) + 1;
        return c;
    }
    __host__ __device__ static constexpr int n_modes() { return off(n, 0); }
    __host__ __device__ static constexpr int size_b() { return TYPE == DGX_ELEM_TET ? n * n * n : n * n; }
    // TET: C(i, j)[r][k] only for the modes (k <= kmax(i, j)), at ((first mode of (i, j)) + k) n + r
    __host__ __device__ static constexpr int size_c() { return TYPE == DGX_ELEM_TET ? n_modes() * n : n * n * n; }
    // (i, j) of the pair-th valid pair
    __device__ static void pair(int pr, int& i, int& j)
    {
        i = 0;
        while (pr > jmax(i)) {
            pr -= jmax(i) + 1;
            ++i;
        }
        j = pr;
    }
    __device__ static int bidx(int i, int q, int j) { return TYPE == DGX_ELEM_TET ? (i * n + q) * n + j : q * n + j; }
    // o = off(i, j), k <= kmax(i, j)
    __device__ static int cidx(int i, int j, int o, int r, int k)
    {
        return TYPE == DGX_ELEM_TET ? (o + k) * n + r : (((i > j ? i : j) * n + r) * n + k);
    }
};

// One volume tile of TET or PYR p = P elements by sum factorization in collapsed coordinates (a, b, c):
// the orthonormal basis is s_m times a product of one factor per coordinate (A(a), B(b), C(c); TET's
// B depends on i and C on (i, j), PYR's C on max(i, j)), the nodes are the (a, b, c) tensor grid (a
// fastest), and the reference gradient follows from the collapse's chain rule at each node:
//   forward   c: lines (v, i, j) -> T1[i, j, r];  b: lines (v, i, r) -> T2[i, q, r];  a: lines (v, q, r)
//   fluxes    per node: Ga = w (da . fr), Gb = w (db . fr), Gc = w fr_t
//   transpose a: Ad^T Ga, A^T Gb, A^T Gc;  b: B^T y_a + Bd^T y_b, B^T y_c;  c: C^T z_ab + Cd^T z_c, times s
// LDS per element: 15 n^3 doubles (nodal work) + 5 Ns (modal work).
template <int P, int TYPE>
__device__ __attribute__((always_inline)) inline void vol_tile_collapsed(const MixArgs& a, const GroupDev& G,
                                                                        const VolTile& t, double* lds)
{
    using K = Collapsed<P, TYPE>;
    constexpr int n = P + 1, n2 = n * n, n3 = n2 * n, ns = K::n_modes(), npr = K::n_pairs();
    const int tpe = kBlock / G.tile;
    const int slot = threadIdx.x / tpe, l = threadIdx.x - slot * tpe;
    const bool act = slot < t.count;
    const int64_t e = t.e0 + slot;
    constexpr int TAB = sf_tab_lds(TYPE, P);
    const double* tab = lds;      // the factor block, staged in LDS
    // A, Ad are read at compile-time indices: scalar operands through the constant address space
    using cdp = const __attribute__((address_space(4))) double*;
    const cdp Al = (cdp)(a.ops + G.tp_off);
    const cdp Adl = Al + n2;
    const double* Bt = tab + 2 * n2;
    const double* Bdt = Bt + K::size_b();
    const double* Ct = Bdt + K::size_b();
    const double* Cdt = Ct + K::size_c();
    const double* xa = Cdt + K::size_c();
    const double* xb = xa + n;
    const double* xc = xb + n;
    const double* wn = xc + n;
    const double* sm = wn + n3;
    double* W = lds + TAB + slot * (15 * n3 + 5 * ns);
    double* Um = W + 15 * n3;
    for (int i = threadIdx.x; i < sf_tab_len(TYPE, P); i += kBlock)
        lds[i] = a.ops[G.tp_off + i];
    if (act) {
        const double* ue = a.u + a.dof[e];
        for (int i = l; i < 5 * ns; i += tpe)
            Um[i] = ue[i];
    }
    __syncthreads();
    if (a.stage != -2) {
        // forward c: T1[v][i + n j + n2 r] = sum_k C(i,j)[r][k] s_m u_m
        if (act) {
            for (int ln = l; ln < 5 * npr; ln += tpe) {
                const int v = ln / npr;
                int i, j;
                K::pair(ln - v * npr, i, j);
                const int o = K::off(i, j), km = K::kmax(i, j);
                double in[n];
#pragma unroll
                for (int k = 0; k < n; ++k)
                    in[k] = k <= km ? sm[o + k] * Um[v * ns + o + k] : 0.0;
#pragma unroll
                for (int r = 0; r < n; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < n; ++k)
                        acc = fma(Ct[K::cidx(i, j, o, r, k <= km ? k : km)], in[k], acc);   // in[k > km] = 0
                    W[v * n3 + i + n * j + n2 * r] = acc;
                }
            }
        }
        __syncthreads();
        // forward b: lines (v, i, r), in place (j <= jmax(i) read, all q written)
        if (act) {
            for (int ln = l; ln < 5 * n2; ln += tpe) {
                const int v = ln / n2, rr = ln - v * n2;
                const int i = rr % n, r = rr / n;
                const int base = v * n3 + i + n2 * r;
                const int jm = K::jmax(i);
                double in[n];
#pragma unroll
                for (int j = 0; j < n; ++j)
                    in[j] = j <= jm ? W[base + n * j] : 0.0;
#pragma unroll
                for (int q = 0; q < n; ++q) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < n; ++j)
                        acc = fma(Bt[K::bidx(i, q, j)], in[j], acc);
                    W[base + n * q] = acc;
                }
            }
        }
        __syncthreads();
        // forward a: lines (v, q, r), in place
        if (act) {
            for (int ln = l; ln < 5 * n2; ln += tpe) {
                const int v = ln / n2, rr = ln - v * n2;
                const int base = v * n3 + n * rr;
                double in[n];
#pragma unroll
                for (int i = 0; i < n; ++i)
                    in[i] = W[base + i];
#pragma unroll
                for (int pa = 0; pa < n; ++pa) {
                    double acc = 0.0;
#pragma unroll
                    for (int i = 0; i < n; ++i)
                        acc = fma(Al[pa * n + i], in[i], acc);
                    W[base + pa] = acc;
                }
            }
        }
        __syncthreads();
        // fluxes at the nodes, combined through the chain rule of the collapse
        if (act) {
            double m[9];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                m[i] = a.met[9 * e + i];
            for (int pt = l; pt < n3; pt += tpe) {
                const int pa = pt % n, q = (pt / n) % n, r = pt / n2;
                const double ca = xa[pa], cb = xb[q], cc = xc[r];
                double da[3], db[3];
                if (TYPE == DGX_ELEM_TET) {
                    const double ib = 1.0 / ((1.0 - cb) * (1.0 - cc));
                    da[0] = 4.0 * ib;
                    da[1] = 2.0 * (1.0 + ca) * ib;
                    da[2] = da[1];
                    db[0] = 0.0;
                    db[1] = 2.0 / (1.0 - cc);
                    db[2] = (1.0 + cb) / (1.0 - cc);
                } else {
                    const double ic = 1.0 / (1.0 - cc);
                    da[0] = 2.0 * ic;
                    da[1] = 0.0;
                    da[2] = ca * ic;
                    db[0] = 0.0;
                    db[1] = 2.0 * ic;
                    db[2] = cb * ic;
                }
                double w[5];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    w[v] = W[v * n3 + pt];
                const double wq = wn[pt];
                double ga[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, gb[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, gc[5];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double col[3] = { m[d], m[3 + d], m[6 + d] };
                    double fd[5];
                    dgx::euler_flux_n<3>(w, col, fd);
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        ga[v] = fma(da[d], fd[v], ga[v]);
                        gb[v] = fma(db[d], fd[v], gb[v]);
                        if (d == 2)
                            gc[v] = fd[v];
                    }
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    W[v * n3 + pt] = wq * ga[v];
                    W[(5 + v) * n3 + pt] = wq * gb[v];
                    W[(10 + v) * n3 + pt] = wq * gc[v];
                }
            }
        }
        __syncthreads();
        // transpose a: lines (comp, v, q, r), in place: Ad^T for comp a, then A^T for b, c (one operator per loop)
        if (act) {
            for (int ln = l; ln < 5 * n2; ln += tpe) {
                const int cv = ln / n2, rr = ln - cv * n2;
                const int base = cv * n3 + n * rr;
                double in[n];
#pragma unroll
                for (int pa = 0; pa < n; ++pa)
                    in[pa] = W[base + pa];
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    double acc = 0.0;
#pragma unroll
                    for (int pa = 0; pa < n; ++pa)
                        acc = fma(Adl[pa * n + i], in[pa], acc);
                    W[base + i] = acc;
                }
            }
            for (int ln = 5 * n2 + l; ln < 15 * n2; ln += tpe) {
                const int cv = ln / n2, rr = ln - cv * n2;
                const int base = cv * n3 + n * rr;
                double in[n];
#pragma unroll
                for (int pa = 0; pa < n; ++pa)
                    in[pa] = W[base + pa];
#pragma unroll
                for (int i = 0; i < n; ++i) {
                    double acc = 0.0;
#pragma unroll
                    for (int pa = 0; pa < n; ++pa)
                        acc = fma(Al[pa * n + i], in[pa], acc);
                    W[base + i] = acc;
                }
            }
        }
        __syncthreads();
        // transpose b: lines (v, i, r): z_ab = B^T y_a + Bd^T y_b -> block 0; z_c = B^T y_c -> block 2
        if (act) {
            for (int ln = l; ln < 10 * n2; ln += tpe) {
                const bool third = ln >= 5 * n2;
                const int l2 = third ? ln - 5 * n2 : ln;
                const int v = l2 / n2, rr = l2 - v * n2;
                const int i = rr % n, r = rr / n;
                const int b0 = ((third ? 10 : 0) + v) * n3 + i + n2 * r;
                const int b1 = (5 + v) * n3 + i + n2 * r;
                const int jm = K::jmax(i);
                double y0[n], y1[n];
#pragma unroll
                for (int q = 0; q < n; ++q) {
                    y0[q] = W[b0 + n * q];
                    y1[q] = third ? 0.0 : W[b1 + n * q];
                }
#pragma unroll
                for (int j = 0; j < n; ++j) {
                    if (j > jm)
                        continue;
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < n; ++q) {
                        acc = fma(Bt[K::bidx(i, q, j)], y0[q], acc);
                        if (!third)
                            acc = fma(Bdt[K::bidx(i, q, j)], y1[q], acc);
                    }
                    W[b0 + n * j] = acc;
                }
            }
        }
        __syncthreads();
        // transpose c: lines (v, i, j) -> modes, times s
        if (act) {
            for (int ln = l; ln < 5 * npr; ln += tpe) {
                const int v = ln / npr;
                int i, j;
                K::pair(ln - v * npr, i, j);
                const int o = K::off(i, j), km = K::kmax(i, j);
                const int b0 = v * n3 + i + n * j, b2 = (10 + v) * n3 + i + n * j;
                double z0[n], z2[n];
#pragma unroll
                for (int r = 0; r < n; ++r) {
                    z0[r] = W[b0 + n2 * r];
                    z2[r] = W[b2 + n2 * r];
                }
#pragma unroll
                for (int k = 0; k < n; ++k) {
                    if (k > km)
                        continue;
                    double acc = 0.0;
#pragma unroll
                    for (int r = 0; r < n; ++r) {
                        acc = fma(Ct[K::cidx(i, j, o, r, k)], z0[r], acc);
                        acc = fma(Cdt[K::cidx(i, j, o, r, k)], z2[r], acc);
                    }
                    Um[v * ns + o + k] = sm[o + k] * acc;
                }
            }
        }
        __syncthreads();
        if (act) {
            const double* ue = a.u + a.dof[e];
            // the runtime Ns keeps the dense lift loops rolled
            const int nsr = G.ns;
            for (int j = l; j < nsr; j += tpe) {
                double r[5], s_old[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    r[v] = Um[v * nsr + j];
                    s_old[v] = ue[v * nsr + j];
                }
                finish_mode(a, e, nsr, j, r, s_old, Um, true, tpe >= 64);
            }
        }
        __syncthreads();
    }
    if (a.stage != -1 && act)
        write_traces(a, e, G.ns, Um, l, tpe);
    __syncthreads();
}

__device__ __attribute__((always_inline)) inline void vol_dispatch(const MixArgs& a, const VolTile& t, double* lds)
{
    const GroupDev G = a.groups[t.group];
    if (G.tp_off >= 0 && G.type == DGX_ELEM_TET) {
        switch (G.p) {
        case 1: vol_tile_collapsed<1, DGX_ELEM_TET>(a, G, t, lds); return;
        case 2: vol_tile_collapsed<2, DGX_ELEM_TET>(a, G, t, lds); return;
        case 3: vol_tile_collapsed<3, DGX_ELEM_TET>(a, G, t, lds); return;
        case 4: vol_tile_collapsed<4, DGX_ELEM_TET>(a, G, t, lds); return;
        case 5: vol_tile_collapsed<5, DGX_ELEM_TET>(a, G, t, lds); return;
        default: break;
        }
    }
    if (G.tp_off >= 0 && G.type == DGX_ELEM_PYR) {
        switch (G.p) {
        case 1: vol_tile_collapsed<1, DGX_ELEM_PYR>(a, G, t, lds); return;
        case 2: vol_tile_collapsed<2, DGX_ELEM_PYR>(a, G, t, lds); return;
        case 3: vol_tile_collapsed<3, DGX_ELEM_PYR>(a, G, t, lds); return;
        case 4: vol_tile_collapsed<4, DGX_ELEM_PYR>(a, G, t, lds); return;
        case 5: vol_tile_collapsed<5, DGX_ELEM_PYR>(a, G, t, lds); return;
        default: break;
        }
    }
    if (G.tp_off >= 0 && G.type == DGX_ELEM_WEDGE) {
        switch (G.p) {
        case 1: vol_tile_wedge<1>(a, G, t, lds); return;
        case 2: vol_tile_wedge<2>(a, G, t, lds); return;
        case 3: vol_tile_wedge<3>(a, G, t, lds); return;
        case 4: vol_tile_wedge<4>(a, G, t, lds); return;
        case 5: vol_tile_wedge<5>(a, G, t, lds); return;
        default: break;
        }
    }
    if (G.tp_off >= 0 && G.type == DGX_ELEM_HEX) {
        switch (G.p) {
        case 1: vol_tile_hex<1>(a, G, t, lds); return;
        case 2: vol_tile_hex<2>(a, G, t, lds); return;
        case 3: vol_tile_hex<3>(a, G, t, lds); return;
        case 4: vol_tile_hex<4>(a, G, t, lds); return;
        case 5: vol_tile_hex<5>(a, G, t, lds); return;
        default: break;
        }
    }
    vol_tile(a, G, t, lds);
}

__global__ void reset_counters(unsigned int* c)
{
    if (threadIdx.x < 2)
        c[threadIdx.x] = 0u;
}

// Face tiles of one size class cost the same, so the persistent face grid takes them round-robin (no work
// counter: one atomic per small tile serializes on its address).
template <int FLUX>
__global__ __launch_bounds__(kBlock) void face_persistent(MixArgs a)
{
    for (int64_t t = blockIdx.x; t < a.n_ftiles; t += gridDim.x)
        face_tile<FLUX>(a, a.ftiles[t]);
}

template <int FLUX>
__global__ __launch_bounds__(kBlock) void face_range(MixArgs a, int64_t t0)
{
    face_tile<FLUX>(a, a.ftiles[t0 + blockIdx.x]);
}

// Volume tiles are taken kVolChunk at a time from the work counter (largest first), so the counter sees
// one atomic per kVolChunk tiles.
constexpr int kVolChunk = 4;

__device__ __attribute__((always_inline)) inline void vol_persistent_body(const MixArgs& a, double* lds, int* s_t)
{
    for (;;) {
        if (threadIdx.x == 0)
            *s_t = (int)atomicAdd(a.counter + 1, (unsigned)kVolChunk);
        __syncthreads();
        const int t0 = *s_t;
        __syncthreads();
        if (t0 >= a.n_vtiles)
            break;
        const int t1 = t0 + kVolChunk < a.n_vtiles ? t0 + kVolChunk : (int)a.n_vtiles;
        for (int t = t0; t < t1; ++t)
            vol_dispatch(a, a.vtiles[a.vorder[t]], lds);
    }
}

// The persistent volume kernel holds every (type, p) path, so its register count is the largest path's
// plus what the tile loop keeps live; three variants trade that for occupancy (2, 3 or 4 waves per SIMD,
// the last two with a few spilled registers) and the autotuner picks one.
__global__ __launch_bounds__(kBlock) void vol_persistent(MixArgs a)
{
    extern __shared__ double lds[];
    __shared__ int s_t;
    vol_persistent_body(a, lds, &s_t);
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 3))) void vol_persistent_w3(MixArgs a)
{
    extern __shared__ double lds[];
    __shared__ int s_t;
    vol_persistent_body(a, lds, &s_t);
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void vol_persistent_w4(MixArgs a)
{
    extern __shared__ double lds[];
    __shared__ int s_t;
    vol_persistent_body(a, lds, &s_t);
}

// The grouped form: one launch over the tiles of every group, one tile per workgroup, in the
// persistent order (largest first); the hardware dispatcher balances the load, and without a work loop
// the kernel keeps the register count of a single tile.
__global__ __launch_bounds__(kBlock) void vol_grouped(MixArgs a, int64_t t0)
{
    extern __shared__ double lds[];
    vol_dispatch(a, a.vtiles[a.vorder[t0 + blockIdx.x]], lds);
}

__global__ __launch_bounds__(kBlock) void vol_range(MixArgs a, int64_t t0)
{
    extern __shared__ double lds[];
    vol_dispatch(a, a.vtiles[t0 + blockIdx.x], lds);
}

// --------------------------------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------------------------------

template <class T>
T* dev_copy(const std::vector<T>& h, hipError_t* err)
{
    T* p = nullptr;
    if (h.empty())
        return nullptr;
    *err = hipMalloc(&p, h.size() * sizeof(T));
    if (*err != hipSuccess)
        return nullptr;
    *err = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return p;
}

int face_lg(int nq)
{
    int lg = 2;
    while ((1 << lg) < nq)
        ++lg;
    return lg;
}

const int kFacesOf[4] = { 6, 5, 4, 5 };

// message of the last failed dgx_mixed_create on this thread (no context to hold it)
thread_local std::string t_create_err;

}  // namespace

struct dgx_mixed_ctx {
    std::string err;
    int flux = 0;
    int64_t E = 0, F = 0, n_dof = 0, fb_len = 0;
    int n_group = 0;
    int mode = DGX_MIXED_LAUNCH_PERSISTENT;
    int bpc = 2;
    int pwaves = 2;                      // occupancy variant of the persistent volume kernel (waves/SIMD)
    std::vector<int32_t> tiles_of[3];    // tile shapes of each launch form
    int n_cu = 0;
    std::vector<GroupDev> groups;
    std::vector<int32_t> face_lg_of;     // per face
    // device
    double *u = nullptr, *up = nullptr, *rhs = nullptr, *fb = nullptr, *tr = nullptr;
    double *ops = nullptr, *jdet = nullptr, *met = nullptr, *fnorm = nullptr, *fjac = nullptr, *fop = nullptr,
           *fw = nullptr;
    int64_t fl_len = 0;
    int64_t *fl_off = nullptr;
    double* fl = nullptr;
    int64_t *dof = nullptr, *fb_off = nullptr, *fop_off = nullptr, *fopt_off = nullptr, *fw_off = nullptr,
            *fopr_off = nullptr, *foprt_off = nullptr, *fte_off = nullptr;
    int32_t *fopr_n = nullptr, *ftc = nullptr;
    int tensor_faces = 1;                // quad face maps in tensor form where they factor
    int n_tensor_ops = 0;                // face operators that have the tensor form
    int32_t *ef = nullptr, *elf = nullptr, *face_elem = nullptr, *face_op = nullptr, *face_wt = nullptr, *fop_nq = nullptr,
            *fop_ns = nullptr, *face_order = nullptr, *vorder = nullptr, *gorder = nullptr;
    GroupDev* d_groups = nullptr;
    VolTile* vtiles = nullptr;
    FaceTile* ftiles = nullptr;
    unsigned int* counter = nullptr;
    int64_t n_vtiles = 0, n_ftiles = 0;
    std::vector<int64_t> vtile_start;    // [n_group+1] per-group ranges of vtiles
    struct GClass {
        int64_t start, count;            // range of gorder
        size_t lds;                      // the launch's LDS: the largest tile of the class
    };
    std::vector<GClass> gclass;          // grouped form: one launch per LDS class
    std::vector<int64_t> ftile_class;    // [k+1] ranges of face tiles of one lanes-per-face class
    size_t lds_max = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    double graph_dt = 0.0;
    std::vector<void*> allocs;
};

namespace {

dgx_status fail(dgx_mixed_ctx* c, dgx_status s, const std::string& m)
{
    if (c)
        c->err = m;
    return s;
}

#define MIX_TRY(c, expr)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(c, DGX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

size_t tile_lds(const GroupDev& g, int tile)
{
    size_t per = 5 * (size_t)g.ns + 15 * (size_t)g.nc;      // dense: coefficients + fluxes
    size_t tab = kTpLds;
    if (g.tp_off >= 0) {
        tab = (size_t)sf_tab_lds(g.type, g.p);               // the staged factor block
        if (g.type == DGX_ELEM_HEX)
            per = (size_t)hex_slot(g.p + 1);                 // in place, n^3 nodes = modes
        else
            per = 15 * (size_t)g.nc + 5 * (size_t)g.ns;      // nodal + modal work
    }
    return (tab + (size_t)tile * per) * sizeof(double);
}

// doubles of a group's factor block (sf_tab_len)
int64_t tp_len(const GroupDev& g)
{
    return sf_tab_len(g.type, g.p);
}

bool tile_ok(const GroupDev& g, int tile)
{
    if (tile < 1 || tile > 64 || (tile & (tile - 1)) || kBlock / tile < 4)
        return false;
    return tile_lds(g, tile) <= (size_t)kMaxTileLds;
}

int default_tile(const GroupDev& g)
{
    int t = 1;
    while (2 * t <= 64 && kBlock / (2 * t) >= std::max(g.nc, g.ns) && tile_ok(g, 2 * t))
        t *= 2;
    // lanes per element >= the larger of Nc, Ns where possible, else one element per tile
    return t;
}

// Rebuilds the volume tiles from the groups' tile sizes (and the persistent order, largest first).
dgx_status build_vtiles(dgx_mixed_ctx* c)
{
    std::vector<VolTile> vt;
    c->vtile_start.assign(c->n_group + 1, 0);
    c->lds_max = 0;
    for (int g = 0; g < c->n_group; ++g) {
        const GroupDev& G = c->groups[g];
        c->vtile_start[g] = (int64_t)vt.size();
        for (int64_t e = 0; e < G.count; e += G.tile)
            vt.push_back({ G.start + e, g, (int32_t)std::min<int64_t>(G.tile, G.count - e) });
        c->lds_max = std::max(c->lds_max, tile_lds(G, G.tile));
    }
    c->vtile_start[c->n_group] = (int64_t)vt.size();
    // cost of a tile ~ its elements x lanes' work (Nc Ns (1 + 3)) / lanes: order largest first
    std::vector<double> cost(vt.size());
    for (size_t i = 0; i < vt.size(); ++i) {
        const GroupDev& G = c->groups[vt[i].group];
        cost[i] = (double)G.nc * G.ns * 4.0 * vt[i].count / (kBlock / G.tile);
    }
    std::vector<int32_t> order(vt.size());
    for (size_t i = 0; i < order.size(); ++i)
        order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    // grouped form: one launch per LDS class, the workgroups a CU can hold at the class's largest tile
    // (at most kGroupedBpc, the register limit), so that one large-tile group does not set the LDS, and
    // with it the occupancy, of every other group's workgroups
    auto bpc_of = [](size_t lds) {
        return std::max(1, std::min(kGroupedBpc, (int)((160u << 10) / std::max<size_t>(lds, 1))));
    };
    std::vector<int32_t> gorder(order);
    std::stable_sort(gorder.begin(), gorder.end(), [&](int32_t x, int32_t y) {
        return bpc_of(tile_lds(c->groups[vt[x].group], c->groups[vt[x].group].tile)) <
               bpc_of(tile_lds(c->groups[vt[y].group], c->groups[vt[y].group].tile));
    });
    c->gclass.clear();
    for (size_t i = 0; i < gorder.size(); ++i) {
        const GroupDev& G = c->groups[vt[gorder[i]].group];
        const size_t l = tile_lds(G, G.tile);
        if (c->gclass.empty() || bpc_of(c->gclass.back().lds) != bpc_of(l))
            c->gclass.push_back({ (int64_t)i, 0, l });
        c->gclass.back().count += 1;
        c->gclass.back().lds = std::max(c->gclass.back().lds, l);
    }
    if (c->vtiles)
        (void)hipFree(c->vtiles);
    if (c->vorder)
        (void)hipFree(c->vorder);
    if (c->gorder)
        (void)hipFree(c->gorder);
    c->vtiles = nullptr;
    c->vorder = nullptr;
    c->gorder = nullptr;
    hipError_t e = hipSuccess;
    c->vtiles = dev_copy(vt, &e);
    MIX_TRY(c, e);
    c->vorder = dev_copy(order, &e);
    MIX_TRY(c, e);
    c->gorder = dev_copy(gorder, &e);
    MIX_TRY(c, e);
    c->n_vtiles = (int64_t)vt.size();
    std::vector<GroupDev> gd = c->groups;
    MIX_TRY(c, hipMemcpy(c->d_groups, gd.data(), gd.size() * sizeof(GroupDev), hipMemcpyHostToDevice));
    return DGX_OK;
}

MixArgs make_args(dgx_mixed_ctx* c, const double* u, const double* reg, double* out, int stage, int minv, double dt)
{
    MixArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = u;
    a.reg = reg;
    a.out = out;
    a.stage = stage;
    a.apply_minv = minv;
    a.dt = dt;
    a.groups = c->d_groups;
    a.ops = c->ops;
    a.dof = c->dof;
    a.jdet = c->jdet;
    a.met = c->met;
    a.ef = c->ef;
    a.face_elem = c->face_elem;
    a.face_op = c->face_op;
    a.face_wt = c->face_wt;
    a.face_normal = c->fnorm;
    a.face_jac = c->fjac;
    a.fb_off = c->fb_off;
    a.fb = c->fb;
    a.fl_off = c->fl_off;
    a.fl = c->fl;
    a.tr = c->tr;
    a.fop_nq = c->fop_nq;
    a.fop_ns = c->fop_ns;
    a.fop_off = c->fop_off;
    a.fopt_off = c->fopt_off;
    a.fopr_off = c->fopr_off;
    a.foprt_off = c->foprt_off;
    a.fopr_n = c->fopr_n;
    a.ftc = c->ftc;
    a.fte_off = c->fte_off;
    a.tensor_faces = c->tensor_faces;
    a.face_lf_of = c->elf;
    a.fop = c->fop;
    a.fw_off = c->fw_off;
    a.fw = c->fw;
    a.face_order = c->face_order;
    a.vtiles = c->vtiles;
    a.vorder = c->vorder;
    a.n_vtiles = c->n_vtiles;
    a.ftiles = c->ftiles;
    a.n_ftiles = c->n_ftiles;
    a.counter = c->counter;
    return a;
}

hipError_t launch_face(dgx_mixed_ctx* c, const MixArgs& a, int mode)
{
    if (mode == DGX_MIXED_LAUNCH_GROUPED && c->n_ftiles > 0) {
        if (c->flux == 0)
            face_range<0><<<(unsigned)c->n_ftiles, kBlock, 0, c->stream>>>(a, 0);
        else
            face_range<1><<<(unsigned)c->n_ftiles, kBlock, 0, c->stream>>>(a, 0);
        return hipGetLastError();
    }
    if (mode == DGX_MIXED_LAUNCH_PERSISTENT) {
        const int grid = (int)std::min<int64_t>((int64_t)c->n_cu * c->bpc, std::max<int64_t>(c->n_ftiles, 1));
        if (c->flux == 0)
            face_persistent<0><<<grid, kBlock, 0, c->stream>>>(a);
        else
            face_persistent<1><<<grid, kBlock, 0, c->stream>>>(a);
        return hipGetLastError();
    }
    for (size_t k = 0; k + 1 < c->ftile_class.size(); ++k) {
        const int64_t t0 = c->ftile_class[k], n = c->ftile_class[k + 1] - t0;
        if (n <= 0)
            continue;
        if (c->flux == 0)
            face_range<0><<<(unsigned)n, kBlock, 0, c->stream>>>(a, t0);
        else
            face_range<1><<<(unsigned)n, kBlock, 0, c->stream>>>(a, t0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_vol(dgx_mixed_ctx* c, const MixArgs& a, int mode, int only_group = -1)
{
    if (mode == DGX_MIXED_LAUNCH_PERSISTENT && only_group < 0) {
        const int grid = (int)std::min<int64_t>((int64_t)c->n_cu * c->bpc, std::max<int64_t>(c->n_vtiles, 1));
        if (c->pwaves == 4)
            vol_persistent_w4<<<grid, kBlock, c->lds_max, c->stream>>>(a);
        else if (c->pwaves == 3)
            vol_persistent_w3<<<grid, kBlock, c->lds_max, c->stream>>>(a);
        else
            vol_persistent<<<grid, kBlock, c->lds_max, c->stream>>>(a);
        return hipGetLastError();
    }
    if (mode == DGX_MIXED_LAUNCH_GROUPED && only_group < 0) {
        MixArgs ag = a;
        ag.vorder = c->gorder;
        for (const auto& k : c->gclass) {
            if (k.count <= 0)
                continue;
            vol_grouped<<<(unsigned)k.count, kBlock, k.lds, c->stream>>>(ag, k.start);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
    for (int g = 0; g < c->n_group; ++g) {
        if (only_group >= 0 && g != only_group)
            continue;
        const int64_t t0 = c->vtile_start[g], n = c->vtile_start[g + 1] - t0;
        if (n <= 0)
            continue;
        vol_range<<<(unsigned)n, kBlock, tile_lds(c->groups[g], c->groups[g].tile), c->stream>>>(a, t0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// One evaluation: counters reset, face kernel, volume kernel.
hipError_t launch_eval(dgx_mixed_ctx* c, const double* u, const double* reg, double* out, int stage, int minv,
                       double dt)
{
    MixArgs a = make_args(c, u, reg, out, stage, minv, dt);
    reset_counters<<<1, 64, 0, c->stream>>>(c->counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    e = launch_face(c, a, c->mode);
    if (e != hipSuccess)
        return e;
    return launch_vol(c, a, c->mode);
}

// The traces of u (after a state upload): the volume kernel's trace phase alone.
hipError_t launch_traces(dgx_mixed_ctx* c, const double* u)
{
    MixArgs a = make_args(c, u, nullptr, nullptr, -2, 0, 0.0);
    reset_counters<<<1, 64, 0, c->stream>>>(c->counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return launch_vol(c, a, c->mode);
}

hipError_t launch_step(dgx_mixed_ctx* c, double dt)
{
    hipError_t e = launch_eval(c, c->u, nullptr, c->up, 0, 1, dt);
    if (e == hipSuccess)
        e = launch_eval(c, c->up, c->u, c->up, 1, 1, dt);
    if (e == hipSuccess)
        e = launch_eval(c, c->up, c->u, c->u, 2, 1, dt);
    return e;
}

void drop_graph(dgx_mixed_ctx* c)
{
    if (c->graph)
        (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
}

dgx_status ensure_graph(dgx_mixed_ctx* c, double dt)
{
    if (c->graph && c->graph_dt == dt)
        return DGX_OK;
    drop_graph(c);
    hipGraph_t g = nullptr;
    MIX_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = launch_step(c, dt);
    hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return fail(c, DGX_ERR_HIP, std::string("stage capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipError_t ie = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        c->graph = nullptr;
        return fail(c, DGX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    c->graph_dt = dt;
    return DGX_OK;
}

void free_all(dgx_mixed_ctx* c)
{
    drop_graph(c);
    for (void* p : c->allocs)
        if (p)
            (void)hipFree(p);
    c->allocs.clear();
    if (c->vtiles)
        (void)hipFree(c->vtiles);
    if (c->vorder)
        (void)hipFree(c->vorder);
    if (c->gorder)
        (void)hipFree(c->gorder);
    c->vtiles = nullptr;
    c->vorder = nullptr;
    c->gorder = nullptr;
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

// Whether a quad face's face-modal rows BF [nq][n^2] (a fastest) are a permuted tensor product,
// BF[q][a + n b] = E[ia][a] E[ib][b] with E [m][n] the 1-D basis at m points per direction and (ia, ib)
// the node q of the face rule's m x m grid under one of the square's 8 symmetries (face_grid_of).  The
// points come from the ratio of each row's linear to its constant mode along a and along b; E from the
// rows' b = 0 entries, with E[.][0] the (constant) value of the constant mode; every entry is checked
// to 1e-12 of the table's scale.  Returns the symmetry code and fills m and E, or returns -1.
int tensor_face_rows(const double* BF, int nq, int n, int* m_out, std::vector<double>& E)
{
    int m = 1;
    while (m * m < nq)
        ++m;
    if (m * m != nq || n < 2 || m < n || m > 8)
        return -1;
    const int n2 = n * n;
    double scale = 0.0;
    for (int64_t i = 0; i < (int64_t)nq * n2; ++i)
        scale = std::max(scale, std::fabs(BF[i]));
    std::vector<double> xa(nq), xb(nq);
    for (int q = 0; q < nq; ++q) {
        const double c0 = BF[(int64_t)q * n2];
        if (!(c0 > 0.0) || !std::isfinite(c0))
            return -1;
        xa[q] = BF[(int64_t)q * n2 + 1] / c0;
        xb[q] = BF[(int64_t)q * n2 + n] / c0;
    }
    std::vector<double> srt(xa);
    std::sort(srt.begin(), srt.end());
    std::vector<double> pts;
    for (double x : srt)
        if (pts.empty() || x - pts.back() > 1e-9)
            pts.push_back(x);
    if ((int)pts.size() != m)
        return -1;
    auto index_of = [&](double x) {
        for (int i = 0; i < m; ++i)
            if (std::fabs(x - pts[i]) <= 1e-9)
                return i;
        return -1;
    };
    std::vector<int> ia(nq), ib(nq);
    for (int q = 0; q < nq; ++q) {
        ia[q] = index_of(xa[q]);
        ib[q] = index_of(xb[q]);
        if (ia[q] < 0 || ib[q] < 0)
            return -1;
    }
    int code = -1;
    for (int cd = 0; cd < 8 && code < 0; ++cd) {
        bool ok = true;
        for (int q = 0; q < nq && ok; ++q) {
            const int qx = q % m, qy = q / m;
            const int u = (cd & 1) ? qy : qx, v = (cd & 1) ? qx : qy;
            ok = ia[q] == ((cd & 2) ? m - 1 - u : u) && ib[q] == ((cd & 4) ? m - 1 - v : v);
        }
        if (ok)
            code = cd;
    }
    if (code < 0)
        return -1;
    E.assign((size_t)m * n, 0.0);
    std::vector<char> have(m, 0);
    for (int q = 0; q < nq; ++q) {
        if (have[ia[q]])
            continue;
        const double e0 = std::sqrt(BF[(int64_t)q * n2]);
        for (int a = 0; a < n; ++a)
            E[(size_t)ia[q] * n + a] = BF[(int64_t)q * n2 + a] / e0;
        have[ia[q]] = 1;
    }
    for (int q = 0; q < nq; ++q)
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a)
                if (!(std::fabs(BF[(int64_t)q * n2 + a + n * b] - E[(size_t)ia[q] * n + a] * E[(size_t)ib[q] * n + b]) <=
                      1e-12 * scale))
                    return -1;
    *m_out = m;
    return code;
}

template <class T>
dgx_status upload(dgx_mixed_ctx* c, T** dst, const std::vector<T>& h)
{
    hipError_t e = hipSuccess;
    *dst = dev_copy(h, &e);
    MIX_TRY(c, e);
    c->allocs.push_back((void*)*dst);
    return DGX_OK;
}

template <class T>
dgx_status alloc(dgx_mixed_ctx* c, T** dst, size_t n)
{
    MIX_TRY(c, hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T)));
    c->allocs.push_back((void*)*dst);
    return DGX_OK;
}

#define MIX_CHECK(c, cond, msg)                                                                   \
    do {                                                                                          \
        if (!(cond))                                                                              \
            return fail(c, DGX_ERR_INVALID_ARG, msg);                                             \
    } while (0)
#define MIX_RUN(expr)                                                                             \
    do {                                                                                          \
        dgx_status s_ = (expr);                                                                   \
        if (s_ != DGX_OK)                                                                         \
            return s_;                                                                            \
    } while (0)

// Host validation and set-up; device memory is allocated by the caller's context afterwards.
dgx_status setup(dgx_mixed_ctx* c, const dgx_mixed_mesh* m)
{
    MIX_CHECK(c, m->n_elem > 0 && m->n_face > 0 && m->n_group > 0 && m->n_elem < (1LL << 30) &&
                     m->n_face < (1LL << 30), "empty mesh or too many elements/faces");
    MIX_CHECK(c, m->group_type && m->group_p && m->group_ns && m->group_nc && m->group_start && m->vol_ops &&
                     m->vol_off && m->jdet && m->metrics && m->face_elem && m->face_op && m->face_wt &&
                     m->face_normal && m->face_jac && m->fop_nq && m->fop_ns && m->fop_off && m->fop_data &&
                     m->fw_n && m->fw_off && m->fw_data, "null table");
    MIX_CHECK(c, m->n_face_op > 0 && m->n_face_wt > 0, "no face operators");
    c->E = m->n_elem;
    c->F = m->n_face;
    c->n_group = m->n_group;
    MIX_CHECK(c, m->group_start[0] == 0 && m->group_start[m->n_group] == m->n_elem, "group_start range");
    // groups and dof offsets
    std::vector<int32_t> ns_of(c->E), type_of(c->E), group_of(c->E);
    std::vector<int64_t> dof(c->E + 1, 0);
    int64_t vol_len = 0;
    c->groups.resize(c->n_group);
    for (int g = 0; g < c->n_group; ++g) {
        GroupDev& G = c->groups[g];
        G.type = m->group_type[g];
        G.p = m->group_p[g];
        G.ns = m->group_ns[g];
        G.nc = m->group_nc[g];
        G.start = m->group_start[g];
        G.count = m->group_start[g + 1] - G.start;
        G.op_off = m->vol_off[g];
        G.tp_off = m->vol_tp_off ? m->vol_tp_off[g] : -1;
        MIX_CHECK(c, G.type >= 0 && G.type <= 3 && G.p >= 1 && G.p <= 5, "group type/order");
        MIX_CHECK(c, G.ns > 0 && G.ns <= 512 && G.nc > 0 && G.nc <= 512, "group Ns/Nc");
        MIX_CHECK(c, G.count > 0, "empty group");
        MIX_CHECK(c, G.op_off >= 0, "vol_off");
        vol_len = std::max<int64_t>(vol_len, G.op_off + 4LL * G.ns * G.nc);
        if (G.tp_off >= 0) {
            const int n = G.p + 1, nt = n * (n + 1) / 2;
            const bool hex = G.type == DGX_ELEM_HEX && G.ns == n * n * n && G.nc == G.ns;
            const bool wedge = G.type == DGX_ELEM_WEDGE && G.ns == n * nt && G.nc == n * n * n;
            const bool tet = G.type == DGX_ELEM_TET && G.ns == n * (n + 1) * (n + 2) / 6 && G.nc == n * n * n;
            const bool pyr = G.type == DGX_ELEM_PYR && G.ns == n * (n + 1) * (2 * n + 1) / 6 && G.nc == n * n * n;
            MIX_CHECK(c, hex || wedge || tet || pyr, "factors given for a group of other sizes than its type's");
            vol_len = std::max<int64_t>(vol_len, G.tp_off + tp_len(G));
        }
        G.tile = default_tile(G);
        MIX_CHECK(c, tile_ok(G, G.tile), "group too large for a volume tile");
        for (int64_t e = G.start; e < G.start + G.count; ++e) {
            ns_of[e] = G.ns;
            type_of[e] = G.type;
            group_of[e] = g;
        }
    }
    for (int64_t e = 0; e < c->E; ++e)
        dof[e + 1] = dof[e] + 5LL * ns_of[e];
    c->n_dof = dof[c->E];
    for (int64_t e = 0; e < c->E; ++e)
        MIX_CHECK(c, std::isfinite(m->jdet[e]) && m->jdet[e] > 0.0, "det J <= 0 or not finite");
    // face operator tables
    std::vector<int32_t> fnq(m->fop_nq, m->fop_nq + m->n_face_op), fns(m->fop_ns, m->fop_ns + m->n_face_op);
    std::vector<int64_t> foff(m->fop_off, m->fop_off + m->n_face_op), fofft(m->n_face_op);
    int64_t fop_len = 0;
    for (int o = 0; o < m->n_face_op; ++o) {
        MIX_CHECK(c, fnq[o] > 0 && fnq[o] <= 64 && fns[o] > 0 && fns[o] <= 512 && foff[o] >= 0, "face operator size");
        fop_len = std::max<int64_t>(fop_len, foff[o] + (int64_t)fnq[o] * fns[o]);
    }
    // face-modal rows of HEX sides: BF [Nq][n^2], n^3 = Ns
    std::vector<int64_t> froff(m->n_face_op, -1), frofft(m->n_face_op, -1);
    std::vector<int32_t> fn2(m->n_face_op, 0);
    if (m->fop_red_off) {
        MIX_CHECK(c, m->fop_red_n, "face-modal rows without their widths");
        for (int o = 0; o < m->n_face_op; ++o) {
            froff[o] = m->fop_red_off[o];
            if (froff[o] < 0)
                continue;
            fn2[o] = m->fop_red_n[o];
            MIX_CHECK(c, fn2[o] > 0 && fn2[o] <= 64 && fn2[o] < fns[o], "face-modal row width");
            fop_len = std::max<int64_t>(fop_len, froff[o] + (int64_t)fnq[o] * fn2[o]);
        }
    }
    // transposed copies after the given data
    std::vector<double> fop(m->fop_data, m->fop_data + fop_len);
    for (int o = 0; o < m->n_face_op; ++o) {
        fofft[o] = (int64_t)fop.size();
        for (int j = 0; j < fns[o]; ++j)
            for (int q = 0; q < fnq[o]; ++q)
                fop.push_back(m->fop_data[foff[o] + (int64_t)q * fns[o] + j]);
        if (froff[o] >= 0) {
            frofft[o] = (int64_t)fop.size();
            for (int j = 0; j < fn2[o]; ++j)
                for (int q = 0; q < fnq[o]; ++q)
                    fop.push_back(m->fop_data[froff[o] + (int64_t)q * fn2[o] + j]);
        }
    }
    // quad face-modal rows in tensor form (HEX faces, WEDGE quad faces): their 1-D factor after the data
    std::vector<int32_t> ftc(m->n_face_op, 0);
    std::vector<int64_t> fte(m->n_face_op, -1);
    for (int o = 0; o < m->n_face_op; ++o) {
        if (froff[o] < 0)
            continue;
        int n = 1;
        while (n * n < fn2[o])
            ++n;
        if (n * n != fn2[o])
            continue;
        int mq = 0;
        std::vector<double> E;
        const int code = tensor_face_rows(m->fop_data + froff[o], fnq[o], n, &mq, E);
        if (code < 0)
            continue;
        ftc[o] = mq | code << 8 | n << 12;
        ++c->n_tensor_ops;
        fte[o] = (int64_t)fop.size();
        fop.insert(fop.end(), E.begin(), E.end());
    }
    int64_t fw_len = 0;
    for (int w = 0; w < m->n_face_wt; ++w) {
        MIX_CHECK(c, m->fw_n[w] > 0 && m->fw_n[w] <= 64 && m->fw_off[w] >= 0, "face weight table");
        fw_len = std::max<int64_t>(fw_len, m->fw_off[w] + m->fw_n[w]);
    }
    // faces: sides, operators, element face lists, flux buffer offsets
    std::vector<int32_t> ef(6 * c->E, -1), elf(6 * c->E, -1), nfe(c->E, 0);
    // elements whose lifts and traces run in face-modal form: sum-factorized HEX and WEDGE groups
    std::vector<char> hex_tp(c->E, 0);
    for (int g = 0; g < c->n_group; ++g)
        if (c->groups[g].tp_off >= 0 && (c->groups[g].type == DGX_ELEM_HEX || c->groups[g].type == DGX_ELEM_WEDGE))
            for (int64_t e = c->groups[g].start; e < c->groups[g].start + c->groups[g].count; ++e)
                hex_tp[e] = (char)(1 + c->groups[g].type);
    std::vector<int64_t> fb_off(c->F), fl_off(2 * c->F, -1);
    int64_t fb = 0, fl = 0;
    c->face_lg_of.resize(c->F);
    for (int64_t f = 0; f < c->F; ++f) {
        const int32_t eL = m->face_elem[2 * f], eR = m->face_elem[2 * f + 1];
        MIX_CHECK(c, eL >= 0 && eL < c->E && eR >= 0 && eR < c->E && eL != eR, "face_elem");
        const int32_t oL = m->face_op[2 * f], oR = m->face_op[2 * f + 1], w = m->face_wt[f];
        MIX_CHECK(c, oL >= 0 && oL < m->n_face_op && oR >= 0 && oR < m->n_face_op, "face_op");
        MIX_CHECK(c, w >= 0 && w < m->n_face_wt, "face_wt");
        MIX_CHECK(c, fns[oL] == ns_of[eL] && fns[oR] == ns_of[eR], "face operator modes differ from the volume's");
        MIX_CHECK(c, fnq[oL] == fnq[oR] && m->fw_n[w] == fnq[oL], "face cubature sizes differ between sides");
        MIX_CHECK(c, std::isfinite(m->face_jac[f]) && m->face_jac[f] > 0.0, "face Jacobian");
        for (int s = 0; s < 2; ++s) {
            const int32_t e = s ? eR : eL;
            const int32_t lf = m->face_lf ? m->face_lf[2 * f + s] : -1;
            MIX_CHECK(c, nfe[e] < kFacesOf[type_of[e]], "element with more faces than its type has");
            if (hex_tp[e]) {
                // the sum-factorized HEX and WEDGE paths lift and trace in face-modal form
                const int nlf = hex_tp[e] == 1 + DGX_ELEM_HEX ? 6 : 5;
                const int o = m->face_op[2 * f + s];
                MIX_CHECK(c, lf >= 0 && lf < nlf && froff[o] >= 0,
                          "HEX / WEDGE sides of a sum-factorized group need face_lf and face-modal rows");
                const int n = c->groups[group_of[e]].p + 1;
                MIX_CHECK(c, fn2[o] == ((hex_tp[e] == 1 + DGX_ELEM_WEDGE && lf >= 3) ? n * (n + 1) / 2 : n * n),
                          "face-modal row width differs from the side's face");
                for (int k = 0; k < nfe[e]; ++k)
                    MIX_CHECK(c, elf[6 * e + k] != lf, "two faces of one element on the same local face");
                // the face kernel projects the side's flux onto its face-modal rows: [5][width]
                fl_off[2 * f + s] = fl;
                fl += 5LL * fn2[o];
            }
            elf[6 * e + nfe[e]] = lf;
            ef[6 * e + nfe[e]++] = (int32_t)(2 * f + s);
        }
        fb_off[f] = fb;
        fb += 5LL * fnq[oL];
        c->face_lg_of[f] = face_lg(fnq[oL]);
    }
    for (int64_t e = 0; e < c->E; ++e)
        MIX_CHECK(c, nfe[e] == kFacesOf[type_of[e]], "element with a face missing (the mesh must be closed)");
    c->fb_len = fb;
    c->fl_len = fl;
    // face tiles: faces ordered by lanes-per-face class
    std::vector<int32_t> order(c->F);
    for (int64_t f = 0; f < c->F; ++f)
        order[f] = (int32_t)f;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return c->face_lg_of[x] > c->face_lg_of[y]; });
    std::vector<FaceTile> ft;
    c->ftile_class.clear();
    int64_t i = 0;
    while (i < c->F) {
        const int lg = c->face_lg_of[order[i]];
        c->ftile_class.push_back((int64_t)ft.size());
        int64_t j = i;
        while (j < c->F && c->face_lg_of[order[j]] == lg)
            ++j;
        const int per = kBlock >> lg;
        for (int64_t k = i; k < j; k += per)
            ft.push_back({ k, (int32_t)std::min<int64_t>(per, j - k), lg });
        i = j;
    }
    c->ftile_class.push_back((int64_t)ft.size());
    c->n_ftiles = (int64_t)ft.size();
    // device copies
    std::vector<double> ops(vol_len);
    // BT [Ns][Nc] in place of B, then G as given
    for (int g = 0; g < c->n_group; ++g) {
        const GroupDev& G = c->groups[g];
        const double* b = m->vol_ops + G.op_off;
        for (int j = 0; j < G.ns; ++j)
            for (int q = 0; q < G.nc; ++q)
                ops[G.op_off + (int64_t)j * G.nc + q] = b[(int64_t)q * G.ns + j];
        std::memcpy(ops.data() + G.op_off + (int64_t)G.ns * G.nc, b + (int64_t)G.ns * G.nc,
                    sizeof(double) * 3 * G.ns * G.nc);
        if (G.tp_off >= 0)
            std::memcpy(ops.data() + G.tp_off, m->vol_ops + G.tp_off, sizeof(double) * tp_len(G));
    }
    MIX_RUN(upload(c, &c->ops, ops));
    MIX_RUN(upload(c, &c->dof, dof));
    MIX_RUN(upload(c, &c->jdet, std::vector<double>(m->jdet, m->jdet + c->E)));
    MIX_RUN(upload(c, &c->met, std::vector<double>(m->metrics, m->metrics + 9 * c->E)));
    MIX_RUN(upload(c, &c->ef, ef));
    MIX_RUN(upload(c, &c->face_elem, std::vector<int32_t>(m->face_elem, m->face_elem + 2 * c->F)));
    MIX_RUN(upload(c, &c->face_op, std::vector<int32_t>(m->face_op, m->face_op + 2 * c->F)));
    MIX_RUN(upload(c, &c->face_wt, std::vector<int32_t>(m->face_wt, m->face_wt + c->F)));
    MIX_RUN(upload(c, &c->fnorm, std::vector<double>(m->face_normal, m->face_normal + 3 * c->F)));
    MIX_RUN(upload(c, &c->fjac, std::vector<double>(m->face_jac, m->face_jac + c->F)));
    MIX_RUN(upload(c, &c->fb_off, fb_off));
    MIX_RUN(upload(c, &c->fl_off, fl_off));
    MIX_RUN(upload(c, &c->fop_nq, fnq));
    MIX_RUN(upload(c, &c->fop_ns, fns));
    MIX_RUN(upload(c, &c->fop_off, foff));
    MIX_RUN(upload(c, &c->fopt_off, fofft));
    MIX_RUN(upload(c, &c->fopr_off, froff));
    MIX_RUN(upload(c, &c->foprt_off, frofft));
    MIX_RUN(upload(c, &c->fopr_n, fn2));
    MIX_RUN(upload(c, &c->ftc, ftc));
    MIX_RUN(upload(c, &c->fte_off, fte));
    MIX_RUN(upload(c, &c->elf, elf));
    MIX_RUN(upload(c, &c->fop, fop));
    MIX_RUN(upload(c, &c->fw_off, std::vector<int64_t>(m->fw_off, m->fw_off + m->n_face_wt)));
    MIX_RUN(upload(c, &c->fw, std::vector<double>(m->fw_data, m->fw_data + fw_len)));
    MIX_RUN(upload(c, &c->face_order, order));
    MIX_RUN(upload(c, &c->ftiles, ft));
    MIX_RUN(alloc(c, &c->d_groups, c->n_group));
    MIX_RUN(alloc(c, &c->u, c->n_dof));
    MIX_RUN(alloc(c, &c->up, c->n_dof));
    MIX_RUN(alloc(c, &c->rhs, c->n_dof));
    MIX_RUN(alloc(c, &c->fb, c->fb_len));
    MIX_RUN(alloc(c, &c->fl, std::max<int64_t>(c->fl_len, 1)));
    MIX_RUN(alloc(c, &c->tr, 2 * c->fb_len));
    MIX_RUN(alloc(c, &c->counter, 2));
    MIX_TRY(c, hipMemset(c->u, 0, c->n_dof * sizeof(double)));
    MIX_TRY(c, hipMemset(c->up, 0, c->n_dof * sizeof(double)));
    MIX_TRY(c, hipMemset(c->tr, 0, 2 * c->fb_len * sizeof(double)));
    for (int k = 0; k < 3; ++k) {
        c->tiles_of[k].resize(c->n_group);
        for (int g = 0; g < c->n_group; ++g)
            c->tiles_of[k][g] = c->groups[g].tile;
    }
    return build_vtiles(c);
}

dgx_status copy_state(dgx_mixed_ctx* c, double* dst, const double* src, int ptr_kind, hipMemcpyKind hk)
{
    MIX_CHECK(c, ptr_kind == DGX_PTR_HOST || ptr_kind == DGX_PTR_DEVICE, "ptr_kind");
    MIX_CHECK(c, dst && src, "null state pointer");
    const hipMemcpyKind k = ptr_kind == DGX_PTR_DEVICE ? hipMemcpyDeviceToDevice : hk;
    MIX_TRY(c, hipMemcpyAsync(dst, src, c->n_dof * sizeof(double), k, c->stream));
    MIX_TRY(c, hipStreamSynchronize(c->stream));
    return DGX_OK;
}

// Mean microseconds per launch of fn over reps launches on the context's stream.
template <class Fn>
dgx_status time_launches(dgx_mixed_ctx* c, int reps, Fn fn, double* us)
{
    hipEvent_t e0, e1;
    MIX_TRY(c, hipEventCreate(&e0));
    MIX_TRY(c, hipEventCreate(&e1));
    hipError_t e = fn();   // warm-up
    if (e == hipSuccess)
        e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r)
        e = fn();
    if (e == hipSuccess)
        e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess)
        e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    MIX_TRY(c, e);
    *us = 1000.0 * ms / reps;
    return DGX_OK;
}

}  // namespace

extern "C" {

dgx_status dgx_mixed_create(const dgx_mixed_mesh* mesh, int32_t num_flux, dgx_mixed_ctx** out)
{
    if (!out)
        return DGX_ERR_INVALID_ARG;
    *out = nullptr;
    if (!mesh || (num_flux != DGX_FLUX_LAX_FRIEDRICHS && num_flux != DGX_FLUX_ROE_PIKE))
        return DGX_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return DGX_ERR_NO_DEVICE;
    dgx_mixed_ctx* c = new dgx_mixed_ctx();
    c->flux = num_flux;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DGX_ERR_HIP;
    }
    c->n_cu = prop.multiProcessorCount;
    dgx_status s = setup(c, mesh);
    if (s != DGX_OK) {
        t_create_err = c->err;
        free_all(c);
        delete c;
        return s;
    }
    *out = c;
    return DGX_OK;
}

dgx_status dgx_mixed_destroy(dgx_mixed_ctx* ctx)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    free_all(ctx);
    delete ctx;
    return DGX_OK;
}

const char* dgx_mixed_last_error(const dgx_mixed_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : t_create_err.c_str();
}

int64_t dgx_mixed_n_dof(const dgx_mixed_ctx* ctx)
{
    return ctx ? ctx->n_dof : -1;
}

int dgx_mixed_persistent_waves(const dgx_mixed_ctx* ctx)
{
    return ctx ? ctx->pwaves : -1;
}

dgx_status dgx_mixed_set_state(dgx_mixed_ctx* ctx, int ptr_kind, const double* u)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_RUN(copy_state(ctx, ctx->u, u, ptr_kind, hipMemcpyHostToDevice));
    MIX_TRY(ctx, launch_traces(ctx, ctx->u));
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

dgx_status dgx_mixed_get_state(dgx_mixed_ctx* ctx, int ptr_kind, double* u)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return copy_state(ctx, u, ctx->u, ptr_kind, hipMemcpyDeviceToHost);
}

dgx_status dgx_mixed_rhs(dgx_mixed_ctx* ctx, int apply_minv, int ptr_kind, double* rhs)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, rhs, "null rhs");
    MIX_TRY(ctx, launch_eval(ctx, ctx->u, nullptr, ctx->rhs, -1, apply_minv ? 1 : 0, 0.0));
    return copy_state(ctx, rhs, ctx->rhs, ptr_kind, hipMemcpyDeviceToHost);
}

dgx_status dgx_mixed_step(dgx_mixed_ctx* ctx, double dt, int nsteps)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, nsteps >= 0 && std::isfinite(dt) && dt > 0.0, "dt or nsteps");
    if (nsteps == 0)
        return DGX_OK;
    MIX_RUN(ensure_graph(ctx, dt));
    for (int s = 0; s < nsteps; ++s)
        MIX_TRY(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

dgx_status dgx_mixed_set_launch(dgx_mixed_ctx* ctx, int mode, const int32_t* tile_elems, int32_t blocks_per_cu)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, mode == DGX_MIXED_LAUNCH_PERSISTENT || mode == DGX_MIXED_LAUNCH_PER_GROUP ||
                       mode == DGX_MIXED_LAUNCH_GROUPED, "launch mode");
    MIX_CHECK(ctx, blocks_per_cu >= 1 && blocks_per_cu <= 8, "blocks_per_cu");
    if (tile_elems)
        for (int g = 0; g < ctx->n_group; ++g)
            MIX_CHECK(ctx, tile_ok(ctx->groups[g], tile_elems[g]), "tile_elems");
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->mode = mode;
    ctx->bpc = blocks_per_cu;
    // each launch form keeps its own tile shapes (the autotuner's, or the last ones given for it)
    if (tile_elems)
        ctx->tiles_of[mode].assign(tile_elems, tile_elems + ctx->n_group);
    for (int g = 0; g < ctx->n_group; ++g)
        ctx->groups[g].tile = ctx->tiles_of[mode][g];
    MIX_RUN(build_vtiles(ctx));
    return DGX_OK;
}

dgx_status dgx_mixed_get_launch(const dgx_mixed_ctx* ctx, int* mode, int32_t* tile_elems, int32_t* blocks_per_cu)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (mode)
        *mode = ctx->mode;
    if (blocks_per_cu)
        *blocks_per_cu = ctx->bpc;
    if (tile_elems)
        for (int g = 0; g < ctx->n_group; ++g)
            tile_elems[g] = ctx->groups[g].tile;
    return DGX_OK;
}

dgx_status dgx_mixed_set_face_maps(dgx_mixed_ctx* ctx, int tensor)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, tensor == 0 || tensor == 1, "face map form");
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->tensor_faces = tensor;
    return DGX_OK;
}

int32_t dgx_mixed_tensor_face_ops(const dgx_mixed_ctx* ctx)
{
    return ctx ? (int32_t)ctx->n_tensor_ops : -1;
}

dgx_status dgx_mixed_autotune(dgx_mixed_ctx* ctx, int reps, double* timings)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, reps >= 1 && reps <= 1000, "reps");
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    // the residual form of the kernels writes rhs only: the state and traces are untouched; the face
    // buffer is filled once
    MIX_TRY(ctx, launch_eval(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0));
    MixArgs a;
    // 1. per-(type, p) launches: every group alone at every legal tile size
    std::vector<int32_t> pg(ctx->n_group);
    for (int g = 0; g < ctx->n_group; ++g) {
        GroupDev& G = ctx->groups[g];
        int best = G.tile;
        double best_us = 1e300;
        for (int t = 1; t <= 64; t *= 2) {
            if (!tile_ok(G, t))
                continue;
            G.tile = t;
            MIX_RUN(build_vtiles(ctx));
            a = make_args(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0);
            double us = 0.0;
            MIX_RUN(time_launches(ctx, reps, [&]() { return launch_vol(ctx, a, DGX_MIXED_LAUNCH_PER_GROUP, g); }, &us));
            if (us < best_us) {
                best_us = us;
                best = t;
            }
        }
        G.tile = best;
        pg[g] = best;
        if (timings)
            timings[g] = best_us;
    }
    ctx->tiles_of[DGX_MIXED_LAUNCH_PER_GROUP] = pg;
    // 2. the persistent grid: one LDS size for every tile, so the tiles are capped to an LDS budget
    //    (the per-group best tile, halved until it fits) and the budget and grid size are searched
    const size_t caps[] = { 24 << 10, 32 << 10, 40 << 10, 48 << 10, (size_t)kMaxTileLds };
    std::vector<int32_t> best_tiles = pg, best_grouped = pg;
    int best_bpc = ctx->bpc, best_w = ctx->pwaves;
    double best_us = 1e300, best_grouped_us = 1e300;
    for (size_t cap : caps) {
        std::vector<int32_t> tl(ctx->n_group);
        bool fits = true;
        for (int g = 0; g < ctx->n_group && fits; ++g) {
            int t = pg[g];
            while (t > 1 && tile_lds(ctx->groups[g], t) > cap)
                t /= 2;
            fits = tile_lds(ctx->groups[g], t) <= cap;
            tl[g] = t;
        }
        if (!fits)
            continue;
        for (int g = 0; g < ctx->n_group; ++g)
            ctx->groups[g].tile = tl[g];
        MIX_RUN(build_vtiles(ctx));
        a = make_args(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0);
        {
            double us = 0.0;
            MIX_RUN(time_launches(ctx, reps, [&]() { return launch_vol(ctx, a, DGX_MIXED_LAUNCH_GROUPED); }, &us));
            if (us < best_grouped_us) {
                best_grouped_us = us;
                best_grouped = tl;
            }
        }
        const int max_bpc = std::max(1, std::min(8, (int)((160u << 10) / ctx->lds_max)));
        for (int w = 2; w <= 4; ++w) {
            ctx->pwaves = w;
            for (int b = 1; b <= max_bpc; ++b) {
                ctx->bpc = b;
                double us = 0.0;
                MIX_RUN(time_launches(ctx, reps, [&]() {
                    reset_counters<<<1, 64, 0, ctx->stream>>>(ctx->counter);
                    return launch_vol(ctx, a, DGX_MIXED_LAUNCH_PERSISTENT);
                }, &us));
                if (us < best_us) {
                    best_us = us;
                    best_bpc = b;
                    best_w = w;
                    best_tiles = tl;
                }
            }
        }
    }
    ctx->tiles_of[DGX_MIXED_LAUNCH_PERSISTENT] = best_tiles;
    ctx->tiles_of[DGX_MIXED_LAUNCH_GROUPED] = best_grouped;
    ctx->bpc = best_bpc;
    ctx->pwaves = best_w;
    for (int g = 0; g < ctx->n_group; ++g)
        ctx->groups[g].tile = ctx->tiles_of[ctx->mode][g];
    MIX_RUN(build_vtiles(ctx));
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

dgx_status dgx_mixed_time_stage(dgx_mixed_ctx* ctx, double dt, int nstages, double* us)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, us && nstages >= 3 && std::isfinite(dt) && dt > 0.0, "time_stage arguments");
    // the timing steps advance the state: keep a copy and restore it
    MIX_TRY(ctx, hipMemcpyAsync(ctx->rhs, ctx->u, ctx->n_dof * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    MIX_RUN(ensure_graph(ctx, dt));
    const int steps = nstages / 3;
    double t = 0.0;
    MIX_RUN(time_launches(ctx, steps, [&]() { return hipGraphLaunch(ctx->graph, ctx->stream); }, &t));
    us[0] = t / 3.0;
    MixArgs a = make_args(ctx, ctx->up, ctx->u, ctx->up, 1, 1, dt);
    MIX_RUN(time_launches(ctx, nstages, [&]() {
        reset_counters<<<1, 64, 0, ctx->stream>>>(ctx->counter);
        return launch_face(ctx, a, ctx->mode);
    }, &us[1]));
    MIX_RUN(time_launches(ctx, nstages, [&]() {
        reset_counters<<<1, 64, 0, ctx->stream>>>(ctx->counter);
        return launch_vol(ctx, a, ctx->mode);
    }, &us[2]));
    MIX_TRY(ctx, hipMemcpyAsync(ctx->u, ctx->rhs, ctx->n_dof * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    MIX_TRY(ctx, launch_traces(ctx, ctx->u));
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

}  // extern "C"
