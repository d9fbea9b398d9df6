// Batching form: explicit DG residual and SSP-RK3 stage on a mixed HEX/WEDGE/TET/PYR mesh with
// per-element orders, elements grouped by (type, p).
//
// A stage is two kernels:
//   face    one lane per face cubature node: both sides' traces (basis rows of the side's volume at the
//           node, the right side's node correspondence folded into its table), the numerical flux,
//           times the cubature weight and face Jacobian -> the face flux buffer
//           (compute_face_rlhs_T.c:166-170, numerical_flux_euler_T.c)
//   volume  one tile of elements of one group per workgroup: coefficients -> LDS, state at the volume
//           cubature nodes and the metric-contracted Euler flux -> LDS, weak divergence (one lane per
//           mode, compute_volume_rlhs_T.c:145-177), lift of the element's faces from the face flux
//           buffer (compute_face_rlhs_T.c:297-359; the right side sees -nnf), M^-1 = 1/det J
//           (orthonormal basis on an affine element, solve_dg.c:265-272) and the SSP-RK3 combination
//           (solve_explicit.c:252-301)
// Launch forms: a persistent grid whose workgroups take tiles from a work counter (tiles of every
// group in one launch, largest first), or one launch per (type, p) group / per face size class.  The
// volume kernel updates its elements in place (an element reads only its own coefficients; neighbour
// data come through the face flux buffer written by the face kernel before it).
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
};

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
    int stage;             // 0..2, or -1: residual
    int apply_minv;
    double dt;
    const GroupDev* groups;
    const double* ops;
    const int64_t* dof;    // [E+1]
    const double* jdet;
    const double* met;     // [E][9]
    const int32_t* ef;     // [E][6]: 2 f + side, -1 = none
    const int32_t* face_elem;
    const int32_t* face_op;
    const int32_t* face_wt;
    const double* face_normal;
    const double* face_jac;
    const int64_t* fb_off;
    double* fb;            // face flux buffer [sum Nq][5]
    const int32_t* fop_nq;
    const int32_t* fop_ns;
    const int64_t* fop_off;    // Bf [Nq][Ns]
    const int64_t* fopt_off;   // BfT [Ns][Nq]
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

template <int FLUX>
__device__ inline void face_tile(const MixArgs& a, const FaceTile& t)
{
    const int lg = t.lg_tpf;
    const int slot = threadIdx.x >> lg, l = threadIdx.x & ((1 << lg) - 1);
    if (slot >= t.count)
        return;
    const int f = a.face_order[t.i0 + slot];
    const int oL = a.face_op[2 * f], oR = a.face_op[2 * f + 1];
    const int nq = a.fop_nq[oL];
    if (l >= nq)
        return;
    const int eL = a.face_elem[2 * f], eR = a.face_elem[2 * f + 1];
    double wl[5], wr[5];
    {
        const int ns = a.fop_ns[oL];
        const double* bt = a.fop + a.fopt_off[oL];
        const double* ue = a.u + a.dof[eL];
#pragma unroll
        for (int v = 0; v < 5; ++v)
            wl[v] = 0.0;
        for (int j = 0; j < ns; ++j) {
            const double b = bt[j * nq + l];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                wl[v] = fma(b, ue[v * ns + j], wl[v]);
        }
    }
    {
        const int ns = a.fop_ns[oR];
        const double* bt = a.fop + a.fopt_off[oR];
        const double* ue = a.u + a.dof[eR];
#pragma unroll
        for (int v = 0; v < 5; ++v)
            wr[v] = 0.0;
        for (int j = 0; j < ns; ++j) {
            const double b = bt[j * nq + l];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                wr[v] = fma(b, ue[v * ns + j], wr[v]);
        }
    }
    const double n[3] = { a.face_normal[3 * f], a.face_normal[3 * f + 1], a.face_normal[3 * f + 2] };
    double nf[5];
    dgx::numflux<3, FLUX>(wl, wr, n, nf);
    const double s = a.fw[a.fw_off[a.face_wt[f]] + l] * a.face_jac[f];
    double* o = a.fb + a.fb_off[f] + 5 * l;
#pragma unroll
    for (int v = 0; v < 5; ++v)
        o[v] = nf[v] * s;
}

__device__ inline double rk_mix(int stage, double dt, double s, double r, const double* reg)
{
    if (stage == 0)
        return s + dt * r;
    if (stage == 1)
        return (1.0 / 4.0) * (3.0 * (*reg) + s + dt * r);
    return (1.0 / 3.0) * ((*reg) + 2.0 * s + 2.0 * dt * r);
}

// One volume tile: `count` elements of group g starting at e0; every lane of the block reaches both
// barriers.
__device__ inline void vol_tile(const MixArgs& a, const VolTile& t, double* lds)
{
    const GroupDev G = a.groups[t.group];
    const int ns = G.ns, nc = G.nc;
    const int tpe = kBlock / G.tile;
    const int slot = threadIdx.x / tpe, l = threadIdx.x - slot * tpe;
    const bool act = slot < t.count;
    const int64_t e = t.e0 + slot;
    double* U = lds + slot * (5 * ns + 15 * nc);
    double* Fr = U + 5 * ns;
    if (act) {
        const double* ue = a.u + a.dof[e];
        for (int i = l; i < 5 * ns; i += tpe)
            U[i] = ue[i];
    }
    __syncthreads();
    if (act) {
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
    if (act) {
        const double* gm = a.ops + G.op_off + (int64_t)ns * nc;
        const double* ue = a.u + a.dof[e];
        double* oe = a.out + a.dof[e];
        const double* re = a.reg ? a.reg + a.dof[e] : nullptr;
        const double ij = a.jdet[e];
        for (int j = l; j < ns; j += tpe) {
            double r[5];
#pragma unroll
            for (int v = 0; v < 5; ++v)
                r[v] = 0.0;
            for (int q = 0; q < nc; ++q) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double g = gm[(d * nc + q) * ns + j];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        r[v] = fma(g, Fr[(q * 3 + d) * 5 + v], r[v]);
                }
            }
            for (int k = 0; k < 6; ++k) {
                const int ent = a.ef[6 * e + k];
                if (ent < 0)
                    break;
                const int f = ent >> 1, side = ent & 1;
                const int o = a.face_op[2 * f + side];
                const int nq = a.fop_nq[o];
                const double* bf = a.fop + a.fop_off[o];
                const double* nf = a.fb + a.fb_off[f];
                double s[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
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
            if (a.stage < 0) {
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    oe[v * ns + j] = a.apply_minv ? r[v] / ij : r[v];
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const int i = v * ns + j;
                    oe[i] = rk_mix(a.stage, a.dt, ue[i], r[v] / ij, re ? re + i : nullptr);
                }
            }
        }
    }
    __syncthreads();
}

__global__ void reset_counters(unsigned int* c)
{
    if (threadIdx.x < 2)
        c[threadIdx.x] = 0u;
}

template <int FLUX>
__global__ __launch_bounds__(kBlock) void face_persistent(MixArgs a)
{
    __shared__ int s_t;
    for (;;) {
        if (threadIdx.x == 0)
            s_t = (int)atomicAdd(a.counter, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= a.n_ftiles)
            break;
        face_tile<FLUX>(a, a.ftiles[t]);
    }
}

template <int FLUX>
__global__ __launch_bounds__(kBlock) void face_range(MixArgs a, int64_t t0)
{
    face_tile<FLUX>(a, a.ftiles[t0 + blockIdx.x]);
}

__global__ __launch_bounds__(kBlock) void vol_persistent(MixArgs a)
{
    extern __shared__ double lds[];
    __shared__ int s_t;
    for (;;) {
        if (threadIdx.x == 0)
            s_t = (int)atomicAdd(a.counter + 1, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= a.n_vtiles)
            break;
        vol_tile(a, a.vtiles[a.vorder[t]], lds);
    }
}

__global__ __launch_bounds__(kBlock) void vol_range(MixArgs a, int64_t t0)
{
    extern __shared__ double lds[];
    vol_tile(a, a.vtiles[t0 + blockIdx.x], lds);
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
    int n_cu = 0;
    std::vector<GroupDev> groups;
    std::vector<int32_t> face_lg_of;     // per face
    // device
    double *u = nullptr, *up = nullptr, *rhs = nullptr, *fb = nullptr;
    double *ops = nullptr, *jdet = nullptr, *met = nullptr, *fnorm = nullptr, *fjac = nullptr, *fop = nullptr,
           *fw = nullptr;
    int64_t *dof = nullptr, *fb_off = nullptr, *fop_off = nullptr, *fopt_off = nullptr, *fw_off = nullptr;
    int32_t *ef = nullptr, *face_elem = nullptr, *face_op = nullptr, *face_wt = nullptr, *fop_nq = nullptr,
            *fop_ns = nullptr, *face_order = nullptr, *vorder = nullptr;
    GroupDev* d_groups = nullptr;
    VolTile* vtiles = nullptr;
    FaceTile* ftiles = nullptr;
    unsigned int* counter = nullptr;
    int64_t n_vtiles = 0, n_ftiles = 0;
    std::vector<int64_t> vtile_start;    // [n_group+1] per-group ranges of vtiles
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
    return (size_t)tile * (5 * g.ns + 15 * g.nc) * sizeof(double);
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
    if (c->vtiles)
        (void)hipFree(c->vtiles);
    if (c->vorder)
        (void)hipFree(c->vorder);
    c->vtiles = nullptr;
    c->vorder = nullptr;
    hipError_t e = hipSuccess;
    c->vtiles = dev_copy(vt, &e);
    MIX_TRY(c, e);
    c->vorder = dev_copy(order, &e);
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
    a.fop_nq = c->fop_nq;
    a.fop_ns = c->fop_ns;
    a.fop_off = c->fop_off;
    a.fopt_off = c->fopt_off;
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
        vol_persistent<<<grid, kBlock, c->lds_max, c->stream>>>(a);
        return hipGetLastError();
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
    c->vtiles = nullptr;
    c->vorder = nullptr;
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
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
    std::vector<int32_t> ns_of(c->E), type_of(c->E);
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
        MIX_CHECK(c, G.type >= 0 && G.type <= 3 && G.p >= 1 && G.p <= 5, "group type/order");
        MIX_CHECK(c, G.ns > 0 && G.ns <= 512 && G.nc > 0 && G.nc <= 512, "group Ns/Nc");
        MIX_CHECK(c, G.count > 0, "empty group");
        MIX_CHECK(c, G.op_off >= 0, "vol_off");
        vol_len = std::max<int64_t>(vol_len, G.op_off + 4LL * G.ns * G.nc);
        G.tile = default_tile(G);
        MIX_CHECK(c, tile_ok(G, G.tile), "group too large for a volume tile");
        for (int64_t e = G.start; e < G.start + G.count; ++e) {
            ns_of[e] = G.ns;
            type_of[e] = G.type;
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
    // transposed copies after the given data
    std::vector<double> fop(m->fop_data, m->fop_data + fop_len);
    for (int o = 0; o < m->n_face_op; ++o) {
        fofft[o] = (int64_t)fop.size();
        for (int j = 0; j < fns[o]; ++j)
            for (int q = 0; q < fnq[o]; ++q)
                fop.push_back(m->fop_data[foff[o] + (int64_t)q * fns[o] + j]);
    }
    int64_t fw_len = 0;
    for (int w = 0; w < m->n_face_wt; ++w) {
        MIX_CHECK(c, m->fw_n[w] > 0 && m->fw_n[w] <= 64 && m->fw_off[w] >= 0, "face weight table");
        fw_len = std::max<int64_t>(fw_len, m->fw_off[w] + m->fw_n[w]);
    }
    // faces: sides, operators, element face lists, flux buffer offsets
    std::vector<int32_t> ef(6 * c->E, -1), nfe(c->E, 0);
    std::vector<int64_t> fb_off(c->F);
    int64_t fb = 0;
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
            MIX_CHECK(c, nfe[e] < kFacesOf[type_of[e]], "element with more faces than its type has");
            ef[6 * e + nfe[e]++] = (int32_t)(2 * f + s);
        }
        fb_off[f] = fb;
        fb += 5LL * fnq[oL];
        c->face_lg_of[f] = face_lg(fnq[oL]);
    }
    for (int64_t e = 0; e < c->E; ++e)
        MIX_CHECK(c, nfe[e] == kFacesOf[type_of[e]], "element with a face missing (the mesh must be closed)");
    c->fb_len = fb;
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
    MIX_RUN(upload(c, &c->fop_nq, fnq));
    MIX_RUN(upload(c, &c->fop_ns, fns));
    MIX_RUN(upload(c, &c->fop_off, foff));
    MIX_RUN(upload(c, &c->fopt_off, fofft));
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
    MIX_RUN(alloc(c, &c->counter, 2));
    MIX_TRY(c, hipMemset(c->u, 0, c->n_dof * sizeof(double)));
    MIX_TRY(c, hipMemset(c->up, 0, c->n_dof * sizeof(double)));
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

dgx_status dgx_mixed_set_state(dgx_mixed_ctx* ctx, int ptr_kind, const double* u)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return copy_state(ctx, ctx->u, u, ptr_kind, hipMemcpyHostToDevice);
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
    MIX_CHECK(ctx, mode == DGX_MIXED_LAUNCH_PERSISTENT || mode == DGX_MIXED_LAUNCH_PER_GROUP, "launch mode");
    MIX_CHECK(ctx, blocks_per_cu >= 1 && blocks_per_cu <= 8, "blocks_per_cu");
    if (tile_elems)
        for (int g = 0; g < ctx->n_group; ++g)
            MIX_CHECK(ctx, tile_ok(ctx->groups[g], tile_elems[g]), "tile_elems");
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->mode = mode;
    ctx->bpc = blocks_per_cu;
    if (tile_elems) {
        for (int g = 0; g < ctx->n_group; ++g)
            ctx->groups[g].tile = tile_elems[g];
        MIX_RUN(build_vtiles(ctx));
    }
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

dgx_status dgx_mixed_autotune(dgx_mixed_ctx* ctx, int reps, double* timings)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    MIX_CHECK(ctx, reps >= 1 && reps <= 1000, "reps");
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    // the residual kernels write rhs only: the state is untouched; fill the face buffer once
    MIX_TRY(ctx, launch_eval(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0));
    MixArgs a = make_args(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0);
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
        if (timings)
            timings[g] = best_us;
    }
    MIX_RUN(build_vtiles(ctx));
    a = make_args(ctx, ctx->u, nullptr, ctx->rhs, -1, 1, 0.0);
    int best_bpc = ctx->bpc;
    double best_us = 1e300;
    for (int b = 1; b <= 4; ++b) {
        ctx->bpc = b;
        double us = 0.0;
        MIX_RUN(time_launches(ctx, reps, [&]() {
            reset_counters<<<1, 64, 0, ctx->stream>>>(ctx->counter);
            return launch_vol(ctx, a, DGX_MIXED_LAUNCH_PERSISTENT);
        }, &us));
        if (us < best_us) {
            best_us = us;
            best_bpc = b;
        }
    }
    ctx->bpc = best_bpc;
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
    MIX_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

}  // extern "C"
