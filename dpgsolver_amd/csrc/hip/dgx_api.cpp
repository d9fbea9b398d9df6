// C-ABI of include/dgx.h: context, uploads, and the stage/residual entry points.
// Host code only; every device computation is a kernel in dgx_stage.hip / dgx_hex3.inc.
//
// Device arrays are in "slot" order: elements are grouped (2x2x2 bricks when the mesh allows it, for
// the HEX p=3 group kernel; the identity otherwise) and slot = group*8 + position.  The caller only
// sees element order; uploads and downloads permute.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dgx.h"
#include "dgx_setup.h"
#include "dgx_internal.hpp"

struct dgx_ctx {
    int d = 0, p = 0, nv = 0, ns = 0, nf = 0, nface = 0;
    int flux = 0, positivity = 1, precision = 0;
    int scheme = 0, n_stages = 3;   // time stepper (dgx::RK_*) and its stages per step
    bool visc = false;           // Navier-Stokes (BR2) instead of Euler
    double mu = 0.0, Pr = 1.0, eta = 0.0;
    double* gtr[2] = { nullptr, nullptr };  // NS: [Ep][nface][d*nv][nf] face traces of the volume gradient;
                                            // HEX p=4: [Ep][nface][nv][nf] viscous normal-flux traces,
                                            // two buffers indexed like tr
    double* lam = nullptr;       // NS: [Ep][nface][nf] lift trace factors
    double* fnv = nullptr;       // NS HEX p=4: [Ep][nface][3][nf] |J_f| n, own outward normal
    bool hex3 = false;           // HEX p=3 group kernel
    bool ns4 = false;            // Navier-Stokes HEX p=4 kernel
    int64_t E = 0, F = 0;
    int64_t Ep = 0;              // number of slots (E plus padding)
    int64_t nmax = 0;            // entries of blockmax
    hipStream_t stream = nullptr;
    double* u = nullptr;         // [nv][Ep][ns]
    double* up = nullptr;        // [nv][Ep][ns]
    double* rhs = nullptr;       // [nv][Ep][ns] residual / staging buffer
    double* stage_buf = nullptr; // [E*nv*ns] caller-layout staging for host transfers
    double* met = nullptr;       // [d*d][Ep][ns]
    double* imass = nullptr;     // [Ep][ns]
    double* tr[2] = { nullptr, nullptr };   // [Ep][nface][nv][nf]
    uint32_t* nbr = nullptr;     // [Ep][nface]
    uint8_t* nord = nullptr;     // [Ep][nface] face node correspondence case (null when every face is aligned)
    int32_t* slot_elem = nullptr;   // [Ep]
    int32_t* elem_slot = nullptr;   // [E]
    uint64_t* face_item = nullptr;  // [G][48] HEX p=3 group face items (dgx::Hex3Args::face_item)
    double* lin_x = nullptr;     // [nv][Ep][ns] tangent / cotangent of dgx_jvp / dgx_vjp (allocated on first use)
    double* lin_tr = nullptr;    // [Ep][nface][nv][nf] its traces
    double* blockmax = nullptr;  // [nmax]
    double* dmax = nullptr;      // [1]
    unsigned long long* bad_elem = nullptr;
    unsigned int* nonfinite = nullptr;
    unsigned int* gate = nullptr;   // [2] stage gate words, indexed by trace parity (dgx::stage_gated)
    int cur = 0;                 // index of the trace buffer holding the current state's traces
    int state_buf = 0;           // HEX p=3: the current state is in u (0) or in up (1) (RK buffer rotation)
    int rk_next = 0;             // next SSP-RK3 stage dgx_rk_stage accepts
    hipGraphExec_t graph[2] = { nullptr, nullptr };
    double graph_dt = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    std::string err;
};

namespace {

dgx_status fail(dgx_ctx* c, dgx_status s, const std::string& msg)
{
    if (c)
        c->err = msg;
    return s;
}

#define HIP_TRY(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail((ctx), DGX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

void free_all(dgx_ctx* c)
{
    for (int i = 0; i < 2; ++i) {
        if (c->graph[i])
            (void)hipGraphExecDestroy(c->graph[i]);
        c->graph[i] = nullptr;
    }
    void* ptrs[] = { c->u, c->up, c->rhs, c->stage_buf, c->met, c->imass, c->tr[0], c->tr[1], c->nbr,
                     c->slot_elem, c->elem_slot, c->face_item, c->blockmax, c->dmax,
                     c->bad_elem, c->nonfinite, c->gate, c->lin_x, c->lin_tr, c->gtr[0], c->gtr[1], c->lam,
                     c->fnv, c->nord };
    for (void* p : ptrs)
        if (p)
            (void)hipFree(p);
    if (c->ev0)
        (void)hipEventDestroy(c->ev0);
    if (c->ev1)
        (void)hipEventDestroy(c->ev1);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
}

dgx_status reset_flags(dgx_ctx* c)
{
    HIP_TRY(c, dgx::launch_fill_u64(c->bad_elem, 1, ~0ull, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->nonfinite, 0, sizeof(unsigned int), c->stream));
    HIP_TRY(c, hipMemsetAsync(c->gate, 0, 2 * sizeof(unsigned int), c->stream));
    return DGX_OK;
}

// Low-storage RK 5-stage 4th order: Carpenter & Kennedy 1994, solution 3 of table 1, rational form
// (the coefficient table of time_step_ls_rk_54, solve_explicit.c:310-321).
const double kLs54A[5] = { 0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                           -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0 };
const double kLs54B[5] = { 1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                           17201463215490.0 / 20902069494980.0, 3134564353537.0 / 4481467310338.0,
                           2277821191437.0 / 14882151754819.0 };

dgx_status check_flags(dgx_ctx* c, int64_t* elem)
{
    unsigned long long bad = 0;
    unsigned int nf = 0;
    HIP_TRY(c, hipMemcpyAsync(&bad, c->bad_elem, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&nf, c->nonfinite, sizeof(nf), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (elem)
        *elem = (bad == ~0ull) ? -1 : (int64_t)bad;
    if (bad != ~0ull)
        return fail(c, DGX_ERR_POSITIVITY,
                    "average density or pressure below EPS_PHYS in element " + std::to_string(bad));
    if (nf)
        return fail(c, DGX_ERR_NONFINITE, "non-finite residual");
    return DGX_OK;
}

// Time-stepper fields of the kernel arguments: scheme, LS-RK54 coefficients, the stage gate words
// (read the one of the trace parity tin, write the other).
template <class Args>
void set_rk(const dgx_ctx* c, int stage, int tin, Args& a)
{
    a.scheme = c->scheme;
    if (c->scheme == dgx::RK_LS54 && stage >= 0) {
        a.rk_a = kLs54A[stage];
        a.rk_b = kLs54B[stage];
    }
    a.gate_in = c->gate + tin;
    a.gate_out = c->gate + (tin ^ 1);
}

// Launches one stage (stage 0..n_stages-1) or the residual (stage -1) reading traces tr[tin].
hipError_t launch_any(dgx_ctx* c, int stage, double dt, int apply_minv, int tin, int precision)
{
    if (c->hex3) {
        dgx::Hex3Args a;
        std::memset(&a, 0, sizeof(a));
        // SSP-RK3 buffer rotation: stage 0 u -> up, stage 1 up (+u^n in u) -> up, stage 2 up (+u^n) -> u;
        // the other steppers update u in place with up as their RK register
        if (stage < 0) {
            a.u = c->state_buf ? c->up : c->u;
        } else if (c->scheme != dgx::RK_SSP33) {
            a.u = c->u;
            a.u_out = c->u;
            a.reg = c->up;
        } else if (stage == 0) {
            a.u = c->u;
            a.u_out = c->up;
        } else {
            a.u = c->up;
            a.up = c->u;
            a.u_out = (stage == 1) ? c->up : c->u;
        }

        a.met = c->met;
        a.imass = c->imass;
        a.tr_in = c->tr[tin];
        a.tr_out = c->tr[tin ^ 1];
        a.nbr = c->nbr;
        a.slot_elem = c->slot_elem;
        a.face_item = c->face_item;
        a.rhs_out = c->rhs;
        a.blockmax = c->blockmax;
        a.bad_elem = c->bad_elem;
        a.n_slots = c->Ep;
        a.dt = dt;
        a.stage = stage;
        a.apply_minv = apply_minv;
        a.positivity = c->positivity;
        set_rk(c, stage, tin, a);
        return dgx::launch_hex3_stage(c->flux, precision == DGX_PREC_MIXED, a, c->stream);
    }
    dgx::StageArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = c->u;
    a.up = c->up;
    a.met = c->met;
    a.imass = c->imass;
    a.tr_in = c->tr[tin];
    a.tr_out = c->tr[tin ^ 1];
    a.nbr = c->nbr;
    a.nord = c->nord;
    a.rhs_out = c->rhs;
    a.blockmax = c->blockmax;
    a.bad_elem = c->bad_elem;
    a.n_elem = c->E;
    a.dt = dt;
    a.stage = stage;
    a.apply_minv = apply_minv;
    a.positivity = c->positivity;
    set_rk(c, stage, tin, a);
    if (c->visc) {
        a.gtr = c->fnv ? c->gtr[tin] : c->gtr[0];
        a.gtr_out = c->fnv ? c->gtr[tin ^ 1] : nullptr;
        a.lam = c->lam;
        a.fnv = c->fnv;
        a.mu = c->mu;
        a.Pr = c->Pr;
        a.eta = c->eta;
        a.mixed = precision == DGX_PREC_MIXED ? 1 : 0;
        return dgx::launch_ns_stage(c->d, c->p, c->flux, a, c->gtr[0], c->stream);
    }
    return dgx::launch_stage(c->d, c->p, c->flux, a, c->stream);
}

dgx_status read_max(dgx_ctx* c, double* max_rhs)
{
    HIP_TRY(c, dgx::launch_reduce_max(c->blockmax, c->nmax, c->dmax, c->nonfinite, c->stream));
    if (max_rhs)
        HIP_TRY(c, hipMemcpyAsync(max_rhs, c->dmax, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return DGX_OK;
}

// caller buffer (element order, `layout`) -> internal slot-order buffer dst
dgx_status upload_state(dgx_ctx* c, int layout, int ptr_kind, const double* src, double* dst)
{
    const size_t bytes = sizeof(double) * (size_t)c->E * c->nv * c->ns;
    const double* s = src;
    if (ptr_kind == DGX_PTR_HOST) {
        HIP_TRY(c, hipMemcpyAsync(c->stage_buf, src, bytes, hipMemcpyHostToDevice, c->stream));
        s = c->stage_buf;
    }
    HIP_TRY(c, dgx::launch_to_internal(s, layout, dst, c->slot_elem, c->E, c->Ep, c->nv, c->ns, c->stream));
    return DGX_OK;
}

// internal slot-order buffer src -> caller buffer (element order, `layout`)
dgx_status download_state(dgx_ctx* c, int layout, int ptr_kind, const double* src, double* dst)
{
    const size_t bytes = sizeof(double) * (size_t)c->E * c->nv * c->ns;
    double* d = (ptr_kind == DGX_PTR_HOST) ? c->stage_buf : dst;
    HIP_TRY(c, dgx::launch_from_internal(src, d, layout, c->elem_slot, c->E, c->Ep, c->nv, c->ns, c->stream));
    if (ptr_kind == DGX_PTR_HOST)
        HIP_TRY(c, hipMemcpyAsync(dst, c->stage_buf, bytes, hipMemcpyDeviceToHost, c->stream));
    return DGX_OK;
}

// Entry-by-entry agreement (1e-13 relative to max(1, |entry|)) of a caller's operator set with the built one.
bool operators_match(const dgx_operators& in, const dgx_ops1d& ref)
{
    if (in.node_type != DGX_NODES_GL || in.kron_order != DGX_KRON_R_FASTEST || in.line.p != ref.p ||
        in.line.n != ref.n)
        return false;
    const int n = ref.n;
    auto near = [](const double* a, const double* b, int k) {
        for (int i = 0; i < k; ++i)
            if (!(std::fabs(a[i] - b[i]) <= 1e-13 * std::fmax(1.0, std::fabs(b[i]))))
                return false;
        return true;
    };
    return near(in.line.x, ref.x, n) && near(in.line.w, ref.w, n) && near(in.line.D, ref.D, n * n) &&
           near(in.line.lm, ref.lm, n) && near(in.line.lp, ref.lp, n) && near(in.line.sb, ref.sb, n * n) &&
           near(in.line.bs, ref.bs, n * n);
}

// Groups of up to 8 elements: 2x2x2 bricks found through the +r/+s/+t neighbours, the remaining
// elements packed in index order.  Returns slot_elem (size multiple of 8, -1 = padding).
std::vector<int32_t> build_groups(int64_t E, const std::vector<int64_t>& nb /* [E][6] */)
{
    std::vector<char> used((size_t)E, 0);
    std::vector<int32_t> slots;
    slots.reserve((size_t)E + 8);
    int64_t scan = 0;
    for (int64_t e = 0; e < E; ++e) {
        if (used[e])
            continue;
        int64_t b[8];
        b[0] = e;
        b[1] = nb[e * 6 + 1];
        b[2] = nb[e * 6 + 3];
        b[3] = nb[b[1] * 6 + 3];
        bool ok = nb[b[2] * 6 + 1] == b[3];
        for (int q = 0; q < 4 && ok; ++q)
            b[4 + q] = nb[b[q] * 6 + 5];
        ok = ok && nb[b[4] * 6 + 1] == b[5] && nb[b[4] * 6 + 3] == b[6] && nb[b[5] * 6 + 3] == b[7] &&
             nb[b[6] * 6 + 1] == b[7];
        for (int q = 0; q < 8 && ok; ++q) {
            if (used[b[q]])
                ok = false;
            for (int r = 0; r < q && ok; ++r)
                if (b[r] == b[q])
                    ok = false;
        }
        if (ok) {
            for (int q = 0; q < 8; ++q) {
                used[b[q]] = 1;
                slots.push_back((int32_t)b[q]);
            }
            continue;
        }
        // packed group of the next unused elements
        int cnt = 0;
        if (scan < e)
            scan = e;
        for (; scan < E && cnt < 8; ++scan)
            if (!used[scan]) {
                used[scan] = 1;
                slots.push_back((int32_t)scan);
                ++cnt;
            }
        while (cnt++ < 8)
            slots.push_back(-1);
    }
    return slots;
}

}  // namespace

extern "C" {

const char* dgx_status_string(dgx_status s)
{
    switch (s) {
    case DGX_OK: return "DGX_OK";
    case DGX_ERR_INVALID_ARG: return "DGX_ERR_INVALID_ARG";
    case DGX_ERR_HIP: return "DGX_ERR_HIP";
    case DGX_ERR_POSITIVITY: return "DGX_ERR_POSITIVITY";
    case DGX_ERR_NONFINITE: return "DGX_ERR_NONFINITE";
    case DGX_ERR_GEOMETRY: return "DGX_ERR_GEOMETRY";
    case DGX_ERR_NO_DEVICE: return "DGX_ERR_NO_DEVICE";
    }
    return "unknown";
}

const char* dgx_last_error(const dgx_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

dgx_status dgx_create(const dgx_config* cfg, const dgx_mesh* mesh, const dgx_geometry* geom,
                      const dgx_operators* ops_in, dgx_ctx** out)
{
    if (!out)
        return DGX_ERR_INVALID_ARG;
    *out = nullptr;
    if (!cfg || !mesh || !geom)
        return DGX_ERR_INVALID_ARG;
    const int d = cfg->d, p = cfg->p;
    const bool visc = cfg->pde == DGX_PDE_NAVIER_STOKES;
    if (visc && !(cfg->mu >= 0.0 && cfg->Pr > 0.0 && std::isfinite(cfg->mu) && std::isfinite(cfg->Pr)))
        return DGX_ERR_INVALID_ARG;
    if (!dgx::supported_order(d, p) || mesh->d != d || (cfg->pde != DGX_PDE_EULER && !visc) ||
        (cfg->num_flux != DGX_FLUX_LAX_FRIEDRICHS && cfg->num_flux != DGX_FLUX_ROE_PIKE) ||
        (cfg->precision != DGX_PREC_FP64 && cfg->precision != DGX_PREC_MIXED) ||
        (cfg->solver_type_e != DGX_SOLVER_E_SSP_RK_33 && cfg->solver_type_e != DGX_SOLVER_E_EULER &&
         cfg->solver_type_e != DGX_SOLVER_E_LS_RK_54))
        return DGX_ERR_INVALID_ARG;
    if (mesh->n_elem < 1 || mesh->n_elem >= (1 << 27) || mesh->n_face < 1 || !mesh->face_elem || !mesh->face_lf ||
        !mesh->face_ord || !geom->metrics_vc || !geom->jdet_vc)
        return DGX_ERR_INVALID_ARG;

    // operator set: built for p, or the caller's after checking that it is the GL/GL set of order p (the
    // kernels' diagonal inverse mass and collocated cubature are exact for it only); its values are applied
    dgx_ops1d o;
    if (dgx_ops1d_build(p, &o) != DGX_OK)
        return DGX_ERR_INVALID_ARG;
    if (ops_in) {
        if (!operators_match(*ops_in, o))
            return DGX_ERR_INVALID_ARG;
        o = ops_in->line;
    }

    const int n1 = p + 1;
    const int ns = (d == 2) ? n1 * n1 : n1 * n1 * n1;
    const int nface = 2 * d;
    const int64_t E = mesh->n_elem, F = mesh->n_face;

    // neighbour table in element order: every (element, local face) slot must be filled exactly once
    std::vector<int64_t> nbe((size_t)E * nface, -1);
    std::vector<int8_t> nbl((size_t)E * nface, -1), nbr_right((size_t)E * nface, 0), nbo((size_t)E * nface, 0);
    bool aligned = true;
    for (int64_t f = 0; f < F; ++f) {
        const int64_t eL = mesh->face_elem[2 * f], eR = mesh->face_elem[2 * f + 1];
        const int lL = mesh->face_lf[2 * f], lR = mesh->face_lf[2 * f + 1];
        if (eL < 0 || eL >= E || eR < 0 || eR >= E || lL < 0 || lL >= nface || lR < 0 || lR >= nface)
            return DGX_ERR_INVALID_ARG;
        if (mesh->face_ord[f] < 0 || mesh->face_ord[f] > (d == 2 ? 1 : 7))
            return DGX_ERR_INVALID_ARG;   // LINE faces have 2 correspondence cases, QUAD faces 8
        if (mesh->face_ord[f] != 0)
            aligned = false;
        const size_t sl = eL * nface + lL, sr = eR * nface + lR;
        if (nbe[sl] >= 0 || nbe[sr] >= 0)
            return DGX_ERR_INVALID_ARG;
        nbe[sl] = eR;
        nbl[sl] = (int8_t)lR;
        nbr_right[sl] = 0;
        nbe[sr] = eL;
        nbl[sr] = (int8_t)lL;
        nbr_right[sr] = 1;
        nbo[sl] = nbo[sr] = mesh->face_ord[f];   // one gather for both sides, as the reference applies nc_fc
    }
    for (int64_t v : nbe)
        if (v < 0)
            return DGX_ERR_INVALID_ARG;   // boundary faces are not part of the periodic explicit path
    if (!aligned && visc)
        return DGX_ERR_INVALID_ARG;       // the Navier-Stokes kernels gather aligned faces only

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return DGX_ERR_NO_DEVICE;

    dgx_ctx* c = new dgx_ctx();
    c->d = d;
    c->p = p;
    c->nv = d + 2;
    c->ns = ns;
    c->nf = (d == 2) ? n1 : n1 * n1;
    c->nface = nface;
    c->flux = cfg->num_flux;
    c->scheme = cfg->solver_type_e;
    c->n_stages = (c->scheme == DGX_SOLVER_E_EULER) ? 1 : (c->scheme == DGX_SOLVER_E_LS_RK_54 ? 5 : 3);
    c->positivity = cfg->positivity ? 1 : 0;
    c->precision = cfg->precision;
    c->E = E;
    c->F = F;
    c->visc = visc;
    if (visc) {
        c->mu = cfg->mu;
        c->Pr = cfg->Pr;
        c->eta = 1.01 * nface;   // PENALTY_SCALING * NFMAX (compute_face_rlhs_dg_T.c:276-296)
    }
    // the HEX p=3 group kernel assumes aligned faces; other meshes run the generic kernel of that order
    c->hex3 = (d == 3 && p == 3 && !visc && aligned);
    // the precision form is built for Euler HEX p=3 (group kernel) and Navier-Stokes HEX p=4
    c->ns4 = (d == 3 && p == 4 && visc);
    if (cfg->precision == DGX_PREC_MIXED && !c->hex3 && !c->ns4) {
        delete c;
        return DGX_ERR_INVALID_ARG;
    }

    std::vector<int32_t> slot_elem;
    if (c->hex3) {
        slot_elem = build_groups(E, nbe);
    } else {
        slot_elem.resize((size_t)E);
        for (int64_t e = 0; e < E; ++e)
            slot_elem[e] = (int32_t)e;
    }
    const int64_t Ep = (int64_t)slot_elem.size();
    c->Ep = Ep;
    std::vector<int32_t> elem_slot((size_t)E, -1);
    for (int64_t s = 0; s < Ep; ++s)
        if (slot_elem[s] >= 0)
            elem_slot[slot_elem[s]] = (int32_t)s;

    // neighbour records in slot order; group face lists
    std::vector<uint32_t> nbr((size_t)Ep * nface, 0u);
    for (int64_t s = 0; s < Ep; ++s) {
        const int64_t e = slot_elem[s];
        if (e < 0)
            continue;
        for (int f = 0; f < nface; ++f) {
            const size_t k = e * nface + f;
            nbr[s * nface + f] = dgx::pack_nbr((uint32_t)elem_slot[nbe[k]], (uint32_t)nbl[k], (uint32_t)nbr_right[k]);
        }
    }
    std::vector<uint64_t> fitem;
    if (c->hex3) {
        const int64_t G = Ep / 8;
        fitem.assign((size_t)G * 48, ~0ull);
        for (int64_t g = 0; g < G; ++g) {
            int cnt = 0;
            for (int sl = 0; sl < 8; ++sl) {
                const int64_t s = g * 8 + sl;
                if (slot_elem[s] < 0)
                    continue;
                for (int f = 0; f < nface; ++f) {
                    const uint32_t rec = nbr[s * nface + f];
                    const int64_t nsl = rec & 0x0FFFFFFFu;
                    const bool right = (rec >> 31) != 0;
                    const bool local = (nsl / 8) == g;
                    if (!local || !right)
                        fitem[g * 48 + cnt++] = ((uint64_t)(sl * 6 + f) << 32) | rec;
                }
            }
        }
    }

    dgx::Ops1D ops;
    std::memset(&ops, 0, sizeof(ops));
    for (int i = 0; i < n1; ++i) {
        ops.w[i] = o.w[i];
        ops.l[0][i] = o.lm[i];
        ops.l[1][i] = o.lp[i];
    }
    for (int i = 0; i < n1 * n1; ++i) {
        ops.D[i] = o.D[i];
        ops.sb[i] = o.sb[i];
        ops.bs[i] = o.bs[i];
    }

    // metrics -> [d*d][Ep][ns]; inverse of the diagonal mass 1/(w J) -> [Ep][ns]
    std::vector<double> met((size_t)d * d * Ep * ns, 0.0), imass((size_t)Ep * ns, 0.0);
    for (int64_t s = 0; s < Ep; ++s) {
        const int64_t e = slot_elem[s];
        if (e < 0)
            continue;
        for (int cc = 0; cc < d * d; ++cc)
            std::memcpy(&met[((size_t)cc * Ep + s) * ns], geom->metrics_vc + ((size_t)e * d * d + cc) * ns,
                        sizeof(double) * ns);
        for (int n = 0; n < ns; ++n) {
            const int i0 = n % n1, i1 = (n / n1) % n1, i2 = n / (n1 * n1);
            double W = o.w[i0] * o.w[i1];
            if (d == 3)
                W *= o.w[i2];
            const double J = geom->jdet_vc[e * ns + n];
            if (!(J > 0.0)) {
                delete c;
                return DGX_ERR_GEOMETRY;
            }
            imass[s * ns + n] = 1.0 / (W * J);
        }
    }

    // NS lift trace factors lam_f(m) = w_f(m) sum_k l_f(k)^2 / (w J)(node(k, m)) (the trace on face f of
    // the lifting of a unit face datum at m)
    const int nf = (d == 2) ? n1 : n1 * n1;
    std::vector<double> lam;
    if (visc) {
        lam.assign((size_t)Ep * nface * nf, 0.0);
        for (int64_t s = 0; s < Ep; ++s) {
            if (slot_elem[s] < 0)
                continue;
            for (int f = 0; f < nface; ++f) {
                const int dn = f / 2;
                const double* l = (f & 1) ? o.lp : o.lm;
                for (int m = 0; m < nf; ++m) {
                    const int m0 = m % n1, m1 = m / n1;
                    const double wf = (d == 2) ? o.w[m] : o.w[m0] * o.w[m1];
                    double acc = 0.0;
                    for (int k = 0; k < n1; ++k) {
                        int node;
                        if (d == 2)
                            node = (dn == 0) ? k + n1 * m : m + n1 * k;
                        else
                            node = (dn == 0) ? k + n1 * (m0 + n1 * m1)
                                             : (dn == 1) ? m0 + n1 * (k + n1 * m1) : m0 + n1 * (m1 + n1 * k);
                        acc += l[k] * l[k] * imass[s * ns + node];
                    }
                    lam[((size_t)s * nface + f) * nf + m] = wf * acc;
                }
            }
        }
    }

    c->nmax = c->hex3 ? Ep : (visc ? dgx::ns_grid(d, p, E) : dgx::stage_grid(d, p, E));
    const size_t sbytes = sizeof(double) * (size_t)Ep * c->nv * ns;
    const size_t tbytes = sizeof(double) * (size_t)Ep * nface * c->nv * c->nf;
    const size_t gbytes = (d == 3 && p == 4) ? tbytes : (size_t)d * tbytes;   // gtr
    dgx_status st = DGX_OK;
    auto step = [&](hipError_t e_, const char* what) {
        if (st == DGX_OK && e_ != hipSuccess)
            st = fail(c, DGX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_));
    };
    step(hipStreamCreate(&c->stream), "hipStreamCreate");
    step(hipEventCreate(&c->ev0), "hipEventCreate");
    step(hipEventCreate(&c->ev1), "hipEventCreate");
    step(hipMalloc(&c->u, sbytes), "hipMalloc(u)");
    step(hipMalloc(&c->up, sbytes), "hipMalloc(up)");
    step(hipMalloc(&c->rhs, sbytes), "hipMalloc(rhs)");
    step(hipMalloc(&c->stage_buf, sizeof(double) * (size_t)E * c->nv * ns), "hipMalloc(stage)");
    step(hipMalloc(&c->met, sizeof(double) * met.size()), "hipMalloc(met)");
    step(hipMalloc(&c->imass, sizeof(double) * imass.size()), "hipMalloc(imass)");
    step(hipMalloc(&c->tr[0], tbytes), "hipMalloc(tr0)");
    step(hipMalloc(&c->tr[1], tbytes), "hipMalloc(tr1)");
    step(hipMalloc(&c->nbr, sizeof(uint32_t) * nbr.size()), "hipMalloc(nbr)");
    step(hipMalloc(&c->slot_elem, sizeof(int32_t) * slot_elem.size()), "hipMalloc(slot_elem)");
    step(hipMalloc(&c->elem_slot, sizeof(int32_t) * elem_slot.size()), "hipMalloc(elem_slot)");
    if (!aligned)
        step(hipMalloc(&c->nord, (size_t)Ep * nface), "hipMalloc(nord)");
    if (c->hex3) {
        step(hipMalloc(&c->face_item, sizeof(uint64_t) * fitem.size()), "hipMalloc(face_item)");
    }
    if (visc) {
        step(hipMalloc(&c->gtr[0], gbytes), "hipMalloc(gtr)");
        step(hipMalloc(&c->lam, sizeof(double) * lam.size()), "hipMalloc(lam)");
        if (d == 3 && p == 4) {
            step(hipMalloc(&c->gtr[1], gbytes), "hipMalloc(gtr1)");
            step(hipMalloc(&c->fnv, sizeof(double) * (size_t)Ep * nface * 3 * nf), "hipMalloc(fnv)");
        }
    }
    step(hipMalloc(&c->blockmax, sizeof(double) * c->nmax), "hipMalloc(blockmax)");
    step(hipMalloc(&c->dmax, sizeof(double)), "hipMalloc(dmax)");
    step(hipMalloc(&c->bad_elem, sizeof(unsigned long long)), "hipMalloc(bad)");
    step(hipMalloc(&c->nonfinite, sizeof(unsigned int)), "hipMalloc(nonfinite)");
    step(hipMalloc(&c->gate, 2 * sizeof(unsigned int)), "hipMalloc(gate)");
    if (st == DGX_OK) {
        step(hipMemcpy(c->met, met.data(), sizeof(double) * met.size(), hipMemcpyHostToDevice), "copy met");
        step(hipMemcpy(c->imass, imass.data(), sizeof(double) * imass.size(), hipMemcpyHostToDevice), "copy imass");
        step(hipMemcpy(c->nbr, nbr.data(), sizeof(uint32_t) * nbr.size(), hipMemcpyHostToDevice), "copy nbr");
        step(hipMemcpy(c->slot_elem, slot_elem.data(), sizeof(int32_t) * slot_elem.size(), hipMemcpyHostToDevice),
             "copy slot_elem");
        step(hipMemcpy(c->elem_slot, elem_slot.data(), sizeof(int32_t) * elem_slot.size(), hipMemcpyHostToDevice),
             "copy elem_slot");
        if (c->nord) {
            // slot order (the identity for the generic kernels)
            std::vector<uint8_t> nord((size_t)Ep * nface, 0);
            for (int64_t sl = 0; sl < Ep; ++sl)
                if (slot_elem[sl] >= 0)
                    for (int f = 0; f < nface; ++f)
                        nord[sl * nface + f] = (uint8_t)nbo[(size_t)slot_elem[sl] * nface + f];
            step(hipMemcpy(c->nord, nord.data(), nord.size(), hipMemcpyHostToDevice), "copy nord");
        }
        if (c->hex3) {
            step(hipMemcpy(c->face_item, fitem.data(), sizeof(uint64_t) * fitem.size(), hipMemcpyHostToDevice),
                 "copy face_item");
        }
        step(hipMemset(c->u, 0, sbytes), "memset u");
        step(hipMemset(c->up, 0, sbytes), "memset up");
        step(hipMemset(c->rhs, 0, sbytes), "memset rhs");
        step(hipMemset(c->tr[0], 0, tbytes), "memset tr0");
        step(hipMemset(c->tr[1], 0, tbytes), "memset tr1");
        step(hipMemset(c->blockmax, 0, sizeof(double) * c->nmax), "memset blockmax");
        if (visc) {
            step(hipMemcpy(c->lam, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice), "copy lam");
            for (int i = 0; i < 2; ++i)
                if (c->gtr[i])
                    step(hipMemset(c->gtr[i], 0, gbytes), "memset gtr");
        }
        step(dgx::upload_ops(p, ops), "upload ops");
        if (c->fnv)
            step(dgx::launch_ns4_face_normals(c->met, c->fnv, E, c->stream), "face normals");
    }
    if (st == DGX_OK)
        st = reset_flags(c);
    if (st == DGX_OK)
        step(hipStreamSynchronize(c->stream), "sync");
    if (st != DGX_OK) {
        fprintf(stderr, "dgx_create: %s\n", c->err.c_str());
        free_all(c);
        delete c;
        return st;
    }
    *out = c;
    return DGX_OK;
}

dgx_status dgx_destroy(dgx_ctx* ctx)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return DGX_OK;
}

void* dgx_get_stream(dgx_ctx* ctx)
{
    return ctx ? (void*)ctx->stream : nullptr;
}

int64_t dgx_state_size(const dgx_ctx* ctx)
{
    return ctx ? ctx->E * ctx->nv * ctx->ns : 0;
}

dgx_status dgx_set_state(dgx_ctx* ctx, int layout, int ptr_kind, const double* u)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (!u || (layout != DGX_LAYOUT_REF && layout != DGX_LAYOUT_SOA) ||
        (ptr_kind != DGX_PTR_HOST && ptr_kind != DGX_PTR_DEVICE))
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_set_state: bad arguments");
    dgx_status st = upload_state(ctx, layout, ptr_kind, u, ctx->u);
    if (st != DGX_OK)
        return st;
    ctx->cur = 0;
    ctx->state_buf = 0;
    ctx->rk_next = 0;
    if (ctx->hex3)
        HIP_TRY(ctx, dgx::launch_hex3_trace(ctx->u, ctx->tr[0], ctx->slot_elem, ctx->Ep, ctx->stream));
    else
        HIP_TRY(ctx, dgx::launch_trace(ctx->d, ctx->p, ctx->u, ctx->tr[0], ctx->E, ctx->stream));
    if (ctx->fnv)
        HIP_TRY(ctx, dgx::launch_ns4_vtrace(ctx->u, ctx->met, ctx->imass, ctx->fnv, ctx->mu, ctx->Pr, ctx->gtr[0],
                                            ctx->E, ctx->stream));
    st = reset_flags(ctx);
    if (st != DGX_OK)
        return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

dgx_status dgx_get_state(dgx_ctx* ctx, int layout, int ptr_kind, double* u)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (!u || (layout != DGX_LAYOUT_REF && layout != DGX_LAYOUT_SOA) ||
        (ptr_kind != DGX_PTR_HOST && ptr_kind != DGX_PTR_DEVICE))
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_get_state: bad arguments");
    dgx_status st = download_state(ctx, layout, ptr_kind, ctx->state_buf ? ctx->up : ctx->u, u);
    if (st != DGX_OK)
        return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DGX_OK;
}

dgx_status dgx_rhs(dgx_ctx* ctx, int precision, int apply_minv, int layout, int ptr_kind, double* rhs,
                   double* max_rhs)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (precision != DGX_PREC_FP64 && precision != DGX_PREC_MIXED)
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_rhs: bad precision");
    if (precision == DGX_PREC_MIXED && !ctx->hex3 && !ctx->ns4)
        return fail(ctx, DGX_ERR_INVALID_ARG,
                    "dgx_rhs: the precision form is built for Euler HEX p=3 and Navier-Stokes HEX p=4");
    if (rhs && ((layout != DGX_LAYOUT_REF && layout != DGX_LAYOUT_SOA) ||
                (ptr_kind != DGX_PTR_HOST && ptr_kind != DGX_PTR_DEVICE)))
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_rhs: bad layout");
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, launch_any(ctx, -1, 0.0, apply_minv ? 1 : 0, ctx->cur, precision));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    dgx_status st = read_max(ctx, max_rhs);
    if (st != DGX_OK)
        return st;
    if (rhs) {
        st = download_state(ctx, layout, ptr_kind, ctx->rhs, rhs);
        if (st != DGX_OK)
            return st;
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->last_ms = ms;
    return check_flags(ctx, nullptr);
}

dgx_status dgx_rk_stage(dgx_ctx* ctx, int stage, double dt, double* max_rhs)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (stage < 0 || stage >= ctx->n_stages || !std::isfinite(dt))
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_rk_stage: bad arguments");
    if (stage != ctx->rk_next)
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_rk_stage: stages run in order (expected stage " +
                                                  std::to_string(ctx->rk_next) + ")");
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, launch_any(ctx, stage, dt, 1, ctx->cur, ctx->precision));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    ctx->cur ^= 1;
    ctx->rk_next = (stage + 1) % ctx->n_stages;
    if (ctx->hex3 && ctx->scheme == dgx::RK_SSP33)
        ctx->state_buf = (stage == 2) ? 0 : 1;
    dgx_status st = read_max(ctx, max_rhs);
    if (st != DGX_OK)
        return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->last_ms = ms;
    return check_flags(ctx, nullptr);
}

static dgx_status build_graph(dgx_ctx* c, int parity, double dt)
{
    if (c->graph[parity]) {
        (void)hipGraphExecDestroy(c->graph[parity]);
        c->graph[parity] = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int cur = parity;
    hipError_t le = hipSuccess;
    for (int s = 0; s < c->n_stages && le == hipSuccess; ++s) {
        le = launch_any(c, s, dt, 1, cur, c->precision);
        cur ^= 1;
    }
    hipError_t ee = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess)
        return fail(c, DGX_ERR_HIP, std::string("capture launch: ") + hipGetErrorString(le));
    if (ee != hipSuccess)
        return fail(c, DGX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
    hipError_t ie = hipGraphInstantiate(&c->graph[parity], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(c, DGX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    return DGX_OK;
}

dgx_status dgx_step(dgx_ctx* ctx, double dt, int nsteps, double* max_rhs)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    if (nsteps < 0 || !std::isfinite(dt))
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_step: bad arguments");
    if (ctx->rk_next != 0)
        return fail(ctx, DGX_ERR_INVALID_ARG, "dgx_step: a step started by dgx_rk_stage is unfinished");
    if (nsteps == 0)
        return DGX_OK;
    if (dt != ctx->graph_dt || !ctx->graph[0] || !ctx->graph[1]) {
        for (int par = 0; par < 2; ++par) {
            dgx_status st = build_graph(ctx, par, dt);
            if (st != DGX_OK)
                return st;
        }
        ctx->graph_dt = dt;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int s = 0; s < nsteps; ++s) {
        HIP_TRY(ctx, hipGraphLaunch(ctx->graph[ctx->cur], ctx->stream));
        ctx->cur ^= 1;   // an odd number of stages (3, 1 or 5) flips the trace parity
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    dgx_status st = read_max(ctx, max_rhs);
    if (st != DGX_OK)
        return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->last_ms = ms;
    return check_flags(ctx, nullptr);
}

int dgx_stages_per_step(const dgx_ctx* ctx)
{
    return ctx ? ctx->n_stages : 0;
}

dgx_status dgx_sync(dgx_ctx* ctx, int64_t* elem)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return check_flags(ctx, elem);
}

// J^T v or J w of the residual at the current state (pre-M^-1 residual, or M^-1 R with apply_minv).
static dgx_status lin_apply(dgx_ctx* c, bool adjoint, int apply_minv, int layout, int ptr_kind, const double* in,
                            double* out, const char* who)
{
    if (!in || !out || (layout != DGX_LAYOUT_REF && layout != DGX_LAYOUT_SOA) ||
        (ptr_kind != DGX_PTR_HOST && ptr_kind != DGX_PTR_DEVICE))
        return fail(c, DGX_ERR_INVALID_ARG, std::string(who) + ": bad arguments");
    if (c->visc)
        return fail(c, DGX_ERR_INVALID_ARG, std::string(who) + ": the linearized residual is built for Euler");
    const size_t sbytes = sizeof(double) * (size_t)c->Ep * c->nv * c->ns;
    const size_t tbytes = sizeof(double) * (size_t)c->Ep * c->nface * c->nv * c->nf;
    if (!c->lin_x) {
        HIP_TRY(c, hipMalloc(&c->lin_x, sbytes));
        HIP_TRY(c, hipMalloc(&c->lin_tr, tbytes));
    }
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    const double* src = in;
    if (ptr_kind == DGX_PTR_HOST) {
        HIP_TRY(c, hipMemcpyAsync(c->stage_buf, in, sizeof(double) * (size_t)c->E * c->nv * c->ns,
                                  hipMemcpyHostToDevice, c->stream));
        src = c->stage_buf;
    }
    // the input in slot order (times M^-1 for the adjoint of M^-1 R) and its face traces, one pass
    HIP_TRY(c, dgx::launch_upload_trace(c->d, c->p, src, layout, c->slot_elem, (adjoint && apply_minv) ? c->imass : nullptr,
                                        c->lin_x, c->lin_tr, c->E, c->Ep, c->stream));
    dgx_status st = DGX_OK;
    dgx::LinArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = c->state_buf ? c->up : c->u;
    a.x = c->lin_x;
    a.met = c->met;
    a.imass = c->imass;
    a.tr_u = c->tr[c->cur];
    a.tr_x = c->lin_tr;
    a.nbr = c->nbr;
    a.nord = c->nord;
    a.slot_elem = c->hex3 ? c->slot_elem : nullptr;
    a.out = c->rhs;
    a.n_elem = c->Ep;
    a.apply_minv = (!adjoint && apply_minv) ? 1 : 0;
    HIP_TRY(c, dgx::launch_lin(c->d, c->p, c->flux, adjoint, a, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    st = download_state(c, layout, ptr_kind, c->rhs, out);
    if (st != DGX_OK)
        return st;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
        c->last_ms = ms;
    return DGX_OK;
}

dgx_status dgx_vjp(dgx_ctx* ctx, int apply_minv, int layout, int ptr_kind, const double* v, double* out)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return lin_apply(ctx, true, apply_minv, layout, ptr_kind, v, out, "dgx_vjp");
}

dgx_status dgx_jvp(dgx_ctx* ctx, int apply_minv, int layout, int ptr_kind, const double* w, double* out)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return lin_apply(ctx, false, apply_minv, layout, ptr_kind, w, out, "dgx_jvp");
}

static dgx_status jacobian_blocks(dgx_ctx* c, int apply_minv, double* bt64, float* bt32)
{
    if (!bt64 && !bt32)
        return fail(c, DGX_ERR_INVALID_ARG, "dgx_jacobian_blocks: bad arguments");
    if (c->visc)
        return fail(c, DGX_ERR_INVALID_ARG, "dgx_jacobian_blocks: the linearized residual is built for Euler");
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    dgx::LinArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u = c->state_buf ? c->up : c->u;
    a.met = c->met;
    a.imass = c->imass;
    a.tr_u = c->tr[c->cur];
    a.nbr = c->nbr;
    a.nord = c->nord;
    a.slot_elem = c->hex3 ? c->slot_elem : nullptr;
    a.n_elem = c->Ep;
    a.apply_minv = apply_minv ? 1 : 0;
    if (bt64)
        HIP_TRY(c, dgx::launch_lin_block(c->d, c->p, c->flux, a, bt64, c->stream));
    else
        HIP_TRY(c, dgx::launch_lin_block_f32(c->d, c->p, c->flux, a, bt32, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess)
        c->last_ms = ms;
    return DGX_OK;
}

dgx_status dgx_jacobian_blocks(dgx_ctx* ctx, int apply_minv, double* blocks_t)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return jacobian_blocks(ctx, apply_minv, blocks_t, nullptr);
}

dgx_status dgx_jacobian_blocks_f32(dgx_ctx* ctx, int apply_minv, float* blocks_t)
{
    if (!ctx)
        return DGX_ERR_INVALID_ARG;
    return jacobian_blocks(ctx, apply_minv, nullptr, blocks_t);
}

double dgx_last_elapsed_ms(const dgx_ctx* ctx)
{
    return ctx ? ctx->last_ms : 0.0;
}

}  // extern "C"
