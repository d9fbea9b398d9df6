// Fused stage kernel, trace kernel, layout kernels and reductions for gfx950 (FP64 plain form).
// See dgx_kernels.hpp for the reference mapping.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dgx_internal.hpp"
#include "dgx_kernels.hpp"

namespace dgx {

// One operator set per order p (index p), written by the host before any launch for that order.
__constant__ Ops1D c_ops[8];

// ---------------------------------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------------------------------

__device__ inline double nan_max(double a, double b)
{
    return (a > b || a != a) ? a : b;
}

template <int D, int P>
__device__ inline void node_index(int n, int* idx)
{
    constexpr int N1 = P + 1;
    idx[0] = n % N1;
    idx[1] = (n / N1) % N1;
    idx[2] = (D == 3) ? n / (N1 * N1) : 0;
}

template <int D, int P>
__device__ inline int tangential_index(int dn, const int* idx)
{
    constexpr int N1 = P + 1;
    if constexpr (D == 2) {
        return dn == 0 ? idx[1] : idx[0];
    } else {
        if (dn == 0) return idx[1] + N1 * idx[2];
        if (dn == 1) return idx[0] + N1 * idx[2];
        return idx[0] + N1 * idx[1];
    }
}

template <int P>
__device__ inline int stride_of(int dir)
{
    constexpr int N1 = P + 1;
    return dir == 0 ? 1 : (dir == 1 ? N1 : N1 * N1);
}

// Apply a 1-D N1 x N1 operator A (row-major) along direction dir to every variable of the element
// tile s[v*NS + n] in place.  All threads of the block must call it (it holds barriers).
template <int D, int P, int NV>
__device__ inline void tp_pass_lds(double* s, const double* A, int dir, int n, bool active)
{
    constexpr int N1 = P + 1;
    constexpr int NS = Shape<D, P>::NS;
    double out[NV];
    if (active) {
        int idx[3];
        node_index<D, P>(n, idx);
        const int st = stride_of<P>(dir);
        const int base = n - idx[dir] * st;
        const double* row = A + idx[dir] * N1;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < N1; ++k)
                acc += row[k] * s[v * NS + base + k * st];
            out[v] = acc;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            s[v * NS + n] = out[v];
    }
    __syncthreads();
}

// Sum and min reductions over the NS nodes of each element tile.  red: [EPB][NQ][NSP] doubles.
// Every thread of the block calls it; results are returned to the element's threads.
template <int D, int P, int NQ>
__device__ inline void element_reduce(double* red, int el, int n, bool in_tile, const double* sum_in, double min_in,
                                      double* sum_out, double* min_out)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NSP = S::NSP;
    double* r = red + el * (NQ + 1) * NSP;
    if (in_tile) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            r[q * NSP + n] = sum_in[q];
        r[NQ * NSP + n] = min_in;
        if (n < NSP - NS) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                r[q * NSP + NS + n] = 0.0;
            r[NQ * NSP + NS + n] = 1.0e300;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = NSP / 2; s > 0; s >>= 1) {
        if (in_tile && n < s) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                r[q * NSP + n] += r[q * NSP + n + s];
            r[NQ * NSP + n] = fmin(r[NQ * NSP + n], r[NQ * NSP + n + s]);
        }
        __syncthreads();
    }
    if (in_tile) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            sum_out[q] = r[q * NSP];
        *min_out = r[NQ * NSP];
    }
    __syncthreads();
}

// Face traces of the element tile s[v*NS + n] written to tr[e][f][v][m].
template <int D, int P>
__device__ inline void write_traces(const double* s, double* tr, int64_t e, int n, bool valid)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NV = S::NV, NF = S::NF, NFACE = S::NFACE, N1 = S::N1;
    const Ops1D& op = c_ops[P];
    if (!valid)
        return;
    for (int q = n; q < NFACE * NF; q += NS) {
        const int f = q / NF, m = q % NF, dn = f >> 1, side = f & 1;
        const double* l = op.l[side];
        double* dst = tr + ((e * NFACE + f) * NV) * NF + m;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < N1; ++a)
                acc += l[a] * s[v * NS + face_node<D, P>(dn, a, m)];
            dst[v * NF] = acc;
        }
    }
}

// Block max of |rhs|, then residual output (stage < 0) or the SSP-RK3 combination, the positivity
// limiter, the state write and the face traces of the updated state.  Every thread of the block calls
// it (it holds barriers); su is the element's [NV][NS] tile (free to overwrite), rhs the node's residual.
template <int D, int P>
__device__ inline void stage_tail(const StageArgs& a, int64_t e, int n, int el, bool in_tile, bool valid,
                                  double* su, double* s_red, double* s_wmax, const double* rhs, const double* uu,
                                  double lmax)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NV = S::NV;
    const Ops1D& op = c_ops[P];
    const int tid = threadIdx.x;
    const int64_t E = a.n_elem;
    // block max of |rhs| -> one value per block (reduced by a second kernel when requested)
    {
        double m = lmax;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            m = nan_max(m, __shfl_xor(m, off, 64));
        if ((tid & 63) == 0)
            s_wmax[tid >> 6] = m;
        __syncthreads();
        if (tid == 0) {
            double b = s_wmax[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
                b = nan_max(b, s_wmax[w]);
            a.blockmax[blockIdx.x] = b;
        }
    }

    if (a.stage < 0) {
        if (valid) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                a.rhs_out[(v * E + e) * NS + n] = rhs[v];
        }
        return;
    }

    // ---- stage combination of the time stepper (solve_explicit.c:226-350) ----------------------------
    double un[NV];
    if (valid) {
        const double dt = a.dt;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int64_t gi = (v * E + e) * NS + n;
            un[v] = rk_combine(a.scheme, a.stage, a.rk_a, a.rk_b, dt, uu[v], rhs[v], a.up + gi);
        }
    }
    __syncthreads();   // sg / snnf reads are complete before su is overwritten
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            su[v * NS + n] = un[v];
    }
    __syncthreads();

    // ---- positivity: enforce_positivity_highorder (solve.c:187-233) -----------------------------
    if (a.positivity) {
        // Lagrange -> Bezier coefficients (ccSB0, one 1-D pass per direction)
#pragma unroll
        for (int dir = 0; dir < D; ++dir)
            tp_pass_lds<D, P, NV>(su, op.sb, dir, n, valid);
        double b[NV];
        if (valid) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                b[v] = su[v * NS + n];
            // conservative -> primitive (solution_euler_T.c:174-189)
            const double rho_inv = 1.0 / b[0];
            double rho2V2 = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                rho2V2 += b[1 + k] * b[1 + k];
                b[1 + k] *= rho_inv;
            }
            b[NV - 1] = kGm1 * (b[NV - 1] - 0.5 * rho2V2 * rho_inv);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                b[v] = 1.0;
        }
        // Every Bezier coefficient of every element of the block at or above the bound: nothing is
        // limited and no element fails (an element average is never below its minimum), so the reductions
        // and the Bezier -> Lagrange transform are skipped and the nodal stage values are kept (the
        // reference's unconditional round trip changes them only by round-off).  NaN takes the full path.
        const bool low = valid && !(b[0] >= kEpsPhys && b[NV - 1] >= kEpsPhys);
        if (__syncthreads_or(low)) {
            bool failed = false, changed = false;
            double sums[NV], mn;
            element_reduce<D, P, NV>(s_red, el, n, in_tile, b, b[0], sums, &mn);
            double avg[NV];
            if (in_tile) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    avg[v] = sums[v] * (1.0 / NS);
                if (avg[0] < kEpsPhys) {
                    failed = true;
                } else if (mn < kEpsPhys) {
                    double p_ho = (avg[0] - kEpsPhys) / (avg[0] - mn);
                    p_ho = fmax(0.0, fmin(1.0, p_ho));
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        double av = avg[v];
                        if ((v == 0 || v == NV - 1) && av < kEpsPhys)
                            av = kEpsPhys;
                        b[v] = p_ho * b[v] + (1.0 - p_ho) * av;
                        avg[v] = p_ho * avg[v] + (1.0 - p_ho) * av;
                    }
                    changed = true;
                }
            }
            double sp_sum[1], mnp;
            element_reduce<D, P, 1>(s_red, el, n, in_tile, &b[NV - 1], b[NV - 1], sp_sum, &mnp);
            if (in_tile && !failed) {
                const double avgp = sp_sum[0] * (1.0 / NS);
                if (avgp < kEpsPhys) {
                    failed = true;
                } else if (mnp < kEpsPhys) {
                    double p_ho = (avgp - kEpsPhys) / (avgp - mnp);
                    p_ho = fmax(0.0, fmin(1.0, p_ho));
                    avg[NV - 1] = avgp;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        double av = avg[v];
                        if ((v == 0 || v == NV - 1) && av < kEpsPhys)
                            av = kEpsPhys;
                        b[v] = p_ho * b[v] + (1.0 - p_ho) * av;
                    }
                    changed = true;
                }
            }
            (void)changed;
            if (valid && failed && n == 0) {
                atomicMin(a.bad_elem, (unsigned long long)e);
                *a.gate_out = 1u;
            }
            if (valid && !failed) {
                // primitive -> conservative (solution_euler_T.c:152-167), then Bezier -> Lagrange (ccBS0)
                double V2 = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    V2 += b[1 + k] * b[1 + k];
                    b[1 + k] *= b[0];
                }
                b[NV - 1] = b[NV - 1] / kGm1 + 0.5 * b[0] * V2;
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    su[v * NS + n] = b[v];
            }
            __syncthreads();
            // the element's state is left as is (no round trip) when it failed the check
            const bool pass_active = valid && !failed;
#pragma unroll
            for (int dir = 0; dir < D; ++dir)
                tp_pass_lds<D, P, NV>(su, op.bs, dir, n, pass_active);
            if (valid && failed) {
                // restore the un-limited stage result for the traces below
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    su[v * NS + n] = un[v];
            }
            __syncthreads();
            if (valid) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    un[v] = su[v * NS + n];
            }
        } else if (valid) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                su[v * NS + n] = un[v];
        }
        __syncthreads();
    }

    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            a.u[(v * E + e) * NS + n] = un[v];
    }
    write_traces<D, P>(su, a.tr_out, e, n, valid);
}

// ---------------------------------------------------------------------------------------------------
// fused stage kernel
// ---------------------------------------------------------------------------------------------------

template <int D, int P, int FLUX>
__global__ __launch_bounds__(256) void stage_kernel(StageArgs a)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NV = S::NV, NF = S::NF, NFACE = S::NFACE, EPB = S::EPB, N1 = S::N1;
    constexpr int NSP = S::NSP, DD = D * D;
    constexpr int REG_A = NV * NS + DD * NS;
    constexpr int REG_B = NV * D * NS;
    constexpr int REGION = REG_A > REG_B ? REG_A : REG_B;
    constexpr int NRED = (NV + 1) * NSP;
    __shared__ double s_reg[EPB * REGION];
    __shared__ double s_nnf[EPB * NFACE * NV * NF];
    __shared__ double s_red[EPB * NRED];
    __shared__ double s_wmax[4];

    const Ops1D& op = c_ops[P];
    const int tid = threadIdx.x;
    const int el = tid / NS;
    const int n = tid - el * NS;
    const bool in_tile = el < EPB;
    const int64_t E = a.n_elem;
    const int64_t e = (int64_t)blockIdx.x * EPB + el;
    const bool valid = in_tile && e < E;
    if (stage_gated(a.stage, a.gate_in)) {   // an earlier stage failed: the state stays as it left it
        if (tid == 0)
            *a.gate_out = 1u;
        return;
    }

    double* su = s_reg + (in_tile ? el : 0) * REGION;   // [NV][NS]
    double* sm = su + NV * NS;                          // [DD][NS]
    double* sg = su;                                    // [NV][D][NS] (after the face phase)
    double* snnf = s_nnf + (in_tile ? el : 0) * NFACE * NV * NF;

    double uu[NV], mm[DD];
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            uu[v] = a.u[(v * E + e) * NS + n];
            su[v * NS + n] = uu[v];
        }
#pragma unroll
        for (int c = 0; c < DD; ++c) {
            mm[c] = a.met[(c * E + e) * NS + n];
            sm[c * NS + n] = mm[c];
        }
    }
    __syncthreads();

    // ---- face phase: numerical flux at every face point of the element -------------------------
    if (valid) {
        for (int q = n; q < NFACE * NF; q += NS) {
            const int f = q / NF, m = q % NF, dn = f >> 1, side = f & 1;
            const double* l = op.l[side];
            double wo[NV], nv[D];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < N1; ++k)
                    acc += l[k] * su[v * NS + face_node<D, P>(dn, k, m)];
                wo[v] = acc;
            }
            double jf = 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < N1; ++k)
                    acc += l[k] * sm[(i + D * dn) * NS + face_node<D, P>(dn, k, m)];
                nv[i] = side ? acc : -acc;
                jf += nv[i] * nv[i];
            }
            jf = sqrt(jf);
            const double jf_inv = 1.0 / jf;
            const uint32_t rec = a.nbr[e * NFACE + f];
            const int64_t en = rec & 0x0FFFFFFFu;
            const int lfn = (rec >> 28) & 7;
            const bool right = (rec >> 31) != 0;
            double wn[NV], nf[NV];
            const double* src = a.tr_in + ((en * NFACE + lfn) * NV) * NF + nbr_node<D, P>(a.nord, e, f, m);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                wn[v] = src[v * NF];
            // the left volume's state first and its outward normal; arguments are selected, not branched
            // on, since the lanes of a wavefront serve left and right sides of faces
            double nh[D], wl[NV], wr[NV];
#pragma unroll
            for (int i = 0; i < D; ++i)
                nh[i] = right ? -nv[i] * jf_inv : nv[i] * jf_inv;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                wl[v] = right ? wn[v] : wo[v];
                wr[v] = right ? wo[v] : wn[v];
            }
            numflux<D, FLUX>(wl, wr, nh, nf);
            const double scale = (right ? -jf : jf) * face_weight<D, P>(op, m);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                snnf[(f * NV + v) * NF + m] = nf[v] * scale;
        }
    }
    __syncthreads();

    // ---- volume: weighted reference flux g(eq, j) = W sum_i m(i,j) F_i(eq) -----------------------
    int idx[3];
    node_index<D, P>(n, idx);
    if (valid) {
        double W = op.w[idx[0]] * op.w[idx[1]];
        if constexpr (D == 3)
            W *= op.w[idx[2]];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double dirv[D];
#pragma unroll
            for (int i = 0; i < D; ++i)
                dirv[i] = mm[i + D * j] * W;
            double fj[NV];
            euler_flux_n<D>(uu, dirv, fj);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                sg[(v * D + j) * NS + n] = fj[v];
        }
    }
    __syncthreads();

    // ---- weak divergence + lift + inverse mass ---------------------------------------------------
    double rhs[NV];
    double lmax = 0.0;
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double r = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const int st = stride_of<P>(j);
                const int base = n - idx[j] * st;
                const double* g = sg + (v * D + j) * NS + base;
#pragma unroll
                for (int q = 0; q < N1; ++q)
                    r += op.D[q * N1 + idx[j]] * g[q * st];
            }
#pragma unroll
            for (int f = 0; f < NFACE; ++f) {
                const int dn = f >> 1, side = f & 1;
                const int m = tangential_index<D, P>(dn, idx);
                r -= op.l[side][idx[dn]] * snnf[(f * NV + v) * NF + m];
            }
            rhs[v] = r;
        }
        const bool minv = (a.stage >= 0) || a.apply_minv;
        if (minv) {
            const double im = a.imass[e * NS + n];
#pragma unroll
            for (int v = 0; v < NV; ++v)
                rhs[v] *= im;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
            lmax = nan_max(lmax, fabs(rhs[v]));
    }

    stage_tail<D, P>(a, e, n, el, in_tile, valid, su, s_red, s_wmax, rhs, uu, lmax);
}

// ---------------------------------------------------------------------------------------------------
// trace kernel, layout kernels, reductions
// ---------------------------------------------------------------------------------------------------

template <int D, int P>
__global__ __launch_bounds__(256) void trace_kernel(const double* u, double* tr, int64_t E)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NV = S::NV, EPB = S::EPB;
    __shared__ double s_u[EPB * NV * NS];
    const int tid = threadIdx.x;
    const int el = tid / NS;
    const int n = tid - el * NS;
    const bool in_tile = el < EPB;
    const int64_t e = (int64_t)blockIdx.x * EPB + el;
    const bool valid = in_tile && e < E;
    double* su = s_u + (in_tile ? el : 0) * NV * NS;
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
            su[v * NS + n] = u[(v * E + e) * NS + n];
    }
    __syncthreads();
    write_traces<D, P>(su, tr, e, n, valid);
}

// Caller layouts (element order): ref [E][NV][NS] (layout 0) or soa [NV][E][NS] (layout 1).
// Internal state: [NV][Ep][NS] in slot order; padding slots hold zeros.
__device__ inline int64_t caller_index(int layout, int64_t e, int64_t v, int64_t n, int64_t E, int NV, int NS)
{
    return layout == 0 ? (e * NV + v) * NS + n : (v * E + e) * NS + n;
}

// The upload of a linearized product's input fused with its face traces: caller layout -> the internal
// slot-order array (times 1/(w J) when imass is given, the adjoint's M^-1) and, from the same LDS tile,
// the traces the linearized kernel reads for the neighbours.  One pass instead of upload + trace.
template <int D, int P>
__global__ __launch_bounds__(256) void upload_trace_kernel(const double* src, int layout, const int32_t* slot_elem,
                                                           const double* imass, double* dst, double* tr, int64_t E,
                                                           int64_t Ep)
{
    using S = Shape<D, P>;
    constexpr int NS = S::NS, NV = S::NV, EPB = S::EPB;
    __shared__ double s_u[EPB * NV * NS];
    const int tid = threadIdx.x;
    const int el = tid / NS;
    const int n = tid - el * NS;
    const bool in_tile = el < EPB;
    const int64_t s = (int64_t)blockIdx.x * EPB + el;
    const bool valid = in_tile && s < Ep;
    double* su = s_u + (in_tile ? el : 0) * NV * NS;
    if (valid) {
        const int64_t e = slot_elem ? (int64_t)slot_elem[s] : s;
        const double sc = imass ? imass[s * NS + n] : 1.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double x = e >= 0 ? src[caller_index(layout, e, v, n, E, NV, NS)] * sc : 0.0;
            su[v * NS + n] = x;
            dst[(v * Ep + s) * NS + n] = x;
        }
    }
    __syncthreads();
    write_traces<D, P>(su, tr, s, n, valid);
}

__global__ void to_internal_kernel(const double* src, int layout, double* dst, const int32_t* slot_elem, int64_t E,
                                   int64_t Ep, int NV, int NS)
{
    const int64_t total = Ep * NV * NS;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t n = i % NS, t = i / NS, s = t % Ep, v = t / Ep;
        const int64_t e = slot_elem[s];
        dst[i] = (e >= 0) ? src[caller_index(layout, e, v, n, E, NV, NS)] : 0.0;
    }
}

__global__ void from_internal_kernel(const double* src, double* dst, int layout, const int32_t* elem_slot, int64_t E,
                                     int64_t Ep, int NV, int NS)
{
    const int64_t total = E * NV * NS;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t e, v, n;
        if (layout == 0) {
            n = i % NS;
            const int64_t t = i / NS;
            v = t % NV;
            e = t / NV;
        } else {
            n = i % NS;
            const int64_t t = i / NS;
            e = t % E;
            v = t / E;
        }
        dst[i] = src[(v * Ep + elem_slot[e]) * NS + n];
    }
}

// max over nb non-negative (or NaN) entries into *out, which must hold +0.0 before the launch: one
// partial max per block, combined with a 64-bit integer atomic max (non-negative doubles order like
// their bit patterns, and every NaN pattern compares above +inf).
__global__ void reduce_max_kernel(const double* bm, int64_t nb, double* out, unsigned int* nonfinite)
{
    __shared__ double s[16];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x)
        m = nan_max(m, bm[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        m = nan_max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0)
        s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = s[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            b = nan_max(b, s[w]);
        atomicMax(reinterpret_cast<unsigned long long*>(out), (unsigned long long)__double_as_longlong(b));
        if (!(b <= 1.0e308))
            atomicOr(nonfinite, 1u);
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// ---------------------------------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------------------------------

template <int D, int P>
static int epb()
{
    return Shape<D, P>::EPB;
}

template <int D, int P>
static hipError_t launch_stage_dp(const StageArgs& args, int flux, hipStream_t st)
{
    const int EPB = Shape<D, P>::EPB;
    const int64_t grid = (args.n_elem + EPB - 1) / EPB;
    if (flux == 0)
        stage_kernel<D, P, 0><<<dim3((unsigned)grid), dim3(256), 0, st>>>(args);
    else
        stage_kernel<D, P, 1><<<dim3((unsigned)grid), dim3(256), 0, st>>>(args);
    return hipGetLastError();
}

template <int D, int P>
static hipError_t launch_trace_dp(const double* u, double* tr, int64_t E, hipStream_t st)
{
    const int EPB = Shape<D, P>::EPB;
    const int64_t grid = (E + EPB - 1) / EPB;
    trace_kernel<D, P><<<dim3((unsigned)grid), dim3(256), 0, st>>>(u, tr, E);
    return hipGetLastError();
}

#define DGX_DISPATCH(D_, P_, CALL)                                                  \
    do {                                                                            \
        if ((D_) == 2) {                                                            \
            switch (P_) {                                                           \
            case 1: return CALL(2, 1);                                              \
            case 2: return CALL(2, 2);                                              \
            case 3: return CALL(2, 3);                                              \
            case 4: return CALL(2, 4);                                              \
            default: break;                                                         \
            }                                                                       \
        } else if ((D_) == 3) {                                                     \
            switch (P_) {                                                           \
            case 1: return CALL(3, 1);                                              \
            case 2: return CALL(3, 2);                                              \
            case 3: return CALL(3, 3);                                              \
            case 4: return CALL(3, 4);                                              \
            default: break;                                                         \
            }                                                                       \
        }                                                                           \
        return hipErrorInvalidValue;                                                \
    } while (0)

bool supported_order(int d, int p)
{
    return (d == 2 || d == 3) && p >= 1 && p <= 4;
}

int64_t stage_grid(int d, int p, int64_t n_elem)
{
    int e = 1;
#define DGX_EPB(DD, PP) (e = epb<DD, PP>(), hipSuccess)
    auto f = [&]() -> hipError_t { DGX_DISPATCH(d, p, DGX_EPB); };
    (void)f();
#undef DGX_EPB
    return (n_elem + e - 1) / e;
}

hipError_t upload_ops(int p, const Ops1D& ops)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_ops), &ops, sizeof(Ops1D), sizeof(Ops1D) * p, hipMemcpyHostToDevice);
}

hipError_t launch_stage(int d, int p, int flux, const StageArgs& args, hipStream_t st)
{
#define DGX_CALL(DD, PP) launch_stage_dp<DD, PP>(args, flux, st)
    DGX_DISPATCH(d, p, DGX_CALL);
#undef DGX_CALL
}

template <int D, int P>
static hipError_t launch_upload_trace_dp(const double* src, int layout, const int32_t* slot_elem, const double* imass,
                                         double* dst, double* tr, int64_t E, int64_t Ep, hipStream_t st)
{
    const int EPB = Shape<D, P>::EPB;
    const int64_t grid = (Ep + EPB - 1) / EPB;
    upload_trace_kernel<D, P><<<dim3((unsigned)grid), dim3(256), 0, st>>>(src, layout, slot_elem, imass, dst, tr, E, Ep);
    return hipGetLastError();
}

hipError_t launch_upload_trace(int d, int p, const double* src, int layout, const int32_t* slot_elem,
                               const double* imass, double* dst, double* tr, int64_t E, int64_t Ep, hipStream_t st)
{
#define DGX_CALL(DD, PP) launch_upload_trace_dp<DD, PP>(src, layout, slot_elem, imass, dst, tr, E, Ep, st)
    DGX_DISPATCH(d, p, DGX_CALL);
#undef DGX_CALL
}

hipError_t launch_trace(int d, int p, const double* u, double* tr, int64_t E, hipStream_t st)
{
#define DGX_CALL(DD, PP) launch_trace_dp<DD, PP>(u, tr, E, st)
    DGX_DISPATCH(d, p, DGX_CALL);
#undef DGX_CALL
}

hipError_t launch_to_internal(const double* src, int layout, double* dst, const int32_t* slot_elem, int64_t E,
                              int64_t Ep, int nv, int ns, hipStream_t st)
{
    to_internal_kernel<<<dim3(2048), dim3(256), 0, st>>>(src, layout, dst, slot_elem, E, Ep, nv, ns);
    return hipGetLastError();
}

hipError_t launch_from_internal(const double* src, double* dst, int layout, const int32_t* elem_slot, int64_t E,
                                int64_t Ep, int nv, int ns, hipStream_t st)
{
    from_internal_kernel<<<dim3(2048), dim3(256), 0, st>>>(src, dst, layout, elem_slot, E, Ep, nv, ns);
    return hipGetLastError();
}

hipError_t launch_reduce_max(const double* bm, int64_t nb, double* out, unsigned int* nonfinite, hipStream_t st)
{
    fill_u64_kernel<<<dim3(1), dim3(64), 0, st>>>(reinterpret_cast<unsigned long long*>(out), 1, 0ull);
    const int64_t blocks = (nb + 1023) / 1024;
    reduce_max_kernel<<<dim3((unsigned)(blocks < 256 ? blocks : 256)), dim3(1024), 0, st>>>(bm, nb, out, nonfinite);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st)
{
    fill_u64_kernel<<<dim3((n + 63) / 64), dim3(64), 0, st>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace dgx

#include "dgx_hex3.inc"
#include "dgx_lin.inc"
#include "dgx_ns.inc"
#include "dgx_ns4.inc"
