// Device kernels of the explicit DG residual path for gfx950 (CDNA4), FP64 plain form.
//
// One fused kernel evaluates a stage: it reads an element's state and metrics, its own face traces
// are recomputed from the state in LDS, the neighbours' face traces come from a trace buffer the
// previous stage wrote, the numerical flux is evaluated at every face point of the element (both
// sides of a face evaluate the same flux with the left side's convention), then the weak divergence
// (sum-factorized: one 1-D derivative contraction per direction), the face lift, the diagonal inverse
// mass, the SSP-RK3 combination and the Bezier positivity limiter are applied, and the new state and
// its face traces are written.  The state is read once and written once per stage.
//
// Reference path (file:line in DPGSolver):
//   compute_rhs                 src/simulation/solvers/solve.c:133-149
//   volume rlhs                 solvers/compute_volume_rlhs_T.c:145-177, compute_rlhs_T.c:71-110
//   face rlhs                   solvers/compute_face_rlhs_T.c:51-146, 297-359; test_case/boundary/boundary_T.c:42-92
//   Euler flux                  test_case/flux/flux_euler_T.c:71-143, 259-300
//   LF / Roe-Pike               solvers/numerical_flux/numerical_flux_euler_T.c:60-327, 470-778
//   inverse mass (GL/GL diag)   solvers/dg/solve_dg.c:265-282, volume_solver_T.c:115-148
//   SSP-RK3                     solvers/solve_explicit.c:252-301
//   positivity                  solvers/solve.c:187-233, 356-374; solution/solution.c:81-95
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgx {

constexpr double kGamma = 1.4;
constexpr double kGm1 = 0.4;
constexpr double kEpsPhys = 1.0e-13;

// 1-D operators of the GL/GL collocated discretization for orders up to 7 (8 points).
struct Ops1D {
    double w[8];        // GL weights
    double D[64];       // D[q*n + a] = l_a'(x_q)
    double l[2][8];     // l_a(-1), l_a(+1)
    double sb[64];      // nodal -> Bezier (ccSB0), sb[i*n + k]
    double bs[64];      // Bezier -> nodal (ccBS0)
};

template <int D, int P>
struct Shape {
    static constexpr int N1 = P + 1;
    static constexpr int NS = (D == 2) ? N1 * N1 : N1 * N1 * N1;
    static constexpr int NF = (D == 2) ? N1 : N1 * N1;
    static constexpr int NFACE = 2 * D;
    static constexpr int NV = D + 2;
    static constexpr int BLOCK = 256;
    static constexpr int EPB = BLOCK / NS;             // elements per block
    static constexpr int NSP = (NS <= 16) ? 16 : (NS <= 32) ? 32 : (NS <= 64) ? 64 : (NS <= 128) ? 128 : 256;
    static_assert(EPB >= 1, "element too large for one block");
};

// Packed neighbour record: bits 0-27 neighbour element, 28-30 its local face, 31 = 1 when this
// element is the face's right volume (neigh_info[1]).
__host__ __device__ inline uint32_t pack_nbr(uint32_t e, uint32_t lf, uint32_t right)
{
    return (e & 0x0FFFFFFFu) | (lf << 28) | (right << 31);
}

template <int D, int P>
__device__ inline int face_node(int dn, int a, int m)
{
    constexpr int N1 = P + 1;
    if constexpr (D == 2) {
        return dn == 0 ? a + N1 * m : m + N1 * a;
    } else {
        // selects, not branches: with dn known only at run time a branch splits every trace loop into
        // one copy per direction, each issuing all its tile loads before the copies merge
        const int m0 = m % N1, m1 = m / N1;
        const int stride = dn == 0 ? 1 : (dn == 1 ? N1 : N1 * N1);
        const int base = dn == 0 ? N1 * m : (dn == 1 ? m0 + N1 * N1 * m1 : m);
        return base + a * stride;
    }
}

// Face node of the neighbour's trace at this side's face node m under the reference's node
// correspondence nc_fc[ord] (constructor_face_corr, src/element/nodes_correspondence.c:87-160; the same
// gather serves both sides, as compute_face_rlhs_T.c:133-146,312-318 applies it): LINE faces reverse
// for ord 1; QUAD faces (m = a + N1 b) start transposed for ord >= 4, reverse a for ord 1, 3, 6, 7 and b
// for ord 2, 3, 5, 7.
template <int D, int P>
__device__ inline int face_perm(int ord, int m)
{
    constexpr int N1 = P + 1;
    if constexpr (D == 2) {
        return (ord & 1) ? N1 - 1 - m : m;
    } else {
        const int a = m % N1, b = m / N1;
        const int a2 = ((0xCA >> ord) & 1) ? N1 - 1 - a : a;
        const int b2 = ((0xAC >> ord) & 1) ? N1 - 1 - b : b;
        return ord >= 4 ? b2 + N1 * a2 : a2 + N1 * b2;
    }
}

// Neighbour face node of face f of element e at own face node m (identity without a correspondence array).
template <int D, int P>
__device__ inline int nbr_node(const uint8_t* nord, int64_t e, int f, int m)
{
    return nord ? face_perm<D, P>(nord[e * (2 * D) + f], m) : m;
}

template <int D, int P>
__device__ inline double face_weight(const Ops1D& op, int m)
{
    constexpr int N1 = P + 1;
    if constexpr (D == 2)
        return op.w[m];
    else
        return op.w[m % N1] * op.w[m / N1];
}

// Scalar types of the pointwise physics: double, and a forward-mode dual number (value + one tangent)
// used by the linearized residual (JVP/VJP).  Comparisons and branch decisions use the value part,
// as the reference's complex instantiation compares real parts (math_functions_T.c:222-236).
struct Dual {
    double v, d;
};
__device__ inline Dual operator+(Dual a, Dual b) { return { a.v + b.v, a.d + b.d }; }
__device__ inline Dual operator-(Dual a, Dual b) { return { a.v - b.v, a.d - b.d }; }
__device__ inline Dual operator-(Dual a) { return { -a.v, -a.d }; }
__device__ inline Dual operator*(Dual a, Dual b) { return { a.v * b.v, fma(a.d, b.v, a.v * b.d) }; }
__device__ inline Dual operator*(double a, Dual b) { return { a * b.v, a * b.d }; }
__device__ inline Dual operator*(Dual a, double b) { return { a.v * b, a.d * b }; }
__device__ inline Dual operator+(Dual a, double b) { return { a.v + b, a.d }; }
__device__ inline Dual operator-(Dual a, double b) { return { a.v - b, a.d }; }
__device__ inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }

__device__ inline double val(double x) { return x; }
__device__ inline double val(Dual x) { return x.v; }
__device__ inline double zero_like(double) { return 0.0; }
__device__ inline Dual zero_like(Dual) { return { 0.0, 0.0 }; }

// 1/x from the hardware reciprocal estimate refined by two Newton steps (within ~1 ulp of the
// correctly rounded quotient; the operands here are positive densities, sound speeds and face
// Jacobians, so the scaling and special-value handling of a full IEEE division is not needed).
__device__ inline double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ inline Dual rcp_nr(Dual x)
{
    const double r = rcp_nr(x.v);
    return { r, -x.d * r * r };
}
// sqrt(x) for x > 0 in the normal range (squared face Jacobians, density ratios, squared sound speeds):
// the hardware reciprocal square-root estimate, one Goldschmidt step for (g, h) ~ (sqrt x, 1/(2 sqrt x))
// and two Newton corrections of g (within ~1 ulp).  It skips the scaling for denormal / huge operands
// and the 0 / inf handling of the library square root, which these operands never need.
__device__ inline double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
__device__ inline double sqrt_t(double x) { return sqrt(x); }
__device__ inline double sqrt_pos_t(double x) { return sqrt_pos(x); }
__device__ inline Dual sqrt_t(Dual x)
{
    const double s = sqrt(x.v);
    return { s, 0.5 * x.d * rcp_nr(s) };
}
__device__ inline Dual sqrt_pos_t(Dual x) { return sqrt_t(x); }
__device__ inline double neg_if_negative(double x) { return x < 0.0 ? -x : x; }
__device__ inline Dual neg_if_negative(Dual x) { return x.v < 0.0 ? -x : x; }

// Euler flux F_k(eq) contracted with a direction: out[eq] = sum_k n_k F_k(eq)  (flux_euler_T.c:259-300)
template <int D, class T, class N>
__device__ inline void euler_flux_n(const T* w, const N* n, T* out)
{
    const T rho = w[0], E = w[D + 1];
    const T rho_inv = rcp_nr(rho);
    T uvw[3], V2 = zero_like(rho), Vn = zero_like(rho);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        uvw[k] = rho_inv * w[1 + k];
        V2 = V2 + uvw[k] * uvw[k];
        Vn = Vn + n[k] * uvw[k];
    }
    const T p = kGm1 * (E - 0.5 * rho * V2);
    out[0] = rho * Vn;
#pragma unroll
    for (int m = 0; m < D; ++m)
        out[1 + m] = w[1 + m] * Vn + n[m] * p;
    out[D + 1] = (E + p) * Vn;
}

// Lax-Friedrichs: 0.5 [n.(F_l + F_r) + max(|V|+c) (w_l - w_r)]  (numerical_flux_euler_T.c:100-181)
template <int D, class T>
__device__ inline void numflux_lf(const T* wl, const T* wr, const double* n, T* nf)
{
    T lam[2];
    const T* ws[2] = { wl, wr };
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const T* w = ws[s];
        const T rho = w[0], rho_inv = rcp_nr(rho);
        T V2 = zero_like(rho);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T u = w[1 + k] * rho_inv;
            V2 = V2 + u * u;
        }
        const T p = kGm1 * (w[D + 1] - 0.5 * rho * V2);
        lam[s] = sqrt_t(V2) + sqrt_pos_t(kGamma * p * rho_inv);
    }
    const T maxV = (val(lam[0]) > val(lam[1])) ? lam[0] : lam[1];
    T fl[D + 2], fr[D + 2];
    euler_flux_n<D>(wl, n, fl);
    euler_flux_n<D>(wr, n, fr);
#pragma unroll
    for (int eq = 0; eq < D + 2; ++eq)
        nf[eq] = 0.5 * ((fl[eq] + fr[eq]) + maxV * (wl[eq] - wr[eq]));
}

// max_abs_real_T / min_abs_real_T (math_functions_T.c:222-236)
template <class T>
__device__ inline T max_abs_real(T a, T b)
{
    return neg_if_negative((fabs(val(a)) > fabs(val(b))) ? a : b);
}

template <class T>
__device__ inline T min_abs_real(T a, T b)
{
    return neg_if_negative((fabs(val(a)) < fabs(val(b))) ? a : b);
}

// Roe-Pike with the entropy fix of Qu et al. 2015 eq. (35)  (numerical_flux_euler_T.c:470-603)
template <int D, class T>
__device__ inline void numflux_roe(const T* wl, const T* wr, const double* n, T* nf)
{
    const T zero = zero_like(wl[0]);
    T ul[3], ur[3];
    const T rhoL = wl[0], rhoR = wr[0], EL = wl[D + 1], ER = wr[D + 1];
    const T rhoL_inv = rcp_nr(rhoL), rhoR_inv = rcp_nr(rhoR);
    T V2L = zero, V2R = zero, VnL = zero, VnR = zero;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ul[k] = wl[1 + k] * rhoL_inv;
        ur[k] = wr[1 + k] * rhoR_inv;
        V2L = V2L + ul[k] * ul[k];
        V2R = V2R + ur[k] * ur[k];
        VnL = VnL + n[k] * ul[k];
        VnR = VnR + n[k] * ur[k];
    }
    const T pL = kGm1 * (EL - 0.5 * rhoL * V2L), pR = kGm1 * (ER - 0.5 * rhoR * V2R);
    const T HL = (EL + pL) * rhoL_inv, HR = (ER + pR) * rhoR_inv;
    const T r = sqrt_pos_t(rhoR * rhoL_inv), rP1_inv = rcp_nr(r + 1.0);
    const T rho = r * rhoL;
    T u[3], Vn = zero, V2 = zero;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        u[k] = (r * ur[k] + ul[k]) * rP1_inv;
        Vn = Vn + n[k] * u[k];
        V2 = V2 + u[k] * u[k];
    }
    const T H = (r * HR + HL) * rP1_inv;
    const T c2 = kGm1 * (H - 0.5 * V2);
    const T c = sqrt_pos_t(c2), c_inv = rcp_nr(c);
    const T l1 = min_abs_real(VnL - c, Vn - c);
    const T l5 = max_abs_real(VnR + c, Vn + c);
    const T l234 = (val(Vn) > 0.0) ? Vn : -Vn;
    const T drho = rhoR - rhoL, dE = ER - EL, dp = pR - pL, dVn = VnR - VnL;
    const T lc1 = 0.5 * (l5 + l1) - l234, lc2 = 0.5 * (l5 - l1);
    const T disInter1 = lc1 * dp * (c_inv * c_inv) + lc2 * rho * dVn * c_inv;
    const T disInter2 = lc1 * rho * dVn + lc2 * dp * c_inv;
    const T rhoVnL = rhoL * VnL, rhoVnR = rhoR * VnR, pLR = pL + pR;
    nf[0] = 0.5 * ((rhoVnL + rhoVnR) - (l234 * drho + disInter1));
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const T nFk = rhoVnL * ul[k] + rhoVnR * ur[k] + n[k] * pLR;
        const T drhou = rhoR * ur[k] - rhoL * ul[k];
        nf[1 + k] = 0.5 * (nFk - (l234 * drhou + disInter1 * u[k] + disInter2 * n[k]));
    }
    const T nFE = VnL * (EL + pL) + VnR * (ER + pR);
    nf[D + 1] = 0.5 * (nFE - (l234 * dE + disInter1 * H + disInter2 * Vn));
}

template <int D, int FLUX, class T>
__device__ inline void numflux(const T* wl, const T* wr, const double* n, T* nf)
{
    if constexpr (FLUX == 0)
        numflux_lf<D>(wl, wr, n, nf);
    else
        numflux_roe<D>(wl, wr, n, nf);
}

// Explicit time steppers (the values of dgx.h DGX_SOLVER_E_*).
enum { RK_SSP33 = 0, RK_EULER = 1, RK_LS54 = 2 };

// Stage combination of one state value s with its residual r (M^-1 applied) for stage `stage` of the
// scheme; reg is the value's RK register (sol_coef_p), read and written as the scheme needs.
//   SSP-RK3  time_step_ssp_rk_33 (solve_explicit.c:271-293): the register keeps u^n from stage 0 on
//   Euler    time_step_euler (solve_explicit.c:240-244)
//   LS-RK54  time_step_ls_rk_54 (solve_explicit.c:338-342): reg = a reg + dt r, s += b reg; the first
//            stage's a is 0, so the register's old content is not read
__device__ inline double rk_combine(int scheme, int stage, double rk_a, double rk_b, double dt, double s, double r,
                                    double* reg)
{
    if (scheme == RK_EULER)
        return s + dt * r;
    if (scheme == RK_LS54) {
        const double sp = (stage == 0) ? dt * r : rk_a * (*reg) + dt * r;
        *reg = sp;
        return s + rk_b * sp;
    }
    if (stage == 0) {
        *reg = s;
        return s + dt * r;
    }
    if (stage == 1)
        return (1.0 / 4.0) * (3.0 * (*reg) + s + dt * r);
    return (1.0 / 3.0) * ((*reg) + 2.0 * s + 2.0 * dt * r);
}

// Stage gate: a stage kernel reads gate_in, the word only earlier stages write, and writes gate_out
// (the other word) when one of its elements fails the positivity check or when it is itself gated.
// gate_in does not change while the kernel runs, so every wavefront of the launch takes the same
// decision without a barrier; once set, both words stay set until the host clears them.
__device__ inline bool stage_gated(int stage, const unsigned int* gate_in)
{
    return stage >= 0 && *gate_in != 0u;
}

struct StageArgs {
    double* u;              // [NV][E][NS] state, updated in place
    double* up;             // [NV][E][NS] RK register (sol_coef_p)
    const double* met;      // [D*D][E][NS] metrics, component c = i + D*j
    const double* imass;    // [E][NS] 1/(w J)
    const double* tr_in;    // [E][NFACE][NV][NF] traces of the current state
    double* tr_out;         // [E][NFACE][NV][NF] traces of the updated state
    const uint32_t* nbr;    // [E][NFACE] packed neighbour records
    const uint8_t* nord;    // [E][NFACE] face node correspondence case (face_perm), null when all aligned
    double* rhs_out;        // [NV][E][NS] residual output (residual mode only)
    double* blockmax;       // [grid] max |rhs| of each block (NaN propagates)
    unsigned long long* bad_elem;  // first element that failed the positivity check
    int64_t n_elem;
    double dt;
    int stage;              // stage of the time stepper; -1: residual only
    int apply_minv;         // residual mode: 1 = M^-1 applied
    int positivity;
    int scheme;             // RK_*
    double rk_a, rk_b;      // LS-RK54 coefficients of this stage
    const unsigned int* gate_in;   // stage gate (stage_gated)
    unsigned int* gate_out;
    // Navier-Stokes only
    const double* gtr;      // [E][NFACE][D*NV][NF] face traces of the volume gradient of the current state
                            // (HEX p=4: [E][NFACE][NV][NF] traces of n.f2(w, grad w), outward unit n)
    const double* lam;      // [E][NFACE][NF] lift trace factor w_f(m) sum_k l_f(k)^2 / (w(k) w_f(m) J)(k, m)
    const double* fnv;      // HEX p=4: [E][NFACE][3][NF] |J_f| n of the own outward normal (set-up)
    double* gtr_out;        // HEX p=4: the n.f2 traces of the state a stage writes
    double mu, Pr, eta;     // viscosity, Prandtl number, BR2 penalty
    int mixed;              // HEX p=4 Navier-Stokes: precision form of the residual's contractions
};

// sum_k c[k] t[base + k st] (k < N) in the precision form: the line's first value t0 goes through FP64
// with the FP64 sum of the coefficients, the FP64 differences t_k - t0 are contracted as FP32 hi/lo parts
// with FP32 coefficients and FP32 accumulation of each part, the two partial sums added in FP64.  The
// FP32 round-off then scales with the variation of the data along the line, not with its size.
template <int N>
__device__ inline double dot_diff_hl(const double* c, const double* t, int base, int st)
{
    const double t0 = t[base];
    double csum = c[0];
    float hi = 0.0f, lo = 0.0f;
#pragma unroll
    for (int k = 1; k < N; ++k) {
        const double d = t[base + k * st] - t0;
        const float h = (float)d;
        const float l = (float)(d - (double)h);
        const float ck = (float)c[k];
        hi = fmaf(ck, h, hi);
        lo = fmaf(ck, l, lo);
        csum += c[k];
    }
    return csum * t0 + ((double)hi + (double)lo);
}

// Second-order (viscous) part of the Navier-Stokes flux at one point, negated as the reference stores
// it (flux_navier_stokes_T.c:1005-1040): f2[i + D*eq] = -F_v(i, eq).  g[v + NV*j] = d(var v)/dx_j.
// tau (compute_tau :778-799), velocity gradient (compute_duvw :708-720), internal-energy gradient
// (compute_dTs :904-925), energy flux gamma/Pr mu grad e + u.tau.
template <int D>
__device__ inline void viscous_flux(const double* w, const double* g, double mu, double Pr, double* f2)
{
    constexpr int NV = D + 2;
    const double rho_inv = rcp_nr(w[0]);
    double uvw[D], duvw[D][D], dTs[D];
#pragma unroll
    for (int k = 0; k < D; ++k)
        uvw[k] = rho_inv * w[1 + k];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            duvw[i][j] = rho_inv * (g[(1 + i) + NV * j] - g[NV * j] * uvw[i]);
    double divV = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k)
        divV += duvw[k][k];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double dE_o_rho = rho_inv * rho_inv * (g[(D + 1) + NV * j] * w[0] - w[D + 1] * g[NV * j]);
        double dV2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k)
            dV2 += uvw[k] * duvw[k][j];
        dTs[j] = dE_o_rho - dV2;
    }
    const double cE = kGamma / Pr * mu;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        f2[i] = 0.0;
        double ut = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            const double tau = (i == m) ? mu * 2.0 * (duvw[i][i] - divV * (1.0 / 3.0)) : mu * (duvw[m][i] + duvw[i][m]);
            f2[i + D * (1 + m)] = -tau;
            ut += uvw[m] * tau;
        }
        f2[i + D * (D + 1)] = -(cE * dTs[i] + ut);
    }
}

// The same flux projected on a direction nh: fn[eq] = sum_i nh_i f2[i + D*eq], formed without the full
// (D, NV) array (tau is symmetric, so sum_i nh_i (u.tau)_i = sum_m u_m (nh.tau)_m).
template <int D>
__device__ inline void viscous_flux_n(const double* w, const double* g, double mu, double Pr, const double* nh,
                                      double* fn)
{
    constexpr int NV = D + 2;
    const double rho_inv = rcp_nr(w[0]);
    double uvw[D], duvw[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k)
        uvw[k] = rho_inv * w[1 + k];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            duvw[i][j] = rho_inv * (g[(1 + i) + NV * j] - g[NV * j] * uvw[i]);
    double divV = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k)
        divV += duvw[k][k];
    double ndT = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double dE_o_rho = rho_inv * rho_inv * (g[(D + 1) + NV * j] * w[0] - w[D + 1] * g[NV * j]);
        double dV2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k)
            dV2 += uvw[k] * duvw[k][j];
        ndT += nh[j] * (dE_o_rho - dV2);
    }
    fn[0] = 0.0;
    double ut = 0.0;
#pragma unroll
    for (int m = 0; m < D; ++m) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double tau = (i == m) ? mu * 2.0 * (duvw[i][i] - divV * (1.0 / 3.0)) : mu * (duvw[m][i] + duvw[i][m]);
            t += nh[i] * tau;
        }
        fn[1 + m] = -t;
        ut += uvw[m] * t;
    }
    fn[D + 1] = -(kGamma / Pr * mu * ndT + ut);
}

// Arguments of the HEX p=3 group kernel; arrays are in slot order (slot = group*8 + wave).
struct Hex3Args {
    const double* u;            // [NV][Ep][64] state the stage starts from (u^n, u^(1) or u^(2))
    const double* up;           // [NV][Ep][64] u^n, read by stages 1 and 2
    double* u_out;              // [NV][Ep][64] stage result: stage 0 -> the register buffer, 1 -> in place,
                                // 2 -> the u^n buffer (in place of u^n, which it no longer needs)
    const double* met;          // [9][Ep][64]
    const double* imass;        // [Ep][64]
    const double* tr_in;        // [Ep][6][NV][16]
    double* tr_out;             // [Ep][6][NV][16] (faces read by another group only)
    const uint32_t* nbr;        // [Ep][6] packed neighbour records (slot of the neighbour)
    const int32_t* slot_elem;   // [Ep] element of the slot, -1 for padding
    const uint64_t* face_item;  // [G][48] faces whose flux the group evaluates: (slot-in-group*6 + f) << 32 | the
                                // face's neighbour record; unused entries all ones, used ones first
    double* rhs_out;            // [NV][Ep][64] residual mode output
    double* blockmax;           // [Ep] max |rhs| per element
    unsigned long long* bad_elem;
    double* reg;                // [NV][Ep][64] RK register of the Euler / LS-RK54 steppers (state in place)
    int64_t n_slots;
    double dt;
    int stage;
    int apply_minv;
    int positivity;
    int scheme;                 // RK_*: SSP-RK3 rotates u / up / u_out as above, the others update u in place
    double rk_a, rk_b;          // LS-RK54 coefficients of this stage
    const unsigned int* gate_in;   // stage gate (stage_gated)
    unsigned int* gate_out;
};

// Linearized residual (J w or J^T v) at the state u; arrays in the internal [v][slot][node] order.
struct LinArgs {
    const double* u;        // [NV][E][NS] linearization point
    const double* x;        // [NV][E][NS] tangent w (JVP) or cotangent v (VJP; already scaled by 1/(wJ) when needed)
    const double* met;      // [D*D][E][NS]
    const double* imass;    // [E][NS]
    const double* tr_u;     // [E][NFACE][NV][NF] traces of u
    const double* tr_x;     // [E][NFACE][NV][NF] traces of x
    const uint32_t* nbr;    // [E][NFACE]
    const uint8_t* nord;    // [E][NFACE] face node correspondence case (face_perm), null when all aligned
    const int32_t* slot_elem;   // [E] element of each slot (-1 = padding) or null when every slot is used
    double* out;            // [NV][E][NS]
    int64_t n_elem;
    int apply_minv;         // JVP: scale the result by 1/(wJ)
};

}  // namespace dgx
