// Batched inversion of the volume-diagonal blocks for the block-Jacobi preconditioner of the implicit
// solver: inv_b = (shift_b I - B_b)^-1 in FP32, one 1024-thread workgroup per block.
//
// The n x n block (n <= 320) lives in registers for the whole inversion: the workgroup is a 32 x 32
// grid of threads and thread (ti, tj) holds the R x R elements (ti + 32 r, tj + 32 c), R = ceil(n/32).
// In-place Gauss-Jordan with partial pivoting that never moves rows: the pivot of column k is the
// largest |a_ik| among rows not yet used, found by the 32 threads that hold column k and reduced by
// every wavefront with cross-lane exchanges.  The element that ends in storage (i, j) belongs to the
// inverse at (step at which row i was the pivot, pivot row of step j), which the final write applies.
//
// Blocked by column tiles: the 32 steps of tile c0 (the panel) run on the panel's columns only, with
// the pivot row and column broadcast through LDS (double-buffered, two barriers per step).  The 32
// elementary row transformations E_k (each the identity but in column p_k, the pivot row) multiply to
// T = E_32 ... E_1 = I + W S^T, where S selects the pivot rows p_k and column k of W is the panel's final
// column k minus e_{p_k} (column k of the panel ends as T e_{p_k}, since E_j e_{p_k} = e_{p_k} for
// j < k).  Every other column x then becomes T x = x + W x[p], one rank-32 update from the pivot rows'
// values before the panel (gathered into LDS as the pivots are chosen) and W (written to LDS after the
// panel): 32 FMAs per element with LDS operands and no per-step selects.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dgx.h"

namespace dgx {
namespace {

constexpr int TG = 32;

// Exchange partners of a 64-bit key within each 32-lane half of the wavefront, as two 32-bit halves:
// DPP quad permutes (xor 1, 2), swizzles (xor 4, 8) and the gfx950 row swap (xor 16).  No LDS round trip,
// unlike a shuffle through ds_bpermute.
template <int CTRL>
__device__ __forceinline__ uint64_t key_dpp(uint64_t k)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)k, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(k >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}

template <int PATTERN>
__device__ __forceinline__ uint64_t key_swz(uint64_t k)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)k, PATTERN);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(k >> 32), PATTERN);
    return ((uint64_t)hi << 32) | lo;
}

// the maximum of the key over the lane's 32-lane half
__device__ __forceinline__ uint64_t key_max32(uint64_t k)
{
    uint64_t o = key_dpp<0xB1>(k);      // xor 1
    k = o > k ? o : k;
    o = key_dpp<0x4E>(k);               // xor 2
    k = o > k ? o : k;
    o = key_swz<0x101F>(k);             // xor 4
    k = o > k ? o : k;
    o = key_swz<0x201F>(k);             // xor 8
    k = o > k ? o : k;
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)k, (uint32_t)k, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(k >> 32), (uint32_t)(k >> 32), false, false);
    const uint64_t a = ((uint64_t)hi[0] << 32) | lo[0], b = ((uint64_t)hi[1] << 32) | lo[1];   // xor 16
    return a > b ? a : b;
}

// The 32 (or, for the last tile, n - 32 CK) pivot steps of column tile CK and the update of every other
// column; CK is a template argument so that every index into the register tile is a constant.
struct InvShared {
    float* w;           // [NP][32]
    float* slab;        // [32][NP]
    float (*prow)[TG];
    float* pcol;        // [2][NP]
    float (*cval)[TG];
    int (*cidx)[TG];
    int* piv;
    int* step;
};

template <int R, int CK>
__device__ __forceinline__ void panel(float (&a)[R][R], unsigned& used, int& flag, int n, int ti, int tj,
                                      const InvShared& sm)
{
    constexpr int NP = TG * R;
    float* const s_w = sm.w;
    float* const s_slab = sm.slab;
    float (*const s_prow)[TG] = sm.prow;
    float (*const s_cval)[TG] = sm.cval;
    int (*const s_cidx)[TG] = sm.cidx;
    int* const s_piv = sm.piv;
    int* const s_step = sm.step;
    const int k0 = TG * CK;
    const int nb = (n - k0 < TG) ? n - k0 : TG;      // panel width (>= 1 since R = ceil(n / 32))
    for (int kk = 0; kk < nb; ++kk) {
        const int k = k0 + kk;
        const int buf = k & 1;
        __syncthreads();

        // every wavefront reduces the 32 candidates (both 32-lane halves hold the same set) as one
        // 64-bit maximum of (bits of |a|, ~row): non-negative floats order as their bit patterns, and
        // equal magnitudes go to the smaller row
        const float cv = s_cval[buf][threadIdx.x & 31];
        uint64_t key = ((uint64_t)(cv > 0.f ? __float_as_uint(cv) : 0u) << 32) |
                       (uint32_t)~(uint32_t)s_cidx[buf][threadIdx.x & 31];
        key = key_max32(key);
        // every lane holds the same maximum: read it as a wavefront-uniform (scalar) value, so the row
        // tests below are scalar compares
        const uint32_t klo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)key);
        const uint32_t khi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(key >> 32));
        const float v = __uint_as_float(khi);
        int p = (int)~klo;
        if (!(v > 0.f))
            flag = k + 1;       // no usable pivot (singular block; non-finite values are voted apart)
        p = (p < n) ? p : k;    // keeps every index in range when the block is singular
        const int rp = p >> 5;
        const bool prow_lane = ti == (p & 31);

        if (prow_lane) {
            // the row's value as an OR of masked bit patterns: every index into the register tile stays
            // a constant (a select between tile elements would be turned into a select of addresses)
            unsigned bits = 0u;
#pragma unroll
            for (int r = 0; r < R; ++r)
                bits |= (r == rp) ? __float_as_uint(a[r][CK]) : 0u;
            s_prow[buf][tj] = __uint_as_float(bits);
            used |= 1u << rp;
        }
        if (tj == kk) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                sm.pcol[buf * NP + ti + TG * r] = a[r][CK];
        }
        if (threadIdx.x == 0) {
            s_piv[k] = p;
            s_step[p] = k;
        }
        __syncthreads();

        const float apk = s_prow[buf][kk];
        // 1 / a_pk: hardware reciprocal (within 1 ulp, exact for powers of two) and one Newton step
        float inv = __builtin_amdgcn_rcpf(apk);
        inv = fmaf(fmaf(-apk, inv, 1.f), inv, inv);
        inv = (v > 0.f) ? inv : 0.f;
        const float pj = s_prow[buf][tj];
        const bool pcol_lane = tj == kk;
        // rows other than the pivot row: a_ij - (a_ik / a_pk) a_pj, and -a_ik / a_pk in the pivot column
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = ti + TG * r;
            const float fo = sm.pcol[buf * NP + i] * inv;
            const float t = fmaf(-fo, pj, a[r][CK]);
            a[r][CK] = pcol_lane ? -fo : t;
        }
        if (prow_lane) {
            // the pivot row (its element here is a_pj itself): a_pj (1 - (1 - inv)) = a_pj inv, and inv in
            // the pivot column; only the half-wavefront holding row p runs this
            const float pv = pcol_lane ? inv : fmaf(-(1.f - inv), pj, pj);
#pragma unroll
            for (int r = 0; r < R; ++r)
                a[r][CK] = (r == rp) ? pv : a[r][CK];
        }

        // pivot candidates of column k + 1 within the panel (unused rows only)
        if (kk + 1 < nb && tj == kk + 1) {
            float best = -1.f;
            int bi = 0x7fffffff;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = ti + TG * r;
                const float av = fabsf(a[r][CK]);
                if (i < n && !((used >> r) & 1u) && av > best) {
                    best = av;
                    bi = i;
                }
            }
            s_cval[buf ^ 1][ti] = best;
            s_cidx[buf ^ 1][ti] = bi;
        }
    }

    if (R > 1) {                // a single tile has no column outside the panel
        // the panel's pivot rows, still as before the panel outside it, by the threads that hold them
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int st = s_step[ti + TG * r] - k0;
            if (((used >> r) & 1u) && st >= 0 && st < nb) {
#pragma unroll
                for (int c = 0; c < R; ++c)
                    if (c != CK)
                        s_slab[st * NP + tj + TG * c] = a[r][c];
            }
        }
        // W = panel - e_{p_kk}; columns kk >= nb (past the order) stay zero
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = ti + TG * r;
            const float wv = (tj < nb) ? a[r][CK] - ((i == s_piv[k0 + tj]) ? 1.f : 0.f) : 0.f;
            s_w[i * TG + tj] = wv;
        }
        __syncthreads();

        // every column outside the panel, earlier tiles included: x += W x[p]
        for (int kk = 0; kk < nb; ++kk) {
            float w[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
                w[r] = s_w[(ti + TG * r) * TG + kk];
#pragma unroll
            for (int c = 0; c < R; ++c) {
                if (c == CK)
                    continue;
                const float x = s_slab[kk * NP + tj + TG * c];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    a[r][c] = fmaf(w[r], x, a[r][c]);
            }
        }
    }

    if (CK + 1 < R) {
        // pivot candidates of the next panel's first column
        if (tj == 0) {
            float best = -1.f;
            int bi = 0x7fffffff;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = ti + TG * r;
                const float av = fabsf(a[r][CK + 1 < R ? CK + 1 : 0]);
                if (i < n && !((used >> r) & 1u) && av > best) {
                    best = av;
                    bi = i;
                }
            }
            s_cval[(k0 + TG) & 1][ti] = best;
            s_cidx[(k0 + TG) & 1][ti] = bi;
        }
    }
}

template <int R, int CK>
struct Panels {
    static __device__ __forceinline__ void run(float (&a)[R][R], unsigned& used, int& flag, int n, int ti, int tj,
                                               const InvShared& sm)
    {
        panel<R, CK>(a, used, flag, n, ti, tj, sm);
        Panels<R, CK + 1>::run(a, used, flag, n, ti, tj, sm);
    }
};

template <int R>
struct Panels<R, R> {
    static __device__ __forceinline__ void run(float (&)[R][R], unsigned&, int&, int, int, int, const InvShared&) {}
};

// TIN = float allows out == B (in place): every block reads all its entries before any is written.
template <int R, class TIN>
__global__ __launch_bounds__(1024) void block_inverse_kernel(const TIN* B, const double* __restrict__ shift, int n,
                                                             float* out, int32_t* __restrict__ singular, int64_t base)
{
    constexpr int NP = TG * R;                 // padded order
    __shared__ float s_w[NP * TG];             // W[i][kk] of the current panel
    __shared__ float s_slab[TG * NP];          // row p_kk before the panel, at every column j: [kk][j]
    __shared__ float s_prow[2][TG];            // pivot row within the panel
    __shared__ float s_pcol[2][NP];            // pivot column
    __shared__ float s_cval[2][TG];
    __shared__ int s_cidx[2][TG];
    __shared__ int s_piv[NP];
    __shared__ int s_step[NP];
    __shared__ int s_bad;                      // a non-finite entry in the block or its inverse

    const int tj = threadIdx.x & (TG - 1);
    const int ti = threadIdx.x >> 5;
    const int64_t blk = base + blockIdx.x;
    const TIN* Bb = B + blk * (int64_t)n * n;
    const double sh = shift[blk];

    float a[R][R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = ti + TG * r;
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const int j = tj + TG * c;
            a[r][c] = (i < n && j < n) ? (float)((i == j ? sh : 0.0) - (double)Bb[(int64_t)i * n + j]) : 0.f;
        }
    }
    if (threadIdx.x < NP) {         // defined permutation entries even when a singular block repeats a row
        s_piv[threadIdx.x] = 0;
        s_step[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0)
        s_bad = 0;
    // non-finite input: NaN never wins a pivot search and would otherwise pass unflagged
    bool bad = false;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < R; ++c)
            bad |= !isfinite(a[r][c]);
    unsigned used = 0;              // bit r: row ti + 32 r has been a pivot row
    int flag = 0;

    // pivot candidates of column 0
    if (tj == 0) {
        float best = -1.f;
        int bi = 0x7fffffff;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = ti + TG * r;
            const float v = fabsf(a[r][0]);
            if (i < n && v > best) {
                best = v;
                bi = i;
            }
        }
        s_cval[0][ti] = best;
        s_cidx[0][ti] = bi;
    }

    const InvShared sm{ s_w, s_slab, s_prow, &s_pcol[0][0], s_cval, s_cidx, s_piv, s_step };
    Panels<R, 0>::run(a, used, flag, n, ti, tj, sm);
    // non-finite inverse: a denormal pivot overflows its reciprocal
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < R; ++c)
            bad |= !isfinite(a[r][c]);
    if (bad)
        atomicOr(&s_bad, 1);
    __syncthreads();

    float* ob = out + blk * (int64_t)n * n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = ti + TG * r;
        if (i >= n)
            continue;
        const int64_t row = (int64_t)s_step[i] * n;
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const int j = tj + TG * c;
            if (j < n)
                ob[row + s_piv[j]] = a[r][c];
        }
    }
    if (threadIdx.x == 0)
        singular[blk] = flag ? flag : (s_bad ? n + 1 : 0);
}

// y_b = A_b x_b for n x n FP32 blocks stored column-major (A_b[j][i] at col[b][j * n + i]), FP64 vectors:
// thread i accumulates row i over the columns, so each load instruction of a wavefront reads 64
// consecutive floats of one column; x_b is staged in LDS and read as a broadcast.
__global__ __launch_bounds__(256) void block_apply_kernel(const float* __restrict__ col, int n,
                                                          const double* __restrict__ x, double* __restrict__ y,
                                                          const int32_t* __restrict__ sel, int64_t base)
{
    __shared__ float sx[320];
    const int64_t blk = sel ? sel[base + blockIdx.x] : base + blockIdx.x;
    const float* A = col + blk * (int64_t)n * n;
    for (int j = threadIdx.x; j < n; j += 256)
        sx[j] = (float)x[blk * n + j];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
        float acc[4] = { 0.f, 0.f, 0.f, 0.f };
        int j = 0;
        for (; j + 4 <= n; j += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                acc[u] = fmaf(A[(int64_t)(j + u) * n + i], sx[j + u], acc[u]);
        }
        for (; j < n; ++j)
            acc[0] = fmaf(A[(int64_t)j * n + i], sx[j], acc[0]);
        y[blk * n + i] = (double)((acc[0] + acc[1]) + (acc[2] + acc[3]));
    }
}

// A dispatch holds at most 2^32 - 1 work-items: blocks go out in launches of at most 2^20 workgroups
// (2^30 work-items at 1024 threads), each with the index of its first block.
constexpr int64_t kMaxGroups = int64_t(1) << 20;

template <int R, class TIN>
hipError_t launch(int64_t nb, int n, const TIN* B, const double* shift, float* out, int32_t* singular,
                  hipStream_t st)
{
    for (int64_t base = 0; base < nb; base += kMaxGroups) {
        const int64_t cnt = (nb - base < kMaxGroups) ? nb - base : kMaxGroups;
        hipLaunchKernelGGL((block_inverse_kernel<R, TIN>), dim3((unsigned)cnt), dim3(1024), 0, st, B, shift, n, out,
                           singular, base);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_apply(int64_t nb, int n, const float* col, const double* x, double* y, const int32_t* sel,
                        hipStream_t st)
{
    for (int64_t base = 0; base < nb; base += kMaxGroups) {
        const int64_t cnt = (nb - base < kMaxGroups) ? nb - base : kMaxGroups;
        hipLaunchKernelGGL(block_apply_kernel, dim3((unsigned)cnt), dim3(256), 0, st, col, n, x, y, sel, base);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

template <class TIN>
dgx_status block_inverse(int64_t n_blocks, int n, const TIN* blocks, const double* shift, float* inv,
                         int32_t* singular, void* stream)
{
    if (n_blocks < 0 || n_blocks > 0x7fffffff || n < 1 || n > 320)
        return DGX_ERR_INVALID_ARG;
    if (n_blocks == 0)
        return DGX_OK;
    if (!blocks || !shift || !inv || !singular)
        return DGX_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int R = (n + 31) / 32;
    hipError_t e;
    switch (R) {
    case 1: e = launch<1>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 2: e = launch<2>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 3: e = launch<3>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 4: e = launch<4>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 5: e = launch<5>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 6: e = launch<6>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 7: e = launch<7>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 8: e = launch<8>(n_blocks, n, blocks, shift, inv, singular, st); break;
    case 9: e = launch<9>(n_blocks, n, blocks, shift, inv, singular, st); break;
    default: e = launch<10>(n_blocks, n, blocks, shift, inv, singular, st); break;
    }
    return e == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}

}  // namespace
}  // namespace dgx

extern "C" dgx_status dgx_block_inverse(int64_t n_blocks, int n, const double* blocks, const double* shift,
                                        float* inv, int32_t* singular, void* stream)
{
    return dgx::block_inverse(n_blocks, n, blocks, shift, inv, singular, stream);
}

extern "C" dgx_status dgx_block_inverse_f32(int64_t n_blocks, int n, const float* blocks, const double* shift,
                                            float* inv, int32_t* singular, void* stream)
{
    return dgx::block_inverse(n_blocks, n, blocks, shift, inv, singular, stream);
}

extern "C" dgx_status dgx_block_apply(int64_t n_blocks, int n, const float* blocks_colmajor, const double* x,
                                      double* y, void* stream)
{
    if (n_blocks < 0 || n_blocks > 0x7fffffff || n < 1 || n > 320)
        return DGX_ERR_INVALID_ARG;
    if (n_blocks == 0)
        return DGX_OK;
    if (!blocks_colmajor || !x || !y)
        return DGX_ERR_INVALID_ARG;
    return dgx::launch_apply(n_blocks, n, blocks_colmajor, x, y, nullptr, (hipStream_t)stream) == hipSuccess
               ? DGX_OK
               : DGX_ERR_HIP;
}

extern "C" dgx_status dgx_block_apply_sel(int64_t n_sel, const int32_t* sel, int n, const float* blocks_colmajor,
                                          const double* x, double* y, void* stream)
{
    if (n_sel < 0 || n_sel > 0x7fffffff || n < 1 || n > 320)
        return DGX_ERR_INVALID_ARG;
    if (n_sel == 0)
        return DGX_OK;
    if (!sel || !blocks_colmajor || !x || !y)
        return DGX_ERR_INVALID_ARG;
    return dgx::launch_apply(n_sel, n, blocks_colmajor, x, y, sel, (hipStream_t)stream) == hipSuccess ? DGX_OK
                                                                                                     : DGX_ERR_HIP;
}
