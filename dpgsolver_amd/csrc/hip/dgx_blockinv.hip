// Batched inversion of the volume-diagonal blocks for the block-Jacobi preconditioner of the implicit
// solver: inv_b = (shift_b I - B_b)^-1 in FP32, one 1024-thread workgroup per block.
//
// The n x n block (n <= 320) lives in registers for the whole inversion: the workgroup is a 32 x 32
// grid of threads and thread (ti, tj) holds the R x R elements (ti + 32 r, tj + 32 c), R = ceil(n/32).
// In-place Gauss-Jordan with partial pivoting that never moves rows: the pivot of column k is the
// largest |a_ik| among rows not yet used, found by the 32 threads that hold column k and reduced by
// every wavefront with cross-lane exchanges; the pivot row and column are broadcast through LDS
// (double-buffered, two barriers per column).  The element that ends in storage (i, j) belongs to
// the inverse at (step at which row i was the pivot, pivot row of step j), which the final write
// applies.  Loads are row-contiguous (32 consecutive doubles per thread row).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dgx.h"

namespace dgx {
namespace {

constexpr int TG = 32;

// TIN = float allows out == B (in place): every block reads all its entries before any is written.
template <int R, class TIN>
__global__ __launch_bounds__(1024) void block_inverse_kernel(const TIN* B, const double* __restrict__ shift, int n,
                                                             float* out, int32_t* __restrict__ singular)
{
    __shared__ float s_prow[2][TG * R];
    __shared__ float s_pcol[2][TG * R];
    __shared__ float s_cval[2][TG];
    __shared__ int s_cidx[2][TG];
    __shared__ int s_piv[TG * R];
    __shared__ int s_step[TG * R];

    const int tj = threadIdx.x & (TG - 1);
    const int ti = threadIdx.x >> 5;
    const int64_t blk = blockIdx.x;
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
    if (threadIdx.x < TG * R) {     // defined permutation entries even when a singular block repeats a row
        s_piv[threadIdx.x] = 0;
        s_step[threadIdx.x] = 0;
    }
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

#pragma unroll
    for (int ck = 0; ck < R; ++ck) {
        for (int kk = 0; kk < TG; ++kk) {
            const int k = TG * ck + kk;
            if (k >= n)
                break;
            const int buf = k & 1;
            __syncthreads();

            // every wavefront reduces the 32 candidates (both 32-lane halves hold the same set)
            float v = s_cval[buf][threadIdx.x & 31];
            int p = s_cidx[buf][threadIdx.x & 31];
#pragma unroll
            for (int off = 16; off >= 1; off >>= 1) {
                const float v2 = __shfl_xor(v, off);
                const int p2 = __shfl_xor(p, off);
                if (v2 > v || (v2 == v && p2 < p)) {
                    v = v2;
                    p = p2;
                }
            }
            if (!(v > 0.f))
                flag = k + 1;       // no usable pivot (singular block or non-finite input)
            p = (p < n) ? p : k;    // keeps every index in range when the block is singular

            if (ti == (p & 31)) {
                const int rp = p >> 5;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (r == rp) {
#pragma unroll
                        for (int c = 0; c < R; ++c)
                            s_prow[buf][tj + TG * c] = a[r][c];
                    }
                used |= 1u << rp;
            }
            if (tj == kk) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    s_pcol[buf][ti + TG * r] = a[r][ck];
            }
            if (threadIdx.x == 0) {
                s_piv[k] = p;
                s_step[p] = k;
            }
            __syncthreads();

            const float apk = s_prow[buf][k];
            const float inv = (v > 0.f) ? 1.f / apk : 0.f;
            float pj[R];
#pragma unroll
            for (int c = 0; c < R; ++c)
                pj[c] = s_prow[buf][tj + TG * c];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = ti + TG * r;
                const float fo = s_pcol[buf][i] * inv;
                // pivot row: a_pj (1 - (1 - inv)) = a_pj inv; other rows: a_ij - (a_ik / a_pk) a_pj
                const float f = (i == p) ? 1.f - inv : fo;
#pragma unroll
                for (int c = 0; c < R; ++c)
                    a[r][c] = fmaf(-f, pj[c], a[r][c]);
                if (tj == kk)
                    a[r][ck] = (i == p) ? inv : -fo;
            }

            // pivot candidates of column k + 1 (unused rows only)
            const int kn = k + 1;
            if (kn < n && tj == (kn & 31)) {
                float best = -1.f;
                int bi = 0x7fffffff;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = ti + TG * r;
                    const float val = (kk < TG - 1) ? a[r][ck] : (ck + 1 < R ? a[r][(ck + 1) % R] : 0.f);
                    const float av = fabsf(val);
                    if (i < n && !((used >> r) & 1u) && av > best) {
                        best = av;
                        bi = i;
                    }
                }
                s_cval[buf ^ 1][ti] = best;
                s_cidx[buf ^ 1][ti] = bi;
            }
        }
    }
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
        singular[blk] = flag;
}

// y_b = A_b x_b for n x n FP32 blocks stored column-major (A_b[j][i] at col[b][j * n + i]), FP64 vectors:
// thread i accumulates row i over the columns, so each load instruction of a wavefront reads 64
// consecutive floats of one column; x_b is staged in LDS and read as a broadcast.
__global__ __launch_bounds__(256) void block_apply_kernel(const float* __restrict__ col, int n,
                                                          const double* __restrict__ x, double* __restrict__ y,
                                                          const int32_t* __restrict__ sel)
{
    __shared__ float sx[320];
    const int64_t blk = sel ? sel[blockIdx.x] : blockIdx.x;
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

template <int R, class TIN>
hipError_t launch(int64_t nb, int n, const TIN* B, const double* shift, float* out, int32_t* singular,
                  hipStream_t st)
{
    hipLaunchKernelGGL((block_inverse_kernel<R, TIN>), dim3((unsigned)nb), dim3(1024), 0, st, B, shift, n, out,
                       singular);
    return hipGetLastError();
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
    hipLaunchKernelGGL(dgx::block_apply_kernel, dim3((unsigned)n_blocks), dim3(256), 0, (hipStream_t)stream,
                       blocks_colmajor, n, x, y, (const int32_t*)nullptr);
    return hipGetLastError() == hipSuccess ? DGX_OK : DGX_ERR_HIP;
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
    hipLaunchKernelGGL(dgx::block_apply_kernel, dim3((unsigned)n_sel), dim3(256), 0, (hipStream_t)stream,
                       blocks_colmajor, n, x, y, sel);
    return hipGetLastError() == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}
