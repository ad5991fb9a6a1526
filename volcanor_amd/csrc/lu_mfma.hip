// Dense LU factorisation with partial pivoting (FP64) for the AIC solve.
//
// Replaces DGETRF (inv2, libMath.f90:48-83; the native Doolittle routine of
// libMath_noBLAS.f90 in no-BLAS builds).  Column-major, 0-based pivots,
// info = first zero pivot (1-based) or 0.  Pivot choice follows IDAMAX:
// the first row of largest magnitude.
//
// n <= LU_SMALL (the AIC of one rotor): one workgroup, unblocked.
// Larger n: right-looking blocked LU with panels of NB columns.
//   1. panel: one workgroup factorises A[k:n, k:k+NB] column by column and
//      applies each row swap across the full row width;
//   2. U12 = L11^-1 A12 (unit lower triangular solve, one lane per column);
//   3. A22 -= L21 U12 on the matrix cores: 64x64 output tiles, four waves
//      of 2x2 v_mfma_f64_16x16x4f64 blocks, operands staged in LDS.
//      Lane layout of the 16x16x4 FP64 MFMA (checked on gfx950): lane l
//      supplies A[l%16][l/16] and B[l/16][l%16]; result register i of lane
//      l is D[4i + l/16][l%16].
#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int LU_THREADS = 1024;
constexpr int LU_WAVES = LU_THREADS / 64;

// IDAMAX order: larger magnitude wins, ties go to the lower row.
__device__ __forceinline__ void better(double &bv, int &bi, double v, int i)
{
    if (v > bv || (v == bv && i < bi)) {
        bv = v;
        bi = i;
    }
}

// Unblocked right-looking LU of columns [k0, k1) over rows [k0, n).  Row
// swaps are applied to columns [s0, s1) (the whole matrix for the
// one-workgroup path, the panel only in the blocked path), the rank-1
// updates to columns (k, k1).
__global__ __launch_bounds__(LU_THREADS) void lu_panel_kernel(int n, double *A, int *ipiv,
                                                              int *info, int k0, int k1, int s0,
                                                              int s1)
{
    __shared__ double sv[LU_WAVES];
    __shared__ int si[LU_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t ld = (size_t)n;
    if (*info != 0) {  // an earlier panel hit a zero pivot: keep pivots in range
        for (int k = k0 + tid; k < k1; k += blockDim.x) ipiv[k] = k;
        return;
    }
    // 2-D split of the trailing update: 64 consecutive rows per wave
    // (coalesced), LU_WAVES column groups.
    for (int k = k0; k < k1; ++k) {
        double bv = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += LU_THREADS) better(bv, bi, fabs(A[i + ld * k]), i);
        for (int off = 32; off > 0; off >>= 1)
            better(bv, bi, __shfl_xor(bv, off), __shfl_xor(bi, off));
        if (lane == 0) {
            sv[w] = bv;
            si[w] = bi;
        }
        __syncthreads();
        bv = sv[0];
        bi = si[0];
        for (int q = 1; q < LU_WAVES; ++q) better(bv, bi, sv[q], si[q]);
        const int p = bi;
        if (tid == 0) ipiv[k] = p;
        if (bv == 0.0) {
            if (tid == 0 && *info == 0) *info = k + 1;
            for (int j = k + 1 + tid; j < k1; j += LU_THREADS) ipiv[j] = j;
            return;
        }
        if (p != k)
            for (int j = s0 + tid; j < s1; j += LU_THREADS) {
                const double t = A[k + ld * j];
                A[k + ld * j] = A[p + ld * j];
                A[p + ld * j] = t;
            }
        __syncthreads();
        const double inv = 1.0 / A[k + ld * k];
        for (int j = k + 1 + w; j < k1; j += LU_WAVES) {
            const double u = A[k + ld * j];
            for (int i = k + 1 + lane; i < n; i += 64) A[i + ld * j] -= (A[i + ld * k] * inv) * u;
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < n; i += LU_THREADS) A[i + ld * k] *= inv;
    }
}

// Row swaps of the panel [k0, k1) applied, in order, to the columns outside
// it; one lane per column.
__global__ void lu_swap_kernel(int n, double *A, const int *ipiv, int k0, int k1)
{
    const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = c0 < k0 ? c0 : c0 + (k1 - k0);
    if (c >= n) return;
    double *col = A + (size_t)n * c;
    for (int k = k0; k < k1; ++k) {
        const int p = ipiv[k];
        if (p != k) {
            const double t = col[k];
            col[k] = col[p];
            col[p] = t;
        }
    }
}

__global__ void zero_info_kernel(int *info) { *info = 0; }

constexpr int LU_SMALL = 256;
constexpr int NB = 32;       // panel width
constexpr int GT = 64;       // output tile of the trailing update
constexpr int AS_LD = NB + 1;
constexpr int BS_LD = GT + 1;

// U12 <- L11^-1 U12 for the columns [k+jb, n): rows k..k+jb of those columns
// hold A12 on entry; L11 is the unit lower triangle of the panel.  Each
// half-wave (32 lanes, one per row) solves one column by forward
// substitution, the solved entry broadcast by a lane shuffle.
constexpr int TRSM_COLS = 8;  // columns per 256-thread block
__global__ __launch_bounds__(256) void lu_trsm_kernel(int n, double *A, int k, int jb)
{
    __shared__ double L[NB][NB + 1];
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, m = e / NB;
        L[i][m] = (i < jb && m < jb) ? A[(size_t)(k + i) + ld * (k + m)] : 0.0;
    }
    __syncthreads();
    const int row = tid & 31;
    const int c = k + jb + blockIdx.x * TRSM_COLS + (tid >> 5);
    const bool live = c < n && row < jb;
    double x = live ? A[(size_t)(k + row) + ld * c] : 0.0;
    const int base = (tid & 63) & ~31;  // first lane of this half-wave
    for (int i = 0; i < jb - 1; ++i) {
        const double xi = __shfl(x, base + i);
        if (row > i) x -= L[row][i] * xi;
    }
    if (live) A[(size_t)(k + row) + ld * c] = x;
}

typedef double dx4 __attribute__((ext_vector_type(4)));

// A22 -= L21 U12 for one 64x64 tile of the trailing matrix.
__global__ __launch_bounds__(256) void lu_gemm_mfma_kernel(int n, double *A, int k, int jb)
{
    __shared__ double As[GT * AS_LD];   // L21 tile: As[r*AS_LD + kk]
    __shared__ double Bs[NB * BS_LD];   // U12 tile: Bs[kk*BS_LD + c]
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x;
    const int t0 = k + jb;
    const int R0 = t0 + blockIdx.y * GT, C0 = t0 + blockIdx.x * GT;
    for (int e = tid; e < GT * NB; e += 256) {
        const int r = e % GT, kk = e / GT;
        const int row = R0 + r;
        As[r * AS_LD + kk] = (row < n && kk < jb) ? A[(size_t)row + ld * (k + kk)] : 0.0;
    }
    for (int e = tid; e < GT * NB; e += 256) {
        const int kk = e % NB, c = e / NB;
        const int col = C0 + c;
        Bs[kk * BS_LD + c] = (col < n && kk < jb) ? A[(size_t)(k + kk) + ld * col] : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6;
    const int rw = (w & 1) * 32, cw = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    dx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dx4{0.0, 0.0, 0.0, 0.0};
    const int kend = (jb + 3) & ~3;
    for (int kk = 0; kk < kend; kk += 4) {
        const double a0 = As[(rw + li) * AS_LD + kk + lk];
        const double a1 = As[(rw + 16 + li) * AS_LD + kk + lk];
        const double b0 = Bs[(kk + lk) * BS_LD + cw + li];
        const double b1 = Bs[(kk + lk) * BS_LD + cw + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = C0 + cw + 16 * b + li;
            if (col >= n) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = R0 + rw + 16 * a + 4 * i + lk;
                if (row < n) A[(size_t)row + ld * col] -= acc[a][b][i];
            }
        }
}

}  // namespace

namespace {

int lu_launch(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info_dev);
    if (n <= LU_SMALL) {
        hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n, A,
                           ipiv, info_dev, 0, n, 0, n);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    for (int k = 0; k < n; k += NB) {
        const int jb = n - k < NB ? n - k : NB;
        hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n, A,
                           ipiv, info_dev, k, k + jb, k, k + jb);
        const int others = n - jb;
        if (others > 0)
            hipLaunchKernelGGL(lu_swap_kernel, dim3((others + 255) / 256), dim3(256), 0,
                               ctx->stream, n, A, ipiv, k, k + jb);
        const int rest = n - k - jb;
        if (rest > 0) {
            hipLaunchKernelGGL(lu_trsm_kernel, dim3((rest + TRSM_COLS - 1) / TRSM_COLS),
                               dim3(256), 0, ctx->stream, n, A, k, jb);
            const int tiles = (rest + GT - 1) / GT;
            hipLaunchKernelGGL(lu_gemm_mfma_kernel, dim3(tiles, tiles), dim3(256), 0,
                               ctx->stream, n, A, k, jb);
        }
        VLM_HIP(hipGetLastError());
    }
    return VLM_OK;
}

}  // namespace

int lu_factor_device(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    if (!ctx->profile) return lu_launch(ctx, n, A, ipiv, info_dev);
    Timer tm;
    VLM_HIP(hipEventCreate(&tm.a));
    VLM_HIP(hipEventCreate(&tm.b));
    tm.pairs = 2.0 * (double)n * n * n / 3.0;
    VLM_HIP(hipEventRecord(tm.a, ctx->stream));
    const int rc = lu_launch(ctx, n, A, ipiv, info_dev);
    VLM_HIP(hipEventRecord(tm.b, ctx->stream));
    ctx->lu_timers.push_back(tm);
    return rc;
}

}  // namespace vlm
