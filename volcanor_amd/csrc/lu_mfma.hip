// Dense LU factorisation with partial pivoting (FP64) for the AIC solve.
//
// Replaces DGETRF (inv2, libMath.f90:48-83; the native Doolittle routine of
// libMath_noBLAS.f90 in no-BLAS builds).  Column-major, 0-based pivots,
// info = first zero pivot (1-based) or 0.
//
// The matrix is factorised by a single workgroup (unblocked, right-looking),
// which covers the AIC of one rotor (N <= a few hundred).
#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int LU_THREADS = 1024;

__device__ __forceinline__ void argmax_block(double v, int i, double *sv, int *si, double &bv,
                                             int &bi)
{
    const int tid = threadIdx.x;
    sv[tid] = v;
    si[tid] = i;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double a = sv[tid], b = sv[tid + s];
            const int ia = si[tid], ib = si[tid + s];
            if (b > a || (b == a && ib < ia)) {
                sv[tid] = b;
                si[tid] = ib;
            }
        }
        __syncthreads();
    }
    bv = sv[0];
    bi = si[0];
    __syncthreads();
}

// Unblocked right-looking LU of columns [k0, k1) of an m-row panel whose
// rows start at r0; row swaps are applied to the full width [0, ncols).
__global__ __launch_bounds__(LU_THREADS) void lu_panel_kernel(int n, double *A, int *ipiv,
                                                              int *info, int k0, int k1,
                                                              int ncols_update)
{
    __shared__ double sv[LU_THREADS];
    __shared__ int si[LU_THREADS];
    const int tid = threadIdx.x;
    const size_t ld = (size_t)n;
    for (int k = k0; k < k1; ++k) {
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += blockDim.x) {
            const double a = fabs(A[i + ld * k]);
            if (a > best) {
                best = a;
                bi = i;
            }
        }
        double bv;
        int p;
        argmax_block(best, bi, sv, si, bv, p);
        if (tid == 0) ipiv[k] = p;
        if (bv == 0.0) {
            if (tid == 0 && *info == 0) *info = k + 1;
            return;
        }
        if (p != k)
            for (int j = tid; j < n; j += blockDim.x) {
                const double t = A[k + ld * j];
                A[k + ld * j] = A[p + ld * j];
                A[p + ld * j] = t;
            }
        __syncthreads();
        const double inv = 1.0 / A[k + ld * k];
        for (int i = k + 1 + tid; i < n; i += blockDim.x) A[i + ld * k] *= inv;
        __syncthreads();
        // rank-1 update of the remaining panel columns (and, for the
        // unblocked path, the whole trailing matrix)
        const int jend = ncols_update;
        const int rows = n - k - 1;
        const int cols = jend - k - 1;
        const int64_t tot = (int64_t)rows * cols;
        for (int64_t e = tid; e < tot; e += blockDim.x) {
            const int i = k + 1 + (int)(e % rows);
            const int j = k + 1 + (int)(e / rows);
            A[i + ld * j] -= A[i + ld * k] * A[k + ld * j];
        }
        __syncthreads();
    }
}

__global__ void zero_info_kernel(int *info) { *info = 0; }

}  // namespace

int lu_factor_device(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info_dev);
    hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n, A, ipiv,
                       info_dev, 0, n, n);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // namespace vlm
