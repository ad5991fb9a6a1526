// AIC assembly and the circulation solve.
//
// rotor_calcAIC (classdef.f90:4151-4179) builds AIC(row,col) = vr_vind of
// wing ring col at collocation point row, projected on nCap(row), and
// inverts it once (inv2 -> DGETRF + DGETRI, libMath.f90:48-83); every step
// then forms gamVec = AIC_inv . RHS (matmulAX, libMath.f90:105-122).
// Here the matrix is factorised once (LU with partial pivoting, kept on the
// device) and every step solves by two triangular sweeps plus iterative
// refinement on an FP64 residual RHS - AIC . gam.
//
// Factorisation: lu_factor_device (lu_mfma.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "vlm_internal.h"

namespace vlm {

int lu_factor_device(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev);

namespace {

constexpr double EPS2 = kEps * kEps;

__device__ __forceinline__ void fil64(const double *P, const double *A, const double *B,
                                      double rc, double *v)
{
    // vf_vind in the reference's own operation order (classdef.f90:476-503)
    const double r1x = P[0] - A[0], r1y = P[1] - A[1], r1z = P[2] - A[2];
    const double r2x = P[0] - B[0], r2y = P[1] - B[1], r2z = P[2] - B[2];
    const double r0x = r1x - r2x, r0y = r1y - r2y, r0z = r1z - r2z;
    const double cx = r1y * r2z - r1z * r2y;
    const double cy = r1z * r2x - r1x * r2z;
    const double cz = r1x * r2y - r1y * r2x;
    const double c2 = cx * cx + cy * cy + cz * cz;
    if (c2 > EPS2) {
        const double n1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
        const double n2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
        double u1x = 0, u1y = 0, u1z = 0, u2x = 0, u2y = 0, u2z = 0;
        if (n1 > kEps) {
            u1x = r1x / n1;
            u1y = r1y / n1;
            u1z = r1z / n1;
        }
        if (n2 > kEps) {
            u2x = r2x / n2;
            u2y = r2y / n2;
            u2z = r2z / n2;
        }
        const double dot = r0x * (u1x - u2x) + r0y * (u1y - u2y) + r0z * (u1z - u2z);
        const double q = rc * sqrt(r0x * r0x + r0y * r0y + r0z * r0z);
        const double den = sqrt(q * q * q * q + c2 * c2);
        const double s = kInv4Pi * dot / den;
        v[0] += cx * s;
        v[1] += cy * s;
        v[2] += cz * s;
    }
}

// One lane per (row, col): unit-strength ring col at CP row, dotted with
// nCap(row).  Column-major A.
__global__ void aic_kernel(int n, const double *wc, const double *wrc, const double *cp,
                           const double *ncap, double *A)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int row = (int)(e % n), col = (int)(e / n);
    const double *c = wc + 12 * (size_t)col;
    const double *rc = wrc + 4 * (size_t)col;
    double v[3] = {0.0, 0.0, 0.0};
    for (int f = 0; f < 4; ++f) fil64(cp + 3 * row, c + 3 * f, c + 3 * ((f + 1) & 3), rc[f], v);
    A[e] = v[0] * ncap[3 * row] + v[1] * ncap[3 * row + 1] + v[2] * ncap[3 * row + 2];
}

// Forward/back substitution with the row pivots, one workgroup.
__global__ __launch_bounds__(1024) void lu_solve_kernel(int n, const double *LU, const int *ipiv,
                                                        double *x)
{
    const int tid = threadIdx.x;
    if (tid == 0)
        for (int k = 0; k < n; ++k)
            if (ipiv[k] != k) {
                const double t = x[k];
                x[k] = x[ipiv[k]];
                x[ipiv[k]] = t;
            }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const double xj = x[j];
        for (int i = j + 1 + tid; i < n; i += blockDim.x) x[i] -= LU[i + (size_t)n * j] * xj;
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        if (tid == 0) x[j] /= LU[j + (size_t)n * j];
        __syncthreads();
        const double xj = x[j];
        for (int i = tid; i < j; i += blockDim.x) x[i] -= LU[i + (size_t)n * j] * xj;
        __syncthreads();
    }
}

// r = b - A x (one lane per row, FP64); also r_inf via per-row |r|.
__global__ void residual_kernel(int n, const double *A, const double *x, const double *b,
                                double *r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = b[i];
    for (int j = 0; j < n; ++j) s -= A[i + (size_t)n * j] * x[j];
    r[i] = s;
}

__global__ void axpy_kernel(int n, const double *d, double *x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

int solve_refine(vlm_ctx *ctx, int n, const double *A, const double *LU, const int *ipiv,
                 const double *rhs_h, double *x_h, int max_iter, double tol, int *iters)
{
    VLM_HIP(ctx->solve_tmp.reserve((size_t)3 * n));
    double *b = ctx->solve_tmp.p, *x = b + n, *r = x + n;
    VLM_HIP(hipMemcpyAsync(b, rhs_h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, LU, ipiv, x);
    VLM_HIP(hipGetLastError());
    double bn = 0.0;
    for (int i = 0; i < n; ++i) bn = std::fmax(bn, std::fabs(rhs_h[i]));
    std::vector<double> rh(n);
    int it = 0;
    for (; it < max_iter; ++it) {
        hipLaunchKernelGGL(residual_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                           A, x, b, r);
        VLM_HIP(hipGetLastError());
        VLM_HIP(hipMemcpyAsync(rh.data(), r, sizeof(double) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
        VLM_HIP(hipStreamSynchronize(ctx->stream));
        double rn = 0.0;
        for (int i = 0; i < n; ++i) rn = std::fmax(rn, std::fabs(rh[i]));
        if (rn <= tol * bn) break;
        hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, LU, ipiv, r);
        hipLaunchKernelGGL(axpy_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, r, x);
        VLM_HIP(hipGetLastError());
    }
    if (iters) *iters = it;
    VLM_HIP(hipMemcpyAsync(x_h, x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    return VLM_OK;
}

}  // namespace
}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_rotor_calc_aic(vlm_ctx *ctx, int ir, const double *cp, const double *ncap, double *Aout)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return VLM_ERR_ARG;
    if (!cp || !ncap) return fail(ctx, VLM_ERR_ARG, "null buffer");
    Rotor &r = *ctx->rotors[ir];
    const int n = r.nring;
    if (ctx->aic_n[ir] != n) {
        if (ctx->aic[ir]) (void)hipFree(ctx->aic[ir]);
        if (ctx->lu[ir]) (void)hipFree(ctx->lu[ir]);
        if (ctx->piv[ir]) (void)hipFree(ctx->piv[ir]);
        ctx->aic[ir] = ctx->lu[ir] = nullptr;
        ctx->piv[ir] = nullptr;
        VLM_HIP(hipMalloc(reinterpret_cast<void **>(&ctx->aic[ir]), sizeof(double) * n * n));
        VLM_HIP(hipMalloc(reinterpret_cast<void **>(&ctx->lu[ir]), sizeof(double) * n * n));
        VLM_HIP(hipMalloc(reinterpret_cast<void **>(&ctx->piv[ir]), sizeof(int) * (n + 1)));
        ctx->aic_n[ir] = n;
    }
    DBuf<double> geo;
    VLM_HIP(geo.reserve((size_t)6 * n));
    VLM_HIP(hipMemcpyAsync(geo.p, cp, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipMemcpyAsync(geo.p + 3 * n, ncap, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                           ctx->stream));
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(aic_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, ctx->stream,
                       n, r.wc.p, r.wrc.p, geo.p, geo.p + 3 * n, ctx->aic[ir]);
    VLM_HIP(hipGetLastError());
    VLM_HIP(hipMemcpyAsync(ctx->lu[ir], ctx->aic[ir], sizeof(double) * nn,
                           hipMemcpyDeviceToDevice, ctx->stream));
    VLM_TRY(lu_factor_device(ctx, n, ctx->lu[ir], ctx->piv[ir], ctx->piv[ir] + n));
    int info = 0;
    VLM_HIP(hipMemcpyAsync(&info, ctx->piv[ir] + n, sizeof(int), hipMemcpyDeviceToHost,
                           ctx->stream));
    if (Aout)
        VLM_HIP(hipMemcpyAsync(Aout, ctx->aic[ir], sizeof(double) * nn, hipMemcpyDeviceToHost,
                               ctx->stream));
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    if (info != 0) return fail(ctx, VLM_ERR_SINGULAR, "AIC is singular (zero pivot)");
    return VLM_OK;
}

int vlm_rotor_solve(vlm_ctx *ctx, int ir, const double *rhs, double *gam, int max_iter,
                    double tol, int *iters)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return VLM_ERR_ARG;
    if (!rhs || !gam) return fail(ctx, VLM_ERR_ARG, "null buffer");
    if (!ctx->lu[ir]) return fail(ctx, VLM_ERR_STATE, "vlm_rotor_calc_aic not called");
    const int n = ctx->aic_n[ir];
    VLM_TRY(solve_refine(ctx, n, ctx->aic[ir], ctx->lu[ir], ctx->piv[ir], rhs, gam, max_iter, tol,
                         iters));
    Rotor &r = *ctx->rotors[ir];
    // rotor_map_gam: gamVec order (ic, is, ib) is the wing ring order
    VLM_HIP(hipMemcpyAsync(r.wg.p, gam, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    return VLM_OK;
}

int vlm_lu_factor(vlm_ctx *ctx, int n, double *A, int *ipiv)
{
    if (!ctx || n <= 0 || !A || !ipiv) return VLM_ERR_ARG;
    DBuf<double> dA;
    DBuf<int> dp;
    VLM_HIP(dA.reserve((size_t)n * n));
    VLM_HIP(dp.reserve((size_t)n + 1));
    VLM_HIP(hipMemcpyAsync(dA.p, A, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_TRY(lu_factor_device(ctx, n, dA.p, dp.p, dp.p + n));
    int info = 0;
    VLM_HIP(hipMemcpyAsync(A, dA.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    VLM_HIP(hipMemcpyAsync(ipiv, dp.p, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    VLM_HIP(hipMemcpyAsync(&info, dp.p + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    if (info != 0) return fail(ctx, VLM_ERR_SINGULAR, "matrix is singular (zero pivot)");
    return VLM_OK;
}

int vlm_lu_solve_refine(vlm_ctx *ctx, int n, const double *A, const double *LU, const int *ipiv,
                        const double *rhs, double *x, int max_iter, double tol, int *iters)
{
    if (!ctx || n <= 0 || !A || !LU || !ipiv || !rhs || !x) return VLM_ERR_ARG;
    DBuf<double> dA, dLU;
    DBuf<int> dp;
    VLM_HIP(dA.reserve((size_t)n * n));
    VLM_HIP(dLU.reserve((size_t)n * n));
    VLM_HIP(dp.reserve((size_t)n));
    VLM_HIP(hipMemcpyAsync(dA.p, A, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipMemcpyAsync(dLU.p, LU, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipMemcpyAsync(dp.p, ipiv, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    return solve_refine(ctx, n, dA.p, dLU.p, dp.p, rhs, x, max_iter, tol, iters);
}

}  // extern "C"
