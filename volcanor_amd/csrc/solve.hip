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
                                                        double *x, const int *done)
{
    if (done && *done) return;
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

__global__ void axpy_kernel(int n, const double *d, double *x, const int *done)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !*done) x[i] += d[i];
}

// rotor_map_gam (classdef.f90:4181-4196): blades ib <= nbConvect take
// their part of gamVec; with axisymmetry (nbConvect = 1) every blade takes
// blade 1's.
__global__ void map_gam_kernel(const double *x, double *wg, int per_blade, int nb, int nbc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_blade * nb) return;
    const int b = e / per_blade, k = e - b * per_blade;
    if (b < nbc) wg[e] = x[e];
    else if (nbc == 1) wg[e] = x[k];
}

__global__ void solve_state_init_kernel(int *st)
{
    st[0] = 0;  // converged
    st[1] = 0;  // corrections applied
}

// Convergence test of the refinement on the device: ||r||_inf against
// tol ||b||_inf; counts the correction that follows when not converged.
__global__ void refine_check_kernel(int n, const double *r, const double *b, double tol, int *st)
{
    __shared__ double sr[256], sb[256];
    double mr = 0.0, mb = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        mr = fmax(mr, fabs(r[i]));
        mb = fmax(mb, fabs(b[i]));
    }
    sr[threadIdx.x] = mr;
    sb[threadIdx.x] = mb;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + w]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && st[0] == 0) {
        if (sr[0] <= tol * sb[0]) st[0] = 1;
        else st[1] += 1;
    }
}

// The whole solve with refinement in one wave for n <= 64 R: lane l holds
// rows l + 64 r (r < R) of x, b and r; LU and A are staged in LDS (both fit
// the 160 KB of a gfx950 CU up to n = 96); values move between lanes by
// readlane (the index is uniform).  Each row takes the same operations in the same order as in
// lu_solve_kernel, residual_kernel, refine_check_kernel and axpy_kernel, so
// the result is bitwise that of the multi-launch path.  The workgroup has
// SOLVE_STAGE threads so that the staging copy keeps many loads in flight
// (one wave alone, waiting on each load, took 136 us for n = 90); after it
// every wave but the first leaves.
constexpr int SOLVE_STAGE = 512;
constexpr int SOLVE_WAVE_N = 64;   // one row per lane
constexpr int SOLVE_WAVE_N2 = 96;  // two rows per lane

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// element i (uniform over the wave) of a vector held as v[r] at row l + 64 r,
// by readlane (no LDS round trip on the substitution's dependency chain)
template <int R>
__device__ __forceinline__ double wave_get(const double (&v)[R], int i)
{
    const int r = i >> 6;
    double x = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q)
        if (r == q) x = v[q];
    const int lane = __builtin_amdgcn_readfirstlane(i & 63);
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
}

template <int R>
__device__ void wave_lu_solve(int n, const double *sLU, const int *spiv, double (&v)[R])
{
    const int l = threadIdx.x;
    for (int k = 0; k < n; ++k) {
        const int p = spiv[k];
        if (p != k) {
            const double vk = wave_get<R>(v, k), vp = wave_get<R>(v, p);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = l + 64 * q;
                if (row == k) v[q] = vp;
                else if (row == p) v[q] = vk;
            }
        }
    }
    // Substitutions in blocks of SU steps whose LDS reads issue first, off
    // the dependency chain through v (a loop holding readlane is not
    // unrolled by the compiler); reads are unconditional, from clamped rows.
    constexpr int SU = 4;
    int j = 0;
    for (; j + SU <= n; j += SU) {
        double L[SU][R];
#pragma unroll
        for (int u = 0; u < SU; ++u)
#pragma unroll
            for (int q = 0; q < R; ++q) L[u][q] = sLU[min(l + 64 * q, n - 1) + n * (j + u)];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const double xj = wave_get<R>(v, j + u);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = l + 64 * q;
                const double t = v[q] - L[u][q] * xj;
                v[q] = (row > j + u && row < n) ? t : v[q];
            }
        }
    }
    for (; j < n; ++j) {
        const double xj = wave_get<R>(v, j);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = l + 64 * q;
            const double t = v[q] - sLU[min(row, n - 1) + n * j] * xj;
            v[q] = (row > j && row < n) ? t : v[q];
        }
    }
    j = n - 1;
    for (; j - SU + 1 >= 0; j -= SU) {
        double U[SU][R], D[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            D[u] = sLU[(j - u) + n * (j - u)];
#pragma unroll
            for (int q = 0; q < R; ++q) U[u][q] = sLU[min(l + 64 * q, n - 1) + n * (j - u)];
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
#pragma unroll
            for (int q = 0; q < R; ++q)
                if (l + 64 * q == j - u) v[q] /= D[u];
            const double xj = wave_get<R>(v, j - u);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = l + 64 * q;
                const double t = v[q] - U[u][q] * xj;
                v[q] = row < j - u ? t : v[q];
            }
        }
    }
    for (; j >= 0; --j) {
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (l + 64 * q == j) v[q] /= sLU[j + n * j];
        const double xj = wave_get<R>(v, j);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = l + 64 * q;
            const double t = v[q] - sLU[min(row, n - 1) + n * j] * xj;
            v[q] = row < j ? t : v[q];
        }
    }
}

template <int R, int NMAX>
__global__ __launch_bounds__(SOLVE_STAGE) void solve_wave_kernel(int n, const double *A,
                                                                 const double *LU,
                                                        const int *ipiv, const double *b,
                                                        double *x, int max_iter, double tol,
                                                        int *st)
{
    __shared__ double sLU[NMAX * NMAX], sA[NMAX * NMAX];
    __shared__ int spiv[NMAX];
    const int l = threadIdx.x;
#pragma unroll 4
    for (int k = l; k < n * n; k += SOLVE_STAGE) {
        sLU[k] = LU[k];
        sA[k] = A[k];
    }
    for (int k = l; k < n; k += SOLVE_STAGE) spiv[k] = ipiv[k];
    __syncthreads();
    if (l >= 64) return;
    double bl[R], xv[R], rl[R];
    double bm = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int row = l + 64 * q;
        bl[q] = row < n ? b[row] : 0.0;
        bm = fmax(bm, fabs(bl[q]));
        xv[q] = bl[q];
    }
    const double bmax = wave_max(bm);
    wave_lu_solve<R>(n, sLU, spiv, xv);
    int conv = 0, corr = 0;
    for (int it = 0; it < max_iter; ++it) {
#pragma unroll
        for (int q = 0; q < R; ++q) rl[q] = l + 64 * q < n ? bl[q] : 0.0;
        int j = 0;
        for (; j + 4 <= n; j += 4) {
            double a[4][R];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < R; ++q) a[u][q] = sA[min(l + 64 * q, n - 1) + n * (j + u)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double xj = wave_get<R>(xv, j + u);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const double t = rl[q] - a[u][q] * xj;
                    rl[q] = l + 64 * q < n ? t : rl[q];
                }
            }
        }
        for (; j < n; ++j) {
            const double xj = wave_get<R>(xv, j);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double t = rl[q] - sA[min(l + 64 * q, n - 1) + n * j] * xj;
                rl[q] = l + 64 * q < n ? t : rl[q];
            }
        }
        double rm = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) rm = fmax(rm, fabs(rl[q]));
        if (wave_max(rm) <= tol * bmax) {
            conv = 1;
            break;
        }
        ++corr;
        wave_lu_solve<R>(n, sLU, spiv, rl);
#pragma unroll
        for (int q = 0; q < R; ++q) xv[q] += rl[q];
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
        if (l + 64 * q < n) x[l + 64 * q] = xv[q];
    if (l == 0) {
        st[0] = conv;
        st[1] = corr;
    }
}

// x = A^-1 rhs by the LU factors plus up to max_iter refinement steps on
// the FP64 residual.  Everything is enqueued: the convergence test runs on
// the device (later corrections are skipped once it passes), so the only
// synchronisation is the final download.  With a rotor given, the solution
// also becomes its device wing circulation (rotor_map_gam).
int solve_refine(vlm_ctx *ctx, int n, const double *A, const double *LU, const int *ipiv,
                 const double *rhs_h, double *x_h, int max_iter, double tol, int *iters,
                 Rotor *rotor = nullptr, bool deferred = false)
{
    VLM_HIP(ctx->solve_tmp.reserve((size_t)3 * n + 1));
    double *b = ctx->solve_tmp.p, *x = b + n, *r = x + n;
    int *st = reinterpret_cast<int *>(r + n);
    if (rhs_h) {
        VLM_TRY(h2d(ctx, b, rhs_h, sizeof(double) * n));
    } else {  // the right-hand side left on the device by vlm_rotor_rhs
        VLM_HIP(hipMemcpyAsync(b, rotor->rhs.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                               ctx->stream));
    }
    if (n <= SOLVE_WAVE_N) {
        hipLaunchKernelGGL((solve_wave_kernel<1, SOLVE_WAVE_N>), dim3(1), dim3(SOLVE_STAGE), 0,
                           ctx->stream, n, A, LU, ipiv, b, x, max_iter, tol, st);
        max_iter = 0;
    } else if (n <= SOLVE_WAVE_N2) {
        hipLaunchKernelGGL((solve_wave_kernel<2, SOLVE_WAVE_N2>), dim3(1), dim3(SOLVE_STAGE), 0,
                           ctx->stream, n, A, LU, ipiv, b, x, max_iter, tol, st);
        max_iter = 0;
    } else {
        VLM_HIP(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(solve_state_init_kernel, dim3(1), dim3(1), 0, ctx->stream, st);
        hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, LU, ipiv, x,
                           nullptr);
    }
    for (int it = 0; it < max_iter; ++it) {
        hipLaunchKernelGGL(residual_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                           A, x, b, r);
        hipLaunchKernelGGL(refine_check_kernel, dim3(1), dim3(256), 0, ctx->stream, n, r, b, tol,
                           st);
        hipLaunchKernelGGL(lu_solve_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, LU, ipiv, r,
                           st);
        hipLaunchKernelGGL(axpy_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, r, x,
                           st);
    }
    VLM_HIP(hipGetLastError());
    if (rotor) {
        const int per = rotor->d.nc * rotor->d.ns;
        hipLaunchKernelGGL(map_gam_kernel, dim3((per * rotor->d.nb + 255) / 256), dim3(256), 0,
                           ctx->stream, x, rotor->wg.p, per, rotor->d.nb, rotor->d.nbConvect);
        VLM_HIP(hipGetLastError());
    }
    if (deferred) {
        VLM_TRY(d2h_deferred(ctx, x_h, x, sizeof(double) * n));
        if (iters) VLM_TRY(d2h_deferred(ctx, iters, st + 1, sizeof(int)));
        return VLM_OK;
    }
    char *px = nullptr, *ps = nullptr;
    VLM_TRY(pin_take(ctx, sizeof(double) * n, &px));
    VLM_TRY(pin_take(ctx, 2 * sizeof(int), &ps));
    VLM_HIP(hipMemcpyAsync(px, x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    VLM_HIP(hipMemcpyAsync(ps, st, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    VLM_TRY(sync_stream(ctx));
    std::memcpy(x_h, px, sizeof(double) * n);
    if (iters) *iters = reinterpret_cast<const int *>(ps)[1];
    return VLM_OK;
}

}  // namespace
}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_rotor_calc_aic(vlm_ctx *ctx, int ir, const double *cp, const double *ncap, double *Aout)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
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
    VLM_TRY(h2d(ctx, geo.p, cp, sizeof(double) * 3 * n));
    VLM_TRY(h2d(ctx, geo.p + 3 * n, ncap, sizeof(double) * 3 * n));
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
    VLM_TRY(sync_stream(ctx));
    if (info != 0) return fail(ctx, VLM_ERR_SINGULAR, "AIC is singular (zero pivot)");
    return VLM_OK;
}

int vlm_rotor_solve(vlm_ctx *ctx, int ir, const double *rhs, double *gam, int max_iter,
                    double tol, int *iters)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (!rhs || !gam) return fail(ctx, VLM_ERR_ARG, "null buffer");
    if (!ctx->lu[ir]) return fail(ctx, VLM_ERR_STATE, "vlm_rotor_calc_aic not called");
    const int n = ctx->aic_n[ir];
    return solve_refine(ctx, n, ctx->aic[ir], ctx->lu[ir], ctx->piv[ir], rhs, gam, max_iter, tol,
                        iters, ctx->rotors[ir]);
}

int vlm_rotor_solve_deferred(vlm_ctx *ctx, int ir, const double *rhs, double *gam, int max_iter,
                             double tol, int *iters)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (!gam) return fail(ctx, VLM_ERR_ARG, "null buffer");
    if (!ctx->lu[ir]) return fail(ctx, VLM_ERR_STATE, "vlm_rotor_calc_aic not called");
    if (!rhs && !ctx->rotors[ir]->panels_set)
        return fail(ctx, VLM_ERR_STATE, "no right-hand side on the device (vlm_rotor_rhs)");
    const int n = ctx->aic_n[ir];
    return solve_refine(ctx, n, ctx->aic[ir], ctx->lu[ir], ctx->piv[ir], rhs, gam, max_iter, tol,
                        iters, ctx->rotors[ir], true);
}

int vlm_lu_factor(vlm_ctx *ctx, int n, double *A, int *ipiv)
{
    if (!ctx || n <= 0 || !A || !ipiv) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
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
    VLM_TRY(sync_stream(ctx));
    if (info != 0) return fail(ctx, VLM_ERR_SINGULAR, "matrix is singular (zero pivot)");
    return VLM_OK;
}

int vlm_lu_solve_refine(vlm_ctx *ctx, int n, const double *A, const double *LU, const int *ipiv,
                        const double *rhs, double *x, int max_iter, double tol, int *iters)
{
    if (!ctx || n <= 0 || !A || !LU || !ipiv || !rhs || !x) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
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
