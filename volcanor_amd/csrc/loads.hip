// Right-hand side and loads of a rotor on the device: the collocation-point
// velocities, the right-hand side handed to the circulation solve, and the
// sectional and rotor loads of forceCalcSwitch = 0.  With the Biot-Savart
// terms already on the device, a step runs from the wake sums through the
// solve to CT without a host round trip; the host only uploads the panel
// geometry after each rigid motion and reads the results at vlm_wait.
#include <cmath>
#include <cstring>

#include "vlm_internal.h"

namespace vlm {
namespace {

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// Rigid-motion data of the right-hand side: velBody, omegaBody, cgCoords,
// omegaSlow, and the shaft and hub of vlm_rotor_set_axis.
struct Kin {
    double vb[3], ob[3], cg[3], osl, shaft[3], hub[3];
};

struct LoadPar {
    double density, dt, sgn, inv;  // sign(1, Omega) and -sign(1, Omega)
    int nb, nbc, nc, ns, span_lift;
    int64_t m;                     // panels (stride of the SoA load velocities)
};

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ double norm3(const double *a) { return sqrt(dot3(a, a)); }
// unit vector, zero below eps (libMath unitVec)
__device__ __forceinline__ void unit3(const double *a, double *u)
{
    const double n = norm3(a);
    for (int i = 0; i < 3; ++i) u[i] = n > kEps ? a[i] / n : 0.0;
}
// projection of a on d, zero for |d|^2 <= eps (libMath projVec)
__device__ __forceinline__ void proj3(const double *a, const double *d, double *p)
{
    const double n2 = dot3(d, d);
    const double t = dot3(a, d);
    for (int i = 0; i < 3; ++i) p[i] = n2 > kEps ? t * d[i] / n2 : 0.0;
}
// a less its projection on d (libMath noProjVec)
__device__ __forceinline__ void noproj3(const double *a, const double *d, double *p)
{
    const double n2 = dot3(d, d);
    const double t = dot3(a, d);
    for (int i = 0; i < 3; ++i) p[i] = n2 > kEps ? a[i] - t * d[i] / n2 : a[i];
}
// angle between two vectors by the tangent formula (libMath.f90:224-236)
__device__ __forceinline__ double angle_tan(const double *a, const double *b)
{
    const double d = dot3(a, b);
    return atan2(sqrt(dot3(a, a) * dot3(b, b) - d * d), d);
}

// Panel records (PANEL_N doubles): CP 0-2, nCap 3-5, tauCapChord 6-8,
// tauCapSpan 9-11, panelArea 12, meanChord 13, meanSpan 14.
// Section records (SECTION_N): secCP 0-2, corner 1 of the section's first
// panel 3-5, secTauCapChord 6-8, secTauCapSpan 9-11, secNormalVec 12-14,
// secArea 15.
// Blade records (BLADE_N): xAxisAziFlap 0-2, yAxisAziFlap 3-5, zAxisAziFlap 6-8.

__global__ void cp_soa_kernel(const double *pan, int64_t n, double *cps)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int i = 0; i < 3; ++i) cps[i * n + e] = pan[PANEL_N * e + i];
}

// main.f90:528-591: velCPm of the rigid motion, velCP = velCPm + the
// induced velocity (SoA, stride n), RHS = -(velCP . nCap), on the
// convected blades.
__global__ void rhs_kernel(const double *pan, const double *vind, int64_t n, int64_t nconv,
                           Kin k, double *vcp, double *rhs)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nconv) return;
    const double *P = pan + PANEL_N * e;
    double r1[3], r2[3], w[3], c1[3], c2[3];
    for (int i = 0; i < 3; ++i) {
        r1[i] = P[i] - k.cg[i];
        r2[i] = P[i] - k.hub[i];
        w[i] = k.osl * k.shaft[i];
    }
    cross3(k.ob, r1, c1);
    cross3(w, r2, c2);
    double v[3];
    for (int i = 0; i < 3; ++i) {
        const double vm = -1.0 * k.vb[i] - c1[i] - c2[i];
        vcp[6 * e + 3 + i] = vm;
        v[i] = vm + vind[i * n + e];
        vcp[6 * e + i] = v[i];
    }
    rhs[e] = -dot3(v, P + 3);
}

// Axisymmetry: the other blades take blade 1's right-hand side
// (main.f90:574-581).
__global__ void rhs_axisym_kernel(double *rhs, int64_t per, int64_t n)
{
    const int64_t e = per + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) rhs[e] = rhs[e % per];
}

// rotor_mirrorVelCP (classdef.f90:4997-5022): velCP and velCPm of every
// panel from the source, component ax negated.
__global__ void mirror_velcp_kernel(const double *src, double *dst, int64_t n, int ax)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 6 * n) return;
    const int c = (int)(e % 6);
    dst[e] = (c % 3 == ax) ? -src[e] : src[e];
}

// Quadratic least-squares fit at xq from the sums of the normal equations
// (libMath lsq2, 577-605): Gaussian elimination with partial pivoting.
__device__ double lsq2_solve(double xq, const double *s, double b0, double b1, double b2)
{
    double A[3][3] = {{s[0], s[1], s[2]}, {s[1], s[2], s[3]}, {s[2], s[3], s[4]}};
    double b[3] = {b0, b1, b2}, c[3];
    for (int k = 0; k < 3; ++k) {
        int p = k;
        for (int i = k + 1; i < 3; ++i)
            if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < 3; ++j) {
                const double t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < 3; ++i) {
            const double f = A[i][k] / A[k][k];
            for (int j = k; j < 3; ++j) A[i][j] = A[i][j] - f * A[k][j];
            b[i] = b[i] - f * b[k];
        }
    }
    for (int i = 2; i >= 0; --i) {
        c[i] = b[i];
        for (int j = i + 1; j < 3; ++j) c[i] = c[i] - A[i][j] * c[j];
        c[i] = c[i] / A[i][i];
    }
    return c[0] + c[1] * xq + c[2] * xq * xq;
}

// Loads of one rotor, forceCalcSwitch = 0 (main.f90:632-668 with
// rotor_calc_secAlpha / blade_calc_secAlpha, classdef.f90:4766-4784 and
// 2234-2265, blade_calc_secChordwiseResVel 2197-2232, rotor_dirLiftDrag,
// blade_calc_force 1704-1896 and rotor_calc_force 4607-4671).  One
// workgroup: one thread per (convected blade, section) computes the
// section's resultant velocity, angles, lift direction and panel pressure
// loop in the reference's order, then one thread per blade sums its
// sections and thread 0 the rotor.  Section results go to out as
// LOADS_SEC_N-double records, the rotor's forceInertial, lift and
// liftUnsteady after them.
__global__ __launch_bounds__(256) void loads_kernel(const double *pan, const double *sec,
                                                    const double *bla, const double *vcp,
                                                    const double *vld, const double *wg,
                                                    double *gprev, double *out, LoadPar p)
{
    extern __shared__ double sh[];  // per section: forceInertial, liftUnsteady; per blade: 9
    const int nsec = p.nbc * p.ns, nc = p.nc, ns = p.ns;
    double *shb = sh + 6 * nsec;
    for (int t = threadIdx.x; t < nsec; t += blockDim.x) {
        const int b = t / ns, is = t - b * ns;
        const double *S = sec + SECTION_N * ((int64_t)b * ns + is);
        const double *B = bla + BLADE_N * b;
        const int64_t e0 = ((int64_t)b * ns + is) * nc;
        // chordwise resultant velocity: a quadratic fit along the chord
        // for three or more panels, else the mean
        double d0[3];
        for (int i = 0; i < 3; ++i) d0[i] = S[i] - S[3 + i];
        const double xq = dot3(d0, S + 6);
        double sx[5] = {(double)nc, 0.0, 0.0, 0.0, 0.0}, by[3] = {0, 0, 0}, bxy[3] = {0, 0, 0},
               bx2y[3] = {0, 0, 0};
        for (int ic = 0; ic < nc; ++ic) {
            const int64_t e = e0 + ic;
            const double *P = pan + PANEL_N * e;
            double vt[3], crv[3], dp[3];
            for (int i = 0; i < 3; ++i) {
                vt[i] = vcp[6 * e + i] + vld[i * p.m + e];
                dp[i] = P[i] - S[3 + i];
            }
            noproj3(vt, P + 9, crv);
            const double x = dot3(dp, S + 6), x2 = x * x;
            sx[1] += x;
            sx[2] += x2;
            sx[3] += x2 * x;
            sx[4] += x2 * x2;
            for (int i = 0; i < 3; ++i) {
                by[i] += crv[i];
                bxy[i] += crv[i] * x;
                bx2y[i] += crv[i] * x2;
            }
        }
        double crv[3];
        for (int i = 0; i < 3; ++i)
            crv[i] = nc >= 3 ? lsq2_solve(xq, sx, by[i], bxy[i], bx2y[i]) : by[i] / nc;
        // angles of attack, inflow and pitch of the section
        const double alpha = atan2(dot3(crv, S + 12), dot3(crv, S + 6));
        const double *vm1 = vcp + 6 * e0 + 3;  // velCPm of the leading panel
        double w[3], vi[3];
        noproj3(vm1, S + 9, w);
        const double phi = angle_tan(crv, w);
        for (int i = 0; i < 3; ++i) vi[i] = crv[i] - vm1[i];
        noproj3(vi, S + 9, w);
        w[2] = w[2] - vi[2];
        const double viz = vi[2], vix = norm3(w);
        double mtc[3] = {-1.0 * S[6], -1.0 * S[7], -1.0 * S[8]};
        const double theta = kPi * 0.5 - angle_tan(mtc, B + 6);
        // lift and drag directions
        double ddir[3], ldir[3], c[3];
        unit3(crv, ddir);
        cross3(ddir, B + 3, c);
        unit3(c, ldir);
        for (int i = 0; i < 3; ++i) ldir[i] = p.sgn * ldir[i];
        // panel pressures (blade_calc_force)
        double fi[3] = {0, 0, 0}, lift[3] = {0, 0, 0}, liftu[3] = {0, 0, 0};
        for (int ic = 0; ic < nc; ++ic) {
            const int64_t e = e0 + ic;
            const double *P = pan + PANEL_N * e;
            const double *vc = vcp + 6 * e;
            const double g = wg[e];
            const double gec = p.inv * (ic == 0 ? g : g - wg[e - 1]);
            const double ges = p.inv * (is == 0 ? g : g - wg[e - nc]);
            const double vtc = dot3(vc, P + 6);
            const double trapz = ic > 0 ? p.inv * 0.5 * (g + wg[e - 1]) : p.inv * 0.5 * g;
            const double dpu = p.density * (trapz - gprev[e]) / p.dt;
            double dpr = dpu + p.density * vtc * gec / P[13];
            if (p.span_lift) {
                const double vts = dot3(vc, P + 9);
                dpr = dpr + p.density * vts * ges / P[14];
            }
            gprev[e] = trapz;
            double nf[3], nfu[3], q[3];
            for (int i = 0; i < 3; ++i) {
                nf[i] = dpr * P[12] * P[3 + i];
                nfu[i] = dpu * P[12] * P[3 + i];
                fi[i] += nf[i];
            }
            proj3(nf, ldir, q);
            for (int i = 0; i < 3; ++i) lift[i] += q[i];
            proj3(nfu, ldir, q);
            for (int i = 0; i < 3; ++i) liftu[i] += q[i];
        }
        double lin[3], lout[3];
        proj3(lift, B, lin);
        for (int i = 0; i < 3; ++i) lin[i] = p.inv * lin[i];
        proj3(lift, B + 6, lout);
        const double vmag = norm3(crv);
        const double q = 0.5 * p.density * (vmag * vmag);
        double cl = 0.0, clu = 0.0;
        if (fabs(q) > kEps) {
            const double sg = copysign(1.0, dot3(lift, B + 6));  // Fortran sign(1, x)
            cl = norm3(lift) * sg / (q * S[15]);
            clu = norm3(liftu) * sg / (q * S[15]);
        }
        double *o = out + LOADS_SEC_N * (int64_t)t;
        for (int i = 0; i < 3; ++i) {
            o[i] = crv[i];
            o[8 + i] = lift[i];
            o[11 + i] = lin[i];
            o[14 + i] = lout[i];
            sh[6 * t + i] = fi[i];
            sh[6 * t + 3 + i] = liftu[i];
        }
        o[3] = alpha;
        o[4] = phi;
        o[5] = viz;
        o[6] = vix;
        o[7] = theta;
        o[17] = cl;
        o[18] = clu;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < p.nbc; b += blockDim.x) {
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int is = 0; is < ns; ++is) {
            const int t = b * ns + is;
            for (int i = 0; i < 3; ++i) {
                s[i] += sh[6 * t + i];
                s[3 + i] += out[LOADS_SEC_N * (int64_t)t + 8 + i];
                s[6 + i] += sh[6 * t + 3 + i];
            }
        }
        for (int k = 0; k < 9; ++k) shb[9 * b + k] = s[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (p.nbc == 1 && p.nb > 1) {
            for (int k = 0; k < 9; ++k) r[k] = (double)p.nb * shb[k];
        } else {
            for (int b = 0; b < p.nbc; ++b)
                for (int k = 0; k < 9; ++k) r[k] = r[k] + shb[9 * b + k];
        }
        double *o = out + LOADS_SEC_N * (int64_t)nsec;
        for (int k = 0; k < 9; ++k) o[k] = r[k];
    }
}

}  // namespace
}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_rotor_set_panels(vlm_ctx *ctx, int ir, const double *panel, const double *section,
                         const double *blade)
{
    Rotor *r = rotor_of(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (!panel || !section || !blade) return fail(ctx, VLM_ERR_ARG, "null buffer");
    VLM_DEVICE(ctx);
    const int64_t n = r->nring, nsec = (int64_t)r->d.nb * r->d.ns;
    VLM_TRY(h2d(ctx, r->pan.p, panel, sizeof(double) * PANEL_N * n));
    VLM_TRY(h2d(ctx, r->sec.p, section, sizeof(double) * SECTION_N * nsec));
    VLM_TRY(h2d(ctx, r->bla.p, blade, sizeof(double) * BLADE_N * r->d.nb));
    hipLaunchKernelGGL(cp_soa_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, r->pan.p, n,
                       r->cps.p);
    VLM_HIP(hipGetLastError());
    r->panels_set = 1;
    return VLM_OK;
}

int vlm_rotor_rhs(vlm_ctx *ctx, int ir, int nterms, const int *irs, const int *sets,
                  const int *flags, const double *kin, double *rhs)
{
    Rotor *r = rotor_of(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (!kin || (nterms > 0 && (!irs || !sets || !flags)) || nterms < 0)
        return fail(ctx, VLM_ERR_ARG, "bad arguments");
    if (!r->panels_set) return fail(ctx, VLM_ERR_STATE, "vlm_rotor_set_panels not called");
    VLM_DEVICE(ctx);
    const int64_t n = r->nring, per = (int64_t)r->d.nc * r->d.ns;
    const int64_t nconv = per * r->d.nbConvect;
    VLM_TRY(induced_terms_soa(ctx, nterms, irs, sets, flags, n, r->cps.p, r->vld.p));
    Kin k;
    std::memcpy(k.vb, kin, 3 * sizeof(double));
    std::memcpy(k.ob, kin + 3, 3 * sizeof(double));
    std::memcpy(k.cg, kin + 6, 3 * sizeof(double));
    k.osl = kin[9];
    std::memcpy(k.shaft, r->shaft, 3 * sizeof(double));
    std::memcpy(k.hub, r->hub, 3 * sizeof(double));
    hipLaunchKernelGGL(rhs_kernel, dim3(nblk(nconv)), dim3(256), 0, ctx->stream, r->pan.p,
                       r->vld.p, n, nconv, k, r->vcp.p, r->rhs.p);
    if (nconv < n)
        hipLaunchKernelGGL(rhs_axisym_kernel, dim3(nblk(n - per)), dim3(256), 0, ctx->stream,
                           r->rhs.p, per, n);
    VLM_HIP(hipGetLastError());
    if (rhs) VLM_TRY(d2h_deferred(ctx, rhs, r->rhs.p, sizeof(double) * n));
    return VLM_OK;
}

int vlm_rotor_mirror_velcp(vlm_ctx *ctx, int ir)
{
    Rotor *r = rotor_of(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (r->image_of < 0) return fail(ctx, VLM_ERR_STATE, "rotor is not an image (vlm_rotor_set_image)");
    VLM_DEVICE(ctx);
    const Rotor &s = *ctx->rotors[r->image_of];
    const int64_t n = r->nring;
    hipLaunchKernelGGL(mirror_velcp_kernel, dim3(nblk(6 * n)), dim3(256), 0, ctx->stream,
                       s.vcp.p, r->vcp.p, n, r->image_axis);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_loads(vlm_ctx *ctx, int ir, int nterms, const int *irs, const int *sets,
                    const int *flags, double density, double dt, double omegaSign,
                    int spanwiseLiftSwitch, double *out)
{
    Rotor *r = rotor_of(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if ((nterms > 0 && (!irs || !sets || !flags)) || nterms < 0 || !out)
        return fail(ctx, VLM_ERR_ARG, "bad arguments");
    if (!r->panels_set) return fail(ctx, VLM_ERR_STATE, "vlm_rotor_set_panels not called");
    VLM_DEVICE(ctx);
    const int64_t n = r->nring;
    LoadPar p;
    p.density = density;
    p.dt = dt;
    p.sgn = copysign(1.0, omegaSign);
    p.inv = -1.0 * p.sgn;
    p.nb = r->d.nb;
    p.nbc = r->d.nbConvect;
    p.nc = r->d.nc;
    p.ns = r->d.ns;
    p.span_lift = spanwiseLiftSwitch != 0;
    p.m = n;
    const size_t shmem = sizeof(double) * (6 * (size_t)p.nbc * p.ns + 9 * (size_t)p.nbc);
    if (shmem > 64 * 1024) return fail(ctx, VLM_ERR_ARG, "too many blade sections for the loads");
    VLM_TRY(induced_terms_soa(ctx, nterms, irs, sets, flags, n, r->cps.p, r->vld.p));
    hipLaunchKernelGGL(loads_kernel, dim3(1), dim3(256), shmem, ctx->stream, r->pan.p, r->sec.p,
                       r->bla.p, r->vcp.p, r->vld.p, r->wg.p, r->gprev.p, r->lout.p, p);
    VLM_HIP(hipGetLastError());
    const size_t nout = (size_t)LOADS_SEC_N * p.nbc * p.ns + LOADS_ROTOR_N;
    return d2h_deferred(ctx, out, r->lout.p, sizeof(double) * nout);
}

}  // extern "C"
