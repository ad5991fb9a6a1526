// Device-resident free wake: rotor objects, segment builders, the N-body
// wake pass and the fused wake-update passes (convection, continuity, age,
// dissipation, strain, shed, rollup, shift, scheme algebra).
//
// HBM layout per rotor (see vlm_c.h for the exported lattice layout):
//   * Near wake as a node lattice: node(b, p, j) with p the PHYSICAL row of
//     a ring buffer; logical row i lives in p = (head + i) mod nNwake.  The
//     reference shifts every near-wake ring down one row per step
//     (rotor_shiftwake, classdef.f90:4481-4498); here the shift is a head
//     move plus one row copy.  Shared ring corners exist once, so the
//     corner copies of blade_wake_continuity (classdef.f90:1609-1702) are
//     implicit.
//   * Velocity arrays keep LOGICAL rows, exactly like velNwake/velNwake1 in
//     the reference, which are not shifted with the wake.
//   * Far wake: per-filament end points (fc1 = trailing end, fc2 = leading).
//
// Sources are built per pass as unique segments: the two ring filaments
// that share a lattice edge are merged into one segment with net
// circulation when their core radii are equal (the common case), otherwise
// both are kept.  Rings with |gam| <= eps are skipped as in
// blade_vind_bywake (classdef.f90:1450-1456).
#include <cmath>
#include <cstring>

#include "vlm_internal.h"

namespace vlm {

int points_to_device(vlm_ctx *ctx, int64_t m, const double *p, double *soa, int64_t stride);
int points_to_host(vlm_ctx *ctx, int64_t m, const double *soa, int64_t stride, double *p,
                   bool deferred = false);

namespace {

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

constexpr double EPS = kEps;

struct RV {  // raw views of a rotor's device state
    int nb, nc, ns, nN, nF, nbc, head, rn, rf;  // rn, rf 0-based
    double *node, *pnode, *le, *ple, *gam, *rc4, *age;
    double *ffc, *pffc, *fgam, *frc, *frc0, *fl0, *flc, *fage;
    double *wc, *wrc, *wg;
    double *vel[4], *velF[4];
};

RV view(Rotor &r)
{
    RV v;
    v.nb = r.d.nb;
    v.nc = r.d.nc;
    v.ns = r.d.ns;
    v.nN = r.d.nNwake;
    v.nF = r.d.nFwake;
    v.nbc = r.d.nbConvect;
    v.head = r.head;
    v.rn = r.rowNear - 1;
    v.rf = r.rowFar - 1;
    v.node = r.node.p;
    v.pnode = r.pnode.p;
    v.le = r.le.p;
    v.ple = r.ple.p;
    v.gam = r.gam.p;
    v.rc4 = r.rc4.p;
    v.age = r.age.p;
    v.ffc = r.ffc.p;
    v.pffc = r.pffc.p;
    v.fgam = r.fgam.p;
    v.frc = r.frc.p;
    v.frc0 = r.frc0.p;
    v.fl0 = r.fl0.p;
    v.flc = r.flc.p;
    v.fage = r.fage.p;
    v.wc = r.wc.p;
    v.wrc = r.wrc.p;
    v.wg = r.wg.p;
    for (int k = 0; k < 4; ++k) {
        v.vel[k] = r.velN[k].p;
        v.velF[k] = r.velF[k].p;
    }
    return v;
}

__device__ __forceinline__ int phys(const RV &v, int i)
{
    int p = v.head + i;
    return p >= v.nN ? p - v.nN : p;
}
__device__ __forceinline__ size_t nidx(const RV &v, int b, int p, int j)
{
    return ((size_t)b * v.nN + p) * (v.ns + 1) + j;
}
__device__ __forceinline__ size_t ridx(const RV &v, int b, int p, int j)
{
    return ((size_t)b * v.nN + p) * v.ns + j;
}
__device__ __forceinline__ size_t vidx(const RV &v, int b, int i, int j)
{
    return ((size_t)b * v.nN + i) * (v.ns + 1) + j;
}

__device__ __forceinline__ void put(Slot *s, int *f, const double *A, const double *B, double g,
                                    double rc)
{
    Slot q;
    q.ax = A[0];
    q.ay = A[1];
    q.az = A[2];
    q.bx = B[0];
    q.by = B[1];
    q.bz = B[2];
    q.gam = g;
    q.rc = rc;
    *s = q;
    *f = g != 0.0 ? 1 : 0;
}

__device__ __forceinline__ double skipg(double g) { return fabs(g) > EPS ? g : 0.0; }

// Wing rings (ring form).  set 0: four filaments per ring; set 1: bound
// vortices vf2 + vf4, without vf2 of the trailing row
// (blade_vind_bywing_boundVortices, classdef.f90:1376-1395).
__global__ void stage_wing_kernel(RV v, int set, Slot *slots, int *flags)
{
    const int nring = v.nb * v.nc * v.ns;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nring) return;
    const double *c = v.wc + 12 * (size_t)k;
    const double *rc = v.wrc + 4 * (size_t)k;
    const double g = v.wg[k];
    if (set == 0) {
        for (int f = 0; f < 4; ++f)
            put(slots + 4 * (size_t)k + f, flags + 4 * (size_t)k + f, c + 3 * f,
                c + 3 * ((f + 1) & 3), g, rc[f]);
    } else {
        const int ic = k % v.nc;
        put(slots + 2 * (size_t)k, flags + 2 * (size_t)k, c + 3 * 3, c, g, rc[3]);
        put(slots + 2 * (size_t)k + 1, flags + 2 * (size_t)k + 1, c + 3, c + 6,
            ic < v.nc - 1 ? g : 0.0, rc[1]);
    }
}

// Near- and far-wake unique segments of every blade.  Edge enumeration per
// blade (R active rows, F active far filaments):
//   [0, R*ns)            trailing spanwise edge of row i: T(i,j) -> T(i,j+1)
//   [.., + ns)           leading edge of the newest row:  LE(j+1) -> LE(j)
//   [.., + R*(ns+1))     streamwise edge: U(i,j) -> T(i,j)
//   [.., + F)            far filament fc1 -> fc2
// Two slots per edge.
__global__ void stage_wake_kernel(RV v, int pred, int far_exists, int64_t per_blade,
                                  Slot *slots, int *flags)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_blade * v.nb) return;
    const int b = (int)(e / per_blade);
    int64_t k = e - (int64_t)b * per_blade;
    Slot *s = slots + 2 * e;
    int *f = flags + 2 * e;
    f[0] = 0;
    f[1] = 0;
    const double *N = pred ? v.pnode : v.node;
    const double *L = (pred ? v.ple : v.le) + 3 * (size_t)b * (v.ns + 1);
    const double *FC = pred ? v.pffc : v.ffc;
    const int R = v.nN - v.rn;
    const int ns = v.ns;
    if (k < (int64_t)R * ns) {
        const int i = v.rn + (int)(k / ns), j = (int)(k % ns);
        const int p = phys(v, i);
        const double *A = N + 3 * nidx(v, b, p, j);
        const double *B = N + 3 * nidx(v, b, p, j + 1);
        const double ga = skipg(v.gam[ridx(v, b, p, j)]);
        const double rca = v.rc4[4 * ridx(v, b, p, j) + 1];
        double gb = 0.0, rcb = rca;
        bool hasb = false;
        if (i + 1 < v.nN) {
            const int q = phys(v, i + 1);
            gb = -skipg(v.gam[ridx(v, b, q, j)]);
            rcb = v.rc4[4 * ridx(v, b, q, j) + 3];
            hasb = true;
        } else if (far_exists) {
            // open horseshoe (classdef.f90:1458-1463): not |gam|-guarded
            gb = -v.gam[ridx(v, b, p, j)];
            rcb = rca;
            hasb = true;
        }
        if (!hasb || rca == rcb) {
            put(s, f, A, B, ga + gb, rca);
        } else {
            put(s, f, A, B, ga, rca);
            put(s + 1, f + 1, A, B, gb, rcb);
        }
        return;
    }
    k -= (int64_t)R * ns;
    if (k < ns) {
        const int j = (int)k;
        const int p = phys(v, v.rn);
        put(s, f, L + 3 * (j + 1), L + 3 * j, skipg(v.gam[ridx(v, b, p, j)]),
            v.rc4[4 * ridx(v, b, p, j) + 3]);
        return;
    }
    k -= ns;
    if (k < (int64_t)R * (ns + 1)) {
        const int i = v.rn + (int)(k / (ns + 1)), j = (int)(k % (ns + 1));
        const int p = phys(v, i);
        const double *T = N + 3 * nidx(v, b, p, j);
        const double *U = (i == v.rn) ? L + 3 * j : N + 3 * nidx(v, b, phys(v, i - 1), j);
        double ga = 0.0, rca = 0.0, gb = 0.0, rcb = 0.0;
        bool hasa = j < ns, hasb = j >= 1;
        if (hasa) {
            ga = skipg(v.gam[ridx(v, b, p, j)]);
            rca = v.rc4[4 * ridx(v, b, p, j) + 0];
        }
        if (hasb) {
            gb = -skipg(v.gam[ridx(v, b, p, j - 1)]);
            rcb = v.rc4[4 * ridx(v, b, p, j - 1) + 2];
        }
        if (hasa && hasb && rca == rcb) {
            put(s, f, U, T, ga + gb, rca);
        } else {
            if (hasa) put(s, f, U, T, ga, rca);
            if (hasb) put(s + 1, f + 1, U, T, gb, rcb);
        }
        return;
    }
    k -= (int64_t)R * (ns + 1);
    {
        const int i = v.rf + (int)k;
        const size_t q = (size_t)b * v.nF + i;
        put(s, f, FC + 6 * q, FC + 6 * q + 3, skipg(v.fgam[q]), v.frc[q]);
    }
}

// Prescribed far-wake filaments as sources (classdef.f90:1471-1476,
// 1501-1507), one slot each, while a far wake exists.
__global__ void stage_presc_kernel(const double *pf, const double *ps, int nb, Slot *slots,
                                   int *flags)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * PF_N) return;
    const int b = e / PF_N;
    put(slots + e, flags + e, pf + 6 * (size_t)e, pf + 6 * (size_t)e + 3, skipg(ps[4 * b]),
        ps[4 * b + 1]);
}

// pFwake_update (classdef.f90:998-1066) for blade blockIdx.x of wake type
// pred, from the far filaments start .. start+n-1 (0-based):
//   radius = mean over them of |(y, x)| of the TE end, pitch = |sum of the
//   z steps between consecutive TE ends| (-2 pi / dpsi) / (n - 1), both
//   relaxed by half against the blade's previous values;
//   a 10-revolution helix of PF_N filaments, clockwise whatever the rotor
//   (isClockwiseRotor is never set, 231 and 1048), turned to the TE end of
//   the last filament (the anchor) and lifted to its height over the hub;
//   filament i joins helix points i (LE end) and i+1 (TE end), filament 0
//   starts at the anchor; every filament takes the last far filament's
//   circulation and core radius.
// The sums run in one lane in the reference's order.
__global__ __launch_bounds__(256) void presc_update_kernel(RV v, int pred, int start, int n,
                                                           double hx, double hy, double hz,
                                                           double dpsi, double *pf, double *ps)
{
    __shared__ double sh[4];
    const int b = blockIdx.x;
    const double *FC = (pred ? v.pffc : v.ffc) + 6 * ((size_t)b * v.nF + start);
    const double *anc = FC + 6 * (size_t)(n - 1);
    constexpr double twoPi = 2.0 * kPi;
    if (threadIdx.x == 0) {
        double pitch = 0.0, radius = 0.0;
        for (int i = 0; i < n; ++i) {
            const double *a = FC + 6 * (size_t)i;
            radius += sqrt(a[1] * a[1] + a[0] * a[0]);
            if (i < n - 1) pitch += a[2] - a[6 + 2];
        }
        pitch = fabs(pitch) * (-twoPi / dpsi) / (n - 1);
        radius = radius / n;
        double *S = ps + 4 * b;
        const size_t q = (size_t)b * v.nF + start + n - 1;
        S[0] = v.fgam[q];
        S[1] = v.frc[q];
        S[2] = 0.5 * pitch + (1 - 0.5) * S[2];
        S[3] = 0.5 * radius + (1 - 0.5) * S[3];
        sh[0] = S[2];
        sh[1] = S[3];
        sh[2] = atan2(anc[1], anc[0]);
        sh[3] = anc[2] - hz;
    }
    __syncthreads();
    const double pitch = sh[0], radius = sh[1], dth = sh[2], dz = sh[3];
    const double dx = (twoPi * 10.0 - 0.0) / (PF_N + 1 - 1);
    double *W = pf + 6 * (size_t)b * PF_N;
    for (int k = threadIdx.x; k <= PF_N; k += blockDim.x) {
        const double th = -(k * dx + 0.0);
        const double c0 = radius * cos(th + dth), c1 = radius * sin(th + dth);
        const double c2 = pitch * fabs(th) / twoPi + dz;
        if (k < PF_N) {   // LE end of filament k (filament 0: the anchor)
            W[6 * k + 3] = k == 0 ? anc[0] : hx + c0;
            W[6 * k + 4] = k == 0 ? anc[1] : hy + c1;
            W[6 * k + 5] = k == 0 ? anc[2] : hz + c2;
        }
        if (k > 0) {      // TE end of filament k-1
            W[6 * (k - 1) + 0] = hx + c0;
            W[6 * (k - 1) + 1] = hy + c1;
            W[6 * (k - 1) + 2] = hz + c2;
        }
    }
}

// Targets: near-wake nodes of rows rn..nN-1 then far trailing ends, for
// blades < nbConvect, written as SoA at t[off..].
__global__ void gather_targets_kernel(RV v, int pred, double *t, int64_t stride, int64_t off)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    const double *P;
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        P = (pred ? v.pnode : v.node) + 3 * nidx(v, b, phys(v, i), j);
    } else {
        const int i = v.rf + (int)(k - (int64_t)R * (v.ns + 1));
        P = (pred ? v.pffc : v.ffc) + 6 * ((size_t)b * v.nF + i);
    }
    t[off + e] = P[0];
    t[stride + off + e] = P[1];
    t[2 * stride + off + e] = P[2];
}

__global__ void scatter_vel_kernel(RV v, int which, const double *o, int64_t stride, int64_t off)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    double *dst;
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        dst = v.vel[which] + 3 * vidx(v, b, i, j);
    } else {
        const int i = v.rf + (int)(k - (int64_t)R * (v.ns + 1));
        dst = v.velF[which] + 3 * ((size_t)b * v.nF + i);
    }
    dst[0] = o[off + e];
    dst[1] = o[stride + off + e];
    dst[2] = o[2 * stride + off + e];
}

// blade_convectwake (classdef.f90:1515-1575): nodes of logical rows
// [i0, i1) move by velNwake*dt; far trailing ends of rows rf.. move by
// velFwake*dt.  Both 'C' and 'P' use velNwake, as the reference does.
__global__ void convect_kernel(RV v, int pred, int i0, int i1, double dt)
{
    const int rows = i1 - i0;
    const int64_t nnear = (int64_t)rows * (v.ns + 1);
    const int F = v.nF - v.rf;
    const int64_t per = nnear + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    if (k < nnear) {
        const int i = i0 + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        double *P = (pred ? v.pnode : v.node) + 3 * nidx(v, b, phys(v, i), j);
        const double *V = v.vel[VLM_VEL] + 3 * vidx(v, b, i, j);
        P[0] += V[0] * dt;
        P[1] += V[1] * dt;
        P[2] += V[2] * dt;
    } else {
        const int i = v.rf + (int)(k - nnear);
        double *P = (pred ? v.pffc : v.ffc) + 6 * ((size_t)b * v.nF + i);
        const double *V = v.velF[VLM_VEL] + 3 * ((size_t)b * v.nF + i);
        P[0] += V[0] * dt;
        P[1] += V[1] * dt;
        P[2] += V[2] * dt;
    }
}

// Far-wake continuity: fc2(i) <- fc1(i-1) for i = rf+1 .. nF-1.
__global__ void far_continuity_kernel(RV v, int pred)
{
    const int F = v.nF - v.rf - 1;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (F <= 0 || e >= F * v.nbc) return;
    const int b = e / F, i = v.rf + 1 + e % F;
    double *FC = pred ? v.pffc : v.ffc;
    const size_t q = (size_t)b * v.nF + i;
    FC[6 * q + 3] = FC[6 * (q - 1) + 0];
    FC[6 * q + 4] = FC[6 * (q - 1) + 1];
    FC[6 * q + 5] = FC[6 * (q - 1) + 2];
}

struct RotMats {
    double T[16][9];  // column-major getTransformAxis per blade
};

__device__ __forceinline__ void rotp(const double *T, const double *o, const double *p, double *q)
{
    const double x = p[0] - o[0], y = p[1] - o[1], z = p[2] - o[2];
    q[0] = T[0] * x + T[3] * y + T[6] * z + o[0];
    q[1] = T[1] * x + T[4] * y + T[7] * z + o[1];
    q[2] = T[2] * x + T[5] * y + T[8] * z + o[2];
}

// Axisymmetric copy of blade 1's wake to blades 2..nb rotated by
// 2 pi (ib-1)/nb about the shaft (rotor_convectwake, classdef.f90:4801-4823).
__global__ void axisym_kernel(RV v, int pred, RotMats M, double ox, double oy, double oz)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + (v.ns + 1) + F + (int64_t)R * v.ns;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * (v.nb - 1)) return;
    const int b = 1 + (int)(e / per);
    int64_t k = e - (int64_t)(b - 1) * per;
    const double o[3] = {ox, oy, oz};
    const double *T = M.T[b];
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        const int p = phys(v, i);
        double *N = pred ? v.pnode : v.node;
        rotp(T, o, N + 3 * nidx(v, 0, p, j), N + 3 * nidx(v, b, p, j));
        return;
    }
    k -= (int64_t)R * (v.ns + 1);
    if (k < v.ns + 1) {
        double *L = pred ? v.ple : v.le;
        rotp(T, o, L + 3 * k, L + 3 * ((size_t)b * (v.ns + 1) + k));
        return;
    }
    k -= v.ns + 1;
    if (k < F) {
        const int i = v.rf + (int)k;
        double *FC = pred ? v.pffc : v.ffc;
        const size_t q0 = (size_t)i, qb = (size_t)b * v.nF + i;
        rotp(T, o, FC + 6 * q0, FC + 6 * qb);
        rotp(T, o, FC + 6 * q0 + 3, FC + 6 * qb + 3);
        if (!pred) {
            v.fgam[qb] = v.fgam[q0];
            v.frc[qb] = v.frc[q0];
            v.frc0[qb] = v.frc0[q0];
            v.fl0[qb] = v.fl0[q0];
            v.flc[qb] = v.flc[q0];
            v.fage[qb] = v.fage[q0];
        }
        return;
    }
    k -= F;
    if (!pred) {
        const int i = v.rn + (int)(k / v.ns), j = (int)(k % v.ns);
        const int p = phys(v, i);
        const size_t r0 = ridx(v, 0, p, j), rb = ridx(v, b, p, j);
        v.gam[rb] = v.gam[r0];
        v.age[rb] = v.age[r0];
        for (int q = 0; q < 4; ++q) v.rc4[4 * rb + q] = v.rc4[4 * r0 + q];
    }
}

// Axisymmetric blades take blade 1's prescribed wake turned by
// 2 pi (b-1)/nb about the shaft (rotor_updatePrescribedWake,
// classdef.f90:5203-5215, pFwake_rot_wake_axis 1068-1086).
__global__ void presc_axisym_kernel(int nb, RotMats M, double ox, double oy, double oz,
                                    double *pf, double *ps)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (nb - 1) * PF_N) return;
    const int b = 1 + e / PF_N, i = e % PF_N;
    const double o[3] = {ox, oy, oz};
    const double *src = pf + 6 * (size_t)i;
    double *dst = pf + 6 * ((size_t)b * PF_N + i);
    rotp(M.T[b], o, src, dst);
    rotp(M.T[b], o, src + 3, dst + 3);
    if (i < 4) ps[4 * b + i] = ps[i];
}

// rotor_assignshed (classdef.f90:4297-4325).  LE: shed line of row rowNear
// from the wing trailing vortex ring corners 2/3, circulation of the wing
// trailing row.  TE: trailing nodes of row max(rowNear-1, 1).
__global__ void assignshed_kernel(RV v, int te, int row)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= v.nb * (v.ns + 1)) return;
    const int b = e / (v.ns + 1), j = e % (v.ns + 1);
    const int js = j < v.ns ? j : v.ns - 1;
    const size_t ring = ((size_t)b * v.ns + js) * v.nc + (v.nc - 1);
    const double *W = v.wc + 12 * ring + (j < v.ns ? 3 : 6);
    const int p = phys(v, row);
    double *dst = te ? v.node + 3 * nidx(v, b, p, j) : v.le + 3 * ((size_t)b * (v.ns + 1) + j);
    dst[0] = W[0];
    dst[1] = W[1];
    dst[2] = W[2];
    if (!te && j < v.ns) v.gam[ridx(v, b, p, j)] = v.wg[ring];
}

// rotor_age_wake (classdef.f90:4331-4354); ageAzimuthal is output-only.
__global__ void age_kernel(RV v, double dt)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * v.ns + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nb) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    if (k < (int64_t)R * v.ns) {
        const int i = v.rn + (int)(k / v.ns), j = (int)(k % v.ns);
        v.age[ridx(v, b, phys(v, i), j)] += dt;
    } else {
        v.fage[(size_t)b * v.nF + v.rf + (k - (int64_t)R * v.ns)] += dt;
    }
}

// rotor_dissipate_wake (classdef.f90:4356-4408), in two fully parallel
// passes: (1) every ring grows vf1 (copied to vf3) and vf2 and decays gam,
// far filaments grow and decay; (2) vf4 of every row but the newest takes
// the grown vf2 of the row upstream.
__global__ void dissipate_kernel(RV v, double inc, double decay)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * v.ns + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nb) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    if (k < (int64_t)R * v.ns) {
        const int i = v.rn + (int)(k / v.ns), j = (int)(k % v.ns);
        const size_t r = ridx(v, b, phys(v, i), j);
        double *rc = v.rc4 + 4 * r;
        rc[0] = sqrt(rc[0] * rc[0] + inc);
        rc[2] = rc[0];
        v.gam[r] = v.gam[r] * decay;
        rc[1] = sqrt(rc[1] * rc[1] + inc);
    } else {
        const size_t q = (size_t)b * v.nF + v.rf + (k - (int64_t)R * v.ns);
        v.frc[q] = sqrt(v.frc[q] * v.frc[q] + inc);
        v.fgam[q] = v.fgam[q] * decay;
    }
}

__global__ void dissipate_vf4_kernel(RV v)
{
    const int R = v.nN - v.rn - 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R <= 0 || e >= (int64_t)R * v.ns * v.nb) return;
    const int b = (int)(e / ((int64_t)R * v.ns));
    const int64_t k = e - (int64_t)b * R * v.ns;
    const int i = v.rn + 1 + (int)(k / v.ns), j = (int)(k % v.ns);
    v.rc4[4 * ridx(v, b, phys(v, i), j) + 3] = v.rc4[4 * ridx(v, b, phys(v, i - 1), j) + 1];
}

// rotor_strain_wake (classdef.f90:4410-4422): far filaments only.
__global__ void strain_kernel(RV v)
{
    const int F = v.nF - v.rf;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (F <= 0 || e >= v.nb * F) return;
    const size_t q = (size_t)(e / F) * v.nF + v.rf + e % F;
    const double *fc = v.ffc + 6 * q;
    const double dx = fc[0] - fc[3], dy = fc[1] - fc[4], dz = fc[2] - fc[5];
    v.flc[q] = sqrt(dx * dx + dy * dy + dz * dz);
    v.frc[q] = v.frc0[q] * sqrt(v.fl0[q] / v.flc[q]);
}

// Skew of the far-wake pair (i, i+1): the angle between filament i
// (fc1 -> fc2) and filament i+1 reversed, as |acos(cos) - pi| / pi
// (blade_burst_wake, classdef.f90:2322-2338; getAngleCos, libMath.f90:238-247).
__device__ double far_pair_skew(const double *ffc, size_t q)
{
    const double *f = ffc + 6 * q, *g = ffc + 6 * (q + 1);
    const double ax = f[3] - f[0], ay = f[4] - f[1], az = f[5] - f[2];
    const double bx = g[0] - g[3], by = g[1] - g[4], bz = g[2] - g[5];
    const double d = ax * bx + ay * by + az * bz;
    const double n = sqrt((ax * ax + ay * ay + az * az) * (bx * bx + by * by + bz * bz));
    return fabs(acos(d / n) - kPi) / kPi;
}

// rotor_burst_wake (classdef.f90:4911-4917): a far filament takes the large
// core radius when the pair with either neighbour reaches the skew limit.
// One thread per filament, so no two threads write one core radius.
__global__ void burst_kernel(RV v, double skew_limit, double large_core)
{
    const int F = v.nF - v.rf;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (F <= 0 || e >= v.nb * F) return;
    const int i = v.rf + e % F;
    const size_t q = (size_t)(e / F) * v.nF + i;
    bool burst = false;
    if (i > v.rf) burst = far_pair_skew(v.ffc, q - 1) >= skew_limit;
    if (i + 1 < v.nF) burst = burst || far_pair_skew(v.ffc, q) >= skew_limit;
    if (burst) v.frc[q] = large_core;
}

// rotor_shiftFwake (classdef.f90:4500-4513) in two passes via staging.
__global__ void far_stage_kernel(RV v, double *tmp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= v.nb * v.nF) return;
    double *t = tmp + 12 * (size_t)e;
    for (int q = 0; q < 6; ++q) t[q] = v.ffc[6 * (size_t)e + q];
    t[6] = v.fgam[e];
    t[7] = v.frc[e];
    t[8] = v.frc0[e];
    t[9] = v.fl0[e];
    t[10] = v.flc[e];
    t[11] = v.fage[e];
}

__global__ void far_shift_kernel(RV v, const double *tmp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= v.nb * v.nF) return;
    const int i = e % v.nF;
    if (i == 0) {
        v.fage[e] = 0.0;
        return;
    }
    const double *t = tmp + 12 * (size_t)(e - 1);
    for (int q = 0; q < 6; ++q) v.ffc[6 * (size_t)e + q] = t[q];
    v.fgam[e] = t[6];
    v.frc[e] = t[7];
    v.frc0[e] = t[8];
    v.fl0[e] = t[9];
    v.flc[e] = t[10];
    v.fage[e] = t[11];
}

// rotor_rollup (classdef.f90:4515-4597): one lane per blade forms the
// circulation-weighted centroid of the oldest near row and writes the new
// far filament at rowFarNext (0-based fi).
__global__ void rollup_kernel(RV v, int rs, int re, int rollup_min, int suppress, int fi)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= v.nb) return;
    const int last = v.nN - 1;
    const int p = phys(v, last);
    const double *N = v.node;
    const double *L = v.le + 3 * (size_t)b * (v.ns + 1);
    // corner 4 of ring (last, j) = node(last-1, j+1) or the shed line
    auto c4 = [&](int j) -> const double * {
        return (last == v.rn) ? L + 3 * (j + 1) : N + 3 * nidx(v, b, phys(v, last - 1), j + 1);
    };
    double gamRollup = v.gam[ridx(v, b, p, v.ns - 1)];
    double cLE[3] = {0, 0, 0}, cTE[3] = {0, 0, 0};
    double radius = 0.0, gamSum = 0.0;
    for (int j = rs; j <= re; ++j) {
        const size_t r = ridx(v, b, p, j);
        const double g = v.gam[r];
        const double *A = c4(j);
        const double *C3 = N + 3 * nidx(v, b, p, j + 1);
        for (int d = 0; d < 3; ++d) {
            cLE[d] += A[d] * g;
            cTE[d] += C3[d] * g;
        }
        gamSum += g;
        if (rollup_min) {
            if (g < gamRollup) gamRollup = g;
        } else {
            if (g > gamRollup) gamRollup = g;
        }
        radius += v.rc4[4 * r + 2] * g;
    }
    const double ageRollup = v.age[ridx(v, b, p, v.ns - 1)];
    if (fabs(gamSum) > EPS) {
        for (int d = 0; d < 3; ++d) {
            cLE[d] /= gamSum;
            cTE[d] /= gamSum;
        }
        radius /= gamSum;
    } else {
        const double *C2 = N + 3 * nidx(v, b, p, re);
        const double *C3 = N + 3 * nidx(v, b, p, re + 1);
        for (int d = 0; d < 3; ++d) {
            cLE[d] = C2[d];
            cTE[d] = C3[d];
        }
        radius = v.rc4[4 * ridx(v, b, p, re) + 2];
    }
    if (suppress) gamRollup = 0.0;
    if (v.nF > 0) {
        const size_t q = (size_t)b * v.nF + fi;
        double *fc = v.ffc + 6 * q;
        for (int d = 0; d < 3; ++d) {
            fc[3 + d] = cLE[d];
            fc[d] = cTE[d];
        }
        v.fgam[q] = gamRollup;
        v.fage[q] = ageRollup;
        v.frc0[q] = radius;
        v.frc[q] = radius;
        const double dx = fc[0] - fc[3], dy = fc[1] - fc[4], dz = fc[2] - fc[5];
        v.flc[q] = v.fl0[q] = sqrt(dx * dx + dy * dy + dz * dz);
        if (fi + 1 < v.nF)
            for (int d = 0; d < 3; ++d) v.ffc[6 * (q + 1) + 3 + d] = cTE[d];
    }
}

// After the head moved back one slot: new logical row 0 <- row 1 (all ring
// attributes), with age reset (rotor_shiftwake, classdef.f90:4488-4495).
__global__ void shift_copy_kernel(RV v)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= v.nb * (v.ns + 1)) return;
    const int b = e / (v.ns + 1), j = e % (v.ns + 1);
    const int p0 = phys(v, 0), p1 = phys(v, 1);
    const double *s = v.node + 3 * nidx(v, b, p1, j);
    double *d = v.node + 3 * nidx(v, b, p0, j);
    d[0] = s[0];
    d[1] = s[1];
    d[2] = s[2];
    if (j < v.ns) {
        const size_t r0 = ridx(v, b, p0, j), r1 = ridx(v, b, p1, j);
        v.gam[r0] = v.gam[r1];
        for (int q = 0; q < 4; ++q) v.rc4[4 * r0 + q] = v.rc4[4 * r1 + q];
        v.age[r0] = 0.0;
    }
}

// Predicted wake := current wake on the active rows (main.f90:1028-1035).
__global__ void copy_pred_kernel(RV v)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + (v.ns + 1) + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    int64_t k = e - b * per;
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        const size_t n = 3 * nidx(v, b, phys(v, i), j);
        v.pnode[n] = v.node[n];
        v.pnode[n + 1] = v.node[n + 1];
        v.pnode[n + 2] = v.node[n + 2];
        return;
    }
    k -= (int64_t)R * (v.ns + 1);
    if (k < v.ns + 1) {
        const size_t n = 3 * ((size_t)b * (v.ns + 1) + k);
        v.ple[n] = v.le[n];
        v.ple[n + 1] = v.le[n + 1];
        v.ple[n + 2] = v.le[n + 2];
        return;
    }
    k -= v.ns + 1;
    const size_t q = 6 * ((size_t)b * v.nF + v.rf + k);
    for (int d = 0; d < 6; ++d) v.pffc[q + d] = v.ffc[q + d];
}

// Scheme algebra over whole velocity arrays of blades < nbConvect, as the
// reference applies it to whole arrays (main.f90:968-1115).
//   op 0: AB2  step = 1.5 v - 0.5 v1; v1 = step; v = step
//   op 1: AM2 predictor  step = v; v = 0.5 (3 v - v1)
//   op 2: AM2 corrector  v = (pred + step) * 0.5
//   op 3: history        v1 = step
//   op 4: first step     v1 = v
__global__ void scheme_kernel(double *v, double *v1, double *vp, double *vs, int64_t n, int op)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    switch (op) {
    case 0: {
        const double s = 0.5 * (3.0 * v[k] - v1[k]);
        vs[k] = s;
        v1[k] = s;
        v[k] = s;
        break;
    }
    case 1:
        vs[k] = v[k];
        v[k] = 0.5 * (3.0 * v[k] - v1[k]);
        break;
    case 2:
        v[k] = (vp[k] + vs[k]) * 0.5;
        break;
    case 3:
        v1[k] = vs[k];
        break;
    default:
        v1[k] = v[k];
        break;
    }
}

// AM3 and AM4 velocity algebra (main.f90:1117-1404) on whole logical-row
// arrays: v = velNwake, v1..v3 = velNwake1..3, vp = velNwakePredicted,
// vs = velNwakeStep.
enum AmOp {
    AM3_PRED, AM3_CORR, AM3_ROTATE,  // 1160-1171, 1219-1235
    AM4_PRED, AM4_CORR, AM4_ROTATE,  // 1307-1320, 1365-1385
    AM_START1, AM_START2, AM_START3  // start-up copies v1|v2|v3 = v
};

__global__ void am_kernel(double *v, double *v1, double *v2, double *v3, const double *vp,
                          double *vs, int64_t n, int op)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    switch (op) {
    case AM3_PRED:
        vs[k] = v[k];
        v[k] = (23.0 * v[k] - 16.0 * v2[k] + 5.0 * v1[k]) / 12.0;
        break;
    case AM3_CORR:
        v[k] = (5.0 * vp[k] + 8.0 * vs[k] - v2[k]) / 12.0;
        break;
    case AM3_ROTATE:
        v1[k] = v2[k];
        v2[k] = vs[k];
        break;
    case AM4_PRED:
        vs[k] = v[k];
        v[k] = (55.0 * v[k] - 59.0 * v3[k] + 37.0 * v2[k] - 9.0 * v1[k]) / 24.0;
        break;
    case AM4_CORR:
        v[k] = (9.0 * vp[k] + 19.0 * vs[k] - 5.0 * v3[k] + v2[k]) / 24.0;
        break;
    case AM4_ROTATE:
        v1[k] = v2[k];
        v2[k] = v3[k];
        v3[k] = vs[k];
        break;
    case AM_START1:
        v1[k] = v[k];
        break;
    case AM_START2:
        v2[k] = v[k];
        break;
    default:
        v3[k] = v[k];
        break;
    }
}

// vel(:, row0:nrows-1, :) += a for blades 0..nbc-1 of a logical-row
// velocity array (cols = ns+1 for the near wake, 1 for the far wake).
__global__ void add_uniform_kernel(double *vel, int nbc, int nrows, int cols, int row0,
                                   double ax, double ay, double az)
{
    const int64_t per = (int64_t)(nrows - row0) * cols;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    double *V = vel + 3 * (((size_t)b * nrows + row0) * cols + k);
    V[0] += ax;
    V[1] += ay;
    V[2] += az;
}

// vel_order2_Nwake / vel_order2_Fwake (libCommon.f90:213-258) on the
// active window, written to `out` (logical rows).
__global__ void order2_kernel(const double *vn, const double *vnp1, double *out, int nbc,
                              int rows, int stride, int cols, int base)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbc * rows * cols * 3) return;
    const int d = (int)(e % 3);
    int64_t r = e / 3;
    const int i = (int)(r % rows);
    r /= rows;
    const int j = (int)(r % cols);
    const int b = (int)(r / cols);
    auto at = [&](int ii) { return (((size_t)b * stride + base + ii) * cols + j) * 3 + d; };
    double val;
    if (i == 0 || i == rows - 1) val = (vnp1[at(i)] + vn[at(i)]) * 0.5;
    else val = (vnp1[at(i)] + vnp1[at(i - 1)] + vn[at(i + 1)] + vn[at(i)]) * 0.25;
    out[at(i)] = val;
}

__global__ void copy_window_kernel(const double *src, double *dst, int nbc, int rows, int stride,
                                   int cols, int base)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nbc * rows * cols * 3) return;
    const int d = (int)(e % 3);
    int64_t r = e / 3;
    const int i = (int)(r % rows);
    r /= rows;
    const int j = (int)(r % cols);
    const int b = (int)(r / cols);
    const size_t k = (((size_t)b * stride + base + i) * cols + j) * 3 + d;
    dst[k] = src[k];
}

__global__ void negate_kernel(double *p, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = -p[e];
}

__global__ void fill_kernel(double *p, int64_t n, double val)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = val;
}

// rotor_mirrorWake (classdef.f90:5024-5131) with vr_mirror / Fwake_mirror
// (722-735, 982-993): image d takes source s's wake on every blade, with
// coordinate ax and the circulation negated.  mode 0 (C): current near
// rows from i0 and far filaments from f0; 1 (P): predicted positions over
// the same rows, with the circulation, cores and ages, which the predicted
// wake shares with the current one (the reference copies them inside its
// predicted rings; the current rows get the same values again at the C
// mirror that follows); 2 (A): every row, i0 = f0 = 0.  Rows are logical
// on both sides (each rotor has its own ring head).  One thread per near node, leading-edge
// node, near ring or far filament of a blade.
__global__ void mirror_wake_kernel(RV d, RV s, int mode, int ax, int i0, int f0)
{
    const int ns = d.ns, R = d.nN - i0, F = d.nF - f0;
    const int64_t per = (int64_t)R * (ns + 1) + (ns + 1) + (int64_t)R * ns + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * d.nb) return;
    const int b = (int)(e / per);
    int64_t k = e - b * per;
    const bool pred = mode == 1;
    auto copy3 = [ax](const double *src, double *dst) {
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
        dst[ax] = -src[ax];
    };
    if (k < (int64_t)R * (ns + 1)) {
        const int i = i0 + (int)(k / (ns + 1)), j = (int)(k % (ns + 1));
        copy3((pred ? s.pnode : s.node) + 3 * nidx(s, b, phys(s, i), j),
              (pred ? d.pnode : d.node) + 3 * nidx(d, b, phys(d, i), j));
        return;
    }
    k -= (int64_t)R * (ns + 1);
    if (k < ns + 1) {
        const size_t q = 3 * ((size_t)b * (ns + 1) + k);
        copy3((pred ? s.ple : s.le) + q, (pred ? d.ple : d.le) + q);
        return;
    }
    k -= ns + 1;
    if (k < (int64_t)R * ns) {
        const int i = i0 + (int)(k / ns), j = (int)(k % ns);
        const size_t rs = ridx(s, b, phys(s, i), j), rd = ridx(d, b, phys(d, i), j);
        d.gam[rd] = -s.gam[rs];
        d.age[rd] = s.age[rs];
        for (int q = 0; q < 4; ++q) d.rc4[4 * rd + q] = s.rc4[4 * rs + q];
        return;
    }
    k -= (int64_t)R * ns;
    const size_t q = (size_t)b * d.nF + f0 + k;
    const double *FS = pred ? s.pffc : s.ffc;
    double *FD = pred ? d.pffc : d.ffc;
    copy3(FS + 6 * q, FD + 6 * q);
    copy3(FS + 6 * q + 3, FD + 6 * q + 3);
    d.fgam[q] = -s.fgam[q];
    d.frc[q] = s.frc[q];
    d.frc0[q] = s.frc0[q];
    d.fl0[q] = s.fl0[q];
    d.flc[q] = s.flc[q];
    d.fage[q] = s.fage[q];
}

// rotor_mirrorGamma + rotor_map_gam: the image's wing circulation is the
// negated source's, ring for ring.
__global__ void mirror_gamma_kernel(const double *src, double *dst, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = -src[e];
}

}  // namespace

// ---------------------------------------------------------------------
// host-side helpers
// ---------------------------------------------------------------------
int64_t wing_slots(const Rotor &r, int set)
{
    return (int64_t)r.nring * (set == VLM_SET_WING ? 4 : 2);
}

static int64_t wake_edges_per_blade(const Rotor &r)
{
    const int64_t R = r.d.nNwake - (r.rowNear - 1);
    const int64_t F = r.d.nFwake - (r.rowFar - 1);
    if (r.d.nNwake <= 0 || R <= 0) return F > 0 ? F : 0;
    return R * r.d.ns + r.d.ns + R * (r.d.ns + 1) + (F > 0 ? F : 0);
}

// prescribed far-wake sources: one slot per filament while a far wake exists
static int64_t presc_slots(const Rotor &r)
{
    return (r.presc_nt > 0 && r.rowFar <= r.d.nFwake) ? (int64_t)PF_N * r.d.nb : 0;
}

int64_t wake_slots(const Rotor &r) { return 2 * wake_edges_per_blade(r) * r.d.nb + presc_slots(r); }

int stage_wing(vlm_ctx *ctx, Rotor &r, int set, int64_t off, int64_t *used)
{
    *used = wing_slots(r, set);
    if (*used == 0) return VLM_OK;
    RV v = view(r);
    hipLaunchKernelGGL(stage_wing_kernel, dim3(nblk(r.nring)), dim3(256), 0, ctx->stream, v,
                       set, ctx->slots.p + off, ctx->flags.p + off);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int stage_wake(vlm_ctx *ctx, Rotor &r, int pred, int64_t off, int64_t *used)
{
    *used = 0;
    if (r.d.nNwake <= 0) return VLM_OK;
    const int64_t per = wake_edges_per_blade(r);
    if (per <= 0) return VLM_OK;
    if (r.rowNear - 1 >= r.d.nNwake) return VLM_OK;  // no active near-wake row yet
    *used = 2 * per * r.d.nb;
    RV v = view(r);
    const int far_exists = r.rowFar <= r.d.nFwake ? 1 : 0;
    hipLaunchKernelGGL(stage_wake_kernel, dim3(nblk(per * r.d.nb)), dim3(256), 0, ctx->stream,
                       v, pred, far_exists, per, ctx->slots.p + off, ctx->flags.p + off);
    const int64_t np = presc_slots(r);
    if (np > 0) {
        const int t = pred ? 1 : 0;
        hipLaunchKernelGGL(stage_presc_kernel, dim3(nblk(np)), dim3(256), 0, ctx->stream,
                           r.pf[t].p, r.ps[t].p, r.d.nb, ctx->slots.p + off + *used,
                           ctx->flags.p + off + *used);
        *used += np;
    }
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

namespace {

int64_t target_count(const Rotor &r)
{
    if (r.d.nNwake <= 0 || r.image_of >= 0) return 0;
    const int64_t R = r.d.nNwake - (r.rowNear - 1);
    const int64_t F = r.d.nFwake - (r.rowFar - 1);
    if (R <= 0) return 0;
    return (R * (r.d.ns + 1) + (F > 0 ? F : 0)) * r.d.nbConvect;
}

Rotor *get_rotor(vlm_ctx *ctx, int ir)
{
    if (!ctx || ir < 0 || ir >= (int)ctx->rotors.size()) return nullptr;
    return ctx->rotors[ir];
}

// Build the source list for a set of (rotor, set) pairs.
int build_sources(vlm_ctx *ctx, const std::vector<std::pair<int, int>> &src)
{
    int64_t total = 0;
    for (auto &s : src) {
        Rotor &r = *ctx->rotors[s.first];
        if (s.second == VLM_SET_WING || s.second == VLM_SET_WING_BOUND)
            total += wing_slots(r, s.second);
        else if (r.rowNear - 1 < r.d.nNwake)
            total += wake_slots(r);
    }
    VLM_TRY(ensure_slots(ctx, total > 0 ? total : 1));
    int64_t off = 0, used = 0;
    for (auto &s : src) {
        Rotor &r = *ctx->rotors[s.first];
        if (s.second == VLM_SET_WING || s.second == VLM_SET_WING_BOUND)
            VLM_TRY(stage_wing(ctx, r, s.second, off, &used));
        else
            VLM_TRY(stage_wake(ctx, r, s.second == VLM_SET_WAKE_PRED, off, &used));
        off += used;
    }
    VLM_TRY(compact_slots(ctx, off));
    ctx->seg.cap = ctx->seg.cap;  // stride of the SoA arrays
    return VLM_OK;
}

int eval_sources(vlm_ctx *ctx, const double *t, int64_t m, int64_t tstride, double *out, int acc,
                 int64_t nmax)
{
    return bs_eval(ctx, ctx->seg.d64.p, ctx->seg.d32.p, ctx->seg.cap, ctx->seg.count.p, nmax, t,
                   m, tstride, out, acc);
}

int64_t staged_upper_bound(vlm_ctx *ctx, const std::vector<std::pair<int, int>> &src)
{
    int64_t total = 0;
    for (auto &s : src) {
        Rotor &r = *ctx->rotors[s.first];
        if (s.second == VLM_SET_WING || s.second == VLM_SET_WING_BOUND)
            total += wing_slots(r, s.second);
        else if (r.rowNear - 1 < r.d.nNwake)
            total += wake_slots(r);
    }
    return total;
}

int launch_scheme(vlm_ctx *ctx, Rotor &r, int op)
{
    const int64_t nn = (int64_t)3 * r.d.nbConvect * r.d.nNwake * (r.d.ns + 1);
    const int64_t nf = (int64_t)3 * r.d.nbConvect * r.d.nFwake;
    if (nn > 0)
        hipLaunchKernelGGL(scheme_kernel, dim3(nblk(nn)), dim3(256), 0, ctx->stream,
                           r.velN[VLM_VEL].p, r.velN[VLM_VEL_PREV].p, r.velN[VLM_VEL_PRED].p,
                           r.velN[VLM_VEL_STEP].p, nn, op);
    if (nf > 0)
        hipLaunchKernelGGL(scheme_kernel, dim3(nblk(nf)), dim3(256), 0, ctx->stream,
                           r.velF[VLM_VEL].p, r.velF[VLM_VEL_PREV].p, r.velF[VLM_VEL_PRED].p,
                           r.velF[VLM_VEL_STEP].p, nf, op);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int launch_am(vlm_ctx *ctx, Rotor &r, int op)
{
    const int64_t nn = (int64_t)3 * r.d.nbConvect * r.d.nNwake * (r.d.ns + 1);
    const int64_t nf = (int64_t)3 * r.d.nbConvect * r.d.nFwake;
    if (nn > 0)
        hipLaunchKernelGGL(am_kernel, dim3(nblk(nn)), dim3(256), 0, ctx->stream,
                           r.velN[VLM_VEL].p, r.velN[VLM_VEL_PREV].p, r.velN[VEL_H2].p,
                           r.velN[VEL_H3].p, r.velN[VLM_VEL_PRED].p, r.velN[VLM_VEL_STEP].p, nn,
                           op);
    if (nf > 0)
        hipLaunchKernelGGL(am_kernel, dim3(nblk(nf)), dim3(256), 0, ctx->stream,
                           r.velF[VLM_VEL].p, r.velF[VLM_VEL_PREV].p, r.velF[VEL_H2].p,
                           r.velF[VEL_H3].p, r.velF[VLM_VEL_PRED].p, r.velF[VLM_VEL_STEP].p, nf,
                           op);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

// initWakeVel of pass 1 (main.f90:829-838: + near, - far) or of pass 2
// (902-913 and the AM blocks: - near, - far) on the active windows.
int add_init_wake_vel(vlm_ctx *ctx, Rotor &r, int iter, bool pass2)
{
    if (iter >= r.init_wake_vel_nt || r.init_wake_vel == 0.0) return VLM_OK;
    const int slot = pass2 ? VLM_VEL_PRED : VLM_VEL;
    const double u = r.init_wake_vel;
    const double sn = pass2 ? -1.0 : 1.0;
    const int nbc = r.d.nbConvect;
    const int rn = r.rowNear - 1, rf = r.rowFar - 1;
    if (r.d.nNwake > rn) {
        const int64_t n = (int64_t)nbc * (r.d.nNwake - rn) * (r.d.ns + 1);
        hipLaunchKernelGGL(add_uniform_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           r.velN[slot].p, nbc, r.d.nNwake, r.d.ns + 1, rn, sn * u * r.shaft[0],
                           sn * u * r.shaft[1], sn * u * r.shaft[2]);
    }
    if (r.d.nFwake > rf) {
        const int64_t n = (int64_t)nbc * (r.d.nFwake - rf);
        hipLaunchKernelGGL(add_uniform_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           r.velF[slot].p, nbc, r.d.nFwake, 1, rf, -u * r.shaft[0],
                           -u * r.shaft[1], -u * r.shaft[2]);
    }
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

// getTransformAxis, libMath.f90:695-726 (column-major)
void transform_axis(double theta, const double *ax, double *T)
{
    const double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const double a0 = ax[0] / n, a1 = ax[1] / n, a2 = ax[2] / n;
    const double ct = std::cos(theta), st = std::sin(theta), om = 1.0 - ct;
    T[0] = ct + a0 * a0 * om;
    T[1] = a2 * st + a1 * a0 * om;
    T[2] = -a1 * st + a2 * a0 * om;
    T[3] = -a2 * st + a0 * a1 * om;
    T[4] = ct + a1 * a1 * om;
    T[5] = a0 * st + a2 * a1 * om;
    T[6] = a1 * st + a0 * a2 * om;
    T[7] = -a0 * st + a1 * a2 * om;
    T[8] = ct + a2 * a2 * om;
}

int copy_predicted(vlm_ctx *ctx, Rotor &r)
{
    const int64_t R = r.d.nNwake - (r.rowNear - 1);
    const int64_t F = r.d.nFwake - (r.rowFar - 1);
    const int64_t per = R * (r.d.ns + 1) + (r.d.ns + 1) + (F > 0 ? F : 0);
    const int64_t n = per * r.d.nbConvect;
    if (R <= 0 || n <= 0) return VLM_OK;
    hipLaunchKernelGGL(copy_pred_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(r));
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // namespace

// rotor_updatePrescribedWake, classdef.f90:5172-5217: the far filaments
// from rowFar (prescWakeGenNt = 0) or the last prescWakeGenNt + 1.
static int presc_update(vlm_ctx *ctx, Rotor &r, int pred)
{
    const int nF = r.d.nFwake;
    const int start = r.presc_gen_nt == 0 ? r.rowFar - 1 : nF - r.presc_gen_nt - 1;
    const int n = nF - start;
    if (start < 0 || n < 1) return VLM_OK;  // no far filament to anchor to
    if (std::fabs(r.shaft[0]) > kEps || std::fabs(r.shaft[1]) > kEps)
        return fail(ctx, VLM_ERR_ARG, "Prescribed far wake only implemented for shaft along Z-axis");
    const int t = pred ? 1 : 0;
    hipLaunchKernelGGL(presc_update_kernel, dim3(r.d.nbConvect), dim3(256), 0, ctx->stream,
                       view(r), pred, start, n, r.hub[0], r.hub[1], r.hub[2], r.presc_dpsi,
                       r.pf[t].p, r.ps[t].p);
    VLM_HIP(hipGetLastError());
    if (r.d.nbConvect == 1 && r.d.nb > 1) {
        if (r.d.nb > 16) return fail(ctx, VLM_ERR_ARG, "axisymmetric copy supports <= 16 blades");
        RotMats M;
        std::memset(&M, 0, sizeof(M));
        for (int b = 1; b < r.d.nb; ++b) transform_axis(2.0 * kPi / r.d.nb * b, r.shaft, M.T[b]);
        hipLaunchKernelGGL(presc_axisym_kernel, dim3(nblk((int64_t)(r.d.nb - 1) * PF_N)),
                           dim3(256), 0, ctx->stream, r.d.nb, M, r.hub[0], r.hub[1], r.hub[2],
                           r.pf[t].p, r.ps[t].p);
        VLM_HIP(hipGetLastError());
    }
    r.presc_built[t] = 1;
    return VLM_OK;
}

// rotor_mirrorWake of image r from its source (mode 0 C, 1 P, 2 A).
int mirror_wake(vlm_ctx *ctx, Rotor &r, int mode)
{
    if (r.image_of < 0 || r.image_of >= (int)ctx->rotors.size())
        return fail(ctx, VLM_ERR_STATE, "rotor is not an image (vlm_rotor_set_image)");
    Rotor &s = *ctx->rotors[r.image_of];
    if (r.d.nNwake <= 0) return VLM_OK;
    const int i0 = mode == 2 ? 0 : r.rowNear - 1, f0 = mode == 2 ? 0 : r.rowFar - 1;
    const int R = r.d.nNwake - i0, F = r.d.nFwake - f0;
    if (R < 0 || F < 0) return fail(ctx, VLM_ERR_STATE, "wake cursors out of range");
    const int64_t per = (int64_t)R * (r.d.ns + 1) + (r.d.ns + 1) + (int64_t)R * r.d.ns + F;
    hipLaunchKernelGGL(mirror_wake_kernel, dim3(nblk(per * r.d.nb)), dim3(256), 0, ctx->stream,
                       view(r), view(s), mode, r.image_axis, i0, f0);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int rotor_convect(vlm_ctx *ctx, Rotor &r, double dt, int pred)
{
    // an image's convection is the mirror of its source's wake
    if (r.image_of >= 0) return mirror_wake(ctx, r, pred ? 1 : 0);
    if (r.d.nNwake <= 0) return VLM_OK;
    RV v = view(r);
    int i0 = r.rowNear - 1, i1 = r.d.nNwake;
    if (pred && ctx->bug_compat) {
        i0 = 0;  // `do i = 1, rowNear, nNwake` runs for row 1 only
        i1 = 1;
    }
    if (i0 >= r.d.nNwake) i1 = i0;  // nothing shed yet
    const int64_t F = r.d.nFwake - (r.rowFar - 1);
    const int64_t per = (int64_t)(i1 - i0) * (r.d.ns + 1) + (F > 0 ? F : 0);
    if (per > 0) {
        hipLaunchKernelGGL(convect_kernel, dim3(nblk(per * r.d.nbConvect)), dim3(256), 0,
                           ctx->stream, v, pred, i0, i1, dt);
        VLM_HIP(hipGetLastError());
    }
    if (F > 1) {
        hipLaunchKernelGGL(far_continuity_kernel, dim3(nblk((F - 1) * r.d.nbConvect)), dim3(256),
                           0, ctx->stream, v, pred);
        VLM_HIP(hipGetLastError());
    }
    if (r.d.nbConvect == 1 && r.d.nb > 1) {
        if (r.d.nb > 16) return fail(ctx, VLM_ERR_ARG, "axisymmetric copy supports <= 16 blades");
        RotMats M;
        std::memset(&M, 0, sizeof(M));
        for (int b = 1; b < r.d.nb; ++b) {
            const double theta = 2.0 * kPi / r.d.nb * b;
            transform_axis(theta, r.shaft, M.T[b]);
        }
        const int64_t R = r.d.nNwake - (r.rowNear - 1);
        const int64_t per2 = (R > 0 ? R : 0) * (r.d.ns + 1) + (r.d.ns + 1) + (F > 0 ? F : 0) +
                             (R > 0 ? R : 0) * r.d.ns;
        hipLaunchKernelGGL(axisym_kernel, dim3(nblk(per2 * (r.d.nb - 1))), dim3(256), 0,
                           ctx->stream, v, pred, M, r.hub[0], r.hub[1], r.hub[2]);
        VLM_HIP(hipGetLastError());
    }
    // the prescribed far wake after the convection (classdef.f90:4826-4829)
    if (r.presc_nt > 0 && r.iter > r.presc_nt) VLM_TRY(presc_update(ctx, r, pred));
    return VLM_OK;
}

int wake_pass(vlm_ctx *ctx, int pred)
{
    std::vector<std::pair<int, int>> src;
    int64_t m = 0;
    for (int ir = 0; ir < (int)ctx->rotors.size(); ++ir) {
        src.push_back({ir, VLM_SET_WING});
        src.push_back({ir, pred ? VLM_SET_WAKE_PRED : VLM_SET_WAKE});
        m += target_count(*ctx->rotors[ir]);
    }
    if (m == 0) return VLM_OK;
    const int64_t nmax = staged_upper_bound(ctx, src);
    VLM_TRY(build_sources(ctx, src));
    VLM_HIP(ctx->tgt.reserve((size_t)3 * m));
    VLM_HIP(ctx->out.reserve((size_t)3 * m));
    int64_t off = 0;
    for (auto *rp : ctx->rotors) {
        const int64_t n = target_count(*rp);
        if (n == 0) continue;
        hipLaunchKernelGGL(gather_targets_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           view(*rp), pred, ctx->tgt.p, m, off);
        VLM_HIP(hipGetLastError());
        off += n;
    }
    VLM_TRY(eval_sources(ctx, ctx->tgt.p, m, m, ctx->out.p, 0, nmax));
    off = 0;
    for (auto *rp : ctx->rotors) {
        const int64_t n = target_count(*rp);
        if (n == 0) continue;
        hipLaunchKernelGGL(scatter_vel_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           view(*rp), pred ? VLM_VEL_PRED : VLM_VEL, ctx->out.p, m, off);
        VLM_HIP(hipGetLastError());
        off += n;
    }
    return VLM_OK;
}


// ---------------------------------------------------------------------
// Learned pass 2 (AM2 predictor stage): the MLP proposes the velocity on
// the predicted wake; a strided sample of targets is evaluated by the full
// Biot-Savart sum every step as judge and training data, and the whole
// pass is recomputed when the sampled error exceeds the tolerance.
// ---------------------------------------------------------------------
namespace {

__constant__ int kNbr[12][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {-1, 1},
                                {1, -1}, {1, 1}, {-2, 0}, {2, 0}, {0, -2}, {0, 2}};

// 45 features per P-pass target in gather order: 12 neighbour offsets on
// the predicted lattice (scaled by 1/lref) and three velocities scaled by
// 1/vref: pass-1 velocity, previous step's velocity, AB2 predictor velocity.
__global__ void features_kernel(RV v, float *X, int64_t off, double lref_inv,
                                const double *vref)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    float *x = X + (off + e) * 45;
    const double vinv = 1.0 / vref[0];
    const double *vs, *vp, *va;
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        const double *P = v.pnode + 3 * nidx(v, b, phys(v, i), j);
        for (int q = 0; q < 12; ++q) {
            const int ii = i + kNbr[q][0], jj = j + kNbr[q][1];
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (jj >= 0 && jj <= v.ns && ii >= v.rn - 1 && ii < v.nN) {
                const double *Q = (ii == v.rn - 1) ? v.ple + 3 * ((size_t)b * (v.ns + 1) + jj)
                                                   : v.pnode + 3 * nidx(v, b, phys(v, ii), jj);
                ox = (float)((Q[0] - P[0]) * lref_inv);
                oy = (float)((Q[1] - P[1]) * lref_inv);
                oz = (float)((Q[2] - P[2]) * lref_inv);
            }
            x[3 * q] = ox;
            x[3 * q + 1] = oy;
            x[3 * q + 2] = oz;
        }
        const size_t vi = 3 * vidx(v, b, i, j);
        vs = v.vel[VLM_VEL_STEP] + vi;
        vp = v.vel[VLM_VEL_PREV] + vi;
        va = v.vel[VLM_VEL] + vi;
    } else {
        const int i = v.rf + (int)(k - (int64_t)R * (v.ns + 1));
        for (int q = 0; q < 36; ++q) x[q] = 0.f;
        const size_t vi = 3 * ((size_t)b * v.nF + i);
        vs = v.velF[VLM_VEL_STEP] + vi;
        vp = v.velF[VLM_VEL_PREV] + vi;
        va = v.velF[VLM_VEL] + vi;
    }
    for (int d = 0; d < 3; ++d) {
        x[36 + d] = (float)(vs[d] * vinv);
        x[39 + d] = (float)(vp[d] * vinv);
        x[42 + d] = (float)(va[d] * vinv);
    }
}

// Pass-1 velocity of every P-pass target (gather order), SoA.
__global__ void gather_step_vel_kernel(RV v, double *o, int64_t stride, int64_t off)
{
    const int R = v.nN - v.rn, F = v.nF - v.rf;
    const int64_t per = (int64_t)R * (v.ns + 1) + F;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per * v.nbc) return;
    const int b = (int)(e / per);
    const int64_t k = e - b * per;
    const double *V;
    if (k < (int64_t)R * (v.ns + 1)) {
        const int i = v.rn + (int)(k / (v.ns + 1)), j = (int)(k % (v.ns + 1));
        V = v.vel[VLM_VEL_STEP] + 3 * vidx(v, b, i, j);
    } else {
        V = v.velF[VLM_VEL_STEP] + 3 * ((size_t)b * v.nF + v.rf + (k - (int64_t)R * (v.ns + 1)));
    }
    o[off + e] = V[0];
    o[stride + off + e] = V[1];
    o[2 * stride + off + e] = V[2];
}

// vref = RMS of the pass-1 speed over all targets (single block).
__global__ __launch_bounds__(1024) void vref_kernel(const double *vn, int64_t m, double *vref)
{
    __shared__ double sh[1024];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < m; k += blockDim.x)
        s += vn[k] * vn[k] + vn[m + k] * vn[m + k] + vn[2 * m + k] * vn[2 * m + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) vref[0] = fmax(sqrt(sh[0] / (double)(m > 0 ? m : 1)), 1e-12);
}

// v = v_pass1 + y * sigma * vref for every target, sigma = scal[4] the RMS
// of the replay buffer's targets: the network predicts the correction in
// units of its typical size.  With sigma = 0 (every recorded pass 2 equal
// to pass 1) the prediction is the pass-1 velocity itself.
__global__ void apply_mlp_kernel(const double *vn, const float *Y, int64_t m, const double *scal,
                                 double *out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const double s = scal[0] * scal[4];
    for (int d = 0; d < 3; ++d)
        out[d * m + e] = vn[d * m + e] + (s > 0.0 ? (double)Y[3 * e + d] * s : 0.0);
}

// sigma = RMS of the n x 3 buffer targets (single block, fixed order) into
// scal[4]; 0 for an empty buffer.
__global__ __launch_bounds__(1024) void target_scale_kernel(const float *bt, int64_t n,
                                                            double *scal)
{
    __shared__ double sh[1024];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < 3 * n; k += blockDim.x) s += (double)bt[k] * bt[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[4] = n > 0 ? sqrt(sh[0] / (double)(3 * n)) : 0.0;
}

// Minibatch targets in units of sigma (zero when sigma is zero).
__global__ void scale_targets_kernel(const float *bt, int64_t n3, const double *scal, float *out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    const double sg = scal[4];
    out[k] = sg > 0.0 ? (float)((double)bt[k] / sg) : 0.f;
}

// Sampled targets: e = phase + k*stride.  Gather their positions.
__global__ void gather_sample_kernel(const double *t, int64_t m, int64_t phase, int64_t stride,
                                     int64_t ms, double *ts)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ms) return;
    const int64_t e = phase + k * stride;
    ts[k] = t[e];
    ts[ms + k] = t[m + e];
    ts[2 * ms + k] = t[2 * m + e];
}

// Judge: per-sample squared error and squared exact speed; then write the
// exact values into the output and the training pair into the buffer.
// The same sample also scores the network-free substitute for pass 2, the
// linear extrapolation 2 v_n - v_(n-1) of the pass-1 velocity (with it the
// AM2 corrector reduces to AB2), so that the network's gain is measured on
// identical targets.
__global__ void judge_kernel(const double *vs, int64_t ms, const double *out, int64_t m,
                             int64_t phase, int64_t stride, const double *vn, const float *X,
                             const double *vref, double *err2, double *ref2, double *base2,
                             float *bx, float *bt, int64_t bcap, int64_t bhead)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ms) return;
    const int64_t e = phase + k * stride;
    double se = 0.0, sr = 0.0, sb = 0.0;
    const double inv = 1.0 / vref[0];
    const int64_t slot = (bhead + k) % bcap;
    for (int d = 0; d < 3; ++d) {
        const double a = vs[d * ms + k], p = out[d * m + e];
        const double ex = 2.0 * vn[d * m + e] - (double)X[45 * e + 39 + d] * vref[0];
        se += (a - p) * (a - p);
        sr += a * a;
        sb += (a - ex) * (a - ex);
        bt[3 * slot + d] = (float)((a - vn[d * m + e]) * inv);
    }
    for (int q = 0; q < 45; ++q) bx[45 * slot + q] = X[45 * e + q];
    err2[k] = se;
    ref2[k] = sr;
    base2[k] = sb;
}

__global__ void put_sample_kernel(const double *vs, int64_t ms, double *out, int64_t m,
                                  int64_t phase, int64_t stride)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ms) return;
    const int64_t e = phase + k * stride;
    for (int d = 0; d < 3; ++d) out[d * m + e] = vs[d * ms + k];
}

// Sums of three length-n arrays a, a+n, a+2n (one block, fixed order).
__global__ void sum3_kernel(const double *a, int64_t n, double *out)
{
    __shared__ double sh[3][256];
    double x[3] = {0.0, 0.0, 0.0};
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x)
        for (int c = 0; c < 3; ++c) x[c] += a[c * n + k];
    for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = x[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) out[c] = sh[c][0];
}

// Training pairs from a full pass: every stride-th target.
__global__ void full_pairs_kernel(const double *vexact, const double *vn, int64_t m,
                                  const float *X, const double *vref, int64_t stride, int64_t np,
                                  float *bx, float *bt, int64_t bcap, int64_t bhead)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const int64_t e = k * stride;
    const int64_t slot = (bhead + k) % bcap;
    const double inv = 1.0 / vref[0];
    for (int d = 0; d < 3; ++d) bt[3 * slot + d] = (float)((vexact[d * m + e] - vn[d * m + e]) * inv);
    for (int q = 0; q < 45; ++q) bx[45 * slot + q] = X[45 * e + q];
}

}  // namespace

// Full Biot-Savart pass 2 of the learned mode, plus training pairs from it
// (every stride-th target, a quarter of the replay buffer).
static int full_pass(vlm_ctx *ctx, int64_t m, int64_t nmax)
{
    Predictor &P = ctx->pred;
    VLM_TRY(eval_sources(ctx, ctx->tgt.p, m, m, ctx->out.p, 0, nmax));
    P.full += 1;
    const int64_t stride = (m + P.buf_cap / 4 - 1) / (P.buf_cap / 4);
    const int64_t np = (m + stride - 1) / stride;
    hipLaunchKernelGGL(full_pairs_kernel, dim3(nblk(np)), dim3(256), 0, ctx->stream, ctx->out.p,
                       P.vn.p, m, P.X.p, P.scal.p, stride, np, P.bx.p, P.bt.p, P.buf_cap,
                       P.buf_head);
    VLM_HIP(hipGetLastError());
    P.buf_head = (P.buf_head + np) % P.buf_cap;
    P.buf_n = P.buf_n + np < P.buf_cap ? P.buf_n + np : P.buf_cap;
    return VLM_OK;
}

int learned_pass(vlm_ctx *ctx)
{
    Predictor &P = ctx->pred;
    std::vector<std::pair<int, int>> src;
    int64_t m = 0;
    for (int ir = 0; ir < (int)ctx->rotors.size(); ++ir) {
        src.push_back({ir, VLM_SET_WING});
        src.push_back({ir, VLM_SET_WAKE_PRED});
        m += target_count(*ctx->rotors[ir]);
    }
    if (m == 0) return VLM_OK;
    const int64_t nmax = staged_upper_bound(ctx, src);
    VLM_TRY(build_sources(ctx, src));
    VLM_HIP(ctx->tgt.reserve((size_t)3 * m));
    VLM_HIP(ctx->out.reserve((size_t)3 * m));
    VLM_HIP(P.vn.reserve((size_t)3 * m));
    VLM_HIP(P.X.reserve((size_t)45 * m));
    VLM_HIP(P.Y.reserve((size_t)3 * m));
    VLM_HIP(P.scal.reserve(8));
    VLM_HIP(P.tb.reserve((size_t)3 * P.batch));
    if (P.bx.n < (size_t)45 * P.buf_cap) {
        VLM_HIP(P.bx.reserve((size_t)45 * P.buf_cap));
        VLM_HIP(P.bt.reserve((size_t)3 * P.buf_cap));
    }
    int64_t off = 0;
    for (auto *rp : ctx->rotors) {
        const int64_t n = target_count(*rp);
        if (n == 0) continue;
        hipLaunchKernelGGL(gather_targets_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           view(*rp), 1, ctx->tgt.p, m, off);
        hipLaunchKernelGGL(gather_step_vel_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           view(*rp), P.vn.p, m, off);
        off += n;
    }
    hipLaunchKernelGGL(vref_kernel, dim3(1), dim3(1024), 0, ctx->stream, P.vn.p, m, P.scal.p);
    off = 0;
    for (auto *rp : ctx->rotors) {
        const int64_t n = target_count(*rp);
        if (n == 0) continue;
        hipLaunchKernelGGL(features_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(*rp),
                           P.X.p, off, 1.0 / P.lref, P.scal.p);
        off += n;
    }
    // scale of the corrections seen so far (the buffer before this step)
    hipLaunchKernelGGL(target_scale_kernel, dim3(1), dim3(1024), 0, ctx->stream, P.bt.p,
                       P.buf_n, P.scal.p);
    VLM_HIP(hipGetLastError());
    const bool warm = P.steps < P.warmup;
    P.steps += 1;
    int64_t ms = 0, phase = 0, stride = 1;
    const double *judged_sums = nullptr;
    if (!warm) {
        VLM_TRY(mlp_forward_dev(ctx, m, P.X.p, P.Y.p));
        hipLaunchKernelGGL(apply_mlp_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, P.vn.p,
                           P.Y.p, m, P.scal.p, ctx->out.p);
        // judge on a strided sample with a rotating phase
        stride = P.stride < 1 ? 1 : P.stride;
        phase = (P.steps * 7) % stride < m ? (P.steps * 7) % stride : 0;
        ms = (m - phase + stride - 1) / stride;
        VLM_HIP(P.ts.reserve((size_t)3 * ms));
        VLM_HIP(P.vs.reserve((size_t)3 * ms));
        VLM_HIP(P.e2.reserve((size_t)3 * ms));
        hipLaunchKernelGGL(gather_sample_kernel, dim3(nblk(ms)), dim3(256), 0, ctx->stream,
                           ctx->tgt.p, m, phase, stride, ms, P.ts.p);
        VLM_TRY(eval_sources(ctx, P.ts.p, ms, ms, P.vs.p, 0, nmax));
        const int64_t take = ms < P.buf_cap ? ms : P.buf_cap;
        hipLaunchKernelGGL(judge_kernel, dim3(nblk(take)), dim3(256), 0, ctx->stream, P.vs.p,
                           take, ctx->out.p, m, phase, stride, P.vn.p, P.X.p, P.scal.p, P.e2.p,
                           P.e2.p + take, P.e2.p + 2 * take, P.bx.p, P.bt.p, P.buf_cap,
                           P.buf_head);
        hipLaunchKernelGGL(sum3_kernel, dim3(1), dim3(256), 0, ctx->stream, P.e2.p, take,
                           P.scal.p + 1);
        VLM_HIP(hipGetLastError());
        // the sums travel to pinned memory behind an event; the host waits
        // for them only after the training below is enqueued, so the GPU
        // trains while the host decides
        if (!ctx->judge_ev)
            VLM_HIP(hipEventCreateWithFlags(&ctx->judge_ev, hipEventDisableTiming));
        if (!ctx->judge_pin)
            VLM_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->judge_pin), 4 * sizeof(double)));
        VLM_HIP(hipMemcpyAsync(ctx->judge_pin, P.scal.p + 1, 3 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
        VLM_HIP(hipEventRecord(ctx->judge_ev, ctx->stream));
        judged_sums = ctx->judge_pin;
        P.buf_head = (P.buf_head + take) % P.buf_cap;
        P.buf_n = P.buf_n + take < P.buf_cap ? P.buf_n + take : P.buf_cap;
    } else {
        VLM_TRY(full_pass(ctx, m, nmax));
    }
    // Adam on minibatches: contiguous windows of the replay buffer at a
    // rotating offset (on a judged step before the decision, so the pairs
    // of a fallback pass are trained on from the next step)
    for (int it = 0; it < P.train_iters && P.buf_n > 0; ++it) {
        const int64_t B = P.buf_n < P.batch ? P.buf_n : P.batch;
        const int64_t start = (P.train_calls * 7919) % (P.buf_n - B + 1);
        P.train_calls += 1;
        hipLaunchKernelGGL(scale_targets_kernel, dim3(nblk(3 * B)), dim3(256), 0, ctx->stream,
                           P.bt.p + 3 * start, 3 * B, P.scal.p, P.tb.p);
        VLM_HIP(hipGetLastError());
        VLM_TRY(mlp_train_dev(ctx, B, P.bx.p + 45 * start, P.tb.p, 1, P.lr, nullptr,
                              it == P.train_iters - 1));
    }
    if (judged_sums) {
        VLM_HIP(hipEventSynchronize(ctx->judge_ev));
        const double h0 = judged_sums[0], h1 = judged_sums[1], h2 = judged_sums[2];
        P.last_err = h1 > 0 ? std::sqrt(h0 / h1) : 0.0;
        P.sum_err2 += h0;
        P.sum_ref2 += h1;
        P.sum_base2 += h2;
        P.judged += 1;
        if (P.last_err > P.tol) {
            P.fallbacks += 1;
            VLM_TRY(full_pass(ctx, m, nmax));
        } else {
            hipLaunchKernelGGL(put_sample_kernel, dim3(nblk(ms)), dim3(256), 0, ctx->stream,
                               P.vs.p, ms, ctx->out.p, m, phase, stride);
            VLM_HIP(hipGetLastError());
            P.learned += 1;
        }
    }
    off = 0;
    for (auto *rp : ctx->rotors) {
        const int64_t n = target_count(*rp);
        if (n == 0) continue;
        hipLaunchKernelGGL(scatter_vel_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                           view(*rp), VLM_VEL_PRED, ctx->out.p, m, off);
        off += n;
    }
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_rotor_create(vlm_ctx *ctx, const vlm_rotor_desc *d, double spanwiseCore,
                     const double *scv, int *ir)
{
    if (!ctx || !d || !ir) return VLM_ERR_ARG;
    if (d->nb <= 0 || d->nc <= 0 || d->ns <= 0 || d->nNwake < 0 || d->nFwake < 0)
        return fail(ctx, VLM_ERR_ARG, "rotor sizes must be positive");
    if (d->nbConvect != 1 && d->nbConvect != d->nb)
        return fail(ctx, VLM_ERR_ARG, "nbConvect must be 1 or nb");
    if (d->nNwake > 0 && d->nNwake < 2)
        return fail(ctx, VLM_ERR_ARG, "at least 2 near wake rows are mandatory");
    if (d->nNwake > 0 && (d->rollupStart < 1 || d->rollupEnd > d->ns || d->rollupStart > d->rollupEnd + 1))
        return fail(ctx, VLM_ERR_ARG, "rollup span range outside 1..ns");
    if (d->nNwake > 0 && !scv) return fail(ctx, VLM_ERR_ARG, "streamwiseCoreVec missing");
    VLM_HIP(hipSetDevice(ctx->device));
    Rotor *r = new Rotor();
    r->d = *d;
    r->nring = d->nb * d->nc * d->ns;
    const size_t nb = d->nb, nN = d->nNwake, ns = d->ns, nF = d->nFwake;
    hipError_t e = hipSuccess;
    auto R = [&](DBuf<double> &b, size_t n) {
        if (e == hipSuccess) e = b.reserve(n);
    };
    R(r->wc, 12 * (size_t)r->nring);
    R(r->wrc, 4 * (size_t)r->nring);
    R(r->wg, r->nring);
    R(r->node, 3 * nb * nN * (ns + 1));
    R(r->pnode, 3 * nb * nN * (ns + 1));
    R(r->le, 3 * nb * (ns + 1));
    R(r->ple, 3 * nb * (ns + 1));
    R(r->gam, nb * nN * ns);
    R(r->rc4, 4 * nb * nN * ns);
    R(r->age, nb * nN * ns);
    R(r->ffc, 6 * nb * nF);
    R(r->pffc, 6 * nb * nF);
    R(r->fgam, nb * nF);
    R(r->frc, nb * nF);
    R(r->frc0, nb * nF);
    R(r->fl0, nb * nF);
    R(r->flc, nb * nF);
    R(r->fage, nb * nF);
    R(r->ftmp, 12 * nb * nF);
    for (int k = 0; k < NVEL; ++k) {
        R(r->velN[k], r->nvn());
        R(r->velF[k], r->nvf());
    }
    for (int t = 0; t < 2; ++t) {
        R(r->pf[t], 6 * nb * PF_N);
        R(r->ps[t], 4 * nb);
    }
    R(r->pan, (size_t)PANEL_N * r->nring);
    R(r->sec, (size_t)SECTION_N * nb * ns);
    R(r->bla, (size_t)BLADE_N * nb);
    R(r->cps, 3 * (size_t)r->nring);
    R(r->vcp, 6 * (size_t)r->nring);
    R(r->rhs, r->nring);
    R(r->vld, 3 * (size_t)r->nring);
    R(r->gprev, r->nring);
    R(r->lout, (size_t)LOADS_SEC_N * nb * ns + LOADS_ROTOR_N);
    if (e != hipSuccess) {
        delete r;
        return hip_fail(ctx, e, "rotor allocation");
    }
    // zero everything (the reference zero-initialises every velocity array,
    // classdef.f90:3737-3820, and all wake circulations, 3830-3831)
    std::vector<DBuf<double> *> all = {&r->wc,   &r->wrc,  &r->wg,   &r->node, &r->pnode,
                                       &r->le,   &r->ple,  &r->gam,  &r->rc4,  &r->age,
                                       &r->ffc,  &r->pffc, &r->fgam, &r->frc,  &r->frc0,
                                       &r->fl0,  &r->flc,  &r->fage, &r->ftmp};
    for (int k = 0; k < NVEL; ++k) {
        all.push_back(&r->velN[k]);
        all.push_back(&r->velF[k]);
    }
    for (int t = 0; t < 2; ++t) {
        all.push_back(&r->pf[t]);
        all.push_back(&r->ps[t]);
    }
    for (auto *b : {&r->pan, &r->sec, &r->bla, &r->cps, &r->vcp, &r->rhs, &r->vld, &r->gprev,
                    &r->lout})
        all.push_back(b);
    for (auto *b : all)
        hipLaunchKernelGGL(fill_kernel, dim3(nblk((int64_t)b->n)), dim3(256), 0, ctx->stream,
                           b->p, (int64_t)b->n, 0.0);
    // initial core radii, classdef.f90:3826-3850
    if (nN > 0) {
        std::vector<double> rc4(4 * nb * nN * ns);
        for (size_t b = 0; b < nb; ++b)
            for (size_t i = 0; i < nN; ++i)
                for (size_t j = 0; j < ns; ++j) {
                    double *q = &rc4[4 * ((b * nN + i) * ns + j)];
                    q[0] = scv[j];
                    q[1] = spanwiseCore;
                    q[2] = scv[j + 1];
                    q[3] = spanwiseCore;
                }
        VLM_HIP(hipMemcpyAsync(r->rc4.p, rc4.data(), rc4.size() * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
        if (nF > 0) {
            std::vector<double> frc(nb * nF, scv[ns]);
            VLM_HIP(hipMemcpyAsync(r->frc.p, frc.data(), frc.size() * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
            VLM_HIP(hipMemcpyAsync(r->frc0.p, frc.data(), frc.size() * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
        }
        VLM_TRY(sync_stream(ctx));
    }
    r->rowNear = (int)nN + 1;
    r->rowFar = (int)nF + 1;
    ctx->rotors.push_back(r);
    ctx->aic.push_back(nullptr);
    ctx->lu.push_back(nullptr);
    ctx->piv.push_back(nullptr);
    ctx->aic_n.push_back(0);
    *ir = (int)ctx->rotors.size() - 1;
    return VLM_OK;
}

int vlm_rotor_set_wing(vlm_ctx *ctx, int ir, const double *corners, const double *rc,
                       const double *gam)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (corners) VLM_TRY(h2d(ctx, r->wc.p, corners, sizeof(double) * 12 * r->nring));
    if (rc) VLM_TRY(h2d(ctx, r->wrc.p, rc, sizeof(double) * 4 * r->nring));
    if (gam) VLM_TRY(h2d(ctx, r->wg.p, gam, sizeof(double) * r->nring));
    return VLM_OK;
}

int vlm_rotor_set_axis(vlm_ctx *ctx, int ir, const double *shaft, const double *hub)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r || !shaft || !hub) return ctx ? fail(ctx, VLM_ERR_ARG, "bad rotor axis") : VLM_ERR_ARG;
    std::memcpy(r->shaft, shaft, sizeof(r->shaft));
    std::memcpy(r->hub, hub, sizeof(r->hub));
    return VLM_OK;
}

int vlm_rotor_set_init_wake_vel(vlm_ctx *ctx, int ir, double initWakeVel, int initWakeVelNt)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (!std::isfinite(initWakeVel)) return fail(ctx, VLM_ERR_ARG, "initWakeVel not finite");
    r->init_wake_vel = initWakeVel;
    r->init_wake_vel_nt = initWakeVelNt;
    return VLM_OK;
}

int vlm_rotor_init_cursors(vlm_ctx *ctx, int ir)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    r->rowFar = r->d.nFwake + 1;
    r->rowNear = r->d.nNwake + 1;
    return VLM_OK;
}

int vlm_rotor_begin_step(vlm_ctx *ctx, int ir, int iter)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    r->iter = iter;
    r->rowNear = r->rowNear - 1 > 1 ? r->rowNear - 1 : 1;
    if (iter > r->d.nNwake) r->rowFar = r->rowFar - 1 > 1 ? r->rowFar - 1 : 1;
    return VLM_OK;
}

int vlm_rotor_cursors(vlm_ctx *ctx, int ir, int *rowNear, int *rowFar)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (rowNear) *rowNear = r->rowNear;
    if (rowFar) *rowFar = r->rowFar;
    return VLM_OK;
}

int vlm_rotor_assignshed(vlm_ctx *ctx, int ir, int edge)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (r->d.nNwake <= 0) return VLM_OK;
    if (edge != VLM_EDGE_LE && edge != VLM_EDGE_TE)
        return fail(ctx, VLM_ERR_ARG, "edge must be VLM_EDGE_LE or VLM_EDGE_TE");
    int row;
    if (edge == VLM_EDGE_LE) {
        if (r->rowNear > r->d.nNwake) return fail(ctx, VLM_ERR_STATE, "assignshed LE before first step");
        row = r->rowNear - 1;
    } else {
        row = (r->rowNear - 1 > 1 ? r->rowNear - 1 : 1) - 1;
    }
    hipLaunchKernelGGL(assignshed_kernel, dim3(nblk((int64_t)r->d.nb * (r->d.ns + 1))), dim3(256),
                       0, ctx->stream, view(*r), edge == VLM_EDGE_TE ? 1 : 0, row);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_age_wake(vlm_ctx *ctx, int ir, double dt)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (r->d.nNwake <= 0 || r->rowNear > r->d.nNwake) return VLM_OK;
    const int64_t R = r->d.nNwake - (r->rowNear - 1);
    const int64_t F = r->d.nFwake - (r->rowFar - 1);
    const int64_t n = (R * r->d.ns + (F > 0 ? F : 0)) * r->d.nb;
    hipLaunchKernelGGL(age_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(*r), dt);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_dissipate_wake(vlm_ctx *ctx, int ir, double dt, double kinematicVisc,
                             double apparentViscCoeff, double decayCoeff)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (r->d.nNwake <= 0 || r->rowNear > r->d.nNwake) return VLM_OK;
    const double oseen = 1.2564;
    const double inc = 4.0 * oseen * apparentViscCoeff * kinematicVisc * dt;
    const double decay = std::exp(-decayCoeff * dt);
    const int64_t R = r->d.nNwake - (r->rowNear - 1);
    const int64_t F = r->d.nFwake - (r->rowFar - 1);
    const int64_t n = (R * r->d.ns + (F > 0 ? F : 0)) * r->d.nb;
    hipLaunchKernelGGL(dissipate_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(*r), inc,
                       decay);
    if (R > 1)
        hipLaunchKernelGGL(dissipate_vf4_kernel, dim3(nblk((R - 1) * r->d.ns * r->d.nb)), dim3(256),
                           0, ctx->stream, view(*r));
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_strain_wake(vlm_ctx *ctx, int ir)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    const int64_t F = r->d.nFwake - (r->rowFar - 1);
    if (r->d.nNwake <= 0 || F <= 0) return VLM_OK;
    hipLaunchKernelGGL(strain_kernel, dim3(nblk(F * r->d.nb)), dim3(256), 0, ctx->stream,
                       view(*r));
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

// Every device array of a rotor's state, in a fixed order (restart files).
static std::vector<DBuf<double> *> state_bufs(Rotor &r)
{
    std::vector<DBuf<double> *> v = {&r.wc,  &r.wrc, &r.wg,  &r.node, &r.pnode, &r.le,
                                     &r.ple, &r.gam, &r.rc4, &r.age,  &r.ffc,   &r.pffc,
                                     &r.fgam, &r.frc, &r.frc0, &r.fl0, &r.flc,  &r.fage};
    for (int k = 0; k < NVEL; ++k) {
        v.push_back(&r.velN[k]);
        v.push_back(&r.velF[k]);
    }
    for (int t = 0; t < 2; ++t) {
        v.push_back(&r.pf[t]);
        v.push_back(&r.ps[t]);
    }
    v.push_back(&r.gprev);  // gamPrev of the loads (blade_calc_force)
    return v;
}

static int64_t state_bytes(Rotor &r)
{
    int64_t n = 4 * (int64_t)sizeof(int64_t);
    for (auto *b : state_bufs(r)) n += (int64_t)(b->n * sizeof(double));
    return n;
}

int vlm_rotor_state_bytes(vlm_ctx *ctx, int ir, int64_t *bytes)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r || !bytes) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor or null pointer")
                                 : VLM_ERR_ARG;
    *bytes = state_bytes(*r);
    return VLM_OK;
}

int vlm_rotor_get_state(vlm_ctx *ctx, int ir, void *buf, int64_t bytes)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r || !buf) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor or null buffer")
                               : VLM_ERR_ARG;
    if (bytes != state_bytes(*r)) return fail(ctx, VLM_ERR_ARG, "state size mismatch");
    VLM_DEVICE(ctx);
    VLM_TRY(sync_stream(ctx));
    int64_t *h = static_cast<int64_t *>(buf);
    h[0] = r->rowNear;
    h[1] = r->rowFar;
    h[2] = r->head;
    h[3] = r->presc_built[0] | (r->presc_built[1] << 1);
    char *p = static_cast<char *>(buf) + 4 * sizeof(int64_t);
    for (auto *b : state_bufs(*r)) {
        if (b->n) VLM_HIP(hipMemcpy(p, b->p, b->n * sizeof(double), hipMemcpyDeviceToHost));
        p += b->n * sizeof(double);
    }
    return VLM_OK;
}

int vlm_rotor_set_state(vlm_ctx *ctx, int ir, const void *buf, int64_t bytes)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r || !buf) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor or null buffer")
                               : VLM_ERR_ARG;
    if (bytes != state_bytes(*r)) return fail(ctx, VLM_ERR_ARG, "state size mismatch");
    const int64_t *h = static_cast<const int64_t *>(buf);
    if (h[0] < 1 || h[0] > r->d.nNwake + 1 || h[1] < 1 || h[1] > r->d.nFwake + 1 || h[2] < 0 ||
        h[2] >= (r->d.nNwake > 0 ? r->d.nNwake : 1))
        return fail(ctx, VLM_ERR_ARG, "state cursors out of range");
    VLM_DEVICE(ctx);
    VLM_TRY(sync_stream(ctx));
    const char *p = static_cast<const char *>(buf) + 4 * sizeof(int64_t);
    for (auto *b : state_bufs(*r)) {
        if (b->n) VLM_HIP(hipMemcpy(b->p, p, b->n * sizeof(double), hipMemcpyHostToDevice));
        p += b->n * sizeof(double);
    }
    r->rowNear = (int)h[0];
    r->rowFar = (int)h[1];
    r->head = (int)h[2];
    r->presc_built[0] = (int)(h[3] & 1);
    r->presc_built[1] = (int)((h[3] >> 1) & 1);
    return VLM_OK;
}

int vlm_rotor_burst_wake(vlm_ctx *ctx, int ir, double skewLimit, double largeCoreRadius)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    const int64_t F = r->d.nFwake - (r->rowFar - 1);
    if (r->d.nNwake <= 0 || F <= 0) return VLM_OK;
    hipLaunchKernelGGL(burst_kernel, dim3(nblk(F * r->d.nb)), dim3(256), 0, ctx->stream,
                       view(*r), skewLimit, largeCoreRadius);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_rollup(vlm_ctx *ctx, int ir)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (r->d.nNwake <= 0) return VLM_OK;
    int fi = r->rowFar - 2;  // rowFarNext - 1, 0-based
    if (r->d.nFwake > 0 && fi < 0) {
        const int64_t n = (int64_t)r->d.nb * r->d.nFwake;
        hipLaunchKernelGGL(far_stage_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(*r),
                           r->ftmp.p);
        hipLaunchKernelGGL(far_shift_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream, view(*r),
                           r->ftmp.p);
        VLM_HIP(hipGetLastError());
        fi = 0;
    }
    hipLaunchKernelGGL(rollup_kernel, dim3(1), dim3(64), 0, ctx->stream, view(*r),
                       r->d.rollupStart - 1, r->d.rollupEnd - 1, r->d.rollupMin,
                       r->d.suppressFwake, fi);
    VLM_HIP(hipGetLastError());
    // shiftwake: logical row i+1 <- row i, new row 0 duplicates row 1
    r->head = (r->head + r->d.nNwake - 1) % r->d.nNwake;
    hipLaunchKernelGGL(shift_copy_kernel, dim3(nblk((int64_t)r->d.nb * (r->d.ns + 1))), dim3(256),
                       0, ctx->stream, view(*r));
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int vlm_rotor_convectwake(vlm_ctx *ctx, int ir, double dt, int wakeType)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (wakeType != VLM_WAKE_C && wakeType != VLM_WAKE_P)
        return fail(ctx, VLM_ERR_ARG, "wakeType must be VLM_WAKE_C or VLM_WAKE_P");
    return rotor_convect(ctx, *r, dt, wakeType == VLM_WAKE_P);
}

int vlm_rotor_induced_velocity(vlm_ctx *ctx, int ir, int set, int64_t m, const double *p,
                               double *v, int accumulate)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (set < VLM_SET_WING || set > VLM_SET_WAKE_PRED) return fail(ctx, VLM_ERR_ARG, "bad set");
    if (m <= 0) return VLM_OK;
    if (!p || !v) return fail(ctx, VLM_ERR_ARG, "null buffer");
    std::vector<std::pair<int, int>> src = {{ir, set}};
    const int64_t nmax = staged_upper_bound(ctx, src);
    VLM_TRY(build_sources(ctx, src));
    VLM_HIP(ctx->tgt.reserve((size_t)3 * m));
    VLM_HIP(ctx->out.reserve((size_t)3 * m));
    VLM_TRY(points_to_device(ctx, m, p, ctx->tgt.p, m));
    if (accumulate) VLM_TRY(points_to_device(ctx, m, v, ctx->out.p, m));
    VLM_TRY(eval_sources(ctx, ctx->tgt.p, m, m, ctx->out.p, accumulate, nmax));
    return points_to_host(ctx, m, ctx->out.p, m, v);
}

static int induced_terms(vlm_ctx *ctx, int nterms, const int *ir, const int *set,
                         const int *flags, int64_t m, const double *p, double *v,
                         int accumulate, bool deferred)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (nterms < 0 || (nterms > 0 && (!ir || !set || !flags)))
        return fail(ctx, VLM_ERR_ARG, "bad term list");
    for (int k = 0; k < nterms; ++k) {
        if (!get_rotor(ctx, ir[k])) return fail(ctx, VLM_ERR_ARG, "unknown rotor");
        if (set[k] < VLM_SET_WING || set[k] > VLM_SET_WAKE_PRED)
            return fail(ctx, VLM_ERR_ARG, "bad set");
    }
    if (m <= 0) return VLM_OK;
    if (!p || !v) return fail(ctx, VLM_ERR_ARG, "null buffer");
    VLM_HIP(ctx->tgt.reserve((size_t)3 * m));
    VLM_HIP(ctx->out.reserve((size_t)3 * m));
    VLM_TRY(points_to_device(ctx, m, p, ctx->tgt.p, m));
    if (accumulate) {
        VLM_TRY(points_to_device(ctx, m, v, ctx->out.p, m));
    } else {
        hipLaunchKernelGGL(fill_kernel, dim3(nblk(3 * m)), dim3(256), 0, ctx->stream,
                           ctx->out.p, 3 * m, 0.0);
        VLM_HIP(hipGetLastError());
    }
    for (int k = 0; k < nterms; ++k) {
        if (flags[k] & VLM_TERM_NEGATE) {
            hipLaunchKernelGGL(negate_kernel, dim3(nblk(3 * m)), dim3(256), 0, ctx->stream,
                               ctx->out.p, 3 * m);
            VLM_HIP(hipGetLastError());
        }
        std::vector<std::pair<int, int>> src = {{ir[k], set[k]}};
        const int64_t nmax = staged_upper_bound(ctx, src);
        VLM_TRY(build_sources(ctx, src));
        VLM_TRY(eval_sources(ctx, ctx->tgt.p, m, m, ctx->out.p, 1, nmax));
    }
    return points_to_host(ctx, m, ctx->out.p, m, v, deferred);
}

}  // extern "C"

// The ordered term sum of induced_terms at m device targets (SoA, stride
// m) into a device SoA buffer out (overwritten): the right-hand side and
// the loads keep their collocation velocities on the device.
int vlm::induced_terms_soa(vlm_ctx *ctx, int nterms, const int *ir, const int *set,
                           const int *flags, int64_t m, const double *tgt, double *out)
{
    for (int k = 0; k < nterms; ++k) {
        if (!get_rotor(ctx, ir[k])) return fail(ctx, VLM_ERR_ARG, "unknown rotor");
        if (set[k] < VLM_SET_WING || set[k] > VLM_SET_WAKE_PRED)
            return fail(ctx, VLM_ERR_ARG, "bad set");
    }
    if (m <= 0) return VLM_OK;
    hipLaunchKernelGGL(fill_kernel, dim3(nblk(3 * m)), dim3(256), 0, ctx->stream, out, 3 * m,
                       0.0);
    VLM_HIP(hipGetLastError());
    for (int k = 0; k < nterms; ++k) {
        if (flags[k] & VLM_TERM_NEGATE) {
            hipLaunchKernelGGL(negate_kernel, dim3(nblk(3 * m)), dim3(256), 0, ctx->stream, out,
                               3 * m);
            VLM_HIP(hipGetLastError());
        }
        std::vector<std::pair<int, int>> src = {{ir[k], set[k]}};
        const int64_t nmax = staged_upper_bound(ctx, src);
        VLM_TRY(build_sources(ctx, src));
        VLM_TRY(eval_sources(ctx, tgt, m, m, out, 1, nmax));
    }
    return VLM_OK;
}

vlm::Rotor *vlm::rotor_of(vlm_ctx *ctx, int ir) { return get_rotor(ctx, ir); }

extern "C" {

int vlm_induced_velocity_terms(vlm_ctx *ctx, int nterms, const int *ir, const int *set,
                               const int *flags, int64_t m, const double *p, double *v,
                               int accumulate)
{
    return induced_terms(ctx, nterms, ir, set, flags, m, p, v, accumulate, false);
}

int vlm_induced_velocity_terms_deferred(vlm_ctx *ctx, int nterms, const int *ir,
                                        const int *set, const int *flags, int64_t m,
                                        const double *p, double *v, int accumulate)
{
    return induced_terms(ctx, nterms, ir, set, flags, m, p, v, accumulate, true);
}

int vlm_wake_velocity(vlm_ctx *ctx, int wakeType)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (wakeType != VLM_WAKE_C && wakeType != VLM_WAKE_P)
        return fail(ctx, VLM_ERR_ARG, "wakeType must be VLM_WAKE_C or VLM_WAKE_P");
    return wake_pass(ctx, wakeType == VLM_WAKE_P);
}

int vlm_convect(vlm_ctx *ctx, int scheme, int iter, double dt)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (scheme < VLM_SCHEME_FT1 || scheme > VLM_SCHEME_AM4)
        return fail(ctx, VLM_ERR_ARG, "fdScheme must be 0..5");
    VLM_TRY(wake_pass(ctx, 0));
    // lifting rotors convect; an image mirrors its source's wake at the
    // same place in the rotor order (rotor_convect does the mirror), and
    // takes no part in the velocity algebra
    auto active = [](Rotor *r) { return r->d.nNwake > 0 && r->image_of < 0; };
    auto image = [](Rotor *r) { return r->d.nNwake > 0 && r->image_of >= 0; };
    auto init_vel = [&](bool pass2) -> int {
        for (auto *r : ctx->rotors)
            if (active(r)) VLM_TRY(add_init_wake_vel(ctx, *r, iter, pass2));
        return VLM_OK;
    };
    VLM_TRY(init_vel(false));
    if (scheme == VLM_SCHEME_AM3 || scheme == VLM_SCHEME_AM4) {
        // start-up steps convect with pass 1 and keep it as history.  AM3's
        // first start-up branch tests iter == 0, which the time loop never
        // reaches (main.f90:1118), so its step 1 runs the full scheme on the
        // zero-initialised histories.
        const bool am3 = scheme == VLM_SCHEME_AM3;
        int start = -1;
        if (am3 && iter == 2) start = AM_START2;
        if (!am3 && iter >= 1 && iter <= 3) start = AM_START1 + (iter - 1);
        if (start >= 0) {
            for (auto *r : ctx->rotors)
                if (active(r)) {
                    VLM_TRY(rotor_convect(ctx, *r, dt, 0));
                    VLM_TRY(launch_am(ctx, *r, start));
                } else if (image(r)) {
                    VLM_TRY(rotor_convect(ctx, *r, dt, 0));
                }
            return VLM_OK;
        }
        for (auto *r : ctx->rotors)
            if (active(r)) {
                VLM_TRY(copy_predicted(ctx, *r));
                VLM_TRY(launch_am(ctx, *r, am3 ? AM3_PRED : AM4_PRED));
                VLM_TRY(rotor_convect(ctx, *r, dt, 1));
            } else if (image(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 1));
            }
        VLM_TRY(wake_pass(ctx, 1));
        VLM_TRY(init_vel(true));
        for (auto *r : ctx->rotors)
            if (active(r)) {
                VLM_TRY(launch_am(ctx, *r, am3 ? AM3_CORR : AM4_CORR));
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
                VLM_TRY(launch_am(ctx, *r, am3 ? AM3_ROTATE : AM4_ROTATE));
            } else if (image(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            }
        return VLM_OK;
    }
    if (scheme == VLM_SCHEME_FT1) {
        for (auto *r : ctx->rotors)
            if (active(r) || image(r)) VLM_TRY(rotor_convect(ctx, *r, dt, 0));
        return VLM_OK;
    }
    if (scheme == VLM_SCHEME_PC2) {
        for (auto *r : ctx->rotors)
            if (active(r)) {
                VLM_TRY(copy_predicted(ctx, *r));
                VLM_TRY(rotor_convect(ctx, *r, dt, 1));
            } else if (image(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 1));
            }
        VLM_TRY(wake_pass(ctx, 1));
        VLM_TRY(init_vel(true));
        for (auto *r : ctx->rotors) {
            if (image(r)) VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            if (!active(r)) continue;
            const int rows = r->d.nNwake - (r->rowNear - 1);
            const int frows = r->d.nFwake - (r->rowFar - 1);
            const int nbc = r->d.nbConvect;
            if (rows > 0) {
                const int64_t n = (int64_t)nbc * rows * (r->d.ns + 1) * 3;
                hipLaunchKernelGGL(order2_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                                   r->velN[VLM_VEL].p, r->velN[VLM_VEL_PRED].p,
                                   r->velN[VLM_VEL_STEP].p, nbc, rows, r->d.nNwake, r->d.ns + 1,
                                   r->rowNear - 1);
                hipLaunchKernelGGL(copy_window_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                                   r->velN[VLM_VEL_STEP].p, r->velN[VLM_VEL].p, nbc, rows,
                                   r->d.nNwake, r->d.ns + 1, r->rowNear - 1);
            }
            if (frows > 0) {
                const int64_t n = (int64_t)nbc * frows * 3;
                hipLaunchKernelGGL(order2_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                                   r->velF[VLM_VEL].p, r->velF[VLM_VEL_PRED].p,
                                   r->velF[VLM_VEL_STEP].p, nbc, frows, r->d.nFwake, 1,
                                   r->rowFar - 1);
                hipLaunchKernelGGL(copy_window_kernel, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                                   r->velF[VLM_VEL_STEP].p, r->velF[VLM_VEL].p, nbc, frows,
                                   r->d.nFwake, 1, r->rowFar - 1);
            }
            VLM_HIP(hipGetLastError());
            VLM_TRY(rotor_convect(ctx, *r, dt, 0));
        }
        return VLM_OK;
    }
    if (iter == 1) {
        for (auto *r : ctx->rotors)
            if (active(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
                VLM_TRY(launch_scheme(ctx, *r, 4));
            } else if (image(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            }
        return VLM_OK;
    }
    if (scheme == VLM_SCHEME_AB2) {
        for (auto *r : ctx->rotors)
            if (active(r)) {
                VLM_TRY(launch_scheme(ctx, *r, 0));
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            } else if (image(r)) {
                VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            }
        return VLM_OK;
    }
    // AM2
    for (auto *r : ctx->rotors)
        if (active(r)) {
            VLM_TRY(copy_predicted(ctx, *r));
            VLM_TRY(launch_scheme(ctx, *r, 1));
            VLM_TRY(rotor_convect(ctx, *r, dt, 1));
        } else if (image(r)) {
            VLM_TRY(rotor_convect(ctx, *r, dt, 1));
        }
    if (ctx->pred.mode == 1) VLM_TRY(learned_pass(ctx));
    else VLM_TRY(wake_pass(ctx, 1));
    VLM_TRY(init_vel(true));
    for (auto *r : ctx->rotors)
        if (active(r)) {
            VLM_TRY(launch_scheme(ctx, *r, 2));
            VLM_TRY(rotor_convect(ctx, *r, dt, 0));
            VLM_TRY(launch_scheme(ctx, *r, 3));
        } else if (image(r)) {
            VLM_TRY(rotor_convect(ctx, *r, dt, 0));
        }
    return VLM_OK;
}

int vlm_rotor_set_image(vlm_ctx *ctx, int ir, int source, int imagePlane)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    Rotor *s = get_rotor(ctx, source);
    if (!s || source == ir) return fail(ctx, VLM_ERR_ARG, "bad source rotor of the image");
    if (imagePlane < 1 || imagePlane > 3) return fail(ctx, VLM_ERR_ARG, "imagePlane must be 1..3");
    if (s->image_of >= 0) return fail(ctx, VLM_ERR_ARG, "the source of an image is an image");
    for (auto *q : ctx->rotors)
        if (q->image_of == ir) return fail(ctx, VLM_ERR_ARG, "rotor is the source of an image");
    if (r->d.nb != s->d.nb || r->d.nc != s->d.nc || r->d.ns != s->d.ns ||
        r->d.nNwake != s->d.nNwake || r->d.nFwake != s->d.nFwake)
        return fail(ctx, VLM_ERR_ARG, "an image needs its source's nb, nc, ns, nNwake and nFwake");
    r->image_of = source;
    r->image_axis = imagePlane - 1;
    return VLM_OK;
}

int vlm_rotor_mirror_gamma(vlm_ctx *ctx, int ir)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (r->image_of < 0) return fail(ctx, VLM_ERR_STATE, "rotor is not an image (vlm_rotor_set_image)");
    VLM_DEVICE(ctx);
    const Rotor &s = *ctx->rotors[r->image_of];
    if (r->nring > 0) {
        hipLaunchKernelGGL(mirror_gamma_kernel, dim3(nblk(r->nring)), dim3(256), 0, ctx->stream,
                           s.wg.p, r->wg.p, r->nring);
        VLM_HIP(hipGetLastError());
    }
    return VLM_OK;
}

int vlm_rotor_mirror_wake(vlm_ctx *ctx, int ir, int wakeType)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (wakeType != VLM_WAKE_C && wakeType != VLM_WAKE_P && wakeType != VLM_WAKE_A)
        return fail(ctx, VLM_ERR_ARG, "wakeType must be VLM_WAKE_C, VLM_WAKE_P or VLM_WAKE_A");
    VLM_DEVICE(ctx);
    return mirror_wake(ctx, *r, wakeType == VLM_WAKE_A ? 2 : wakeType == VLM_WAKE_P ? 1 : 0);
}

int vlm_rotor_set_prescribed_wake(vlm_ctx *ctx, int ir, int prescWakeNt, int prescWakeGenNt,
                                  double deltaPsi)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (prescWakeNt < 0 || prescWakeGenNt < 0)
        return fail(ctx, VLM_ERR_ARG, "prescWakeNt and prescWakeGenNt must be >= 0");
    if (prescWakeNt > 0 && (r->d.nFwake < 1 || prescWakeGenNt >= r->d.nFwake))
        return fail(ctx, VLM_ERR_ARG,
                    "the prescribed far wake needs a far wake longer than prescWakeGenNt");
    r->presc_nt = prescWakeNt;
    r->presc_gen_nt = prescWakeGenNt;
    r->presc_dpsi = deltaPsi;
    return VLM_OK;
}

int vlm_rotor_get_prescribed_wake(vlm_ctx *ctx, int ir, int wakeType, double *fc, double *gam,
                                  double *rc, int *built)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    if (wakeType != 0 && wakeType != 1) return fail(ctx, VLM_ERR_ARG, "wakeType must be 0 or 1");
    VLM_DEVICE(ctx);
    VLM_TRY(sync_stream(ctx));
    const size_t nb = (size_t)r->d.nb;
    if (fc)
        VLM_HIP(hipMemcpy(fc, r->pf[wakeType].p, 6 * nb * PF_N * sizeof(double),
                          hipMemcpyDeviceToHost));
    if (gam || rc) {
        std::vector<double> S(4 * nb);
        VLM_HIP(hipMemcpy(S.data(), r->ps[wakeType].p, 4 * nb * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (size_t b = 0; b < nb; ++b)
            for (int i = 0; i < PF_N; ++i) {
                if (gam) gam[b * PF_N + i] = S[4 * b];
                if (rc) rc[b * PF_N + i] = S[4 * b + 1];
            }
    }
    if (built) *built = r->presc_built[wakeType];
    return VLM_OK;
}

int vlm_rotor_get_wake(vlm_ctx *ctx, int ir, int wakeType, double *nodes, double *le,
                       double *gam, double *rc4, double *far_fc, double *far_gam, double *far_rc)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    const int pred = wakeType == VLM_WAKE_P;
    const size_t nb = r->d.nb, nN = r->d.nNwake, ns = r->d.ns, nF = r->d.nFwake;
    VLM_TRY(sync_stream(ctx));
    std::vector<double> tn(3 * nb * nN * (ns + 1)), tg(nb * nN * ns), tr(4 * nb * nN * ns);
    if (nN > 0) {
        VLM_HIP(hipMemcpy(tn.data(), pred ? r->pnode.p : r->node.p, tn.size() * sizeof(double),
                          hipMemcpyDeviceToHost));
        VLM_HIP(hipMemcpy(tg.data(), r->gam.p, tg.size() * sizeof(double), hipMemcpyDeviceToHost));
        VLM_HIP(hipMemcpy(tr.data(), r->rc4.p, tr.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t b = 0; b < nb; ++b)
            for (size_t i = 0; i < nN; ++i) {
                const size_t p = (r->head + i) % nN;
                if (nodes)
                    std::memcpy(nodes + 3 * (b * nN + i) * (ns + 1), &tn[3 * (b * nN + p) * (ns + 1)],
                                3 * (ns + 1) * sizeof(double));
                if (gam)
                    std::memcpy(gam + (b * nN + i) * ns, &tg[(b * nN + p) * ns], ns * sizeof(double));
                if (rc4)
                    std::memcpy(rc4 + 4 * (b * nN + i) * ns, &tr[4 * (b * nN + p) * ns],
                                4 * ns * sizeof(double));
            }
        if (le)
            VLM_HIP(hipMemcpy(le, pred ? r->ple.p : r->le.p, 3 * nb * (ns + 1) * sizeof(double),
                              hipMemcpyDeviceToHost));
    }
    if (nF > 0) {
        if (far_fc)
            VLM_HIP(hipMemcpy(far_fc, pred ? r->pffc.p : r->ffc.p, 6 * nb * nF * sizeof(double),
                              hipMemcpyDeviceToHost));
        if (far_gam)
            VLM_HIP(hipMemcpy(far_gam, r->fgam.p, nb * nF * sizeof(double), hipMemcpyDeviceToHost));
        if (far_rc)
            VLM_HIP(hipMemcpy(far_rc, r->frc.p, nb * nF * sizeof(double), hipMemcpyDeviceToHost));
    }
    return VLM_OK;
}

int vlm_rotor_set_wake(vlm_ctx *ctx, int ir, const double *nodes, const double *le,
                       const double *gam, const double *rc4, const double *far_fc,
                       const double *far_gam, const double *far_rc, int rowNear, int rowFar)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    const size_t nb = r->d.nb, nN = r->d.nNwake, ns = r->d.ns, nF = r->d.nFwake;
    if (rowNear < 1 || rowNear > (int)nN + 1 || rowFar < 1 || rowFar > (int)nF + 1)
        return fail(ctx, VLM_ERR_ARG, "cursor out of range");
    if (nN > 0 && (!nodes || !le || !gam || !rc4)) return fail(ctx, VLM_ERR_ARG, "null buffer");
    if (nF > 0 && (!far_fc || !far_gam || !far_rc)) return fail(ctx, VLM_ERR_ARG, "null buffer");
    VLM_TRY(sync_stream(ctx));
    r->head = 0;
    r->rowNear = rowNear;
    r->rowFar = rowFar;
    if (nN > 0) {
        std::vector<double> age(nb * nN * ns, 0.0);
        VLM_HIP(hipMemcpy(r->node.p, nodes, 3 * nb * nN * (ns + 1) * sizeof(double),
                          hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->le.p, le, 3 * nb * (ns + 1) * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->gam.p, gam, nb * nN * ns * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->rc4.p, rc4, 4 * nb * nN * ns * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->age.p, age.data(), age.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (nF > 0) {
        std::vector<double> l(nb * nF), z(nb * nF, 0.0);
        for (size_t k = 0; k < nb * nF; ++k) {
            const double *f = far_fc + 6 * k;
            const double dx = f[0] - f[3], dy = f[1] - f[4], dz = f[2] - f[5];
            l[k] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        VLM_HIP(hipMemcpy(r->ffc.p, far_fc, 6 * nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->fgam.p, far_gam, nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->frc.p, far_rc, nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->frc0.p, far_rc, nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->fl0.p, l.data(), nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->flc.p, l.data(), nb * nF * sizeof(double), hipMemcpyHostToDevice));
        VLM_HIP(hipMemcpy(r->fage.p, z.data(), nb * nF * sizeof(double), hipMemcpyHostToDevice));
    }
    return VLM_OK;
}

int vlm_rotor_get_velocity(vlm_ctx *ctx, int ir, int which, double *velN, double *velF)
{
    Rotor *r = get_rotor(ctx, ir);
    if (!r) return ctx ? fail(ctx, VLM_ERR_ARG, "unknown rotor") : VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (which < VLM_VEL || which > VLM_VEL_STEP) return fail(ctx, VLM_ERR_ARG, "bad which");
    VLM_TRY(sync_stream(ctx));
    if (velN && r->nvn() > 0)
        VLM_HIP(hipMemcpy(velN, r->velN[which].p, r->nvn() * sizeof(double), hipMemcpyDeviceToHost));
    if (velF && r->nvf() > 0)
        VLM_HIP(hipMemcpy(velF, r->velF[which].p, r->nvf() * sizeof(double), hipMemcpyDeviceToHost));
    return VLM_OK;
}

}  // extern "C"

extern "C" {

int vlm_predictor_errors(vlm_ctx *ctx, double *mlp_rel_rms, double *base_rel_rms,
                         int64_t *judged_steps)
{
    if (!ctx) return VLM_ERR_ARG;
    const Predictor &P = ctx->pred;
    const double r = P.sum_ref2 > 0 ? P.sum_ref2 : 1.0;
    if (mlp_rel_rms) *mlp_rel_rms = std::sqrt(P.sum_err2 / r);
    if (base_rel_rms) *base_rel_rms = std::sqrt(P.sum_base2 / r);
    if (judged_steps) *judged_steps = P.judged;
    return VLM_OK;
}

int vlm_set_predictor(vlm_ctx *ctx, int mode, double tol, int sample_stride, int warmup_steps,
                      int train_iters, double lref)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    if (mode != 0 && mode != 1) return fail(ctx, VLM_ERR_ARG, "predictor mode must be 0 or 1");
    if (sample_stride < 1 || warmup_steps < 0 || train_iters < 0 || !(lref > 0) || !(tol > 0))
        return fail(ctx, VLM_ERR_ARG, "bad predictor settings");
    Predictor &P = ctx->pred;
    P.mode = mode;
    P.tol = tol;
    P.stride = sample_stride;
    P.warmup = warmup_steps;
    P.train_iters = train_iters;
    P.lref = lref;
    P.sum_err2 = P.sum_base2 = P.sum_ref2 = 0.0;
    P.judged = 0;
    return VLM_OK;
}

int vlm_predictor_stats(vlm_ctx *ctx, int64_t *full, int64_t *learned, int64_t *fallbacks,
                        double *last_err, double *last_loss)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    Predictor &P = ctx->pred;
    if (full) *full = P.full;
    if (learned) *learned = P.learned;
    if (fallbacks) *fallbacks = P.fallbacks;
    if (last_err) *last_err = P.last_err;
    if (last_loss) {
        float l = 0.f;
        VLM_TRY(mlp_read_loss(ctx, &l));
        *last_loss = l;
    }
    return VLM_OK;
}

}  // extern "C"
