// Biot-Savart N-body kernels for MI355X (gfx950).
//
// Velocity induced by straight vortex segments with the Vatistas n=2 core,
// the leaf vf_vind of the reference (classdef.f90:476-503):
//   v = gam/(4 pi) * (r1 x r2) * [r0 . (r1/|r1| - r2/|r2|)]
//       / sqrt((rc |r0|)^4 + |r1 x r2|^4),
// zero when |r1 x r2|^2 <= eps^2, r/|r| = 0 when |r| <= eps
// (unitVec, libMath.f90:249-262).  Per segment the library precomputes
// r0 = b - a, g = gam/(4 pi) and q = rc^2 |r0|^2 (stored squared in the
// FP64 record); per pair r1 = p - a, r2 = r1 - r0 and r1 x r2 = r0 x r1.
//
// Work decomposition: one target point per lane; a workgroup walks one
// (target tile, segment chunk) work item at a time, staging 256 segments in
// LDS and reading them back as wave-uniform broadcasts.  Work items are
// handed out by an atomic counter to a grid sized to the resident capacity
// of the device, so every CU stays busy to the end of the launch whatever
// the target count.  Each item writes its own partial sums, which are
// reduced in a fixed order afterwards: results are bitwise repeatable.  No
// MFMA: the pair is not a contraction.
//
// Precision modes
//   FP64  : FP64 throughout; 1/sqrt via v_rsq_f64 plus one third-order
//           correction.
//   MIXED : FP32 pair arithmetic on hi/lo-split coordinates
//           (r1 = (p_hi - a_hi) + (p_lo - a_lo), accurate to FP32 relative
//           precision of r1 itself), v_rsq_f32, per-tile FP32 sums folded
//           into a Kahan-compensated accumulator.  Pairs with the target
//           inside the core (within one core radius of the segment) are
//           recomputed in FP64 (the regularised core and target-on-filament
//           terms, where the reference's exact-zero guards act).
//           Launches with fewer than 1024 targets (collocation points, judge
//           samples) run the FP64 small-target kernel in both modes.
#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int BS_BLOCK = 256;       // wake passes: many targets
constexpr double EPS2 = kEps * kEps;
constexpr float EPS2F = 4.93038065763132e-32f;
constexpr float NEAR_CORE2 = 1.0f;  // h < rc, and |r1|^2 + |r2|^2 < 4 |r0|^2 (stored)

__device__ __forceinline__ double rsq64(double x)
{
    // v_rsq_f64 is accurate to about 2^-24; one third-order correction
    // y (1 + e/2 + 3 e^2/8) with e = 1 - x y^2 brings it to FP64 rounding
    // (the dropped 5 e^3/16 term is below 2^-69).
    const double y = __builtin_amdgcn_rsq(x);
    const double h = x * y;
    const double e = __builtin_fma(-h, y, 1.0);
    const double c = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y, e * c, y);
}

// x where c holds, else a value that acts as zero here: only the high dword
// (sign, exponent, top of the mantissa) is cleared, one select instead of
// two.  What is left is a positive number below 2^-1042.  It only enters
// products with the components of c = r0 x r1, below 2^-51 when
// |r1 x r2| <= eps, which underflow to exactly 0, so the sums are bitwise
// those of a true zero.
__device__ __forceinline__ double zero_unless(bool c, double x)
{
    const int hi = __double2hiint(x);
    return __hiloint2double(c ? hi : 0, __double2loint(x));
}

// FP64 pair; segment record (ax ay az g | r0x r0y r0z qq | n0), qq = (rc |r0|)^4,
// n0 = |r0|^2 (so r0 . r2 = r0 . r1 - n0).
// The reference's r0 . (r1/|r1| - r2/|r2|) / sqrt(qq + c2^2) is evaluated as
// d1 / sqrt(n1 den) - d2 / sqrt(n2 den): two reciprocal square roots per
// pair instead of three.  The pair guard |r1 x r2|^2 <= eps^2 is exact.  The
// unitVec guard (|r| <= eps -> that term 0) is a clamp n >= eps^2, one
// v_max instead of a compare and a select: on an end point (r = 0) the term
// is d * j = 0 exactly, as in the reference; for 0 < |r| <= eps,
// |r1 x r2| <= |r0| |r| lets the pair guard zero the pair whenever
// |r0| <= 1, and for longer segments the term left over is of the order of
// the pair's own (eps-sized) contribution.
__device__ __forceinline__ void pair64(double px, double py, double pz, double4 sa, double4 sb,
                                       double n0, double &vx, double &vy, double &vz)
{
    const double r1x = px - sa.x, r1y = py - sa.y, r1z = pz - sa.z;
    const double r0x = sb.x, r0y = sb.y, r0z = sb.z;
    const double r2x = r1x - r0x, r2y = r1y - r0y, r2z = r1z - r0z;
    const double cx = r0y * r1z - r0z * r1y;
    const double cy = r0z * r1x - r0x * r1z;
    const double cz = r0x * r1y - r0y * r1x;
    const double c2 = cx * cx + cy * cy + cz * cz;
    const double n1 = r1x * r1x + r1y * r1y + r1z * r1z;
    const double n2 = r2x * r2x + r2y * r2y + r2z * r2z;
    const double d1 = r0x * r1x + r0y * r1y + r0z * r1z;
    const double d2 = d1 - n0;
    const double den = __builtin_fma(c2, c2, sb.w);
    const double j1 = rsq64(__builtin_fmax(n1, EPS2) * den);
    const double j2 = rsq64(__builtin_fmax(n2, EPS2) * den);
    const double s = zero_unless(c2 > EPS2, sa.w * (d1 * j1 - d2 * j2));
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
}

// FP32 pair on split coordinates; record (ahx ahy ahz g | alx aly alz q |
// r0x r0y r0z 4|r0|^2).  Returns true for near-singular pairs, whose
// contribution is left to the FP64 recompute (zero here).  Same evaluation
// order as pair64: d1 / sqrt(n1 den) - d2 / sqrt(n2 den).
__device__ __forceinline__ bool pair32(float phx, float phy, float phz, float plx, float ply,
                                       float plz, float4 s0, float4 s1, float4 s2, float &vx,
                                       float &vy, float &vz)
{
    const float r1x = (phx - s0.x) + (plx - s1.x);
    const float r1y = (phy - s0.y) + (ply - s1.y);
    const float r1z = (phz - s0.z) + (plz - s1.z);
    const float r0x = s2.x, r0y = s2.y, r0z = s2.z;
    const float r2x = r1x - r0x, r2y = r1y - r0y, r2z = r1z - r0z;
    const float cx = r0y * r1z - r0z * r1y;
    const float cy = r0z * r1x - r0x * r1z;
    const float cz = r0x * r1y - r0y * r1x;
    const float c2 = cx * cx + cy * cy + cz * cz;
    const float n1 = r1x * r1x + r1y * r1y + r1z * r1z;
    const float n2 = r2x * r2x + r2y * r2y + r2z * r2z;
    const bool nearp = (c2 < NEAR_CORE2 * s1.w) && (n1 + n2 < s2.w);
    const float d1 = r0x * r1x + r0y * r1y + r0z * r1z;
    const float d2 = __builtin_fmaf(s2.w, -0.25f, d1);  // d1 - |r0|^2
    const float den = __builtin_fmaf(c2, c2, s1.w * s1.w);
    const float j1 = __builtin_amdgcn_rsqf(__builtin_fmaxf(n1, EPS2F) * den);
    const float j2 = __builtin_amdgcn_rsqf(__builtin_fmaxf(n2, EPS2F) * den);
    float s = s0.w * (d1 * j1 - d2 * j2);
    s = (c2 > EPS2F && !nearp) ? s : 0.0f;
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
    return nearp;
}

// ---------------------------------------------------------------------
// Launches with 1024 targets or more: stream-K over XCD-resident rounds.
//
// Targets form tiles of SK_TILE (one per lane of a 4-wave workgroup) and
// segments batches of SK_BATCH (one LDS staging).  The batch list is cut
// into NX contiguous ranges, one per XCD: the dispatcher deals workgroups
// round-robin over the 8 XCDs, so workgroup b runs on XCD b % NX.  Each
// range is cut into R rounds of about RB_TARGET batches, few enough that
// the records of a round stay in the XCD's 4 MB L2 while all of the XCD's
// workgroups sweep it.  Within an XCD the work units (round, tile, batch)
// are ordered round-major, then tile-major, and split into B = grid / NX
// equal contiguous ranges, one per workgroup: every workgroup evaluates
// the same number of pairs whatever the tile and batch counts (stream-K).
//
// A workgroup's range is a sequence of runs, one per (round, tile) it
// touches; each run's partial sums go to a piece slot of their own.  A
// planning launch writes every workgroup's run list (tile, first batch,
// batch count), so the evaluation kernel's control is wave-uniform scalar
// loads and its registers go to the pair arithmetic; a fix-up launch adds
// each tile's pieces in a fixed order (XCD, round, workgroup) into the
// output.  Results are bitwise repeatable, and each target has a few
// pieces where a split over segment chunks gave one per chunk.
// ---------------------------------------------------------------------
constexpr int NX = 8;
constexpr int SK_TILE = 256;
constexpr int SK_BATCH = 256;
constexpr int RB_TARGET = 96;  // batches per round: 1.8 MB of FP64 records
constexpr int SK_RMAX = 64;    // beyond it rounds grow instead of multiplying

struct SkPlan {
    int64_t nb;  // batches
    int64_t ns;  // slices in all, NX * R; XCD x owns slices x*R .. x*R+R-1
    int R;
    __host__ __device__ int64_t sb(int64_t s) const { return s * nb / ns; }  // first batch
};

__host__ __device__ inline SkPlan sk_plan(int64_t nseg)
{
    SkPlan p;
    p.nb = (nseg + SK_BATCH - 1) / SK_BATCH;
    int64_t R = (p.nb + (int64_t)NX * RB_TARGET - 1) / ((int64_t)NX * RB_TARGET);
    if (R < 1) R = 1;
    if (R > SK_RMAX) R = SK_RMAX;
    p.R = (int)R;
    p.ns = (int64_t)NX * R;
    return p;
}

// Piece slots per workgroup: a range of U/B units touches at most
// 2 ntiles R / B + R + 2 (round, tile) pairs (every non-empty round holds
// at least half the batches of the largest).
__host__ __device__ inline int sk_kmax(int ntiles, int R, int B)
{
    return (int)((2LL * ntiles * R + B - 1) / B) + R + 2;
}

struct SkArgs {
    const double *s64;  // SEG64_N x stride SoA
    const float *s32;   // SEG32_N x stride SoA
    int64_t sstride;
    const long long *nseg;
    const double *t;
    int64_t m;
    int64_t tstride;
    double *piece;      // [grid][kmax][3][SK_TILE]
    const int4 *runs;   // [grid][kmax]: tile, first batch, batch count
    const int *nruns;   // [grid]
    int kmax;
};

// The XCD-local unit space of XCD x.
struct SkXcd {
    SkPlan p;
    int64_t s0;      // first slice
    int64_t b0;      // first batch
    int64_t ntiles;
    int64_t U;       // units
    int B;           // workgroups of the XCD
    __device__ int64_t base(int r) const { return ntiles * (p.sb(s0 + r) - b0); }
    __device__ int64_t nsb(int r) const { return p.sb(s0 + r + 1) - p.sb(s0 + r); }
    __device__ int64_t lo(int64_t j) const { return j * U / B; }
    // workgroup whose range holds unit u (ranges may be empty when U < B)
    __device__ int64_t jof(int64_t u) const { return ((u + 1) * B + U - 1) / U - 1; }
    __device__ int round_of(int64_t u, int r) const
    {
        while (r + 1 < p.R && base(r + 1) <= u) ++r;
        return r;
    }
    // (round, tile) pair index of unit u
    __device__ int64_t pair_of(int64_t u) const
    {
        const int r = round_of(u, 0);
        return (int64_t)r * ntiles + (u - base(r)) / nsb(r);
    }
};

__device__ inline SkXcd sk_xcd(const SkPlan &p, int x, int ntiles, int B)
{
    SkXcd X;
    X.p = p;
    X.s0 = (int64_t)x * p.R;
    X.b0 = p.sb(X.s0);
    X.ntiles = ntiles;
    X.U = (int64_t)ntiles * (p.sb(X.s0 + p.R) - X.b0);
    X.B = B;
    return X;
}

// Run lists: one thread per evaluation workgroup b writes its runs
// (tile, first batch, batch count) in range order; run k of b owns piece
// slot b * kmax + k.
__global__ void sk_plan_kernel(const long long *nseg_p, int ntiles, int grid, int kmax, int4 *runs,
                               int *nruns)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= grid) return;
    const SkPlan p = sk_plan(*nseg_p);
    const int B = grid / NX;
    const SkXcd X = sk_xcd(p, b % NX, ntiles, B);
    int k = 0;
    if (p.nb > 0 && X.U > 0) {
        const int64_t j = b / NX;
        int64_t u = X.lo(j);
        const int64_t u1 = X.lo(j + 1);
        int r = 0;
        while (u < u1 && k < kmax) {
            r = X.round_of(u, r);
            const int64_t nsb = X.nsb(r), off = u - X.base(r);
            const int64_t tile = off / nsb;
            const int64_t ue = min(u1, X.base(r) + (tile + 1) * nsb);
            runs[(int64_t)b * kmax + k] =
                make_int4((int)tile, (int)(p.sb(X.s0 + r) + (off - tile * nsb)), (int)(ue - u), 0);
            ++k;
            u = ue;
        }
    }
    nruns[b] = k;
}

// Fix-up: one workgroup per tile, one target per lane; the tile's pieces
// in the fixed order (XCD, round, workgroup), added after the given
// values when accumulating.  With no segment at all the result is zero
// (or the given values).
__global__ __launch_bounds__(SK_TILE) void sk_fix_kernel(const long long *nseg_p, const double *piece,
                                                         int kmax, int ntiles, int grid,
                                                         int64_t m, double *out, int64_t ostride,
                                                         int acc)
{
    const int64_t t = blockIdx.x;
    const int64_t e = t * SK_TILE + threadIdx.x;
    if (e >= m) return;
    const SkPlan p = sk_plan(*nseg_p);
    const int B = grid / NX;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (acc) {
        s0 = out[e];
        s1 = out[ostride + e];
        s2 = out[2 * ostride + e];
    }
    for (int x = 0; x < NX && p.nb > 0; ++x) {
        const SkXcd X = sk_xcd(p, x, ntiles, B);
        if (X.U == 0) continue;
        for (int r = 0; r < p.R; ++r) {
            const int64_t nsb = X.nsb(r);
            if (nsb == 0) continue;
            const int64_t us = X.base(r) + t * nsb, ue = us + nsb;
            // the non-empty workgroups whose ranges meet [us, ue), in order
            for (int64_t j = X.jof(us);;) {
                const int64_t k = (int64_t)r * ntiles + t - X.pair_of(X.lo(j));
                const double *q = piece + ((j * NX + x) * (int64_t)kmax + k) * 3 * SK_TILE +
                                  threadIdx.x;
                s0 += q[0];
                s1 += q[SK_TILE];
                s2 += q[2 * SK_TILE];
                const int64_t next = X.lo(j + 1);
                if (next >= ue) break;
                j = X.jof(next);
            }
        }
    }
    out[e] = s0;
    out[ostride + e] = s1;
    out[2 * ostride + e] = s2;
}

// __launch_bounds__'s second argument asks for 8 waves per SIMD (64
// VGPRs): the pair loop fits, and only loop-invariant addresses of the run
// loop are spilled.
template <int TILE>
__global__ __launch_bounds__(TILE, 8) void bs_fp64_kernel(SkArgs a)
{
    static_assert(TILE == SK_TILE && SK_BATCH == SK_TILE, "one record staged per lane");
    __shared__ double4 sa[SK_BATCH];
    __shared__ double4 sb[SK_BATCH];
    __shared__ double sn[SK_BATCH];
    const int tid = threadIdx.x;
    const int64_t nseg = *a.nseg;
    const int64_t S = a.sstride;
    const int nr = a.nruns[blockIdx.x];
    for (int k = 0; k < nr; ++k) {
        const int4 run = a.runs[(int64_t)blockIdx.x * a.kmax + k];
        const int64_t e = (int64_t)run.x * SK_TILE + tid;
        const bool live = e < a.m;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (live) {
            px = a.t[e];
            py = a.t[a.tstride + e];
            pz = a.t[2 * a.tstride + e];
        }
        double vx = 0.0, vy = 0.0, vz = 0.0;
        for (int ib = 0; ib < run.z; ++ib) {
            const int64_t base = (int64_t)(run.y + ib) * SK_BATCH;
            const int cnt = (int)min<int64_t>(SK_BATCH, nseg - base);
            __syncthreads();
            if (tid < cnt) {
                const int64_t s = base + tid;
                sa[tid] = make_double4(a.s64[s], a.s64[S + s], a.s64[2 * S + s], a.s64[3 * S + s]);
                sb[tid] = make_double4(a.s64[4 * S + s], a.s64[5 * S + s], a.s64[6 * S + s],
                                       a.s64[7 * S + s]);
                sn[tid] = a.s64[8 * S + s];
            }
            __syncthreads();
#pragma unroll 2
            for (int q = 0; q < cnt; ++q) pair64(px, py, pz, sa[q], sb[q], sn[q], vx, vy, vz);
        }
        double *o = a.piece + (((int64_t)blockIdx.x * a.kmax + k) * 3) * SK_TILE + tid;
        o[0] = vx;
        o[SK_TILE] = vy;
        o[2 * SK_TILE] = vz;
    }
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    const float y = x - c;
    const float tt = s + y;
    c = (tt - s) - y;
    s = tt;
}

// FP32 pair without the near-field handling, for a group of segments that
// every target of the wave is far from (see the group test in
// bs_mixed_kernel): no target lies within two segment lengths of a
// segment, so no pair is near-singular, |r1| and |r2| are bounded away
// from zero (no unitVec clamp) and every core term q is positive (no
// zero-denominator guard).
__device__ __forceinline__ void pair32_far(float phx, float phy, float phz, float plx,
                                           float ply, float plz, float4 s0, float4 s1,
                                           float4 s2, float &vx, float &vy, float &vz)
{
    const float r1x = (phx - s0.x) + (plx - s1.x);
    const float r1y = (phy - s0.y) + (ply - s1.y);
    const float r1z = (phz - s0.z) + (plz - s1.z);
    const float r0x = s2.x, r0y = s2.y, r0z = s2.z;
    const float r2x = r1x - r0x, r2y = r1y - r0y, r2z = r1z - r0z;
    const float cx = r0y * r1z - r0z * r1y;
    const float cy = r0z * r1x - r0x * r1z;
    const float cz = r0x * r1y - r0y * r1x;
    const float c2 = cx * cx + cy * cy + cz * cz;
    const float n1 = r1x * r1x + r1y * r1y + r1z * r1z;
    const float n2 = r2x * r2x + r2y * r2y + r2z * r2z;
    const float d1 = r0x * r1x + r0y * r1y + r0z * r1z;
    const float d2 = __builtin_fmaf(s2.w, -0.25f, d1);
    const float den = __builtin_fmaf(c2, c2, s1.w * s1.w);
    const float j1 = __builtin_amdgcn_rsqf(n1 * den);
    const float j2 = __builtin_amdgcn_rsqf(n2 * den);
    const float s = s0.w * (d1 * j1 - d2 * j2);
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
}

// MIXED: the staged batch is cut into groups of 32 segments.  At staging,
// each group gets a far radius around its first segment's start point:
// the largest |a_i - c| + 2 |r0_i| over the group (plus a rounding
// margin), or infinity for a short group or one holding a segment with
// q = rc^2 |r0|^2 below 1e-18.  A wave whose targets all lie outside that
// radius runs the group with pair32_far; otherwise it runs pair32, which
// flags near-singular pairs in a 32-bit mask, and recomputes the flagged
// ones in FP64 after the group.  Staged groups are padded with zero
// records, which contribute exactly zero and are never near (r0 = 0).
// 7 waves per SIMD (72 VGPRs): both pair loops fit; a few per-batch
// addresses are spilled.
template <int TILE>
__global__ __launch_bounds__(TILE, 7) void bs_mixed_kernel(SkArgs a)
{
    static_assert(TILE == SK_TILE && SK_BATCH == SK_TILE && SK_BATCH % 32 == 0, "groups of 32");
    __shared__ float4 s0v[SK_BATCH];
    __shared__ float4 s1v[SK_BATCH];
    __shared__ float4 s2v[SK_BATCH];
    __shared__ float4 far_sh[SK_BATCH / 32];  // centre, squared far radius
    const int tid = threadIdx.x;
    const int64_t nseg = *a.nseg;
    const int64_t S = a.sstride;
    const int nr = a.nruns[blockIdx.x];
    for (int k = 0; k < nr; ++k) {
        const int4 run = a.runs[(int64_t)blockIdx.x * a.kmax + k];
        const int64_t e = (int64_t)run.x * SK_TILE + tid;
        const bool live = e < a.m;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (live) {
            px = a.t[e];
            py = a.t[a.tstride + e];
            pz = a.t[2 * a.tstride + e];
        }
        const float phx = (float)px, phy = (float)py, phz = (float)pz;
        const float plx = (float)(px - (double)phx), ply = (float)(py - (double)phy),
                    plz = (float)(pz - (double)phz);
        float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
        double nx = 0.0, ny = 0.0, nz = 0.0;
        for (int ib = 0; ib < run.z; ++ib) {
            const int64_t base = (int64_t)(run.y + ib) * SK_BATCH;
            const int cnt = (int)min<int64_t>(SK_BATCH, nseg - base);
            const int cnt32 = (cnt + 31) & ~31;
            __syncthreads();
            {
                float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0, v2 = v0;
                if (tid < cnt) {
                    const int64_t s = base + tid;
                    v0 = make_float4(a.s32[s], a.s32[S + s], a.s32[2 * S + s], a.s32[3 * S + s]);
                    v1 = make_float4(a.s32[4 * S + s], a.s32[5 * S + s], a.s32[6 * S + s],
                                     a.s32[7 * S + s]);
                    v2 = make_float4(a.s32[8 * S + s], a.s32[9 * S + s], a.s32[10 * S + s],
                                     a.s32[11 * S + s]);
                }
                if (tid < cnt32) {
                    s0v[tid] = v0;
                    s1v[tid] = v1;
                    s2v[tid] = v2;
                }
                // far radius of this lane's group (lanes 32k .. 32k+31)
                const float cx = __shfl(v0.x, 0, 32), cy = __shfl(v0.y, 0, 32),
                            cz = __shfl(v0.z, 0, 32);
                const float dx = v0.x - cx, dy = v0.y - cy, dz = v0.z - cz;
                float ex = sqrtf(dx * dx + dy * dy + dz * dz) + sqrtf(v2.w);  // + 2 |r0|
                if (tid >= cnt || !(v1.w > 1e-18f)) ex = __builtin_inff();
                for (int o = 16; o > 0; o >>= 1) ex = fmaxf(ex, __shfl_xor(ex, o, 32));
                if ((tid & 31) == 0 && tid < cnt32) {
                    const float cm = fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)));
                    const float rr = ex * 1.0001f + 1e-5f * cm + 1e-30f;
                    far_sh[tid >> 5] = make_float4(cx, cy, cz, rr * rr);
                }
            }
            __syncthreads();
            float tx = 0.f, ty = 0.f, tz = 0.f;
            for (int g32 = 0; g32 < cnt32; g32 += 32) {
                const float4 f = far_sh[g32 >> 5];
                const float qx = phx - f.x, qy = phy - f.y, qz = phz - f.z;
                const bool far = !live || (qx * qx + qy * qy + qz * qz > f.w);
                if (__builtin_amdgcn_readfirstlane(__all(far))) {
#pragma unroll 4
                    for (int jj = 0; jj < 32; ++jj)
                        pair32_far(phx, phy, phz, plx, ply, plz, s0v[g32 + jj], s1v[g32 + jj],
                                   s2v[g32 + jj], tx, ty, tz);
                    continue;
                }
                unsigned mask = 0u;
#pragma unroll 4
                for (int jj = 0; jj < 32; ++jj) {
                    const bool nearp = pair32(phx, phy, phz, plx, ply, plz, s0v[g32 + jj],
                                              s1v[g32 + jj], s2v[g32 + jj], tx, ty, tz);
                    mask |= (nearp ? 1u : 0u) << jj;
                }
                if (!live) mask = 0u;
                while (mask) {
                    const int jj = __builtin_ctz(mask);
                    mask &= mask - 1u;
                    const int64_t s = base + g32 + jj;
                    pair64(px, py, pz,
                           make_double4(a.s64[s], a.s64[S + s], a.s64[2 * S + s], a.s64[3 * S + s]),
                           make_double4(a.s64[4 * S + s], a.s64[5 * S + s], a.s64[6 * S + s],
                                        a.s64[7 * S + s]),
                           a.s64[8 * S + s], nx, ny, nz);
                }
            }
            kahan_add(Sx, Cx, tx);
            kahan_add(Sy, Cy, ty);
            kahan_add(Sz, Cz, tz);
        }
        double *o = a.piece + (((int64_t)blockIdx.x * a.kmax + k) * 3) * SK_TILE + tid;
        o[0] = ((double)Sx - (double)Cx) + nx;
        o[SK_TILE] = ((double)Sy - (double)Cy) + ny;
        o[2 * SK_TILE] = ((double)Sz - (double)Cz) + nz;
    }
}

// Few targets (collocation points, judge samples: m < 1024): each block
// takes SM_T targets and one segment chunk; its 256 lanes stride over the
// chunk (coalesced SoA loads, each record shared by the SM_T targets) and
// the partial sums are reduced in a fixed order (butterfly within the wave,
// waves in order).  Always FP64: these launches are a small fraction of the
// pairs.
constexpr int SM_T = 4;
constexpr int SM_BLOCK = 256;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(SM_BLOCK) void bs_small_kernel(const double *s64, int64_t S,
                                                            const long long *nseg_p, int nchunks,
                                                            const double *t, int64_t m,
                                                            int64_t tstride, double *part)
{
    __shared__ double red[3 * SM_T][SM_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t nseg = *nseg_p;
    const int64_t len = (nseg + nchunks - 1) / nchunks;
    const int64_t s0 = (int64_t)blockIdx.x * len;
    const int64_t s1 = min(nseg, s0 + len);
    const int64_t g0 = (int64_t)blockIdx.y * SM_T;
    double px[SM_T], py[SM_T], pz[SM_T], vx[SM_T], vy[SM_T], vz[SM_T];
#pragma unroll
    for (int q = 0; q < SM_T; ++q) {
        const int64_t tq = g0 + q < m ? g0 + q : m - 1;
        px[q] = t[tq];
        py[q] = t[tstride + tq];
        pz[q] = t[2 * tstride + tq];
        vx[q] = vy[q] = vz[q] = 0.0;
    }
    for (int64_t sg = s0 + tid; sg < s1; sg += SM_BLOCK) {
        const double4 sa = make_double4(s64[sg], s64[S + sg], s64[2 * S + sg], s64[3 * S + sg]);
        const double4 sb =
            make_double4(s64[4 * S + sg], s64[5 * S + sg], s64[6 * S + sg], s64[7 * S + sg]);
        const double n0 = s64[8 * S + sg];
#pragma unroll
        for (int q = 0; q < SM_T; ++q)
            pair64(px[q], py[q], pz[q], sa, sb, n0, vx[q], vy[q], vz[q]);
    }
#pragma unroll
    for (int q = 0; q < SM_T; ++q) {
        const double a = wave_sum(vx[q]), b = wave_sum(vy[q]), c = wave_sum(vz[q]);
        if (lane == 0) {
            red[3 * q][w] = a;
            red[3 * q + 1][w] = b;
            red[3 * q + 2][w] = c;
        }
    }
    __syncthreads();
    if (tid < 3 * SM_T) {
        const int q = tid / 3, d = tid % 3;
        if (g0 + q < m) {
            double v = 0.0;
            for (int k = 0; k < SM_BLOCK / 64; ++k) v += red[tid][k];
            part[((size_t)blockIdx.x * 3 + d) * m + g0 + q] = v;
        }
    }
}

// Fixed-order sum of the per-chunk partials: one lane per (target,
// component), chunks added in increasing order; loads are issued eight
// chunks ahead of the adds.
constexpr int RED_BLOCK = 64;

__global__ __launch_bounds__(RED_BLOCK) void bs_reduce_kernel(const double *part, int nchunks,
                                                              int64_t m, double *out,
                                                              int64_t ostride, int acc)
{
    const int64_t t = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x;
    const int d = blockIdx.y;
    if (t >= m) return;
    const double *p = part + (size_t)d * m + t;
    const size_t cs = 3 * (size_t)m;
    double s = acc ? out[d * ostride + t] : 0.0;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = p[(size_t)(c + k) * cs];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; c < nchunks; ++c) s += p[(size_t)c * cs];
    out[d * ostride + t] = s;
}

__global__ void zero_soa3_kernel(double *o, int64_t m, int64_t stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    o[i] = 0.0;
    o[stride + i] = 0.0;
    o[2 * stride + i] = 0.0;
}

template <typename K>
int resident_blocks(vlm_ctx *ctx, K kernel, int block, int *cache)
{
    if (*cache > 0) return *cache;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess)
        return -1;
    *cache = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    return *cache;
}

// Records one profiled launch: the pair count is m times the device
// segment count, copied to the pinned slot of this record without a stream
// synchronisation and resolved at read-out (vlm_profile_get).
int profile_count(vlm_ctx *ctx, const long long *nseg_dev, int64_t m, int kind, const Timer *tm)
{
    Timer r{};
    if (tm) r = *tm;
    r.kind = kind;
    r.pairs = -1.0;
    const size_t slot = ctx->timers.size();
    if (slot < ctx->prof_cap) {
        VLM_HIP(hipMemcpyAsync(ctx->prof_nseg + slot, nseg_dev, sizeof(long long),
                               hipMemcpyDeviceToHost, ctx->stream));
        r.pairs = (double)m;
    }
    ctx->timers.push_back(r);
    if (kind == 0) ctx->prof_launches += 1;
    return VLM_OK;
}

}  // namespace

int bs_eval(vlm_ctx *ctx, const double *seg64, const float *seg32, int64_t segstride,
            const long long *nseg_dev, int64_t nseg_max, const double *t, int64_t m,
            int64_t tstride, double *out, int accumulate)
{
    if (m <= 0) return VLM_OK;
    if (m < 1024) {
        // about 2048 blocks, at least 8 segments per lane
        const int64_t groups = (m + SM_T - 1) / SM_T;
        int64_t nc = (2048 + groups - 1) / groups;
        const int64_t maxc = nseg_max / (8 * SM_BLOCK);
        if (nc > maxc) nc = maxc;
        if (nc < 1) nc = 1;
        const int nchunks = nseg_max > 0 ? (int)nc : 0;
        if (nchunks > 0) {
            VLM_HIP(ctx->part.reserve((size_t)nchunks * 3 * (size_t)m));
            hipLaunchKernelGGL(bs_small_kernel, dim3(nchunks, (unsigned)groups), dim3(SM_BLOCK),
                               0, ctx->stream, seg64, segstride, nseg_dev, nchunks, t, m, tstride,
                               ctx->part.p);
            VLM_HIP(hipGetLastError());
            if (ctx->profile) VLM_TRY(profile_count(ctx, nseg_dev, m, 1, nullptr));
        }
        hipLaunchKernelGGL(bs_reduce_kernel, dim3((unsigned)((m + RED_BLOCK - 1) / RED_BLOCK), 3),
                           dim3(RED_BLOCK), 0,
                           ctx->stream, ctx->part.p, nchunks, m, out, tstride, accumulate);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    const int blk = BS_BLOCK;
    const int ntiles = (int)((m + SK_TILE - 1) / SK_TILE);
    const bool mixed = ctx->precision == VLM_MIXED;
    int cap = mixed ? resident_blocks(ctx, bs_mixed_kernel<BS_BLOCK>, blk, &ctx->occ[1])
                    : resident_blocks(ctx, bs_fp64_kernel<BS_BLOCK>, blk, &ctx->occ[3]);
    if (cap <= 0) return fail(ctx, VLM_ERR_HIP, "occupancy query failed");
    if (nseg_max <= 0) {
        // no source segment can exist: the sum is empty
        if (!accumulate) {
            hipLaunchKernelGGL(zero_soa3_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                               ctx->stream, out, m, tstride);
            VLM_HIP(hipGetLastError());
        }
        return VLM_OK;
    }
    // the resident grid, a multiple of the XCD count
    const int grid = cap >= NX ? cap - cap % NX : NX;
    const int B = grid / NX;
    const SkPlan plan = sk_plan(nseg_max);   // bounds the device count's plan
    const int kmax = sk_kmax(ntiles, plan.R, B);
    VLM_HIP(ctx->part.reserve((size_t)grid * kmax * 3 * SK_TILE));
    VLM_HIP(ctx->skruns.reserve((size_t)grid * kmax));
    VLM_HIP(ctx->counter.reserve((size_t)grid));
    hipLaunchKernelGGL(sk_plan_kernel, dim3((unsigned)((grid + 255) / 256)), dim3(256), 0,
                       ctx->stream, nseg_dev, ntiles, grid, kmax, ctx->skruns.p,
                       ctx->counter.p);
    VLM_HIP(hipGetLastError());
    SkArgs a;
    a.s64 = seg64;
    a.s32 = seg32;
    a.sstride = segstride;
    a.nseg = nseg_dev;
    a.t = t;
    a.m = m;
    a.tstride = tstride;
    a.piece = ctx->part.p;
    a.runs = ctx->skruns.p;
    a.nruns = ctx->counter.p;
    a.kmax = kmax;
    Timer tm{};
    const bool timed = ctx->profile;
    if (timed) {
        VLM_HIP(hipEventCreate(&tm.a));
        VLM_HIP(hipEventCreate(&tm.b));
        VLM_HIP(hipEventRecord(tm.a, ctx->stream));
    }
    if (mixed)
        hipLaunchKernelGGL(bs_mixed_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(bs_fp64_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0, ctx->stream, a);
    VLM_HIP(hipGetLastError());
    if (timed) {
        VLM_HIP(hipEventRecord(tm.b, ctx->stream));
        VLM_TRY(profile_count(ctx, nseg_dev, m, 0, &tm));
    }
    hipLaunchKernelGGL(sk_fix_kernel, dim3((unsigned)ntiles), dim3(SK_TILE), 0, ctx->stream,
                       nseg_dev, ctx->part.p, kmax, ntiles, grid, m, out, tstride, accumulate);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // namespace vlm
