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

struct BsArgs {
    const double *s64;   // 8 x stride SoA
    const float *s32;    // 12 x stride SoA
    int64_t sstride;
    const long long *nseg;
    int64_t chunk_len;
    const double *t;
    int64_t m;
    int64_t tstride;
    double *part;
    int ntiles;
    int nitems;
    int *counter;
};

__device__ __forceinline__ int next_item(int *counter, int *sh)
{
    if (threadIdx.x == 0) *sh = atomicAdd(counter, 1);
    __syncthreads();
    const int it = *sh;
    __syncthreads();
    return it;
}

template <int TILE>
__global__ __launch_bounds__(TILE) void bs_fp64_kernel(BsArgs a)
{
    __shared__ double4 sa[TILE];
    __shared__ double4 sb[TILE];
    __shared__ double sn[TILE];
    __shared__ int item_sh;
    const int tid = threadIdx.x;
    const int64_t nseg = *a.nseg;
    const int64_t S = a.sstride;
    for (;;) {
        const int item = next_item(a.counter, &item_sh);
        if (item >= a.nitems) break;
        const int chunk = item / a.ntiles;
        const int tile = item - chunk * a.ntiles;
        const int64_t s0 = (int64_t)chunk * a.chunk_len;
        const int64_t s1 = min(nseg, s0 + a.chunk_len);
        const int64_t t = (int64_t)tile * TILE + tid;
        const bool live = t < a.m;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (live) {
            px = a.t[t];
            py = a.t[a.tstride + t];
            pz = a.t[2 * a.tstride + t];
        }
        double vx = 0.0, vy = 0.0, vz = 0.0;
        for (int64_t base = s0; base < s1; base += TILE) {
            const int cnt = (int)min<int64_t>(TILE, s1 - base);
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
            for (int k = 0; k < cnt; ++k) pair64(px, py, pz, sa[k], sb[k], sn[k], vx, vy, vz);
        }
        if (live) {
            double *o = a.part + (size_t)chunk * 3 * a.m;
            o[t] = vx;
            o[a.m + t] = vy;
            o[2 * a.m + t] = vz;
        }
    }
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    const float y = x - c;
    const float tt = s + y;
    c = (tt - s) - y;
    s = tt;
}

// FP32 pair for a group of segments that every target of the wave is far
// from (see the group test in bs_mixed_kernel), without the near-field
// handling: no target lies within two segment lengths of a segment, so no
// pair is near-singular, |r1| and |r2| are bounded away from zero (no
// unitVec clamp) and every core term q is positive (no zero-denominator
// guard).  Coordinates are relative to the group's centre c: q = p - c
// (formed once per group), al = (a - c, (rc^2 |r0|^2)^2), g the
// circulation factor, s2 = (r0, 4 |r0|^2).  The rounding of q and al is
// about eps (|p - c| + |a - c|) <= eps (|r1| + 2 R) for a group of radius
// R; the far test keeps |r1| >= 2 |r0| for every segment, and groups wider
// than 64 times their longest segment are never far.  A segment's error
// and its contribution both scale with its |r0|, so no pair's error
// exceeds about 64 eps of the contribution of the group's longest segment
// at the same distance, without the hi/lo split.
__device__ __forceinline__ void pair32_local(float qx, float qy, float qz, float4 al, float g,
                                             float4 s2, float &vx, float &vy, float &vz)
{
    const float r1x = qx - al.x, r1y = qy - al.y, r1z = qz - al.z;
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
    const float den = __builtin_fmaf(c2, c2, al.w);
    const float j1 = __builtin_amdgcn_rsqf(n1 * den);
    const float j2 = __builtin_amdgcn_rsqf(n2 * den);
    const float s = g * (d1 * j1 - d2 * j2);
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
}

// MIXED: the staged batch is cut into groups of 32 segments.  At staging,
// each group gets a far radius around its first segment's start point:
// the largest |a_i - c| + 2 |r0_i| over the group (plus a rounding
// margin), or infinity for a short group, one holding a segment with
// q = rc^2 |r0|^2 below 1e-18, or one wider than 64 of its longest segments.  A wave whose targets all lie outside that
// radius runs the group with pair32_local (coordinates relative to the
// group's centre, staged beside the split ones); otherwise it runs pair32, which
// flags near-singular pairs in a 32-bit mask, and recomputes the flagged
// ones in FP64 after the group (records read from the FP64 list in global
// memory), so the FP32 loops stay free of the FP64 branch.  Staged groups
// are padded with zero records, which contribute exactly zero and are
// never near (r0 = 0).  __launch_bounds__'s second argument asks for 7
// waves per SIMD (72 VGPRs), which both pair loops fit.
template <int TILE>
__global__ __launch_bounds__(TILE, 7) void bs_mixed_kernel(BsArgs a)
{
    static_assert(TILE % 32 == 0, "groups of 32 segments");
    __shared__ float4 s0v[TILE];
    __shared__ float4 s1v[TILE];
    __shared__ float4 s2v[TILE];
    __shared__ float4 s3v[TILE];          // a - centre of the group, q^2
    __shared__ float4 far_sh[TILE / 32];  // centre, squared far radius
    __shared__ int item_sh;
    const int tid = threadIdx.x;
    const int64_t nseg = *a.nseg;
    const int64_t S = a.sstride;
    for (;;) {
        const int item = next_item(a.counter, &item_sh);
        if (item >= a.nitems) break;
        const int chunk = item / a.ntiles;
        const int tile = item - chunk * a.ntiles;
        const int64_t sb0 = (int64_t)chunk * a.chunk_len;
        const int64_t sb1 = min(nseg, sb0 + a.chunk_len);
        const int64_t t = (int64_t)tile * TILE + tid;
        const bool live = t < a.m;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (live) {
            px = a.t[t];
            py = a.t[a.tstride + t];
            pz = a.t[2 * a.tstride + t];
        }
        const float phx = (float)px, phy = (float)py, phz = (float)pz;
        const float plx = (float)(px - (double)phx), ply = (float)(py - (double)phy),
                    plz = (float)(pz - (double)phz);
        float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
        double nx = 0.0, ny = 0.0, nz = 0.0;
        for (int64_t base = sb0; base < sb1; base += TILE) {
            const int cnt = (int)min<int64_t>(TILE, sb1 - base);
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
                // far radius of this lane's group (lanes 32k .. 32k+31)
                const float cx = __shfl(v0.x, 0, 32), cy = __shfl(v0.y, 0, 32),
                            cz = __shfl(v0.z, 0, 32);
                if (tid < cnt32) {
                    s0v[tid] = v0;
                    s1v[tid] = v1;
                    s2v[tid] = v2;
                    s3v[tid] = make_float4((v0.x - cx) + v1.x, (v0.y - cy) + v1.y,
                                           (v0.z - cz) + v1.z, v1.w * v1.w);
                }
                const float dx = v0.x - cx, dy = v0.y - cy, dz = v0.z - cz;
                float ex = sqrtf(dx * dx + dy * dy + dz * dz) + sqrtf(v2.w);  // + 2 |r0|
                float lmax = sqrtf(v2.w);                                     // 2 |r0|
                if (tid >= cnt || !(v1.w > 1e-18f)) ex = __builtin_inff();
                for (int o = 16; o > 0; o >>= 1) {
                    ex = fmaxf(ex, __shfl_xor(ex, o, 32));
                    lmax = fmaxf(lmax, __shfl_xor(lmax, o, 32));
                }
                // a group wider than 64 of its longest segments (records that
                // jump between distant wake rows; a chain of 32 connected
                // segments spans at most 34) takes the near path, whose split
                // coordinates do not lose accuracy with the group's width
                if (ex > 32.0f * lmax) ex = __builtin_inff();
                if ((tid & 31) == 0 && tid < cnt32) {
                    const float cm = fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)));
                    const float rr = ex * 1.0001f + 1e-5f * cm + 1e-30f;
                    far_sh[tid >> 5] = make_float4(cx, cy, cz, rr * rr);
                }
            }
            __syncthreads();
            float tx = 0.f, ty = 0.f, tz = 0.f;
            for (int g = 0; g < cnt32; g += 32) {
                const float4 f = far_sh[g >> 5];
                const float qx = phx - f.x, qy = phy - f.y, qz = phz - f.z;
                const bool far = !live || (qx * qx + qy * qy + qz * qz > f.w);
                if (__builtin_amdgcn_readfirstlane(__all(far))) {
                    const float lx = qx + plx, ly = qy + ply, lz = qz + plz;
#pragma unroll 4
                    for (int j = 0; j < 32; ++j)
                        pair32_local(lx, ly, lz, s3v[g + j], s0v[g + j].w, s2v[g + j], tx, ty,
                                     tz);
                    continue;
                }
                unsigned mask = 0u;
#pragma unroll 4
                for (int j = 0; j < 32; ++j) {
                    const bool nearp = pair32(phx, phy, phz, plx, ply, plz, s0v[g + j],
                                              s1v[g + j], s2v[g + j], tx, ty, tz);
                    mask |= (nearp ? 1u : 0u) << j;
                }
                if (!live) mask = 0u;
                while (mask) {
                    const int j = __builtin_ctz(mask);
                    mask &= mask - 1u;
                    const int64_t s = base + g + j;
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
        if (live) {
            double *o = a.part + (size_t)chunk * 3 * a.m;
            o[t] = ((double)Sx - (double)Cx) + nx;
            o[a.m + t] = ((double)Sy - (double)Cy) + ny;
            o[2 * a.m + t] = ((double)Sz - (double)Cz) + nz;
        }
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

__global__ void zero_int_kernel(int *p) { *p = 0; }

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
            int64_t tstride, double *out, int accumulate, const BsWork *ws)
{
    if (m <= 0) return VLM_OK;
    // chunk partial sums: the context's growing buffer, or a fixed
    // workspace that caps the chunk count
    const int64_t max_chunks = ws ? (int64_t)(ws->part_cap / (3 * (size_t)m)) : INT32_MAX;
    if (ws && max_chunks < 1) return fail(ctx, VLM_ERR_ARG, "workspace too small for the targets");
    const bool prof = ctx->profile && !ws;
    if (m < 1024) {
        // about 2048 blocks, at least 8 segments per lane
        const int64_t groups = (m + SM_T - 1) / SM_T;
        int64_t nc = (2048 + groups - 1) / groups;
        const int64_t maxc = nseg_max / (8 * SM_BLOCK);
        if (nc > maxc) nc = maxc;
        if (nc > max_chunks) nc = max_chunks;
        if (nc < 1) nc = 1;
        const int nchunks = nseg_max > 0 ? (int)nc : 0;
        double *part = ws ? ws->part : nullptr;
        if (nchunks > 0) {
            if (!ws) {
                VLM_HIP(ctx->part.reserve((size_t)nchunks * 3 * (size_t)m));
                part = ctx->part.p;
            }
            hipLaunchKernelGGL(bs_small_kernel, dim3(nchunks, (unsigned)groups), dim3(SM_BLOCK),
                               0, ctx->stream, seg64, segstride, nseg_dev, nchunks, t, m, tstride,
                               part);
            VLM_HIP(hipGetLastError());
            if (prof) VLM_TRY(profile_count(ctx, nseg_dev, m, 1, nullptr));
        }
        hipLaunchKernelGGL(bs_reduce_kernel, dim3((unsigned)((m + RED_BLOCK - 1) / RED_BLOCK), 3),
                           dim3(RED_BLOCK), 0,
                           ctx->stream, part, nchunks, m, out, tstride, accumulate);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    const int blk = BS_BLOCK;
    const int ntiles = (int)((m + blk - 1) / blk);
    const bool mixed = ctx->precision == VLM_MIXED;
    int cap = mixed ? resident_blocks(ctx, bs_mixed_kernel<BS_BLOCK>, blk, &ctx->occ[1])
                    : resident_blocks(ctx, bs_fp64_kernel<BS_BLOCK>, blk, &ctx->occ[3]);
    if (cap <= 0) return fail(ctx, VLM_ERR_HIP, "occupancy query failed");
    int nchunks = 0;
    int64_t chunk_len = 1;
    if (nseg_max > 0) {
        // about 16 work items per resident workgroup, chunks of at least
        // four tiles
        const int64_t min_len = 4 * BS_BLOCK;
        int64_t nc = ((int64_t)16 * cap + ntiles - 1) / ntiles;
        const int64_t maxc = (nseg_max + min_len - 1) / min_len;
        if (nc > maxc) nc = maxc;
        if (nc > max_chunks) nc = max_chunks;
        if (nc < 1) nc = 1;
        chunk_len = (nseg_max + nc - 1) / nc;
        nchunks = (int)((nseg_max + chunk_len - 1) / chunk_len);
    }
    double *part = ws ? ws->part : nullptr;
    if (nchunks > 0) {
        int *counter = ws ? ws->counter : nullptr;
        if (!ws) {
            VLM_HIP(ctx->part.reserve((size_t)nchunks * 3 * (size_t)m));
            VLM_HIP(ctx->counter.reserve(1));
            part = ctx->part.p;
            counter = ctx->counter.p;
        }
        BsArgs a;
        a.s64 = seg64;
        a.s32 = seg32;
        a.sstride = segstride;
        a.nseg = nseg_dev;
        a.chunk_len = chunk_len;
        a.t = t;
        a.m = m;
        a.tstride = tstride;
        a.part = part;
        a.ntiles = ntiles;
        a.nitems = ntiles * nchunks;
        a.counter = counter;
        const int grid = a.nitems < cap ? a.nitems : cap;
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, ctx->stream, counter);
        Timer tm{};
        const bool timed = prof;
        if (timed) {
            VLM_HIP(hipEventCreate(&tm.a));
            VLM_HIP(hipEventCreate(&tm.b));
            VLM_HIP(hipEventRecord(tm.a, ctx->stream));
        }
        if (mixed)
            hipLaunchKernelGGL(bs_mixed_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0, ctx->stream,
                               a);
        else
            hipLaunchKernelGGL(bs_fp64_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0, ctx->stream,
                               a);
        VLM_HIP(hipGetLastError());
        if (timed) {
            VLM_HIP(hipEventRecord(tm.b, ctx->stream));
            VLM_TRY(profile_count(ctx, nseg_dev, m, 0, &tm));
        }
    }
    hipLaunchKernelGGL(bs_reduce_kernel, dim3((unsigned)((m + RED_BLOCK - 1) / RED_BLOCK), 3),
                       dim3(RED_BLOCK), 0,
                       ctx->stream, part, nchunks, m, out, tstride, accumulate);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // namespace vlm
