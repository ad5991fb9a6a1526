// Biot-Savart N-body kernels for MI355X (gfx950).
//
// Velocity induced by straight vortex segments with the Vatistas n=2 core,
// the leaf vf_vind of the reference (classdef.f90:476-503):
//   v = gam/(4 pi) * (r1 x r2) * [r0 . (r1/|r1| - r2/|r2|)]
//       / sqrt((rc |r0|)^4 + |r1 x r2|^4),
// zero when |r1 x r2|^2 <= eps^2, r/|r| = 0 when |r| <= eps
// (unitVec, libMath.f90:249-262).
//
// Work decomposition: one target point per lane; a workgroup of 256 lanes
// (four wave64s) walks one chunk of the segment list, staging 256 segments
// at a time in LDS and reading them back as wave-uniform broadcasts.  The
// segment list is split into chunks so that the grid holds several
// workgroups per CU even for a few thousand targets; per-chunk partial sums
// are reduced in a fixed order afterwards, so results are bitwise
// repeatable.  Block ids are remapped so that all tiles of one chunk run on
// one XCD and share its L2.  No MFMA: the pair is not a contraction.
//
// Precision modes
//   FP64  : all arithmetic in FP64; 1/sqrt via v_rsq_f64 plus two Newton
//           steps.
//   MIXED : FP32 pair arithmetic (v_rsq_f32), r1 x r2 formed as r0 x r1 to
//           avoid cancellation, per-tile FP32 sums folded into a Kahan-
//           compensated FP32 accumulator; pairs whose target lies within
//           kappa segment lengths of an end point are recomputed in FP64
//           from the FP64 copy of the segment (near-singular terms).
#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int BS_BLOCK = 256;      // wake passes: many targets
constexpr int BS_BLOCK_SMALL = 64; // collocation points: tens to hundreds of targets
constexpr int XCDS = 8;
constexpr double EPS2 = kEps * kEps;
constexpr float EPS2F = 4.93038065763132e-32f;
constexpr float KAPPA2 = 9.0f;  // near field: |r1| or |r2| < 3 |r0|

__device__ __forceinline__ double rsq64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    double r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    h = x * y;
    r = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, r, y);
    return y;
}

__device__ __forceinline__ void pair64(double px, double py, double pz,
                                       double ax, double ay, double az,
                                       double bx, double by, double bz,
                                       double g, double rc2, double &vx,
                                       double &vy, double &vz)
{
    const double r1x = px - ax, r1y = py - ay, r1z = pz - az;
    const double r2x = px - bx, r2y = py - by, r2z = pz - bz;
    const double r0x = r1x - r2x, r0y = r1y - r2y, r0z = r1z - r2z;
    const double cx = r1y * r2z - r1z * r2y;
    const double cy = r1z * r2x - r1x * r2z;
    const double cz = r1x * r2y - r1y * r2x;
    const double c2 = cx * cx + cy * cy + cz * cz;
    const double n1 = r1x * r1x + r1y * r1y + r1z * r1z;
    const double n2 = r2x * r2x + r2y * r2y + r2z * r2z;
    const double n0 = r0x * r0x + r0y * r0y + r0z * r0z;
    const double i1 = n1 > EPS2 ? rsq64(fmax(n1, EPS2)) : 0.0;
    const double i2 = n2 > EPS2 ? rsq64(fmax(n2, EPS2)) : 0.0;
    const double dot = (r0x * r1x + r0y * r1y + r0z * r1z) * i1 -
                       (r0x * r2x + r0y * r2y + r0z * r2z) * i2;
    const double q = rc2 * n0;
    const double den = q * q + c2 * c2;
    double s = g * dot * rsq64(den > 0.0 ? den : 1.0);
    s = c2 > EPS2 ? s : 0.0;
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
}

// FP32 pair.  Returns true when the pair belongs to the FP64 near field
// (its FP32 contribution is then zero).
__device__ __forceinline__ bool pair32(float px, float py, float pz,
                                       float4 s0, float4 s1, float &vx,
                                       float &vy, float &vz)
{
    const float r1x = px - s0.x, r1y = py - s0.y, r1z = pz - s0.z;
    const float r0x = s0.w, r0y = s1.x, r0z = s1.y;
    const float r2x = r1x - r0x, r2y = r1y - r0y, r2z = r1z - r0z;
    const float cx = r0y * r1z - r0z * r1y;
    const float cy = r0z * r1x - r0x * r1z;
    const float cz = r0x * r1y - r0y * r1x;
    const float c2 = cx * cx + cy * cy + cz * cz;
    const float n1 = r1x * r1x + r1y * r1y + r1z * r1z;
    const float n2 = r2x * r2x + r2y * r2y + r2z * r2z;
    const float n0 = r0x * r0x + r0y * r0y + r0z * r0z;
    const bool nearp = (n1 < KAPPA2 * n0) || (n2 < KAPPA2 * n0);
    const float d1 = r0x * r1x + r0y * r1y + r0z * r1z;
    const float d2 = d1 - n0;
    const float i1 = n1 > EPS2F ? __builtin_amdgcn_rsqf(n1) : 0.0f;
    const float i2 = n2 > EPS2F ? __builtin_amdgcn_rsqf(n2) : 0.0f;
    const float dot = d1 * i1 - d2 * i2;
    const float den = s1.w + c2 * c2;
    float s = s1.z * dot * __builtin_amdgcn_rsqf(den > 1e-37f ? den : 1.0f);
    s = (c2 > EPS2F && !nearp) ? s : 0.0f;
    vx += cx * s;
    vy += cy * s;
    vz += cz * s;
    return nearp;
}

struct BsArgs {
    const double *s64;
    const float *s32;
    int64_t sstride;
    const long long *nseg;
    int64_t chunk_len;
    const double *t;
    int64_t m;
    int64_t tstride;
    double *part;
    int ntiles;
    int nlogical;
};

template <int BS_TILE>
__global__ __launch_bounds__(BS_TILE) void bs_fp64_kernel(BsArgs a)
{
    const int per = gridDim.x / XCDS;
    const int lid = (blockIdx.x % XCDS) * per + blockIdx.x / XCDS;
    if (lid >= a.nlogical) return;
    const int chunk = lid / a.ntiles;
    const int tile = lid - chunk * a.ntiles;
    const int64_t nseg = *a.nseg;
    const int64_t s0 = (int64_t)chunk * a.chunk_len;
    const int64_t s1 = min(nseg, s0 + a.chunk_len);
    const int tid = threadIdx.x;
    const int64_t t = (int64_t)tile * BS_TILE + tid;
    const bool live = t < a.m;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (live) {
        px = a.t[t];
        py = a.t[a.tstride + t];
        pz = a.t[2 * a.tstride + t];
    }
    __shared__ double4 sa[BS_TILE];
    __shared__ double4 sb[BS_TILE];
    double vx = 0.0, vy = 0.0, vz = 0.0;
    const int64_t S = a.sstride;
    for (int64_t base = s0; base < s1; base += BS_TILE) {
        const int cnt = (int)min<int64_t>(BS_TILE, s1 - base);
        __syncthreads();
        if (tid < cnt) {
            const int64_t s = base + tid;
            sa[tid] = make_double4(a.s64[s], a.s64[S + s], a.s64[2 * S + s], a.s64[3 * S + s]);
            sb[tid] = make_double4(a.s64[4 * S + s], a.s64[5 * S + s], a.s64[6 * S + s],
                                   a.s64[7 * S + s]);
        }
        __syncthreads();
#pragma unroll 2
        for (int k = 0; k < cnt; ++k) {
            const double4 A = sa[k];
            const double4 B = sb[k];
            pair64(px, py, pz, A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, vx, vy, vz);
        }
    }
    if (live) {
        double *o = a.part + (size_t)chunk * 3 * a.m;
        o[t] = vx;
        o[a.m + t] = vy;
        o[2 * a.m + t] = vz;
    }
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    const float y = x - c;
    const float tt = s + y;
    c = (tt - s) - y;
    s = tt;
}

template <int BS_TILE>
__global__ __launch_bounds__(BS_TILE) void bs_mixed_kernel(BsArgs a)
{
    const int per = gridDim.x / XCDS;
    const int lid = (blockIdx.x % XCDS) * per + blockIdx.x / XCDS;
    if (lid >= a.nlogical) return;
    const int chunk = lid / a.ntiles;
    const int tile = lid - chunk * a.ntiles;
    const int64_t nseg = *a.nseg;
    const int64_t s0 = (int64_t)chunk * a.chunk_len;
    const int64_t s1 = min(nseg, s0 + a.chunk_len);
    const int tid = threadIdx.x;
    const int64_t t = (int64_t)tile * BS_TILE + tid;
    const bool live = t < a.m;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (live) {
        px = a.t[t];
        py = a.t[a.tstride + t];
        pz = a.t[2 * a.tstride + t];
    }
    const float fx = (float)px, fy = (float)py, fz = (float)pz;
    __shared__ float4 sa[BS_TILE];
    __shared__ float4 sb[BS_TILE];
    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    const int64_t S = a.sstride;
    for (int64_t base = s0; base < s1; base += BS_TILE) {
        const int cnt = (int)min<int64_t>(BS_TILE, s1 - base);
        __syncthreads();
        if (tid < cnt) {
            const int64_t s = base + tid;
            sa[tid] = make_float4(a.s32[s], a.s32[S + s], a.s32[2 * S + s], a.s32[3 * S + s]);
            sb[tid] = make_float4(a.s32[4 * S + s], a.s32[5 * S + s], a.s32[6 * S + s],
                                  a.s32[7 * S + s]);
        }
        __syncthreads();
        float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll 4
        for (int k = 0; k < cnt; ++k) {
            const bool nearp = pair32(fx, fy, fz, sa[k], sb[k], tx, ty, tz);
            if (nearp && live) {
                const int64_t s = base + k;
                pair64(px, py, pz, a.s64[s], a.s64[S + s], a.s64[2 * S + s], a.s64[3 * S + s],
                       a.s64[4 * S + s], a.s64[5 * S + s], a.s64[6 * S + s], a.s64[7 * S + s],
                       nx, ny, nz);
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

__global__ void bs_reduce_kernel(const double *part, int nchunks, int64_t m,
                                 double *out, int64_t ostride, int acc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    for (int d = 0; d < 3; ++d) {
        double s = acc ? out[d * ostride + t] : 0.0;
        for (int c = 0; c < nchunks; ++c) s += part[((size_t)c * 3 + d) * m + t];
        out[d * ostride + t] = s;
    }
}

}  // namespace

int bs_eval(vlm_ctx *ctx, const double *seg64, const float *seg32,
            int64_t segstride, const long long *nseg_dev, int64_t nseg_max,
            const double *t, int64_t m, int64_t tstride, double *out,
            int accumulate)
{
    if (m <= 0) return VLM_OK;
    const bool small = m < 1024;
    const int blk = small ? BS_BLOCK_SMALL : BS_BLOCK;
    const int ntiles = (int)((m + blk - 1) / blk);
    int nchunks = 0;
    int64_t chunk_len = 1;
    if (nseg_max > 0) {
        const int64_t want = 4096;
        int64_t nc = (want + ntiles - 1) / ntiles;
        const int64_t maxc = (nseg_max + blk - 1) / blk;
        if (nc > maxc) nc = maxc;
        if (nc < 1) nc = 1;
        chunk_len = (nseg_max + nc - 1) / nc;
        nchunks = (int)((nseg_max + chunk_len - 1) / chunk_len);
    }
    if (nchunks > 0) {
        VLM_HIP(ctx->part.reserve((size_t)nchunks * 3 * (size_t)m));
        BsArgs a;
        a.s64 = seg64;
        a.s32 = seg32;
        a.sstride = segstride;
        a.nseg = nseg_dev;
        a.chunk_len = chunk_len;
        a.t = t;
        a.m = m;
        a.tstride = tstride;
        a.part = ctx->part.p;
        a.ntiles = ntiles;
        a.nlogical = ntiles * nchunks;
        const int grid = ((a.nlogical + XCDS - 1) / XCDS) * XCDS;
        Timer tm{};
        const bool timed = ctx->profile && !small;
        if (timed) {
            VLM_HIP(hipEventCreate(&tm.a));
            VLM_HIP(hipEventCreate(&tm.b));
            VLM_HIP(hipEventRecord(tm.a, ctx->stream));
        }
        if (ctx->precision == VLM_MIXED) {
            if (small)
                hipLaunchKernelGGL(bs_mixed_kernel<BS_BLOCK_SMALL>, dim3(grid), dim3(blk), 0,
                                   ctx->stream, a);
            else
                hipLaunchKernelGGL(bs_mixed_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0,
                                   ctx->stream, a);
        } else {
            if (small)
                hipLaunchKernelGGL(bs_fp64_kernel<BS_BLOCK_SMALL>, dim3(grid), dim3(blk), 0,
                                   ctx->stream, a);
            else
                hipLaunchKernelGGL(bs_fp64_kernel<BS_BLOCK>, dim3(grid), dim3(blk), 0,
                                   ctx->stream, a);
        }
        VLM_HIP(hipGetLastError());
        if (timed) {
            VLM_HIP(hipEventRecord(tm.b, ctx->stream));
            tm.pairs = -1.0;  // resolved from the device count at read-out
            ctx->timers.push_back(tm);
            ctx->prof_launches += 1;
        }
    }
    hipLaunchKernelGGL(bs_reduce_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                       ctx->stream, ctx->part.p, nchunks, m, out, tstride, accumulate);
    VLM_HIP(hipGetLastError());
    if (ctx->profile && nchunks > 0 && !small) {
        // pair count = m * device segment count, copied to pinned memory
        // without a stream synchronisation and resolved at read-out
        Timer &tb = ctx->timers.back();
        const size_t slot = ctx->timers.size() - 1;
        if (slot < ctx->prof_cap) {
            VLM_HIP(hipMemcpyAsync(ctx->prof_nseg + slot, nseg_dev, sizeof(long long),
                                   hipMemcpyDeviceToHost, ctx->stream));
            tb.pairs = (double)m;
        }
    }
    return VLM_OK;
}

}  // namespace vlm
