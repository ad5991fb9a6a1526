// Context, error handling, segment compaction, explicit filament sets and
// kernel timing for libvlm.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "vlm_internal.h"

namespace vlm {

int fail(vlm_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(vlm_ctx *ctx, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? VLM_ERR_ALLOC : VLM_ERR_HIP, m);
}

namespace {

// One record per staged slot -> SoA FP64 and FP32 copies at the compacted
// position.  The last thread publishes the count.
__global__ void compact_kernel(const Slot *slots, const int *flags, const int *idx,
                               int64_t n, double *d64, float *d32, int64_t cap,
                               long long *count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == n - 1) *count = (long long)idx[i] + (long long)flags[i];
    if (!flags[i]) return;
    write_segment_record(slots[i], idx[i], d64, d32, cap);
}

__global__ void zero_count_kernel(long long *count) { *count = 0; }

}  // namespace

int ensure_slots(vlm_ctx *ctx, int64_t n)
{
    VLM_HIP(ctx->slots.reserve((size_t)n));
    VLM_HIP(ctx->flags.reserve((size_t)n));
    VLM_HIP(ctx->idx.reserve((size_t)n));
    return VLM_OK;
}

int compact_into(vlm_ctx *ctx, const Slot *slots, const int *flags, int64_t nslots,
                 DBuf<double> &d64, DBuf<float> &d32, DBuf<long long> &count, int64_t cap)
{
    VLM_HIP(count.reserve(1));
    if (nslots <= 0) {
        hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, ctx->stream, count.p);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    if (nslots > 0x7fffffff) return fail(ctx, VLM_ERR_ARG, "segment list too large");
    size_t bytes = 0;
    VLM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flags, ctx->idx.p, (int)nslots,
                                             ctx->stream));
    if (bytes > ctx->scan_tmp.n) VLM_HIP(ctx->scan_tmp.reserve(bytes));
    bytes = ctx->scan_tmp.n;
    VLM_HIP(hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, bytes, flags, ctx->idx.p,
                                             (int)nslots, ctx->stream));
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0,
                       ctx->stream, slots, flags, ctx->idx.p, nslots, d64.p, d32.p, cap,
                       count.p);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int compact_slots(vlm_ctx *ctx, int64_t nslots)
{
    const int64_t cap = nslots > 0 ? nslots : 1;
    if (ctx->seg.cap < cap) {
        VLM_HIP(ctx->seg.d64.reserve((size_t)SEG64_N * cap));
        VLM_HIP(ctx->seg.d32.reserve((size_t)SEG32_N * cap));
        ctx->seg.cap = cap;
    }
    return compact_into(ctx, ctx->slots.p, ctx->flags.p, nslots, ctx->seg.d64, ctx->seg.d32,
                        ctx->seg.count, ctx->seg.cap);
}

namespace {

__global__ void stage_fil_kernel(int64_t n, const double *a, const double *b, const double *g,
                                 const double *rc, Slot *slots, int *flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Slot s;
    s.ax = a[3 * i];
    s.ay = a[3 * i + 1];
    s.az = a[3 * i + 2];
    s.bx = b[3 * i];
    s.by = b[3 * i + 1];
    s.bz = b[3 * i + 2];
    s.gam = g[i];
    s.rc = rc[i];
    slots[i] = s;
    flags[i] = 1;  // explicit lists keep every filament (no |gam| skip)
}

// AoS host layout (3 x m) -> SoA device layout
__global__ void aos_to_soa_kernel(int64_t m, const double *aos, double *soa, int64_t stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    soa[i] = aos[3 * i];
    soa[stride + i] = aos[3 * i + 1];
    soa[2 * stride + i] = aos[3 * i + 2];
}

__global__ void soa_to_aos_kernel(int64_t m, const double *soa, int64_t stride, double *aos)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    aos[3 * i] = soa[i];
    aos[3 * i + 1] = soa[stride + i];
    aos[3 * i + 2] = soa[2 * stride + i];
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

int upload_fset(vlm_ctx *ctx, int64_t nfil, const double *a, const double *b,
                const double *gam, const double *rc, DBuf<double> &d64, DBuf<float> &d32,
                DBuf<long long> &count)
{
    DBuf<double> raw;
    VLM_HIP(raw.reserve((size_t)8 * nfil));
    VLM_HIP(hipMemcpyAsync(raw.p, a, sizeof(double) * 3 * nfil, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_HIP(hipMemcpyAsync(raw.p + 3 * nfil, b, sizeof(double) * 3 * nfil,
                           hipMemcpyHostToDevice, ctx->stream));
    VLM_HIP(hipMemcpyAsync(raw.p + 6 * nfil, gam, sizeof(double) * nfil, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_HIP(hipMemcpyAsync(raw.p + 7 * nfil, rc, sizeof(double) * nfil, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_TRY(ensure_slots(ctx, nfil));
    hipLaunchKernelGGL(stage_fil_kernel, dim3(nblk(nfil)), dim3(256), 0, ctx->stream, nfil,
                       raw.p, raw.p + 3 * nfil, raw.p + 6 * nfil, raw.p + 7 * nfil,
                       ctx->slots.p, ctx->flags.p);
    VLM_HIP(hipGetLastError());
    VLM_HIP(d64.reserve((size_t)SEG64_N * nfil));
    VLM_HIP(d32.reserve((size_t)SEG32_N * nfil));
    VLM_TRY(compact_into(ctx, ctx->slots.p, ctx->flags.p, nfil, d64, d32, count, nfil));
    VLM_TRY(sync_stream(ctx));  // raw is released on return
    return VLM_OK;
}

}  // namespace

static void flush_deferred(vlm_ctx *ctx)
{
    for (const auto &d : ctx->deferred) std::memcpy(d.dst, d.src, d.bytes);
    ctx->deferred.clear();
}

int sync_stream(vlm_ctx *ctx)
{
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    flush_deferred(ctx);
    ctx->pin_off = 0;
    return VLM_OK;
}

int d2h_deferred(vlm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return VLM_OK;
    char *p = nullptr;
    VLM_TRY(pin_take(ctx, bytes, &p));
    VLM_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (!ctx->defer_ev) VLM_HIP(hipEventCreateWithFlags(&ctx->defer_ev, hipEventDisableTiming));
    VLM_HIP(hipEventRecord(ctx->defer_ev, ctx->stream));
    ctx->deferred.push_back({dst, p, bytes});
    return VLM_OK;
}

// Waits only for the work enqueued up to the last deferred copy; the arena
// is not recycled here because later work on the stream may still use it.
int wait_deferred(vlm_ctx *ctx)
{
    if (ctx->deferred.empty()) return VLM_OK;
    VLM_HIP(hipEventSynchronize(ctx->defer_ev));
    flush_deferred(ctx);
    return VLM_OK;
}

// Reserve bytes of the pinned arena (64-byte aligned), synchronising and
// recycling it when full, growing it when a request exceeds its size.
int pin_take(vlm_ctx *ctx, size_t bytes, char **out)
{
    const size_t need = (bytes + 63) & ~(size_t)63;
    if (ctx->pin_off + need > ctx->pin_cap) {
        VLM_TRY(sync_stream(ctx));
        if (need > ctx->pin_cap) {
            if (ctx->pin) VLM_HIP(hipHostFree(ctx->pin));
            ctx->pin = nullptr;
            ctx->pin_cap = 0;
            size_t cap = need > ((size_t)8 << 20) ? need : ((size_t)8 << 20);
            VLM_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->pin), cap,
                                  hipHostMallocDefault));
            ctx->pin_cap = cap;
        }
    }
    *out = ctx->pin + ctx->pin_off;
    ctx->pin_off += need;
    return VLM_OK;
}

int h2d(vlm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return VLM_OK;
    char *p = nullptr;
    VLM_TRY(pin_take(ctx, bytes, &p));
    std::memcpy(p, src, bytes);
    VLM_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    return VLM_OK;
}

int d2h(vlm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return sync_stream(ctx);
    char *p = nullptr;
    VLM_TRY(pin_take(ctx, bytes, &p));
    VLM_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VLM_TRY(sync_stream(ctx));
    std::memcpy(dst, p, bytes);
    return VLM_OK;
}

int points_to_device(vlm_ctx *ctx, int64_t m, const double *p, double *soa, int64_t stride)
{
    VLM_HIP(ctx->hostio.reserve((size_t)3 * m));
    VLM_TRY(h2d(ctx, ctx->hostio.p, p, sizeof(double) * 3 * m));
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, m,
                       ctx->hostio.p, soa, stride);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

int points_to_host(vlm_ctx *ctx, int64_t m, const double *soa, int64_t stride, double *p,
                   bool deferred = false)
{
    VLM_HIP(ctx->hostio.reserve((size_t)3 * m));
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, m, soa,
                       stride, ctx->hostio.p);
    VLM_HIP(hipGetLastError());
    if (deferred) return d2h_deferred(ctx, p, ctx->hostio.p, sizeof(double) * 3 * m);
    return d2h(ctx, p, ctx->hostio.p, sizeof(double) * 3 * m);
}

}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_create(vlm_ctx **out, int device)
{
    if (!out) return VLM_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return VLM_ERR_HIP;
    if (device < 0 || device >= ndev) return VLM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return VLM_ERR_HIP;
    vlm_ctx *ctx = new vlm_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return VLM_ERR_HIP;
    }
    ctx->prof_cap = 1 << 16;
    if (hipHostMalloc(reinterpret_cast<void **>(&ctx->prof_nseg),
                      ctx->prof_cap * sizeof(long long), hipHostMallocDefault) != hipSuccess) {
        ctx->prof_nseg = nullptr;
        ctx->prof_cap = 0;
    }
    *out = ctx;
    return VLM_OK;
}

int vlm_destroy(vlm_ctx *ctx)
{
    if (!ctx) return VLM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    flush_deferred(ctx);
    mlp_release(ctx);
    for (auto *r : ctx->rotors) delete r;
    for (auto *f : ctx->fsets) delete f;
    for (auto *p : ctx->psets) delete p;
    for (auto *p : ctx->aic) (void)hipFree(p);
    for (auto *p : ctx->lu) (void)hipFree(p);
    for (auto *p : ctx->piv) (void)hipFree(p);
    for (auto *v : {&ctx->timers, &ctx->lu_timers})
        for (auto &t : *v) {
            if (t.a) (void)hipEventDestroy(t.a);
            if (t.b) (void)hipEventDestroy(t.b);
        }
    if (ctx->prof_nseg) (void)hipHostFree(ctx->prof_nseg);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->judge_ev) (void)hipEventDestroy(ctx->judge_ev);
    if (ctx->defer_ev) (void)hipEventDestroy(ctx->defer_ev);
    if (ctx->judge_pin) (void)hipHostFree(ctx->judge_pin);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VLM_OK;
}

const char *vlm_last_error(const vlm_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int vlm_set_precision(vlm_ctx *ctx, int precision)
{
    if (!ctx) return VLM_ERR_ARG;
    if (precision != VLM_FP64 && precision != VLM_MIXED)
        return fail(ctx, VLM_ERR_ARG, "precision must be VLM_FP64 or VLM_MIXED");
    ctx->precision = precision;
    return VLM_OK;
}

int vlm_set_bug_compat(vlm_ctx *ctx, int bug_compat)
{
    if (!ctx) return VLM_ERR_ARG;
    ctx->bug_compat = bug_compat ? 1 : 0;
    return VLM_OK;
}

int vlm_synchronize(vlm_ctx *ctx)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_TRY(sync_stream(ctx));
    return VLM_OK;
}

int vlm_wait(vlm_ctx *ctx)
{
    if (!ctx) return VLM_ERR_ARG;
    return wait_deferred(ctx);
}

int vlm_filament_vind(vlm_ctx *ctx, int64_t nfil, const double *a, const double *b,
                      const double *gam, const double *rc, int64_t m, const double *p,
                      double *v, int accumulate)
{
    if (!ctx) return VLM_ERR_ARG;
    if (nfil < 0 || m < 0) return fail(ctx, VLM_ERR_ARG, "negative size");
    if (m == 0) return VLM_OK;
    if (!p || !v || (nfil > 0 && (!a || !b || !gam || !rc)))
        return fail(ctx, VLM_ERR_ARG, "null buffer");
    VLM_HIP(hipSetDevice(ctx->device));
    FSet fs;
    if (nfil > 0) VLM_TRY(upload_fset(ctx, nfil, a, b, gam, rc, fs.d64, fs.d32, fs.count));
    else {
        VLM_HIP(fs.d64.reserve(SEG64_N));
        VLM_HIP(fs.d32.reserve(SEG32_N));
        VLM_TRY(compact_into(ctx, nullptr, nullptr, 0, fs.d64, fs.d32, fs.count, 1));
    }
    VLM_HIP(ctx->tgt.reserve((size_t)3 * m));
    VLM_HIP(ctx->out.reserve((size_t)3 * m));
    VLM_TRY(points_to_device(ctx, m, p, ctx->tgt.p, m));
    if (accumulate) VLM_TRY(points_to_device(ctx, m, v, ctx->out.p, m));
    const int64_t stride = nfil > 0 ? nfil : 1;
    VLM_TRY(bs_eval(ctx, fs.d64.p, fs.d32.p, stride, fs.count.p, nfil, ctx->tgt.p, m, m,
                    ctx->out.p, accumulate));
    VLM_TRY(points_to_host(ctx, m, ctx->out.p, m, v));
    return VLM_OK;
}

int vlm_fset_create(vlm_ctx *ctx, int64_t nfil, const double *a, const double *b,
                    const double *gam, const double *rc, int *fset)
{
    if (!ctx || !fset) return VLM_ERR_ARG;
    if (nfil <= 0 || !a || !b || !gam || !rc) return fail(ctx, VLM_ERR_ARG, "empty filament set");
    VLM_HIP(hipSetDevice(ctx->device));
    FSet *fs = new FSet();
    int rc_ = upload_fset(ctx, nfil, a, b, gam, rc, fs->d64, fs->d32, fs->count);
    if (rc_ != VLM_OK) {
        delete fs;
        return rc_;
    }
    fs->n = nfil;
    ctx->fsets.push_back(fs);
    *fset = (int)ctx->fsets.size() - 1;
    return VLM_OK;
}

int vlm_pset_create(vlm_ctx *ctx, int64_t m, const double *p, int *pset)
{
    if (!ctx || !pset) return VLM_ERR_ARG;
    if (m <= 0 || !p) return fail(ctx, VLM_ERR_ARG, "empty point set");
    VLM_HIP(hipSetDevice(ctx->device));
    PSet *ps = new PSet();
    if (ps->p.reserve((size_t)3 * m) != hipSuccess || ps->v.reserve((size_t)3 * m) != hipSuccess) {
        delete ps;
        return fail(ctx, VLM_ERR_ALLOC, "point set allocation failed");
    }
    ps->m = m;
    int rc_ = points_to_device(ctx, m, p, ps->p.p, m);
    if (rc_ != VLM_OK) {
        delete ps;
        return rc_;
    }
    ctx->psets.push_back(ps);
    *pset = (int)ctx->psets.size() - 1;
    return VLM_OK;
}

int vlm_fset_vind(vlm_ctx *ctx, int fset, int pset)
{
    if (!ctx) return VLM_ERR_ARG;
    if (fset < 0 || fset >= (int)ctx->fsets.size() || pset < 0 ||
        pset >= (int)ctx->psets.size())
        return fail(ctx, VLM_ERR_ARG, "unknown filament or point set");
    VLM_HIP(hipSetDevice(ctx->device));
    FSet *fs = ctx->fsets[fset];
    PSet *ps = ctx->psets[pset];
    return bs_eval(ctx, fs->d64.p, fs->d32.p, fs->n, fs->count.p, fs->n, ps->p.p, ps->m, ps->m,
                   ps->v.p, 0);
}

int vlm_pset_velocity(vlm_ctx *ctx, int pset, double *v)
{
    if (!ctx || !v) return VLM_ERR_ARG;
    if (pset < 0 || pset >= (int)ctx->psets.size())
        return fail(ctx, VLM_ERR_ARG, "unknown point set");
    VLM_HIP(hipSetDevice(ctx->device));
    PSet *ps = ctx->psets[pset];
    return points_to_host(ctx, ps->m, ps->v.p, ps->m, v);
}

int vlm_profile_enable(vlm_ctx *ctx, int enable)
{
    if (!ctx) return VLM_ERR_ARG;
    ctx->profile = enable ? 1 : 0;
    return VLM_OK;
}

int vlm_profile_reset(vlm_ctx *ctx)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_TRY(sync_stream(ctx));
    for (auto &t : ctx->timers) {
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
    }
    ctx->timers.clear();
    for (auto &t : ctx->lu_timers) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    ctx->lu_timers.clear();
    ctx->prof_ms = 0.0;
    ctx->prof_pairs = 0.0;
    ctx->prof_launches = 0;
    return VLM_OK;
}

int vlm_profile_get(vlm_ctx *ctx, double *bs_ms, int64_t *bs_launches, double *pairs)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_TRY(sync_stream(ctx));
    double ms = 0.0, pr = 0.0;
    int64_t n = 0;
    for (size_t i = 0; i < ctx->timers.size(); ++i) {
        if (ctx->timers[i].kind != 0) continue;
        float e = 0.f;
        VLM_HIP(hipEventElapsedTime(&e, ctx->timers[i].a, ctx->timers[i].b));
        ms += e;
        ++n;
        if (i < ctx->prof_cap && ctx->timers[i].pairs > 0)
            pr += ctx->timers[i].pairs * (double)ctx->prof_nseg[i];
    }
    if (bs_ms) *bs_ms = ms;
    if (bs_launches) *bs_launches = n;
    if (pairs) *pairs = pr;
    return VLM_OK;
}

int vlm_profile_get_all(vlm_ctx *ctx, double *pairs, int64_t *launches)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_TRY(sync_stream(ctx));
    double pr = 0.0;
    for (size_t i = 0; i < ctx->timers.size(); ++i)
        if (i < ctx->prof_cap && ctx->timers[i].pairs > 0)
            pr += ctx->timers[i].pairs * (double)ctx->prof_nseg[i];
    if (pairs) *pairs = pr;
    if (launches) *launches = (int64_t)ctx->timers.size();
    return VLM_OK;
}

int vlm_profile_get_lu(vlm_ctx *ctx, double *ms, int64_t *factorisations, double *flops)
{
    if (!ctx) return VLM_ERR_ARG;
    VLM_TRY(sync_stream(ctx));
    double t = 0.0, f = 0.0;
    for (auto &tm : ctx->lu_timers) {
        float e = 0.f;
        VLM_HIP(hipEventElapsedTime(&e, tm.a, tm.b));
        t += e;
        f += tm.pairs;
    }
    if (ms) *ms = t;
    if (factorisations) *factorisations = (int64_t)ctx->lu_timers.size();
    if (flops) *flops = f;
    return VLM_OK;
}

}  // extern "C"

// Calibration streams for vlm_fp64_issue_rate: every lane runs 8 independent
// chains of CALIB_ITERS dependent operations (FMA, or v_rsq_f64) and stores
// one value at its global index, so the output holds exactly one double per
// launched lane.
namespace {
constexpr int CALIB_ITERS = 4096;
constexpr int CALIB_CHAINS = 8;

__global__ void __launch_bounds__(256) calib_fma_kernel(double *out, double s)
{
    double a[CALIB_CHAINS];
    for (int c = 0; c < CALIB_CHAINS; ++c) a[c] = s + c + threadIdx.x * 1e-9;
    for (int i = 0; i < CALIB_ITERS; ++i)
#pragma unroll
        for (int c = 0; c < CALIB_CHAINS; ++c) a[c] = __builtin_fma(a[c], 0.999999, 1e-7);
    double t = 0.0;
    for (int c = 0; c < CALIB_CHAINS; ++c) t += a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}

__global__ void __launch_bounds__(256) calib_rsq_kernel(double *out, double s)
{
    double a[CALIB_CHAINS];
    for (int c = 0; c < CALIB_CHAINS; ++c) a[c] = s + c + threadIdx.x * 1e-9;
    for (int i = 0; i < CALIB_ITERS; ++i)
#pragma unroll
        for (int c = 0; c < CALIB_CHAINS; ++c) a[c] = __builtin_amdgcn_rsq(a[c]);
    double t = 0.0;
    for (int c = 0; c < CALIB_CHAINS; ++c) t += a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}
}  // namespace

extern "C" {

int vlm_fp64_issue_rate(vlm_ctx *ctx, double *fma_tflops, double *rsq_over_fma)
{
    if (!ctx || !fma_tflops || !rsq_over_fma) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
    VLM_TRY(sync_stream(ctx));
    int cus = 0;
    VLM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int threads = 256, blocks = cus * 16;
    double *out = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float ms_f = 0.f, ms_r = 0.f;
    hipError_t err = hipMalloc(&out, sizeof(double) * (size_t)blocks * threads);
    for (int i = 0; i < 3 && err == hipSuccess; ++i) err = hipEventCreate(&ev[i]);
    // the first round is untimed (clocks ramp up), the second is timed
    for (int round = 0; round < 2 && err == hipSuccess; ++round) {
        err = hipEventRecord(ev[0], ctx->stream);
        if (err != hipSuccess) break;
        hipLaunchKernelGGL(calib_fma_kernel, dim3(blocks), dim3(threads), 0, ctx->stream, out, 1.5);
        err = hipEventRecord(ev[1], ctx->stream);
        if (err != hipSuccess) break;
        hipLaunchKernelGGL(calib_rsq_kernel, dim3(blocks), dim3(threads), 0, ctx->stream, out, 1.5);
        err = hipEventRecord(ev[2], ctx->stream);
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventSynchronize(ev[2]);
    }
    if (err == hipSuccess) err = hipEventElapsedTime(&ms_f, ev[0], ev[1]);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms_r, ev[1], ev[2]);
    for (int i = 0; i < 3; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (out) (void)hipFree(out);
    if (err != hipSuccess) return vlm::hip_fail(ctx, err, "vlm_fp64_issue_rate");
    const double fmas = 64.0 * (double)blocks * (threads / 64) * CALIB_ITERS * CALIB_CHAINS;
    *fma_tflops = 2.0 * fmas / (ms_f * 1e-3) / 1e12;
    *rsq_over_fma = (double)ms_r / (double)ms_f;
    return VLM_OK;
}

}  // extern "C"
