// vlm_filament_vind_noalloc: the filament sum of vlm_filament_vind on
// device-resident inputs, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// Every intermediate buffer of the call lives in one file-scope device
// array, `scratch`, of NOALLOC_SCRATCH_BYTES: the segment records (FP64
// and FP32 structure of arrays, the count), the targets and sums in
// structure-of-arrays form, the chunk partial sums of the Biot-Savart
// launch and its work counter.  The call enqueues kernels on the context's
// stream and returns; it calls no hipMalloc, no hipFree, no allocating
// library routine and no hipMemsetAsync (counters are cleared by kernels).
// The array's device address is taken with hipGetSymbolAddress once per
// device, at the first call on that device, and kept per device.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: calls that overlap on one GPU (two contexts, two streams, or
// two host threads) race and must be serialised by the caller.
#include <mutex>

#include "vlm_internal.h"

namespace vlm {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;  // 64 MiB per GPU
}

static __device__ __align__(256) unsigned char scratch[vlm::NOALLOC_SCRATCH_BYTES];

namespace vlm {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the context's device (which is
// current), looked up on the first call per device.
int scratch_address(vlm_ctx *ctx, unsigned char **ptr)
{
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return fail(ctx, VLM_ERR_ARG, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[ctx->device] = static_cast<unsigned char *>(p);
    }
    *ptr = g_scratch_ptr[ctx->device];
    return VLM_OK;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

// Records of an explicit list (every filament kept, as in
// vlm_filament_vind) at their own index, from 3 x n AoS device arrays.
__global__ void fil_records_kernel(int64_t n, const double *a, const double *b, const double *g,
                                   const double *rc, double *d64, float *d32, int64_t cap)
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
    write_segment_record(s, i, d64, d32, cap);
}

__global__ void set_count_kernel(long long *count, int64_t n) { *count = (long long)n; }

__global__ void aos_to_soa3_kernel(int64_t m, const double *aos, double *soa)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    soa[i] = aos[3 * i];
    soa[m + i] = aos[3 * i + 1];
    soa[2 * m + i] = aos[3 * i + 2];
}

__global__ void soa3_to_aos_kernel(int64_t m, const double *soa, double *aos)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    aos[3 * i] = soa[i];
    aos[3 * i + 1] = soa[m + i];
    aos[3 * i + 2] = soa[2 * m + i];
}

}  // namespace
}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_filament_vind_noalloc(vlm_ctx *ctx, int64_t nfil, const double *a, const double *b,
                              const double *gam, const double *rc, int64_t m, const double *p,
                              double *v, int accumulate)
{
    if (!ctx) return VLM_ERR_ARG;
    if (nfil < 0 || m < 0) return fail(ctx, VLM_ERR_ARG, "negative size");
    if (m == 0) return VLM_OK;
    if (!p || !v || (nfil > 0 && (!a || !b || !gam || !rc)))
        return fail(ctx, VLM_ERR_ARG, "null buffer");
    // workspace layout: d64 | d32 | count | counter | targets | sums | partials
    const int64_t cap = nfil > 0 ? nfil : 1;
    const size_t o64 = 0;
    const size_t o32 = o64 + up256((size_t)SEG64_N * cap * sizeof(double));
    const size_t ocnt = o32 + up256((size_t)SEG32_N * cap * sizeof(float));
    const size_t octr = ocnt + 256;
    const size_t otgt = octr + 256;
    const size_t oout = otgt + up256((size_t)3 * m * sizeof(double));
    const size_t opart = oout + up256((size_t)3 * m * sizeof(double));
    if (nfil > ((int64_t)1 << 32) || m > ((int64_t)1 << 32) ||
        opart + (size_t)3 * m * sizeof(double) > NOALLOC_SCRATCH_BYTES)
        return fail(ctx, VLM_ERR_ARG,
                    "vlm_filament_vind_noalloc: " + std::to_string(nfil) + " filaments and " +
                        std::to_string(m) + " points exceed the " +
                        std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                        " MiB workspace (about 120 B per filament and 72 B per point, plus "
                        "24 B per point for each chunk of partial sums)");
    VLM_DEVICE(ctx);
    unsigned char *w = nullptr;
    VLM_TRY(scratch_address(ctx, &w));
    double *d64 = reinterpret_cast<double *>(w + o64);
    float *d32 = reinterpret_cast<float *>(w + o32);
    long long *count = reinterpret_cast<long long *>(w + ocnt);
    int *counter = reinterpret_cast<int *>(w + octr);
    double *tgt = reinterpret_cast<double *>(w + otgt);
    double *out = reinterpret_cast<double *>(w + oout);
    BsWork ws;
    ws.part = reinterpret_cast<double *>(w + opart);
    ws.part_cap = (NOALLOC_SCRATCH_BYTES - opart) / sizeof(double);
    ws.counter = counter;
    if (nfil > 0)
        hipLaunchKernelGGL(fil_records_kernel, dim3(nblk(nfil)), dim3(256), 0, ctx->stream, nfil,
                           a, b, gam, rc, d64, d32, cap);
    hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(1), 0, ctx->stream, count, nfil);
    hipLaunchKernelGGL(aos_to_soa3_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, m, p, tgt);
    if (accumulate)
        hipLaunchKernelGGL(aos_to_soa3_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, m, v,
                           out);
    VLM_HIP(hipGetLastError());
    VLM_TRY(bs_eval(ctx, d64, d32, cap, count, nfil, tgt, m, m, out, accumulate, &ws));
    hipLaunchKernelGGL(soa3_to_aos_kernel, dim3(nblk(m)), dim3(256), 0, ctx->stream, m, out, v);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

}  // extern "C"
