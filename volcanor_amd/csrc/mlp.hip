// Learned wake-velocity predictor: a 45 -> 64 -> 64 -> 64 -> 3 MLP (tanh
// hidden layers, FP32), with hand-written inference, backward and Adam
// kernels.  No reference counterpart: it stands in for pass 2 of the AM2
// predictor-corrector (the Biot-Savart pass on the predicted wake,
// main.f90:1052-1086) and is trained online on the pairs that full passes
// produce.
//
// Parameter packing (row-major W[out][in]):
//   W1 64x45 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 3x64 | b4 3
//
// Inference: one sample per lane, all weights staged in LDS (45.8 KB) and
// read as wave-uniform broadcasts.  Training: kernel A runs the forward and
// backward sweep per sample and stores activations and deltas; kernel B
// forms each parameter's gradient as a fixed-order sum over the batch
// (deterministic, no atomics) and applies the Adam update.
#include <cmath>
#include <vector>

#include "vlm_internal.h"

namespace vlm {

constexpr int MLP_IN = 45, MLP_H = 64, MLP_OUT = 3;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_H;
constexpr int OFF_W4 = OFF_B3 + MLP_H;
constexpr int OFF_B4 = OFF_W4 + MLP_OUT * MLP_H;
constexpr int MLP_NP = OFF_B4 + MLP_OUT;  // 11459
constexpr int ACT = MLP_IN + 3 * MLP_H;    // stored activations per sample
constexpr int DEL = 3 * MLP_H + MLP_OUT;   // stored deltas per sample

namespace {

__device__ __forceinline__ void dense_tanh(const float *W, const float *b, const float *x, int nin,
                                           float *y)
{
    for (int o = 0; o < MLP_H; ++o) {
        float a = b[o];
        for (int i = 0; i < nin; ++i) a = __builtin_fmaf(W[o * nin + i], x[i], a);
        y[o] = tanhf(a);
    }
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *params, int64_t m,
                                                          const float *X, float *Y)
{
    extern __shared__ float sw[];
    for (int k = threadIdx.x; k < MLP_NP; k += blockDim.x) sw[k] = params[k];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    float x[MLP_IN], h1[MLP_H], h2[MLP_H];
    for (int i = 0; i < MLP_IN; ++i) x[i] = X[s * MLP_IN + i];
    dense_tanh(sw + OFF_W1, sw + OFF_B1, x, MLP_IN, h1);
    dense_tanh(sw + OFF_W2, sw + OFF_B2, h1, MLP_H, h2);
    dense_tanh(sw + OFF_W3, sw + OFF_B3, h2, MLP_H, h1);
    for (int o = 0; o < MLP_OUT; ++o) {
        float a = sw[OFF_B4 + o];
        for (int i = 0; i < MLP_H; ++i) a = __builtin_fmaf(sw[OFF_W4 + o * MLP_H + i], h1[i], a);
        Y[s * MLP_OUT + o] = a;
    }
}

// Forward + backward of one sample per lane; stores activations
// [x | h1 | h2 | h3] and deltas [d1 | d2 | d3 | dy], and the sample loss.
__global__ __launch_bounds__(256) void mlp_backprop_kernel(const float *params, int64_t m,
                                                           const float *X, const float *T,
                                                           float *acts, float *dels,
                                                           float *loss)
{
    extern __shared__ float sw[];
    for (int k = threadIdx.x; k < MLP_NP; k += blockDim.x) sw[k] = params[k];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    float *a = acts + s * ACT;
    float *d = dels + s * DEL;
    float *x = a, *h1 = a + MLP_IN, *h2 = h1 + MLP_H, *h3 = h2 + MLP_H;
    for (int i = 0; i < MLP_IN; ++i) x[i] = X[s * MLP_IN + i];
    float r1[MLP_H], r2[MLP_H], r3[MLP_H];
    dense_tanh(sw + OFF_W1, sw + OFF_B1, x, MLP_IN, r1);
    dense_tanh(sw + OFF_W2, sw + OFF_B2, r1, MLP_H, r2);
    dense_tanh(sw + OFF_W3, sw + OFF_B3, r2, MLP_H, r3);
    float dy[MLP_OUT], l = 0.f;
    const float inv_m = 1.0f / (float)m;
    for (int o = 0; o < MLP_OUT; ++o) {
        float y = sw[OFF_B4 + o];
        for (int i = 0; i < MLP_H; ++i) y = __builtin_fmaf(sw[OFF_W4 + o * MLP_H + i], r3[i], y);
        const float e = y - T[s * MLP_OUT + o];
        l += 0.5f * e * e * inv_m;
        dy[o] = e * inv_m;
    }
    float d3[MLP_H], d2[MLP_H];
    for (int i = 0; i < MLP_H; ++i) {
        float acc = 0.f;
        for (int o = 0; o < MLP_OUT; ++o) acc = __builtin_fmaf(sw[OFF_W4 + o * MLP_H + i], dy[o], acc);
        d3[i] = acc * (1.f - r3[i] * r3[i]);
    }
    for (int i = 0; i < MLP_H; ++i) {
        float acc = 0.f;
        for (int o = 0; o < MLP_H; ++o) acc = __builtin_fmaf(sw[OFF_W3 + o * MLP_H + i], d3[o], acc);
        d2[i] = acc * (1.f - r2[i] * r2[i]);
    }
    for (int i = 0; i < MLP_H; ++i) {
        float acc = 0.f;
        for (int o = 0; o < MLP_H; ++o) acc = __builtin_fmaf(sw[OFF_W2 + o * MLP_H + i], d2[o], acc);
        d[i] = acc * (1.f - r1[i] * r1[i]);  // d1
    }
    for (int i = 0; i < MLP_H; ++i) {
        h1[i] = r1[i];
        h2[i] = r2[i];
        h3[i] = r3[i];
        d[MLP_H + i] = d2[i];
        d[2 * MLP_H + i] = d3[i];
    }
    for (int o = 0; o < MLP_OUT; ++o) d[3 * MLP_H + o] = dy[o];
    loss[s] = l;
}

// One lane per parameter: gradient as a fixed-order sum over the batch,
// then (optionally) an Adam step.
__global__ void mlp_grad_kernel(float *params, int64_t m, const float *acts, const float *dels,
                                float *grad, float *mom, float *var, int adam, float lr,
                                float b1, float b2, float bc1, float bc2, float epsa)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NP) return;
    // locate (layer, out, in)
    int dsel, asel, nin, isbias;
    int o, i;
    if (k < OFF_B1) {
        o = (k - OFF_W1) / MLP_IN; i = (k - OFF_W1) % MLP_IN; dsel = 0; asel = 0; nin = MLP_IN; isbias = 0;
    } else if (k < OFF_W2) {
        o = k - OFF_B1; i = 0; dsel = 0; asel = 0; nin = 0; isbias = 1;
    } else if (k < OFF_B2) {
        o = (k - OFF_W2) / MLP_H; i = (k - OFF_W2) % MLP_H; dsel = MLP_H; asel = MLP_IN; nin = MLP_H; isbias = 0;
    } else if (k < OFF_W3) {
        o = k - OFF_B2; i = 0; dsel = MLP_H; asel = 0; nin = 0; isbias = 1;
    } else if (k < OFF_B3) {
        o = (k - OFF_W3) / MLP_H; i = (k - OFF_W3) % MLP_H; dsel = 2 * MLP_H; asel = MLP_IN + MLP_H; nin = MLP_H; isbias = 0;
    } else if (k < OFF_W4) {
        o = k - OFF_B3; i = 0; dsel = 2 * MLP_H; asel = 0; nin = 0; isbias = 1;
    } else if (k < OFF_B4) {
        o = (k - OFF_W4) / MLP_H; i = (k - OFF_W4) % MLP_H; dsel = 3 * MLP_H; asel = MLP_IN + 2 * MLP_H; nin = MLP_H; isbias = 0;
    } else {
        o = k - OFF_B4; i = 0; dsel = 3 * MLP_H; asel = 0; nin = 0; isbias = 1;
    }
    (void)nin;
    float g = 0.f;
    for (int64_t s = 0; s < m; ++s) {
        const float dv = dels[s * DEL + dsel + o];
        g = isbias ? g + dv : __builtin_fmaf(dv, acts[s * ACT + asel + i], g);
    }
    grad[k] = g;
    if (adam) {
        const float mk = b1 * mom[k] + (1.f - b1) * g;
        const float vk = b2 * var[k] + (1.f - b2) * g * g;
        mom[k] = mk;
        var[k] = vk;
        params[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + epsa);
    }
}

__global__ void sum_kernel(const float *x, int64_t n, float *out)
{
    __shared__ float sh[256];
    float s = 0.f;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += x[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace

struct Mlp {
    DBuf<float> params, mom, var, grad;
    DBuf<float> acts, dels, loss, lsum;
    DBuf<float> xin, yout;
    int64_t steps = 0;
    bool ready = false;
};

static Mlp *get_mlp(vlm_ctx *ctx);

int mlp_forward_dev(vlm_ctx *ctx, int64_t m, const float *X, float *Y)
{
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    if (m <= 0) return VLM_OK;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256),
                       MLP_NP * sizeof(float), ctx->stream, M->params.p, m, X, Y);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

// Gradient (and optionally one Adam step) on a device batch.
int mlp_train_dev(vlm_ctx *ctx, int64_t m, const float *X, const float *T, int adam, float lr,
                  float *loss_host)
{
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    if (m <= 0) return VLM_OK;
    VLM_HIP(M->acts.reserve((size_t)m * ACT));
    VLM_HIP(M->dels.reserve((size_t)m * DEL));
    VLM_HIP(M->loss.reserve((size_t)m));
    hipLaunchKernelGGL(mlp_backprop_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256),
                       MLP_NP * sizeof(float), ctx->stream, M->params.p, m, X, T, M->acts.p,
                       M->dels.p, M->loss.p);
    VLM_HIP(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        M->steps += 1;
        bc1 = 1.f - std::pow(b1, (float)M->steps);
        bc2 = 1.f - std::pow(b2, (float)M->steps);
    }
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, ctx->stream,
                       M->params.p, m, M->acts.p, M->dels.p, M->grad.p, M->mom.p, M->var.p, adam,
                       lr, b1, b2, bc1, bc2, 1e-8f);
    VLM_HIP(hipGetLastError());
    if (loss_host) {
        hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, ctx->stream, M->loss.p, m,
                           M->lsum.p);
        VLM_HIP(hipMemcpyAsync(loss_host, M->lsum.p, sizeof(float), hipMemcpyDeviceToHost,
                               ctx->stream));
        VLM_HIP(hipStreamSynchronize(ctx->stream));
    }
    return VLM_OK;
}

// Deterministic small-scale initialisation (uniform +-sqrt(6/(in+out)),
// zero biases) from a fixed-seed linear congruential stream.
static void mlp_init_host(std::vector<float> &p, unsigned long long seed)
{
    p.assign(MLP_NP, 0.f);
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (float)((s >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
    };
    auto fill = [&](int off, int nout, int nin) {
        const float a = std::sqrt(6.f / (float)(nin + nout));
        for (int k = 0; k < nout * nin; ++k) p[off + k] = a * uni();
    };
    fill(OFF_W1, MLP_H, MLP_IN);
    fill(OFF_W2, MLP_H, MLP_H);
    fill(OFF_W3, MLP_H, MLP_H);
    fill(OFF_W4, MLP_OUT, MLP_H);
    for (int k = 0; k < MLP_OUT * MLP_H; ++k) p[OFF_W4 + k] *= 0.1f;
}

static Mlp *get_mlp(vlm_ctx *ctx)
{
    if (ctx->mlp) return ctx->mlp;
    Mlp *M = new Mlp();
    if (M->params.reserve(MLP_NP) != hipSuccess || M->mom.reserve(MLP_NP) != hipSuccess ||
        M->var.reserve(MLP_NP) != hipSuccess || M->grad.reserve(MLP_NP) != hipSuccess ||
        M->lsum.reserve(1) != hipSuccess) {
        delete M;
        return nullptr;
    }
    std::vector<float> p;
    mlp_init_host(p, 20261202ULL);
    std::vector<float> z(MLP_NP, 0.f);
    if (hipMemcpy(M->params.p, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->mom.p, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->var.p, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        delete M;
        return nullptr;
    }
    ctx->mlp = M;
    return M;
}

void mlp_release(vlm_ctx *ctx)
{
    delete ctx->mlp;
    ctx->mlp = nullptr;
}

}  // namespace vlm

using namespace vlm;

extern "C" {

int vlm_mlp_num_params(void) { return MLP_NP; }

int vlm_mlp_set_params(vlm_ctx *ctx, const float *params)
{
    if (!ctx || !params) return VLM_ERR_ARG;
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    VLM_HIP(hipMemcpy(M->params.p, params, MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> z(MLP_NP, 0.f);
    VLM_HIP(hipMemcpy(M->mom.p, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    VLM_HIP(hipMemcpy(M->var.p, z.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    M->steps = 0;
    return VLM_OK;
}

int vlm_mlp_get_params(vlm_ctx *ctx, float *params)
{
    if (!ctx || !params) return VLM_ERR_ARG;
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    VLM_HIP(hipMemcpy(params, M->params.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
    return VLM_OK;
}

int vlm_mlp_forward(vlm_ctx *ctx, int64_t m, const float *x, float *y)
{
    if (!ctx || m < 0 || (m > 0 && (!x || !y))) return VLM_ERR_ARG;
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    if (m == 0) return VLM_OK;
    VLM_HIP(M->xin.reserve((size_t)m * MLP_IN));
    VLM_HIP(M->yout.reserve((size_t)m * MLP_OUT));
    VLM_HIP(hipMemcpyAsync(M->xin.p, x, sizeof(float) * m * MLP_IN, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_TRY(mlp_forward_dev(ctx, m, M->xin.p, M->yout.p));
    VLM_HIP(hipMemcpyAsync(y, M->yout.p, sizeof(float) * m * MLP_OUT, hipMemcpyDeviceToHost,
                           ctx->stream));
    VLM_HIP(hipStreamSynchronize(ctx->stream));
    return VLM_OK;
}

int vlm_mlp_loss_grad(vlm_ctx *ctx, int64_t m, const float *x, const float *t, float *grad,
                      float *loss)
{
    if (!ctx || m <= 0 || !x || !t) return VLM_ERR_ARG;
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    VLM_HIP(M->xin.reserve((size_t)m * (MLP_IN + MLP_OUT)));
    VLM_HIP(hipMemcpyAsync(M->xin.p, x, sizeof(float) * m * MLP_IN, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_HIP(hipMemcpyAsync(M->xin.p + m * MLP_IN, t, sizeof(float) * m * MLP_OUT,
                           hipMemcpyHostToDevice, ctx->stream));
    float l = 0.f;
    VLM_TRY(mlp_train_dev(ctx, m, M->xin.p, M->xin.p + m * MLP_IN, 0, 0.f, &l));
    if (loss) *loss = l;
    if (grad) {
        VLM_HIP(hipMemcpy(grad, M->grad.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
    }
    return VLM_OK;
}

int vlm_mlp_train_step(vlm_ctx *ctx, int64_t m, const float *x, const float *t, float lr,
                       float *loss)
{
    if (!ctx || m <= 0 || !x || !t) return VLM_ERR_ARG;
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    VLM_HIP(M->xin.reserve((size_t)m * (MLP_IN + MLP_OUT)));
    VLM_HIP(hipMemcpyAsync(M->xin.p, x, sizeof(float) * m * MLP_IN, hipMemcpyHostToDevice,
                           ctx->stream));
    VLM_HIP(hipMemcpyAsync(M->xin.p + m * MLP_IN, t, sizeof(float) * m * MLP_OUT,
                           hipMemcpyHostToDevice, ctx->stream));
    float l = 0.f;
    VLM_TRY(mlp_train_dev(ctx, m, M->xin.p, M->xin.p + m * MLP_IN, 1, lr, &l));
    if (loss) *loss = l;
    return VLM_OK;
}

}  // extern "C"
