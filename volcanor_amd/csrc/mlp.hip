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
// Every layer is a small matrix product on the matrix cores (FP32 MFMA):
// a workgroup takes a tile of 64 samples, each of its four waves 16 of
// them, with the parameters and activations in LDS.  Training runs the
// forward pass, the loss, the backward deltas and the tile's weight
// gradients in one kernel per tile; a second kernel adds the tile
// gradients in a fixed order (deterministic, no atomics) and applies Adam.
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

namespace {

// Matrix-core layout of v_mfma_f32_16x16x4f32 (checked on gfx950): lane l
// supplies A[l%16][l/16] and B[l/16][l%16] of the 16x16x4 step; result
// register r of lane l is D[4*(l/16) + r][l%16].
typedef float fx4 __attribute__((ext_vector_type(4)));

constexpr int TS = 64;            // samples per tile (one workgroup, four waves)
constexpr int K1 = 48;            // MLP_IN padded to a multiple of 4
constexpr int LW1 = K1 + 1;       // LDS row strides
constexpr int LH = MLP_H + 1;
constexpr int LDY = 16 + 1;
// LDS image of the parameters: W1 64 x LW1 | W2 64 x LH | W3 64 x LH |
// W4 16 x LH (rows 3..15 zero) | b1 b2 b3 (64 each) | b4 (16, padded)
constexpr int S_W1 = 0;
constexpr int S_W2 = S_W1 + MLP_H * LW1;
constexpr int S_W3 = S_W2 + MLP_H * LH;
constexpr int S_W4 = S_W3 + MLP_H * LH;
constexpr int S_B1 = S_W4 + 16 * LH;
constexpr int S_B2 = S_B1 + MLP_H;
constexpr int S_B3 = S_B2 + MLP_H;
constexpr int S_B4 = S_B3 + MLP_H;
constexpr int S_NP = S_B4 + 16;

__device__ void stage_params(const float *params, float *sp)
{
    for (int k = threadIdx.x; k < S_NP; k += blockDim.x) {
        float v = 0.f;
        if (k < S_W2) {
            const int o = k / LW1, i = k % LW1;
            if (i < MLP_IN) v = params[OFF_W1 + o * MLP_IN + i];
        } else if (k < S_W3) {
            const int o = (k - S_W2) / LH, i = (k - S_W2) % LH;
            if (i < MLP_H) v = params[OFF_W2 + o * MLP_H + i];
        } else if (k < S_W4) {
            const int o = (k - S_W3) / LH, i = (k - S_W3) % LH;
            if (i < MLP_H) v = params[OFF_W3 + o * MLP_H + i];
        } else if (k < S_B1) {
            const int o = (k - S_W4) / LH, i = (k - S_W4) % LH;
            if (o < MLP_OUT && i < MLP_H) v = params[OFF_W4 + o * MLP_H + i];
        } else if (k < S_B2) {
            v = params[OFF_B1 + (k - S_B1)];
        } else if (k < S_B3) {
            v = params[OFF_B2 + (k - S_B2)];
        } else if (k < S_B4) {
            v = params[OFF_B3 + (k - S_B3)];
        } else if (k - S_B4 < MLP_OUT) {
            v = params[OFF_B4 + (k - S_B4)];
        }
        sp[k] = v;
    }
}

// X tile (TS x MLP_IN, row-major in global) into LDS, zero padded.
__device__ void stage_x(const float *X, int64_t s0, int64_t m, float *Xs)
{
    for (int e = threadIdx.x; e < TS * LW1; e += blockDim.x) {
        const int s = e / LW1, k = e % LW1;
        Xs[e] = (k < MLP_IN && s0 + s < m) ? X[(s0 + s) * MLP_IN + k] : 0.f;
    }
}

// Rows [16w, 16w+16) of O = act(A . W^T + b): A is TS x K (stride LA), W is
// NOB*16 x K (stride LWt), O stride LO.
template <int K, int LA, int LWt, int NOB, bool TANH>
__device__ __forceinline__ void dense_fwd(const float *A, const float *W, const float *b, float *O,
                                          int LO, int w, int l)
{
    const int li = l & 15, lk = l >> 4;
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) {
        fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 4)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(16 * w + li) * LA + k0 + lk],
                                                       W[(16 * ob + li) * LWt + k0 + lk], acc, 0,
                                                       0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int srow = 16 * w + 4 * lk + r, j = 16 * ob + li;
            const float v = acc[r] + b[j];
            O[srow * LO + j] = TANH ? tanhf(v) : v;
        }
    }
}

// Rows [16w, 16w+16) of the backward delta: H <- (D . W) * (1 - H^2), where
// D is TS x K (stride LD) and W is K x 64 (stride LWt); H holds tanh
// outputs on entry and the deltas on exit.
template <int K, int LD, int LWt>
__device__ __forceinline__ void dense_bwd(const float *D, const float *W, float *H, int w, int l)
{
    const int li = l & 15, lk = l >> 4;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
        fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 4)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(D[(16 * w + li) * LD + k0 + lk],
                                                       W[(k0 + lk) * LWt + 16 * ib + li], acc, 0,
                                                       0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int srow = 16 * w + 4 * lk + r, j = 16 * ib + li;
            const float h = H[srow * LH + j];
            H[srow * LH + j] = acc[r] * (1.f - h * h);
        }
    }
}

// Tile gradient of one layer: dW[o][i] = sum_s D[s][o] Act[s][i] for the
// 16 outputs of block ob and NIB input blocks, written to the tile's
// partial-gradient row; bias gradients by lanes of wave 0.
template <int NIB, int LD, int LA>
__device__ __forceinline__ void dense_grad(const float *D, const float *Act, int ob, int nout,
                                           int nin, float *gw, int l)
{
    const int li = l & 15, lk = l >> 4;
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib) {
        fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < TS; k0 += 4)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(D[(k0 + lk) * LD + 16 * ob + li],
                                                       Act[(k0 + lk) * LA + 16 * ib + li], acc, 0,
                                                       0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 16 * ob + 4 * lk + r, i = 16 * ib + li;
            if (o < nout && i < nin) gw[o * nin + i] = acc[r];
        }
    }
}

__device__ __forceinline__ void bias_grad(const float *D, int LD, int nout, float *gb)
{
    const int o = threadIdx.x;
    if (o < nout) {
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g += D[s * LD + o];
        gb[o] = g;
    }
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *params, int64_t m,
                                                          const float *X, float *Y)
{
    __shared__ float sp[S_NP];
    __shared__ float Xs[TS * LW1];
    __shared__ float Ha[TS * LH], Hb[TS * LH];
    __shared__ float Ys[TS * LDY];
    stage_params(params, sp);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t ntiles = (m + TS - 1) / TS;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t s0 = t * TS;
        __syncthreads();
        stage_x(X, s0, m, Xs);
        __syncthreads();
        dense_fwd<K1, LW1, LW1, 4, true>(Xs, sp + S_W1, sp + S_B1, Ha, LH, w, l);
        __syncthreads();
        dense_fwd<MLP_H, LH, LH, 4, true>(Ha, sp + S_W2, sp + S_B2, Hb, LH, w, l);
        __syncthreads();
        dense_fwd<MLP_H, LH, LH, 4, true>(Hb, sp + S_W3, sp + S_B3, Ha, LH, w, l);
        __syncthreads();
        dense_fwd<MLP_H, LH, LH, 1, false>(Ha, sp + S_W4, sp + S_B4, Ys, LDY, w, l);
        __syncthreads();
        for (int e = threadIdx.x; e < TS * MLP_OUT; e += blockDim.x) {
            const int sl = e / MLP_OUT, o = e % MLP_OUT;
            if (s0 + sl < m) Y[(s0 + sl) * MLP_OUT + o] = Ys[sl * LDY + o];
        }
    }
}

// One tile of a training batch: forward, loss 0.5 |y - t|^2 / m, backward
// deltas and the tile's partial gradient (row `tile` of part).
__global__ __launch_bounds__(256) void mlp_train_tile_kernel(const float *params, int64_t m,
                                                             const float *X, const float *T,
                                                             float *part, float *loss)
{
    __shared__ float sp[S_NP];
    __shared__ float Xs[TS * LW1];
    __shared__ float H1[TS * LH], H2[TS * LH], H3[TS * LH];
    __shared__ float DY[TS * LDY];
    __shared__ float red[256];
    stage_params(params, sp);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t tile = blockIdx.x, s0 = tile * TS;
    float *g = part + (size_t)tile * MLP_NP;
    stage_x(X, s0, m, Xs);
    __syncthreads();
    dense_fwd<K1, LW1, LW1, 4, true>(Xs, sp + S_W1, sp + S_B1, H1, LH, w, l);
    __syncthreads();
    dense_fwd<MLP_H, LH, LH, 4, true>(H1, sp + S_W2, sp + S_B2, H2, LH, w, l);
    __syncthreads();
    dense_fwd<MLP_H, LH, LH, 4, true>(H2, sp + S_W3, sp + S_B3, H3, LH, w, l);
    __syncthreads();
    dense_fwd<MLP_H, LH, LH, 1, false>(H3, sp + S_W4, sp + S_B4, DY, LDY, w, l);
    __syncthreads();
    // dy = (y - t) / m on real samples and outputs, zero elsewhere
    const float inv_m = 1.0f / (float)m;
    float lsum = 0.f;
    for (int e = threadIdx.x; e < TS * 16; e += blockDim.x) {
        const int sl = e / 16, o = e % 16;
        float d = 0.f;
        if (o < MLP_OUT && s0 + sl < m) {
            const float err = DY[sl * LDY + o] - T[(s0 + sl) * MLP_OUT + o];
            lsum += 0.5f * err * err * inv_m;
            d = err * inv_m;
        }
        DY[sl * LDY + o] = d;
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[tile] = red[0];
    // layer 4
    if (w == 0) dense_grad<4, LDY, LH>(DY, H3, 0, MLP_OUT, MLP_H, g + OFF_W4, l);
    bias_grad(DY, LDY, MLP_OUT, g + OFF_B4);
    __syncthreads();
    dense_bwd<16, LDY, LH>(DY, sp + S_W4, H3, w, l);       // H3 <- delta3
    __syncthreads();
    dense_grad<4, LH, LH>(H3, H2, w, MLP_H, MLP_H, g + OFF_W3, l);
    bias_grad(H3, LH, MLP_H, g + OFF_B3);
    __syncthreads();
    dense_bwd<MLP_H, LH, LH>(H3, sp + S_W3, H2, w, l);     // H2 <- delta2
    __syncthreads();
    dense_grad<4, LH, LH>(H2, H1, w, MLP_H, MLP_H, g + OFF_W2, l);
    bias_grad(H2, LH, MLP_H, g + OFF_B2);
    __syncthreads();
    dense_bwd<MLP_H, LH, LH>(H2, sp + S_W2, H1, w, l);     // H1 <- delta1
    __syncthreads();
    dense_grad<3, LH, LW1>(H1, Xs, w, MLP_H, MLP_IN, g + OFF_W1, l);
    bias_grad(H1, LH, MLP_H, g + OFF_B1);
}

// Sum of the tile gradients in tile order, then (optionally) Adam.
__global__ void mlp_grad_kernel(float *params, const float *part, int ntiles, float *grad,
                                float *mom, float *var, int adam, float lr, float b1, float b2,
                                float bc1, float bc2, float epsa)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NP) return;
    float g = 0.f;
    for (int c = 0; c < ntiles; ++c) g += part[(size_t)c * MLP_NP + k];
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
    DBuf<float> params, mom, var, grad, part;
    DBuf<float> loss, lsum;
    DBuf<float> xin, yout;
    int64_t steps = 0;
    bool ready = false;
    bool have_loss = false;
};

static Mlp *get_mlp(vlm_ctx *ctx);

int mlp_forward_dev(vlm_ctx *ctx, int64_t m, const float *X, float *Y)
{
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    if (m <= 0) return VLM_OK;
    const int64_t ntiles = (m + TS - 1) / TS;
    const unsigned grid = (unsigned)(ntiles < 1024 ? ntiles : 1024);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid), dim3(256), 0, ctx->stream, M->params.p,
                       m, X, Y);
    VLM_HIP(hipGetLastError());
    return VLM_OK;
}

// Gradient (and optionally one Adam step) on a device batch.
int mlp_train_dev(vlm_ctx *ctx, int64_t m, const float *X, const float *T, int adam, float lr,
                  float *loss_host, int keep_loss)
{
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    if (m <= 0) return VLM_OK;
    const int64_t ntiles = (m + TS - 1) / TS;
    VLM_HIP(M->part.reserve((size_t)ntiles * MLP_NP));
    VLM_HIP(M->loss.reserve((size_t)ntiles));
    hipLaunchKernelGGL(mlp_train_tile_kernel, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream,
                       M->params.p, m, X, T, M->part.p, M->loss.p);
    VLM_HIP(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        M->steps += 1;
        bc1 = 1.f - std::pow(b1, (float)M->steps);
        bc2 = 1.f - std::pow(b2, (float)M->steps);
    }
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, ctx->stream,
                       M->params.p, M->part.p, (int)ntiles, M->grad.p, M->mom.p, M->var.p, adam,
                       lr, b1, b2, bc1, bc2, 1e-8f);
    VLM_HIP(hipGetLastError());
    if (loss_host || keep_loss) {
        VLM_HIP(M->lsum.reserve(1));
        hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, ctx->stream, M->loss.p, ntiles,
                           M->lsum.p);
        VLM_HIP(hipGetLastError());
        M->have_loss = true;
    }
    if (loss_host) VLM_TRY(d2h(ctx, loss_host, M->lsum.p, sizeof(float)));
    return VLM_OK;
}

// Loss of the last training call that kept it on the device (0 if none).
int mlp_read_loss(vlm_ctx *ctx, float *loss)
{
    *loss = 0.f;
    Mlp *M = ctx->mlp;
    if (!M || !M->have_loss) return VLM_OK;
    return d2h(ctx, loss, M->lsum.p, sizeof(float));
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
    VLM_DEVICE(ctx);
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
    VLM_DEVICE(ctx);
    Mlp *M = get_mlp(ctx);
    if (!M) return fail(ctx, VLM_ERR_ALLOC, "mlp allocation failed");
    VLM_TRY(sync_stream(ctx));
    VLM_HIP(hipMemcpy(params, M->params.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
    return VLM_OK;
}

int vlm_mlp_forward(vlm_ctx *ctx, int64_t m, const float *x, float *y)
{
    if (!ctx || m < 0 || (m > 0 && (!x || !y))) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
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
    VLM_TRY(sync_stream(ctx));
    return VLM_OK;
}

int vlm_mlp_loss_grad(vlm_ctx *ctx, int64_t m, const float *x, const float *t, float *grad,
                      float *loss)
{
    if (!ctx || m <= 0 || !x || !t) return VLM_ERR_ARG;
    VLM_DEVICE(ctx);
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
    VLM_DEVICE(ctx);
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
