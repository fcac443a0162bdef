// Learned surrogate of the SLQ path: a 2 -> 128 -> 128 -> 1 MLP with SiLU
// activations, trained on the GPU by hand-written forward / backward / Adam
// kernels (FP32) and evaluated by an inference kernel whose 128 x 128 layer
// runs on the matrix cores (BF16 inputs, FP32 accumulation).
//
// Parameter layout (flat FP32, kMlpParams = 17025 values):
//   W1[128][2] | b1[128] | W2[128][128] (row = output unit) | b2[128] |
//   W3[128] | b3
// Loss: mean squared error over the batch.
//
// No reference counterpart: imate's own surrogates over t are the
// interpolators (imate/interpolator, restated in imate_amd/interpolator).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "slq_internal.hpp"

namespace slq {

namespace {

constexpr int H = 128;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + 2 * H;
constexpr int kW2 = kB1 + H;
constexpr int kB2 = kW2 + H * H;
constexpr int kW3 = kB2 + H;
constexpr int kB3 = kW3 + H;
constexpr int kMlpParams = kB3 + 1;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s + z * s * (1.0f - s);
}

// ---------------------------------------------------------------- training

// One workgroup of H threads per sample: z1, z2 (pre-activations) and y.
__global__ __launch_bounds__(H) void mlp_forward_kernel(const float* __restrict__ P,
                                                        const float* __restrict__ x, int B,
                                                        float* __restrict__ z1,
                                                        float* __restrict__ z2,
                                                        float* __restrict__ y)
{
    __shared__ float h1[H];
    __shared__ float red[H];
    const int b = blockIdx.x;
    const int j = threadIdx.x;
    if (b >= B) return;
    const float x0 = x[2 * b], x1 = x[2 * b + 1];
    const float a1 = P[kW1 + 2 * j] * x0 + P[kW1 + 2 * j + 1] * x1 + P[kB1 + j];
    z1[(size_t) b * H + j] = a1;
    h1[j] = silu(a1);
    __syncthreads();
    float a2 = P[kB2 + j];
    const float* w = P + kW2 + (size_t) j * H;
    for (int k = 0; k < H; ++k) a2 += w[k] * h1[k];
    z2[(size_t) b * H + j] = a2;
    red[j] = P[kW3 + j] * silu(a2);
    __syncthreads();
    for (int s = H / 2; s > 0; s >>= 1)
    {
        if (j < s) red[j] += red[j + s];
        __syncthreads();
    }
    if (j == 0) y[b] = red[0] + P[kB3];
}

// Per-sample back-propagation: d3 = dL/dy, d2 = dL/dz2, d1 = dL/dz1.
__global__ __launch_bounds__(H) void mlp_backward_kernel(
    const float* __restrict__ P, const float* __restrict__ target, int B,
    const float* __restrict__ z1, const float* __restrict__ z2,
    const float* __restrict__ y, float* __restrict__ d1, float* __restrict__ d2,
    float* __restrict__ d3)
{
    __shared__ float s2[H];
    const int b = blockIdx.x;
    const int j = threadIdx.x;
    if (b >= B) return;
    const float g3 = 2.0f * (y[b] - target[b]) / (float) B;
    if (j == 0) d3[b] = g3;
    const float g2 = g3 * P[kW3 + j] * dsilu(z2[(size_t) b * H + j]);
    d2[(size_t) b * H + j] = g2;
    s2[j] = g2;
    __syncthreads();
    // d1_k = dsilu(z1_k) sum_j W2[j][k] d2_j (thread k reads column k)
    float acc = 0.0f;
    for (int jj = 0; jj < H; ++jj) acc += P[kW2 + (size_t) jj * H + j] * s2[jj];
    d1[(size_t) b * H + j] = acc * dsilu(z1[(size_t) b * H + j]);
}

// Gradient of every parameter: one thread per (parameter, batch chunk),
// each summing its chunk of samples in order into partial[chunk][p]; the
// chunks are then added in order by mlp_reduce_kernel (deterministic, and
// the batch loop is split over blockIdx.y so that a small network still
// fills the GPU). The last parameter's thread of each chunk also writes the
// chunk's sum of squared errors.
__global__ __launch_bounds__(256) void mlp_grad_kernel(
    const float* __restrict__ x, const float* __restrict__ target, int B, int chunk,
    const float* __restrict__ z1, const float* __restrict__ z2,
    const float* __restrict__ y, const float* __restrict__ d1,
    const float* __restrict__ d2, const float* __restrict__ d3,
    float* __restrict__ partial, float* __restrict__ loss_partial)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const int c = blockIdx.y;
    const int b0 = c * chunk;
    const int b1 = b0 + chunk < B ? b0 + chunk : B;
    float g = 0.0f;
    if (p < kB1)
    {
        const int j = p / 2, i = p % 2;
        for (int b = b0; b < b1; ++b) g += d1[(size_t) b * H + j] * x[2 * b + i];
    }
    else if (p < kW2)
    {
        const int j = p - kB1;
        for (int b = b0; b < b1; ++b) g += d1[(size_t) b * H + j];
    }
    else if (p < kB2)
    {
        const int j = (p - kW2) / H, k = (p - kW2) % H;
        for (int b = b0; b < b1; ++b)
            g += d2[(size_t) b * H + j] * silu(z1[(size_t) b * H + k]);
    }
    else if (p < kW3)
    {
        const int j = p - kB2;
        for (int b = b0; b < b1; ++b) g += d2[(size_t) b * H + j];
    }
    else if (p < kB3)
    {
        const int j = p - kW3;
        for (int b = b0; b < b1; ++b) g += d3[b] * silu(z2[(size_t) b * H + j]);
    }
    else
    {
        for (int b = b0; b < b1; ++b) g += d3[b];
        float l = 0.0f;
        for (int b = b0; b < b1; ++b)
        {
            const float e = y[b] - target[b];
            l += e * e;
        }
        loss_partial[c] = l;
    }
    partial[(size_t) c * kMlpParams + p] = g;
}

__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ partial,
                                                         const float* __restrict__ loss_partial,
                                                         int chunks, int B,
                                                         float* __restrict__ grad,
                                                         float* __restrict__ loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = 0.0f;
    for (int c = 0; c < chunks; ++c) g += partial[(size_t) c * kMlpParams + p];
    grad[p] = g;
    if (p == kMlpParams - 1 && loss_out)
    {
        float l = 0.0f;
        for (int c = 0; c < chunks; ++c) l += loss_partial[c];
        *loss_out = l / (float) B;
    }
}

// Batch chunks of the gradient kernel: about 64 samples each, at most 32.
inline int grad_chunks(int B)
{
    const int c = (B + 63) / 64;
    return c < 32 ? c : 32;
}

// Adam (Kingma & Ba 2015) with bias correction; the step counter lives in
// device memory so that a captured step can be replayed.
__global__ __launch_bounds__(256) void mlp_adam_kernel(
    float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ m1,
    float* __restrict__ m2, const int* __restrict__ step, float lr, float beta1,
    float beta2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const int t = *step + 1;
    const float g = grad[p];
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = v;
    const float ah = a / (1.0f - powf(beta1, (float) t));
    const float vh = v / (1.0f - powf(beta2, (float) t));
    P[p] -= lr * ah / (sqrtf(vh) + eps);
}

// Advances the step counter and records the loss of the step just taken.
__global__ void mlp_step_kernel(int* step, const float* loss, float* losses, int max_steps)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        const int t = *step;
        if (t < max_steps) losses[t] = *loss;
        *step = t + 1;
    }
}

__global__ void zero_i32_kernel(int* p) { *p = 0; }

__global__ void zero_f32_kernel(float* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---------------------------------------------------------------- inference

// BF16 matrix-core inference. One wavefront per 16 queries: layer 1 in FP32
// (h1 = silu(W1 x + b1)), layer 2 as 8 x 8 products of 16x16x16 BF16 tiles
// on v_mfma_f32_16x16x16_bf16 with FP32 accumulation, then silu, the dot
// with W3 and b3. Operand layout of the 16x16x16 MFMA (wave64): lane l holds
// A[l % 16][4 (l / 16) + 0..3] and B[4 (l / 16) + 0..3][l % 16]; the result
// lane l holds D[4 (l / 16) + 0..3][l % 16].
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f)
{
    // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (short) (u >> 16);
}

__global__ __launch_bounds__(256) void mlp_infer_bf16_kernel(const float* __restrict__ P,
                                                             const short* __restrict__ W2b,
                                                             const float* __restrict__ x,
                                                             int Q, float* __restrict__ y)
{
    // per wavefront: h1 tile [16 queries][128] as BF16 in LDS
    __shared__ short h1s[4][16][H];
    __shared__ float part[4][16][16];
    const int lane = threadIdx.x % 64;
    const int w = threadIdx.x / 64;
    const int q0 = (blockIdx.x * 4 + w) * 16;
    // every wavefront reaches both barriers; inactive ones skip the work
    const bool active = q0 < Q;
    if (active)
    {
        // layer 1: lane handles (query r = lane / 4, units (lane % 4) * 32 .. +31)
        const int r = lane / 4;
        const int u0 = (lane % 4) * 32;
        const int q = q0 + r;
        const float x0 = q < Q ? x[2 * q] : 0.0f;
        const float x1 = q < Q ? x[2 * q + 1] : 0.0f;
        for (int u = u0; u < u0 + 32; ++u)
        {
            const float a = P[kW1 + 2 * u] * x0 + P[kW1 + 2 * u + 1] * x1 + P[kB1 + u];
            h1s[w][r][u] = to_bf16(silu(a));
        }
    }
    __syncthreads();
    const int row = lane % 16;        // A row (query) / B,D column
    const int kq = 4 * (lane / 16);   // k offset within the 16-wide slice
    float ysum[4] = {0.f, 0.f, 0.f, 0.f};
    if (active)
    {
        for (int nb = 0; nb < H / 16; ++nb)          // output column block
        {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < H / 16; ++kb)      // reduction block
            {
                bf16x4 a, bm;
                for (int t = 0; t < 4; ++t)
                {
                    a[t] = h1s[w][row][kb * 16 + kq + t];
                    // B[k][n] = W2[n][k]: output unit n = nb*16 + row, input k
                    bm[t] = W2b[(size_t) (nb * 16 + row) * H + kb * 16 + kq + t];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bm, acc, 0, 0, 0);
            }
            // lane holds z2[query 4*(lane/16)+i][unit nb*16 + row]
            const int unit = nb * 16 + row;
            const float b2 = P[kB2 + unit], w3 = P[kW3 + unit];
            for (int i = 0; i < 4; ++i) ysum[i] += w3 * silu(acc[i] + b2);
        }
        // sum over the 16 columns held by lanes with the same lane / 16
        for (int i = 0; i < 4; ++i) part[w][4 * (lane / 16) + i][row] = ysum[i];
    }
    __syncthreads();
    if (active && lane < 16)
    {
        float s = P[kB3];
        for (int c = 0; c < 16; ++c) s += part[w][lane][c];
        const int q = q0 + lane;
        if (q < Q) y[q] = s;
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ in, short* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = to_bf16(in[i]);
}

// FP32 inference: forward pass without storing pre-activations.
__global__ __launch_bounds__(H) void mlp_infer_f32_kernel(const float* __restrict__ P,
                                                          const float* __restrict__ x, int Q,
                                                          float* __restrict__ y)
{
    __shared__ float h1[H];
    __shared__ float red[H];
    const int b = blockIdx.x;
    const int j = threadIdx.x;
    if (b >= Q) return;
    const float x0 = x[2 * b], x1 = x[2 * b + 1];
    h1[j] = silu(P[kW1 + 2 * j] * x0 + P[kW1 + 2 * j + 1] * x1 + P[kB1 + j]);
    __syncthreads();
    float a2 = P[kB2 + j];
    const float* w = P + kW2 + (size_t) j * H;
    for (int k = 0; k < H; ++k) a2 += w[k] * h1[k];
    red[j] = P[kW3 + j] * silu(a2);
    __syncthreads();
    for (int s = H / 2; s > 0; s >>= 1)
    {
        if (j < s) red[j] += red[j + s];
        __syncthreads();
    }
    if (j == 0) y[b] = red[0] + P[kB3];
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void) hipFree(p);
    }
};

}  // namespace

}  // namespace slq

using namespace slq;

extern "C" {

int slq_mlp_num_params(void) { return kMlpParams; }

int slq_mlp_grad(const float* params, const float* x, const float* target, int32_t B,
                 float* loss, float* grad)
{
    if (!params || !x || !target || B < 1 || !grad)
    {
        set_error("slq_mlp_grad: invalid argument");
        return SLQ_E_ARG;
    }
    DevBuf dp, dx, dt, dz1, dz2, dy, dd1, dd2, dd3, dg, dl, dpart, dlpart;
    const int C = grad_chunks(B);
    const int chunk = (B + C - 1) / C;
    SLQ_CHECK(hipMalloc(&dpart.p, sizeof(float) * kMlpParams * C));
    SLQ_CHECK(hipMalloc(&dlpart.p, sizeof(float) * C));
    SLQ_CHECK(hipMalloc(&dp.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dx.p, sizeof(float) * 2 * B));
    SLQ_CHECK(hipMalloc(&dt.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dz1.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dz2.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dy.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dd1.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dd2.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dd3.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dg.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dl.p, sizeof(float)));
    SLQ_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(dx.p, x, sizeof(float) * 2 * B, hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(dt.p, target, sizeof(float) * B, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(B), dim3(H), 0, 0, (const float*) dp.p,
                       (const float*) dx.p, B, (float*) dz1.p, (float*) dz2.p, (float*) dy.p);
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(B), dim3(H), 0, 0, (const float*) dp.p,
                       (const float*) dt.p, B, (const float*) dz1.p, (const float*) dz2.p,
                       (const float*) dy.p, (float*) dd1.p, (float*) dd2.p, (float*) dd3.p);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((kMlpParams + 255) / 256, C), dim3(256), 0, 0,
                       (const float*) dx.p, (const float*) dt.p, B, chunk,
                       (const float*) dz1.p, (const float*) dz2.p, (const float*) dy.p,
                       (const float*) dd1.p, (const float*) dd2.p, (const float*) dd3.p,
                       (float*) dpart.p, (float*) dlpart.p);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, 0,
                       (const float*) dpart.p, (const float*) dlpart.p, C, B, (float*) dg.p,
                       (float*) dl.p);
    SLQ_CHECK(hipGetLastError());
    SLQ_CHECK(hipDeviceSynchronize());
    SLQ_CHECK(hipMemcpy(grad, dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (loss) SLQ_CHECK(hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    return SLQ_OK;
}

int slq_mlp_train(float* params, const float* x, const float* target, int32_t B,
                  int32_t steps, float lr, float* losses)
{
    if (!params || !x || !target || B < 1 || steps < 1 || !(lr > 0.f))
    {
        set_error("slq_mlp_train: invalid argument");
        return SLQ_E_ARG;
    }
    DevBuf dp, dx, dt, dz1, dz2, dy, dd1, dd2, dd3, dg, dl, dm1, dm2, dstep, dlosses, dpart,
        dlpart;
    const int C = grad_chunks(B);
    const int chunk = (B + C - 1) / C;
    SLQ_CHECK(hipMalloc(&dpart.p, sizeof(float) * kMlpParams * C));
    SLQ_CHECK(hipMalloc(&dlpart.p, sizeof(float) * C));
    SLQ_CHECK(hipMalloc(&dp.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dx.p, sizeof(float) * 2 * B));
    SLQ_CHECK(hipMalloc(&dt.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dz1.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dz2.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dy.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dd1.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dd2.p, sizeof(float) * H * B));
    SLQ_CHECK(hipMalloc(&dd3.p, sizeof(float) * B));
    SLQ_CHECK(hipMalloc(&dg.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dl.p, sizeof(float)));
    SLQ_CHECK(hipMalloc(&dm1.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dm2.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dstep.p, sizeof(int)));
    SLQ_CHECK(hipMalloc(&dlosses.p, sizeof(float) * steps));
    SLQ_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(dx.p, x, sizeof(float) * 2 * B, hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(dt.p, target, sizeof(float) * B, hipMemcpyHostToDevice));
    hipStream_t s = nullptr;
    SLQ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const int pg = (kMlpParams + 255) / 256;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(pg), dim3(256), 0, s, (float*) dm1.p, kMlpParams);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(pg), dim3(256), 0, s, (float*) dm2.p, kMlpParams);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, s, (int*) dstep.p);
    // one training step, captured once and replayed `steps` times
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int r = SLQ_OK;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(B), dim3(H), 0, s, (const float*) dp.p,
                           (const float*) dx.p, B, (float*) dz1.p, (float*) dz2.p,
                           (float*) dy.p);
        hipLaunchKernelGGL(mlp_backward_kernel, dim3(B), dim3(H), 0, s, (const float*) dp.p,
                           (const float*) dt.p, B, (const float*) dz1.p, (const float*) dz2.p,
                           (const float*) dy.p, (float*) dd1.p, (float*) dd2.p,
                           (float*) dd3.p);
        hipLaunchKernelGGL(mlp_grad_kernel, dim3(pg, C), dim3(256), 0, s, (const float*) dx.p,
                           (const float*) dt.p, B, chunk, (const float*) dz1.p,
                           (const float*) dz2.p, (const float*) dy.p, (const float*) dd1.p,
                           (const float*) dd2.p, (const float*) dd3.p, (float*) dpart.p,
                           (float*) dlpart.p);
        hipLaunchKernelGGL(mlp_reduce_kernel, dim3(pg), dim3(256), 0, s,
                           (const float*) dpart.p, (const float*) dlpart.p, C, B,
                           (float*) dg.p, (float*) dl.p);
        hipLaunchKernelGGL(mlp_adam_kernel, dim3(pg), dim3(256), 0, s, (float*) dp.p,
                           (const float*) dg.p, (float*) dm1.p, (float*) dm2.p,
                           (const int*) dstep.p, lr, 0.9f, 0.999f, 1e-8f);
        hipLaunchKernelGGL(mlp_step_kernel, dim3(1), dim3(1), 0, s, (int*) dstep.p,
                           (const float*) dl.p, (float*) dlosses.p, steps);
        e = hipStreamEndCapture(s, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int i = 0; e == hipSuccess && i < steps; ++i) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) r = hip_status(e, "slq_mlp_train", __FILE__, __LINE__);
    if (exec) (void) hipGraphExecDestroy(exec);
    if (graph) (void) hipGraphDestroy(graph);
    (void) hipStreamDestroy(s);
    if (r != SLQ_OK) return r;
    SLQ_CHECK(hipMemcpy(params, dp.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    if (losses)
        SLQ_CHECK(hipMemcpy(losses, dlosses.p, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return SLQ_OK;
}

int slq_mlp_predict(const float* params, const float* x, int32_t Q, int32_t bf16,
                    float* y, double* seconds)
{
    if (!params || !x || Q < 1 || !y)
    {
        set_error("slq_mlp_predict: invalid argument");
        return SLQ_E_ARG;
    }
    DevBuf dp, dx, dy, dw;
    SLQ_CHECK(hipMalloc(&dp.p, sizeof(float) * kMlpParams));
    SLQ_CHECK(hipMalloc(&dx.p, sizeof(float) * 2 * Q));
    SLQ_CHECK(hipMalloc(&dy.p, sizeof(float) * Q));
    SLQ_CHECK(hipMalloc(&dw.p, sizeof(short) * H * H));
    SLQ_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(dx.p, x, sizeof(float) * 2 * Q, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((H * H + 255) / 256), dim3(256), 0, 0,
                       (const float*) dp.p + kW2, (short*) dw.p, H * H);
    SLQ_CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    SLQ_CHECK(hipEventCreate(&e0));
    SLQ_CHECK(hipEventCreate(&e1));
    (void) hipEventRecord(e0, 0);
    if (bf16)
    {
        const int waves = (Q + 15) / 16;
        hipLaunchKernelGGL(mlp_infer_bf16_kernel, dim3((waves + 3) / 4), dim3(256), 0, 0,
                           (const float*) dp.p, (const short*) dw.p, (const float*) dx.p, Q,
                           (float*) dy.p);
    }
    else
        hipLaunchKernelGGL(mlp_infer_f32_kernel, dim3(Q), dim3(H), 0, 0, (const float*) dp.p,
                           (const float*) dx.p, Q, (float*) dy.p);
    (void) hipEventRecord(e1, 0);
    SLQ_CHECK(hipGetLastError());
    SLQ_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    (void) hipEventElapsedTime(&ms, e0, e1);
    if (seconds) *seconds = ms * 1e-3;
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    SLQ_CHECK(hipMemcpy(y, dy.p, sizeof(float) * Q, hipMemcpyDeviceToHost));
    return SLQ_OK;
}

}  // extern "C"
