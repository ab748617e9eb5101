// Kernels of the learned schedule: a 32 -> 128 -> 128 -> 10 multilayer
// perceptron with GELU activations (forward, backward, Adam), and the
// diagonal-energy histogram that feeds it.  FP32, batch-major row layouts,
// every reduction in a fixed order (no floating-point atomics), so training
// is bitwise repeatable.  The matrices are tiny (batch x 128 x 128); each
// output element is one thread.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "mqw_internal.h"

namespace mqw {

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// Z = X W + b, A = act ? gelu(Z) : Z.  X: B x I, W: I x O (row-major), b: O.
__global__ void __launch_bounds__(256)
linear_fwd_kernel(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ bias,
                  float* __restrict__ Z, float* __restrict__ A, int B, int I, int O, int act)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * O) return;
    const int b = idx / O, o = idx % O;
    float z = bias[o];
    for (int i = 0; i < I; i++) z += X[b * I + i] * W[i * O + o];
    Z[idx] = z;
    A[idx] = act ? gelu(z) : z;
}

// Masked mean-squared error: dY = 2 M (Y - T) / sum(M); per-row loss partials.
__global__ void __launch_bounds__(256)
mse_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ M,
                float* __restrict__ dY, float* __restrict__ row_loss, int B, int O, float inv_count)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    float l = 0.0f;
    for (int o = 0; o < O; o++) {
        const int k = b * O + o;
        const float d = M[k] * (Y[k] - T[k]);
        dY[k] = 2.0f * d * inv_count;
        l += d * (Y[k] - T[k]);
    }
    row_loss[b] = l * inv_count;
}

// dX = (dZ W^T) .* (act ? gelu'(Zprev) : 1).  dZ: B x O, W: I x O, Zprev: B x I.
__global__ void __launch_bounds__(256)
linear_bwd_data_kernel(const float* __restrict__ dZ, const float* __restrict__ W, const float* __restrict__ Zprev,
                       float* __restrict__ dX, int B, int I, int O, int act)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= B * I) return;
    const int b = idx / I, i = idx % I;
    float s = 0.0f;
    for (int o = 0; o < O; o++) s += dZ[b * O + o] * W[i * O + o];
    dX[idx] = act ? s * gelu_grad(Zprev[idx]) : s;
}

// dW = X^T dZ, db = sum_b dZ.  One thread per weight, batch summed in order.
__global__ void __launch_bounds__(256)
linear_bwd_weight_kernel(const float* __restrict__ X, const float* __restrict__ dZ, float* __restrict__ dW,
                         float* __restrict__ db, int B, int I, int O)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < I * O) {
        const int i = idx / O, o = idx % O;
        float s = 0.0f;
        for (int b = 0; b < B; b++) s += X[b * I + i] * dZ[b * O + o];
        dW[idx] = s;
    }
    if (idx < O) {
        float s = 0.0f;
        for (int b = 0; b < B; b++) s += dZ[b * O + idx];
        db[idx] = s;
    }
}

__global__ void __launch_bounds__(256)
sum_kernel(const float* __restrict__ x, int n, float* __restrict__ out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int k = threadIdx.x; k < n; k += 256) s += x[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void __launch_bounds__(256)
adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2, int n,
            float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.0f - b1) * gk;
    const float v = b2 * m2[k] + (1.0f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = v;
    p[k] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// Counts of the shifted diagonal in `bins` equal bins over [-r, r].
__global__ void __launch_bounds__(256)
hist_kernel(const float* __restrict__ hp, uint64_t N, float r, int bins, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t loc[64];
    for (int k = threadIdx.x; k < bins; k += 256) loc[k] = 0u;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (uint64_t)gridDim.x * 256) {
        const float x = hp[i];
        int b = (int)floorf((x + r) / (2.0f * r) * (float)bins);
        b = b < 0 ? 0 : (b >= bins ? bins - 1 : b);
        atomicAdd(&loc[b], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < bins; k += 256)
        if (loc[k]) atomicAdd(&counts[k], loc[k]);
}

// Power sums of (x + shift - mu) for k = 1..4 over the diagonal, one
// deterministic partial per workgroup (grid-stride, fixed tree reduction).
__global__ void __launch_bounds__(256)
moments_kernel(const float* __restrict__ hp, uint64_t N, double shift, double mu, double* __restrict__ part)
{
    __shared__ double red[4][256];
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (uint64_t)gridDim.x * 256) {
        const double d = (double)hp[i] + shift - mu;
        const double d2 = d * d;
        s1 += d;
        s2 += d2;
        s3 += d2 * d;
        s4 += d2 * d2;
    }
    const int t = threadIdx.x;
    red[0][t] = s1;
    red[1][t] = s2;
    red[2][t] = s3;
    red[3][t] = s4;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 4; k++) red[k][t] += red[k][t + w];
        __syncthreads();
    }
    if (t < 4) part[(uint64_t)blockIdx.x * 4 + t] = red[t][0];
}

static int ok() { return hipGetLastError() == hipSuccess ? 0 : -1; }
static unsigned blocks(int n) { return (unsigned)((n + 255) / 256); }

int mlp_linear_fwd(const float* X, const float* W, const float* b, float* Z, float* A, int B, int I, int O, int act,
                   void* stream)
{
    hipLaunchKernelGGL(linear_fwd_kernel, dim3(blocks(B * O)), dim3(256), 0, (hipStream_t)stream, X, W, b, Z, A, B,
                       I, O, act);
    return ok();
}

int mlp_mse_grad(const float* Y, const float* T, const float* M, float* dY, float* row_loss, float* loss, int B,
                 int O, float inv_count, void* stream)
{
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks(B)), dim3(256), 0, (hipStream_t)stream, Y, T, M, dY, row_loss,
                       B, O, inv_count);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, row_loss, B, loss);
    return ok();
}

int mlp_linear_bwd(const float* X, const float* dZ, const float* W, const float* Zprev, float* dX, float* dW,
                   float* db, int B, int I, int O, int act_prev, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(linear_bwd_weight_kernel, dim3(blocks(I * O > O ? I * O : O)), dim3(256), 0, st, X, dZ, dW, db,
                       B, I, O);
    if (dX)
        hipLaunchKernelGGL(linear_bwd_data_kernel, dim3(blocks(B * I)), dim3(256), 0, st, dZ, W, Zprev, dX, B, I, O,
                           act_prev);
    return ok();
}

int mlp_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps, int step,
             void* stream)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, p, g, m1, m2, n, lr, b1, b2,
                       eps, c1, c2);
    return ok();
}

int diag_histogram(const float* hp, uint64_t N, float r, int bins, uint32_t* counts, void* stream)
{
    uint64_t g = (N + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(hist_kernel, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, hp, N, r, bins, counts);
    return ok();
}

} // namespace mqw

namespace mqw {

int diag_moment_blocks(uint64_t N)
{
    uint64_t g = (N + 255) / 256;
    return (int)(g > 1024 ? 1024 : g);
}

int diag_power_sums(const float* hp, uint64_t N, double shift, double mu, double* part, void* stream)
{
    hipLaunchKernelGGL(moments_kernel, dim3((unsigned)diag_moment_blocks(N)), dim3(256), 0, (hipStream_t)stream, hp,
                       N, shift, mu, part);
    return ok();
}

} // namespace mqw
