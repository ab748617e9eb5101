// Learned schedule: the MLP object of the C ABI (parameters, Adam state and
// activation buffers on the device) driving the kernels of mqw_mlp.hip.
//
// Network: 32 features -> 128 -> 128 -> 10, GELU after the two hidden layers.
// The 10 outputs are per-stage offsets (log gamma_i, log t_i), i < 5, added
// to the heuristic schedule of MultiQW.cpp:386-423; stages beyond m are
// masked out of the loss.  Parameters are one flat float array:
//   W1[32x128] b1[128] W2[128x128] b2[128] W3[128x10] b3[10]   (row-major W[in][out])
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "mqw_internal.h"

using namespace mqw;

namespace mqw {
int mlp_linear_fwd(const float* X, const float* W, const float* b, float* Z, float* A, int B, int I, int O, int act,
                   void* stream);
int mlp_mse_grad(const float* Y, const float* T, const float* M, float* dY, float* row_loss, float* loss, int B,
                 int O, float inv_count, void* stream);
int mlp_linear_bwd(const float* X, const float* dZ, const float* W, const float* Zprev, float* dX, float* dW,
                   float* db, int B, int I, int O, int act_prev, void* stream);
int mlp_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps, int step,
             void* stream);
} // namespace mqw

namespace {
constexpr int kIn = MQW_MLP_IN, kHid = MQW_MLP_HIDDEN, kOut = MQW_MLP_OUT;
constexpr size_t oW1 = 0, ob1 = oW1 + kIn * kHid, oW2 = ob1 + kHid, ob2 = oW2 + kHid * kHid, oW3 = ob2 + kHid,
                 ob3 = oW3 + kHid * kOut, kParams = ob3 + kOut;
static_assert(kParams == MQW_MLP_PARAMS, "parameter layout");
} // namespace

struct mqw_mlp {
    mqw_ctx* ctx = nullptr;
    float* p = nullptr;  // parameters
    float* g = nullptr;  // gradients
    float* m1 = nullptr; // Adam first moments
    float* m2 = nullptr; // Adam second moments
    int cap = 0;         // batch capacity of the activation buffers
    float* act = nullptr;
    float *X, *T, *M, *Z1, *A1, *Z2, *A2, *Y, *dY, *dZ2, *dZ1, *rows, *loss;
};

static mqw_status mlp_reserve(mqw_mlp* n, int B)
{
    if (B <= n->cap) return MQW_OK;
    mqw_ctx* c = n->ctx;
    if (n->act) hipFree(n->act);
    n->act = nullptr;
    const size_t per = (size_t)B * (3 * kIn + 0) + (size_t)B * (4 * kHid) + (size_t)B * (5 * kOut) + (size_t)B * 2 * kHid +
                       (size_t)B + 1;
    if (hipMalloc(&n->act, sizeof(float) * (per + 64)) != hipSuccess) {
        hipGetLastError();
        return fail(c, MQW_ENOMEM, "device allocation failed for an MLP batch of %d", B);
    }
    float* q = n->act;
    auto take = [&](size_t k) {
        float* r = q;
        q += k;
        return r;
    };
    n->X = take((size_t)B * kIn);
    n->T = take((size_t)B * kOut);
    n->M = take((size_t)B * kOut);
    n->Z1 = take((size_t)B * kHid);
    n->A1 = take((size_t)B * kHid);
    n->Z2 = take((size_t)B * kHid);
    n->A2 = take((size_t)B * kHid);
    n->Y = take((size_t)B * kOut);
    n->dY = take((size_t)B * kOut);
    n->dZ2 = take((size_t)B * kHid);
    n->dZ1 = take((size_t)B * kHid);
    n->rows = take((size_t)B);
    n->loss = take(1);
    n->cap = B;
    return MQW_OK;
}

static mqw_status mlp_forward_dev(mqw_mlp* n, int B)
{
    void* s = n->ctx->stream;
    int rc = mlp_linear_fwd(n->X, n->p + oW1, n->p + ob1, n->Z1, n->A1, B, kIn, kHid, 1, s);
    rc |= mlp_linear_fwd(n->A1, n->p + oW2, n->p + ob2, n->Z2, n->A2, B, kHid, kHid, 1, s);
    rc |= mlp_linear_fwd(n->A2, n->p + oW3, n->p + ob3, n->Y, n->Y, B, kHid, kOut, 0, s);
    return rc ? fail(n->ctx, MQW_EDEVICE, "MLP forward launch failed") : MQW_OK;
}

extern "C" {

mqw_status mqw_mlp_create(mqw_ctx* c, uint64_t seed, mqw_mlp** out)
{
    if (!c || !out) return MQW_EINVAL;
    hipSetDevice(c->device);
    mqw_mlp* n = new mqw_mlp();
    n->ctx = c;
    float* all = nullptr;
    if (hipMalloc(&all, sizeof(float) * 4 * kParams) != hipSuccess) {
        hipGetLastError();
        delete n;
        return fail(c, MQW_ENOMEM, "device allocation failed for MLP parameters");
    }
    n->p = all;
    n->g = all + kParams;
    n->m1 = all + 2 * kParams;
    n->m2 = all + 3 * kParams;
    // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, from the SK generator's stream.
    std::vector<float> h(kParams, 0.0f);
    std::vector<double> z((size_t)((kParams + 2) / 3) * 3);
    generate_sk(2, (int)(z.size() / 3), seed, z.data());
    auto fill = [&](size_t off, int fan_in, int count) {
        const double a = 1.0 / sqrt((double)fan_in);
        for (int k = 0; k < count; k++) h[off + k] = (float)(a * tanh(z[off + k]));
    };
    fill(oW1, kIn, kIn * kHid);
    fill(oW2, kHid, kHid * kHid);
    fill(oW3, kHid, kHid * kOut);
    HIPCHK(c, hipMemcpy(n->p, h.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(n->g, 0, sizeof(float) * 3 * kParams));
    *out = n;
    return MQW_OK;
}

void mqw_mlp_destroy(mqw_mlp* n)
{
    if (!n) return;
    hipSetDevice(n->ctx->device);
    hipStreamSynchronize(n->ctx->stream);
    if (n->p) hipFree(n->p);
    if (n->act) hipFree(n->act);
    delete n;
}

mqw_status mqw_mlp_get_params(mqw_mlp* n, float* out, size_t count)
{
    if (!n || !out || count != kParams) return MQW_EINVAL;
    HIPCHK(n->ctx, hipMemcpyAsync(out, n->p, sizeof(float) * kParams, hipMemcpyDeviceToHost, n->ctx->stream));
    HIPCHK(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return MQW_OK;
}

mqw_status mqw_mlp_set_params(mqw_mlp* n, const float* in, size_t count)
{
    if (!n || !in || count != kParams) return MQW_EINVAL;
    for (size_t k = 0; k < count; k++)
        if (!isfinite(in[k])) return fail(n->ctx, MQW_EINVAL, "non-finite MLP parameter at %zu", k);
    HIPCHK(n->ctx, hipMemcpyAsync(n->p, in, sizeof(float) * kParams, hipMemcpyHostToDevice, n->ctx->stream));
    HIPCHK(n->ctx, hipMemsetAsync(n->m1, 0, sizeof(float) * 2 * kParams, n->ctx->stream));
    HIPCHK(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return MQW_OK;
}

mqw_status mqw_mlp_forward(mqw_mlp* n, const float* X, int B, float* Y)
{
    if (!n || !X || !Y || B < 1) return MQW_EINVAL;
    hipSetDevice(n->ctx->device);
    mqw_status st = mlp_reserve(n, B);
    if (st) return st;
    HIPCHK(n->ctx, hipMemcpyAsync(n->X, X, sizeof(float) * B * kIn, hipMemcpyHostToDevice, n->ctx->stream));
    st = mlp_forward_dev(n, B);
    if (st) return st;
    HIPCHK(n->ctx, hipMemcpyAsync(Y, n->Y, sizeof(float) * B * kOut, hipMemcpyDeviceToHost, n->ctx->stream));
    HIPCHK(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return MQW_OK;
}

mqw_status mqw_mlp_loss_grad(mqw_mlp* n, const float* X, const float* T, const float* M, int B, float* loss,
                             float* grads)
{
    if (!n || !X || !T || !M || B < 1) return MQW_EINVAL;
    hipSetDevice(n->ctx->device);
    mqw_ctx* c = n->ctx;
    mqw_status st = mlp_reserve(n, B);
    if (st) return st;
    double msum = 0.0;
    for (int k = 0; k < B * kOut; k++) msum += M[k];
    if (!(msum > 0.0)) return fail(c, MQW_EINVAL, "empty loss mask");
    HIPCHK(c, hipMemcpyAsync(n->X, X, sizeof(float) * B * kIn, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(n->T, T, sizeof(float) * B * kOut, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(n->M, M, sizeof(float) * B * kOut, hipMemcpyHostToDevice, c->stream));
    st = mlp_forward_dev(n, B);
    if (st) return st;
    void* s = c->stream;
    int rc = mlp_mse_grad(n->Y, n->T, n->M, n->dY, n->rows, n->loss, B, kOut, (float)(1.0 / msum), s);
    rc |= mlp_linear_bwd(n->A2, n->dY, n->p + oW3, n->Z2, n->dZ2, n->g + oW3, n->g + ob3, B, kHid, kOut, 1, s);
    rc |= mlp_linear_bwd(n->A1, n->dZ2, n->p + oW2, n->Z1, n->dZ1, n->g + oW2, n->g + ob2, B, kHid, kHid, 1, s);
    rc |= mlp_linear_bwd(n->X, n->dZ1, n->p + oW1, nullptr, nullptr, n->g + oW1, n->g + ob1, B, kIn, kHid, 0, s);
    if (rc) return fail(c, MQW_EDEVICE, "MLP backward launch failed");
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (grads) HIPCHK(c, hipMemcpyAsync(grads, n->g, sizeof(float) * kParams, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MQW_OK;
}

mqw_status mqw_mlp_adam_step(mqw_mlp* n, float lr, float beta1, float beta2, float eps, int step)
{
    if (!n || step < 1 || !(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return MQW_EINVAL;
    hipSetDevice(n->ctx->device);
    if (mlp_adam(n->p, n->g, n->m1, n->m2, (int)kParams, lr, beta1, beta2, eps, step, n->ctx->stream))
        return fail(n->ctx, MQW_EDEVICE, "Adam launch failed");
    HIPCHK(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return MQW_OK;
}

} // extern "C"
