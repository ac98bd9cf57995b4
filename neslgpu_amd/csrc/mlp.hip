// mlp.hip -- the learned dispatcher's network: forward, backward and Adam
// as HIP kernels (fp32).  No counterpart in the reference.
//
// Network: x (F features) -> ReLU(W1 x + b1) (H) -> ReLU(W2 h1 + b2) (H)
//          -> W3 h2 + b3 (A outputs, one predicted log-time per schedule).
// Parameters live in one flat device array in the order W1 b1 W2 b2 W3 b3
// (row-major W[out][in]).  Training minimises the masked mean squared error
// between predicted and measured log-times (a mask of 0 skips a missing
// measurement).  Activations for the backward pass are kept in `acts`
// (B*H for h1, then B*H for h2).
#include "cvl_hip.hpp"
#include "mlp.hpp"
#include "../../include/cvl_hip_ext.h"

namespace cvlhip {

// Y[r][o] = act(b[o] + sum_i X[r][i] W[o][i])
__global__ void k_dense_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ b,
                            float *__restrict__ Y, int B, int in, int out, int relu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * out) return;
    const int r = t / out, o = t % out;
    float acc = b[o];
    for (int i = 0; i < in; i++) acc += X[r * in + i] * W[o * in + i];
    Y[t] = relu && acc < 0.f ? 0.f : acc;
}

// dP = 2 M (P - T) / max(1, sum M); loss += sum M (P - T)^2 / max(1, sum M)
__global__ void k_mse_grad(const float *__restrict__ P, const float *__restrict__ T, const float *__restrict__ M,
                           float *__restrict__ dP, float *__restrict__ loss, int count, float inv_weight) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float e = (P[t] - T[t]) * M[t];
    dP[t] = 2.f * e * inv_weight;
    atomicAdd(loss, e * e * inv_weight);
}

// dX[r][i] = (sum_o dY[r][o] W[o][i]) * (Xact[r][i] > 0)   (Xact: the ReLU output feeding W)
__global__ void k_dense_bwd_x(const float *__restrict__ dY, const float *__restrict__ W,
                              const float *__restrict__ Xact, float *__restrict__ dX, int B, int in, int out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * in) return;
    const int r = t / in, i = t % in;
    float acc = 0.f;
    for (int o = 0; o < out; o++) acc += dY[r * out + o] * W[o * in + i];
    dX[t] = Xact[t] > 0.f ? acc : 0.f;
}

// dW[o][i] = sum_r dY[r][o] X[r][i];  db[o] = sum_r dY[r][o]
__global__ void k_dense_bwd_w(const float *__restrict__ dY, const float *__restrict__ X, float *__restrict__ dW,
                              float *__restrict__ db, int B, int in, int out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < out * in) {
        const int o = t / in, i = t % in;
        float acc = 0.f;
        for (int r = 0; r < B; r++) acc += dY[r * out + o] * X[r * in + i];
        dW[t] = acc;
    } else if (t < out * in + out) {
        const int o = t - out * in;
        float acc = 0.f;
        for (int r = 0; r < B; r++) acc += dY[r * out + o];
        db[o] = acc;
    }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int count, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

static int g256(int n) { return (n + 255) / 256; }

static void forward(const float *P, const float *X, int B, float *out, float *acts) {
    float *h1 = acts, *h2 = acts + (size_t)B * kH;
    k_dense_fwd<<<g256(B * kH), 256, 0, g_stream>>>(X, P + kW1, P + kB1, h1, B, kF, kH, 1);
    k_dense_fwd<<<g256(B * kH), 256, 0, g_stream>>>(h1, P + kW2, P + kB2, h2, B, kH, kH, 1);
    k_dense_fwd<<<g256(B * kA), 256, 0, g_stream>>>(h2, P + kW3, P + kB3, out, B, kH, kA, 0);
    LAUNCHCHECK();
}

// grads: kParams floats; work: 2*B*kH floats
static void backward(const float *P, const float *X, int B, const float *acts, const float *dOut, float *grads,
                     float *work) {
    const float *h1 = acts, *h2 = acts + (size_t)B * kH;
    float *dh2 = work, *dh1 = work + (size_t)B * kH;
    k_dense_bwd_w<<<g256(kA * kH + kA), 256, 0, g_stream>>>(dOut, h2, grads + kW3, grads + kB3, B, kH, kA);
    k_dense_bwd_x<<<g256(B * kH), 256, 0, g_stream>>>(dOut, P + kW3, h2, dh2, B, kH, kA);
    k_dense_bwd_w<<<g256(kH * kH + kH), 256, 0, g_stream>>>(dh2, h1, grads + kW2, grads + kB2, B, kH, kH);
    k_dense_bwd_x<<<g256(B * kH), 256, 0, g_stream>>>(dh2, P + kW2, h1, dh1, B, kH, kH);
    k_dense_bwd_w<<<g256(kH * kF + kH), 256, 0, g_stream>>>(dh1, X, grads + kW1, grads + kB1, B, kF, kH);
    LAUNCHCHECK();
}

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

void cvlhip_mlp_dims(int *features, int *hidden, int *arms, int *params) {
    *features = kF;
    *hidden = kH;
    *arms = kA;
    *params = kParams;
}

// All pointers are device pointers.  acts: 2*B*H floats.
void cvlhip_mlp_forward(const float *params, const float *X, int B, float *out, float *acts) {
    if (B <= 0) return;
    forward(params, X, B, out, acts);
}

// Gradients of the masked MSE for one batch (no update): grads gets kParams
// floats; loss (one float, device) gets the loss; work: B*(2H + 2A) floats.
void cvlhip_mlp_gradients(const float *params, const float *X, const float *T, const float *M, int B,
                          float mask_sum, float *grads, float *loss, float *acts, float *work) {
    if (B <= 0) return;
    float *pred = work, *dpred = work + (size_t)B * kA, *bwork = work + (size_t)2 * B * kA;
    forward(params, X, B, pred, acts);
    fill_u32(loss, 0, 1);
    const float inv = 1.f / (mask_sum > 1.f ? mask_sum : 1.f);
    k_mse_grad<<<g256(B * kA), 256, 0, g_stream>>>(pred, T, M, dpred, loss, B * kA, inv);
    LAUNCHCHECK();
    backward(params, X, B, acts, dpred, grads, bwork);
}

// One Adam step (step number t >= 1) on params with gradients grads.
void cvlhip_mlp_adam(float *params, const float *grads, float *m, float *v, int t, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
    k_adam<<<g256(kParams), 256, 0, g_stream>>>(params, grads, m, v, kParams, lr, b1, b2, 1e-8f, bc1, bc2);
    LAUNCHCHECK();
}

}  // extern "C"
