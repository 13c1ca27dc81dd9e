// mlp.hip -- the autotuner's FP32 cost model: a 3-layer MLP (F -> H -> H -> 1,
// ReLU) trained with MSE on log-latency and Adam, forward and backward as
// plain HIP kernels.  It only ever sees configuration features and measured
// times; it never touches field values.
//
// Parameter buffer (floats): W1[H][F] b1[H] W2[H][H] b2[H] W3[H] b3[1].
// Activations are kept for the backward pass: A1 = relu(X W1^T + b1) [B][H],
// A2 = relu(A1 W2^T + b2) [B][H], Y = A2 W3^T + b3 [B].
#include "common.h"

namespace imp {

struct MlpDims {
  u32 F, H;
  u32 oW1, ob1, oW2, ob2, oW3, ob3, n;
};

__host__ __device__ inline MlpDims mlp_dims(u32 F, u32 H) {
  MlpDims d;
  d.F = F;
  d.H = H;
  d.oW1 = 0;
  d.ob1 = d.oW1 + H * F;
  d.oW2 = d.ob1 + H;
  d.ob2 = d.oW2 + H * H;
  d.oW3 = d.ob2 + H;
  d.ob3 = d.oW3 + H;
  d.n = d.ob3 + 1;
  return d;
}

// Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o]), one thread per (b, o)
__global__ void k_linear_fwd(const float* X, const float* W, const float* bias, float* Y, u32 B,
                             u32 in, u32 out, int relu) {
  u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * out) return;
  u32 b = t / out, o = t % out;
  float acc = bias[o];
  for (u32 i = 0; i < in; i++) acc = fmaf(X[b * in + i], W[o * in + i], acc);
  Y[t] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dX[b][i] = sum_o dY[b][o] W[o][i], then masked by the forward ReLU (A > 0)
__global__ void k_linear_bwd_x(const float* dY, const float* W, const float* A, float* dX, u32 B,
                               u32 in, u32 out) {
  u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * in) return;
  u32 b = t / in, i = t % in;
  float acc = 0.0f;
  for (u32 o = 0; o < out; o++) acc = fmaf(dY[b * out + o], W[o * in + i], acc);
  dX[t] = A[t] > 0.0f ? acc : 0.0f;
}

// dW[o][i] = sum_b dY[b][o] X[b][i] (i < in); db[o] = sum_b dY[b][o] (i == in).
// One work-group per (o, i): 256 strided partial sums over the batch, then a
// fixed-order tree, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_linear_bwd_w(const float* X, const float* dY, float* dW,
                                                      float* db, u32 B, u32 in, u32 out) {
  __shared__ float red[256];
  const u32 o = blockIdx.x / (in + 1), i = blockIdx.x % (in + 1), t = threadIdx.x;
  float acc = 0.0f;
  if (i < in) {
    for (u32 b = t; b < B; b += 256) acc = fmaf(dY[b * out + o], X[b * in + i], acc);
  } else {
    for (u32 b = t; b < B; b += 256) acc += dY[b * out + o];
  }
  red[t] = acc;
  __syncthreads();
  for (u32 d = 128; d > 0; d >>= 1) {
    if (t < d) red[t] += red[t + d];
    __syncthreads();
  }
  if (t == 0) {
    if (i < in) dW[o * in + i] = red[0];
    else db[o] = red[0];
  }
}

// dY = 2 (Y - T) / B and per-sample squared error
__global__ void k_mse_grad(const float* Y, const float* T, float* dY, float* sq, u32 B) {
  u32 b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float d = Y[b] - T[b];
  dY[b] = 2.0f * d / (float)B;
  sq[b] = d * d;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, u32 n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

static inline dim3 g1(u64 n) { return dim3((u32)((n + 255) / 256)); }

static hipError_t mlp_forward(const float* params, const float* X, u32 B, const MlpDims& d,
                              float* A1, float* A2, float* Y, hipStream_t s) {
  hipLaunchKernelGGL(k_linear_fwd, g1((u64)B * d.H), dim3(256), 0, s, X, params + d.oW1,
                     params + d.ob1, A1, B, d.F, d.H, 1);
  hipLaunchKernelGGL(k_linear_fwd, g1((u64)B * d.H), dim3(256), 0, s, A1, params + d.oW2,
                     params + d.ob2, A2, B, d.H, d.H, 1);
  hipLaunchKernelGGL(k_linear_fwd, g1(B), dim3(256), 0, s, A2, params + d.oW3, params + d.ob3, Y,
                     B, d.H, 1u, 0);
  return hipGetLastError();
}

}  // namespace imp

using namespace imp;

extern "C" unsigned int impeller_mlp_param_count(unsigned int F, unsigned int H) {
  return mlp_dims(F, H).n;
}

// workspace floats needed by impeller_mlp_train_step / forward for batch B
extern "C" size_t impeller_mlp_workspace_floats(unsigned int F, unsigned int H, unsigned int B) {
  MlpDims d = mlp_dims(F, H);
  return (size_t)B * H * 4 + (size_t)B * 3 + d.n;
}

extern "C" hipError_t impeller_mlp_forward(const float* params, const float* X, unsigned int B,
                                           unsigned int F, unsigned int H, float* Y,
                                           float* workspace, hipStream_t stream) {
  IMP_REQUIRE(params && X && Y && workspace && F > 0 && H > 0, "mlp args");
  MlpDims d = mlp_dims(F, H);
  float* A1 = workspace;
  float* A2 = A1 + (size_t)B * H;
  return mlp_forward(params, X, B, d, A1, A2, Y, stream);
}

// Gradient of the mean squared error w.r.t. all parameters (grads: n floats),
// and the loss (1 float, device).  Used by the train step and by the tests.
extern "C" hipError_t impeller_mlp_grad(const float* params, const float* X, const float* T,
                                        unsigned int B, unsigned int F, unsigned int H,
                                        float* grads, float* loss_sum, float* workspace,
                                        hipStream_t stream) {
  IMP_REQUIRE(params && X && T && grads && workspace && B > 0, "mlp args");
  MlpDims d = mlp_dims(F, H);
  float* A1 = workspace;
  float* A2 = A1 + (size_t)B * H;
  float* D1 = A2 + (size_t)B * H;
  float* D2 = D1 + (size_t)B * H;
  float* Y = D2 + (size_t)B * H;
  float* dY = Y + B;
  float* sq = dY + B;
  IMP_CHECK(mlp_forward(params, X, B, d, A1, A2, Y, stream));
  hipLaunchKernelGGL(k_mse_grad, g1(B), dim3(256), 0, stream, Y, T, dY, sq, B);
  // layer 3: dW3, db3; dA2 = dY W3 masked by A2
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(d.H + 1), dim3(256), 0, stream, A2, dY, grads + d.oW3,
                     grads + d.ob3, B, d.H, 1u);
  hipLaunchKernelGGL(k_linear_bwd_x, g1((u64)B * d.H), dim3(256), 0, stream, dY, params + d.oW3,
                     A2, D2, B, d.H, 1u);
  // layer 2
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(d.H * (d.H + 1)), dim3(256), 0, stream, A1, D2,
                     grads + d.oW2, grads + d.ob2, B, d.H, d.H);
  hipLaunchKernelGGL(k_linear_bwd_x, g1((u64)B * d.H), dim3(256), 0, stream, D2, params + d.oW2,
                     A1, D1, B, d.H, d.H);
  // layer 1
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(d.H * (d.F + 1)), dim3(256), 0, stream, X, D1,
                     grads + d.oW1, grads + d.ob1, B, d.F, d.H);
  if (loss_sum) {
    // sum of squared errors / B via the same column-sum kernel (1 output)
    hipLaunchKernelGGL(k_linear_bwd_w, dim3(1), dim3(256), 0, stream, sq, sq, loss_sum,
                       loss_sum, B, 0u, 1u);
  }
  return hipGetLastError();
}

// One Adam step on (X, T); m and v are n-float state buffers; step >= 1.
extern "C" hipError_t impeller_mlp_train_step(float* params, float* m, float* v,
                                              const float* X, const float* T, unsigned int B,
                                              unsigned int F, unsigned int H, float lr,
                                              unsigned int step, float* loss_sum,
                                              float* workspace, hipStream_t stream) {
  IMP_REQUIRE(params && m && v && step >= 1, "mlp args");
  MlpDims d = mlp_dims(F, H);
  float* grads = workspace + (size_t)B * H * 4 + (size_t)B * 3;
  IMP_CHECK(impeller_mlp_grad(params, X, T, B, F, H, grads, loss_sum, workspace, stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, g1(d.n), dim3(256), 0, stream, params, grads, m, v, d.n, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError();
}
