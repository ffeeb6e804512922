// Cost-model MLP of the autotuner: in -> H -> H -> 1, ReLU, FP32, trained with Adam on
// mean squared error.  Every step of training and inference runs in the kernels below
// (one thread per output element; the reductions over the batch are plain loops, so a
// gradient is computed in a fixed order and training is reproducible).
//
// Parameter vector layout (row-major, W[i][o] = weight from input i to output o):
//   W1[in][H], b1[H], W2[H][H], b2[H], W3[H][1], b3[1]
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace cospdnn {

namespace {

// Y[n][out] = act(X[n][in] . W[in][out] + b)
__global__ void dense_fwd_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                 const float* __restrict__ b, float* __restrict__ Y, int n, int in, int out,
                                 int relu) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(n) * out) return;
  const int r = static_cast<int>(idx / out), o = static_cast<int>(idx % out);
  float acc = b[o];
  for (int i = 0; i < in; ++i) acc = fmaf(X[static_cast<int64_t>(r) * in + i], W[i * out + o], acc);
  Y[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

// d/dpred of mean((pred - y)^2) = 2 (pred - y) / n, and the squared errors
__global__ void mse_grad_kernel(const float* __restrict__ pred, const float* __restrict__ y, float* __restrict__ dpred,
                                float* __restrict__ sq, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float d = pred[r] - y[r];
  dpred[r] = 2.0f * d / static_cast<float>(n);
  sq[r] = d * d;
}

// dW[i][o] = sum_r X[r][i] dY[r][o];  db[o] = sum_r dY[r][o]  (thread per (i, o); i == in: bias)
__global__ void dense_wgrad_kernel(const float* __restrict__ X, const float* __restrict__ dY, float* __restrict__ dW,
                                   float* __restrict__ db, int n, int in, int out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (in + 1) * out) return;
  const int i = idx / out, o = idx % out;
  float acc = 0.0f;
  if (i < in) {
    for (int r = 0; r < n; ++r) acc = fmaf(X[static_cast<int64_t>(r) * in + i], dY[static_cast<int64_t>(r) * out + o], acc);
    dW[i * out + o] = acc;
  } else {
    for (int r = 0; r < n; ++r) acc += dY[static_cast<int64_t>(r) * out + o];
    db[o] = acc;
  }
}

// dX[r][i] = (sum_o dY[r][o] W[i][o]) * (Xact[r][i] > 0)  (ReLU of the layer that produced X)
__global__ void dense_xgrad_kernel(const float* __restrict__ dY, const float* __restrict__ W,
                                   const float* __restrict__ Xact, float* __restrict__ dX, int n, int in, int out) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(n) * in) return;
  const int r = static_cast<int>(idx / in), i = static_cast<int>(idx % in);
  float acc = 0.0f;
  for (int o = 0; o < out; ++o) acc = fmaf(dY[static_cast<int64_t>(r) * out + o], W[i * out + o], acc);
  dX[idx] = Xact[idx] > 0.0f ? acc : 0.0f;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int np, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

hipError_t mlp_forward(const MlpDims& d, const float* params, const float* X, int n, MlpWork& w, hipStream_t st) {
  const float* W1 = params;
  const float* b1 = W1 + d.in * d.hidden;
  const float* W2 = b1 + d.hidden;
  const float* b2 = W2 + d.hidden * d.hidden;
  const float* W3 = b2 + d.hidden;
  const float* b3 = W3 + d.hidden;
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(blocks(static_cast<int64_t>(n) * d.hidden)), dim3(256), 0, st, X, W1, b1,
                     w.h1, n, d.in, d.hidden, 1);
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(blocks(static_cast<int64_t>(n) * d.hidden)), dim3(256), 0, st, w.h1, W2,
                     b2, w.h2, n, d.hidden, d.hidden, 1);
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(blocks(n)), dim3(256), 0, st, w.h2, W3, b3, w.pred, n, d.hidden, 1, 0);
  return hipGetLastError();
}

hipError_t mlp_backward(const MlpDims& d, const float* params, const float* X, const float* y, int n, MlpWork& w,
                        float* grad, hipStream_t st) {
  const float* W2 = params + d.in * d.hidden + d.hidden;
  const float* W3 = W2 + d.hidden * d.hidden + d.hidden;
  float* gW1 = grad;
  float* gb1 = gW1 + d.in * d.hidden;
  float* gW2 = gb1 + d.hidden;
  float* gb2 = gW2 + d.hidden * d.hidden;
  float* gW3 = gb2 + d.hidden;
  float* gb3 = gW3 + d.hidden;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks(n)), dim3(256), 0, st, w.pred, y, w.dpred, w.sq, n);
  // layer 3 (linear)
  hipLaunchKernelGGL(dense_wgrad_kernel, dim3(blocks((d.hidden + 1) * 1)), dim3(256), 0, st, w.h2, w.dpred, gW3, gb3, n,
                     d.hidden, 1);
  hipLaunchKernelGGL(dense_xgrad_kernel, dim3(blocks(static_cast<int64_t>(n) * d.hidden)), dim3(256), 0, st, w.dpred,
                     W3, w.h2, w.d2, n, d.hidden, 1);
  // layer 2
  hipLaunchKernelGGL(dense_wgrad_kernel, dim3(blocks((d.hidden + 1) * d.hidden)), dim3(256), 0, st, w.h1, w.d2, gW2,
                     gb2, n, d.hidden, d.hidden);
  hipLaunchKernelGGL(dense_xgrad_kernel, dim3(blocks(static_cast<int64_t>(n) * d.hidden)), dim3(256), 0, st, w.d2, W2,
                     w.h1, w.d1, n, d.hidden, d.hidden);
  // layer 1
  hipLaunchKernelGGL(dense_wgrad_kernel, dim3(blocks((d.in + 1) * d.hidden)), dim3(256), 0, st, X, w.d1, gW1, gb1, n,
                     d.in, d.hidden);
  return hipGetLastError();
}

hipError_t mlp_adam(float* params, const float* grad, float* m, float* v, int np, float lr, int t, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f, c2 = 1.0f;
  for (int i = 0; i < t; ++i) {
    c1 *= b1;
    c2 *= b2;
  }
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(np)), dim3(256), 0, st, params, grad, m, v, np, lr, b1, b2, eps, 1.0f - c1,
                     1.0f - c2);
  return hipGetLastError();
}

}  // namespace cospdnn
