// Cost-model MLP kernels (in -> H -> H -> 1, ReLU, FP32) and their device work buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cospdnn {

struct MlpDims {
  int in = 0, hidden = 0;
  int num_params() const { return in * hidden + hidden + hidden * hidden + hidden + hidden + 1; }
};

// Per-batch activations and gradients ([n][hidden] each, [n] for the output side).
struct MlpWork {
  float *h1 = nullptr, *h2 = nullptr, *pred = nullptr;
  float *dpred = nullptr, *sq = nullptr, *d1 = nullptr, *d2 = nullptr;
};

// pred = MLP(X); X is [n][in] on the device.
hipError_t mlp_forward(const MlpDims& d, const float* params, const float* X, int n, MlpWork& w, hipStream_t st);
// Gradient of mean((pred - y)^2) w.r.t. the parameters into grad (after mlp_forward on the
// same batch); per-sample squared errors in w.sq.
hipError_t mlp_backward(const MlpDims& d, const float* params, const float* X, const float* y, int n, MlpWork& w,
                        float* grad, hipStream_t st);
// One Adam step (t = 1-based step count, default betas 0.9 / 0.999).
hipError_t mlp_adam(float* params, const float* grad, float* m, float* v, int np, float lr, int t, hipStream_t st);

}  // namespace cospdnn
