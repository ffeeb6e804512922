// Calibration kernel for the HBM traffic counters: one plain copy of a known size.
// rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over this program gives the factor that turns
// each counter's per-dispatch reading into bytes (read = write = 512 MiB per dispatch).
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void hbm_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

int main() {
  const size_t bytes = size_t{512} << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  float best = 1e30f;
  for (int it = 0; it < 3; ++it) {
    if (hipEventRecord(e0, 0) != hipSuccess) return 1;
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipEventRecord(e1, 0) != hipSuccess) return 1;
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    if (ms < best) best = ms;
  }
  printf("hbm_copy: %zu KiB read + %zu KiB written per dispatch, best %.3f ms = %.0f GB/s\n", bytes >> 10,
         bytes >> 10, best, 2.0 * bytes / (best * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
