// gfx950 kernels of the sparse-DNN inference path.
//
//   spmm_mfma_kernel   FP16 gathered-BCSC SpMM on v_mfma_f32_16x16x32_f16 with a
//                      fused bias + clamp(0,32) + liveness / guard flags + compacted
//                      store epilogue (replaces cospdnn_single, kernels/kernel.cu:240-379,
//                      and check_active_f_col / matrix_reshape, kernel.cu:129-172).
//   spmm_simt_kernel   plain SIMT per-output gather SpMM (FP32 baseline path and the
//                      FP16 path for layers without a compact tile image).
//   scan_kernel        one work-group: live flags -> prefix sum -> position map, id
//                      compaction, guard bookkeeping (replaces the host scan of
//                      cospdnn.cu:249-265).
#include <hip/hip_runtime.h>

#include "gen.hpp"
#include "kernels.hpp"

namespace cospdnn {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr_t;

constexpr int kChunk = 32;  // K per MFMA chunk = gathered input rows per LDS tile

// --------------------------------------------------------------------------
// input generation and small fills (fill kernels instead of hipMemsetAsync so
// that a captured graph holds only kernel nodes)
// --------------------------------------------------------------------------
constexpr int kGenNeuronsPerThread = 64;

__global__ __launch_bounds__(256) void gen_input_kernel(_Float16* __restrict__ y0, int64_t ld, int N, int64_t S,
                                                        uint64_t seed) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint64_t key = stream_key(seed, kTagInput, static_cast<uint64_t>(s), 0);
  const uint32_t thr = input_row_threshold(seed, static_cast<uint64_t>(s));
  const int n0 = blockIdx.y * kGenNeuronsPerThread;
  const int n1 = min(N, n0 + kGenNeuronsPerThread);
  for (int n = n0; n < n1; ++n) {
    const bool one = (stream_draw(key, static_cast<uint64_t>(n) + 1) >> 32) < thr;
    y0[static_cast<int64_t>(n) * ld + s] = one ? static_cast<_Float16>(1.0f) : static_cast<_Float16>(0.0f);
  }
}

__global__ void fill_u8_kernel(uint8_t* p, uint8_t v, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void iota_i32_kernel(int32_t* p, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<int32_t>(i);
}
__global__ void set_i32_kernel(int32_t* p, int32_t v) { *p = v; }

// --------------------------------------------------------------------------
// scan: one work-group of 1024 threads; thread t owns a contiguous run of
// positions, a wave64 shuffle scan plus a 16-entry LDS scan gives its base.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void scan_kernel(ScanArgs a) {
  __shared__ int wsum[16];
  __shared__ int total;
  const int n = a.counts[a.layer];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int per = (n + 1023) / 1024;
  const int p0 = min(n, tid * per), p1 = min(n, p0 + per);
  int cnt = 0;
  for (int p = p0; p < p1; ++p) cnt += a.alive[p] ? 1 : 0;
  int x = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int i = 0; i < 16; ++i) {
      const int t = wsum[i];
      wsum[i] = s;
      s += t;
    }
    total = s;
  }
  __syncthreads();
  int q = wsum[wv] + x - cnt;
  for (int p = p0; p < p1; ++p) {
    const bool al = a.alive[p] != 0;
    const int32_t id = a.ids_in[p];
    if (al) {
      a.qmap[p] = q;
      a.ids_out[q] = id;
      ++q;
    } else {
      a.qmap[p] = -1;
    }
    const bool fl = (a.band != nullptr && a.band[p] != 0) || (a.flag_undecided && al && a.unsat[p] != 0);
    if (fl && a.flag_mark[id] == 0) {
      a.flag_mark[id] = 1;
      const int k = atomicAdd(a.flag_count, 1);
      a.flag_ids[k] = id;
    }
    a.clear0[p] = 0;
    a.clear1[p] = 0;
    a.clear2[p] = 0;
  }
  if (tid == 0) a.counts[a.layer + 1] = total;
}

// --------------------------------------------------------------------------
// MFMA SpMM.  Work item = (sample tile t of 256, block column c of 32 outputs).
// Four waves; wave w owns samples [64w, 64w+64) of the tile as four 16-sample
// N-tiles and both 16-output M-tiles: D[out][sample] = W^T[out][k] * X[k][sample].
// Per K chunk the 32 gathered input rows x 256 samples (16 KB) are copied
// global -> LDS with global_load_lds_dwordx4 (lane i writes bytes 16i..16i+15
// of the wave's LDS window), then each lane gathers its B fragment
// X[8g..8g+7][sample] with eight 16-bit LDS reads.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void spmm_mfma_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) _Float16 xs[kChunk * kTileS];
  const int n = a.counts[a.layer];
  const int ntiles = (n + kTileS - 1) / kTileS;
  const int64_t nitems = static_cast<int64_t>(ntiles) * w.nblk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);

  for (int64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int t = static_cast<int>(item / w.nblk);
    const int c = static_cast<int>(item - static_cast<int64_t>(t) * w.nblk);
    const int s0 = t * kTileS;
    float4_t acc[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[m][nt] = float4_t{0.f, 0.f, 0.f, 0.f};

    const int ch_end = w.kptr[c + 1];
    for (int ch = w.kptr[c]; ch < ch_end; ++ch) {
      __syncthreads();  // previous readers of xs are done
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wave * 8 + i * 2;  // wave-uniform first row of the pair
        const int row = w.gidx[ch * kChunk + r + (lane >> 5)];
        const _Float16* src = X + static_cast<int64_t>(row) * a.ldx + s0 + (lane & 31) * 8;
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(xs + r * kTileS), 16, 0, 0);
      }
      const half8_t a0 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(ch) * 2 + 0) * 512 + lane * 8);
      const half8_t a1 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(ch) * 2 + 1) * 512 + lane * 8);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const int col = wave * 64 + nt * 16 + j;
        half8_t b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = xs[(g * 8 + e) * kTileS + col];
        acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b, acc[0][nt], 0, 0, 0);
        acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b, acc[1][nt], 0, 0, 0);
      }
    }

    // Epilogue. D[i][j] of a 16x16 tile sits in lane j + 16*(i/4), element i%4:
    // this lane holds outputs 16m + 4g + r of sample 16nt + j.
    int outs[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) outs[m][r] = w.out_idx[c * 32 + m * 16 + g * 4 + r];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int p = s0 + wave * 64 + nt * 16 + j;
      const int q = p < n ? a.qmap[p] : -1;
      int fl = 0;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc[m][nt][r] + a.bias;
          if (outs[m][r] >= 0) {
            fl |= (v > 0.0f) ? 1 : 0;
            fl |= (v < 32.0f) ? 2 : 0;
            fl |= (fabsf(v) < a.band || fabsf(v - 32.0f) < a.band) ? 4 : 0;
            const float y = fminf(fmaxf(v, 0.0f), 32.0f);
            if (q >= 0) Y[static_cast<int64_t>(outs[m][r]) * a.ldy + q] = static_cast<_Float16>(y);
          }
        }
      }
      fl |= __shfl_xor(fl, 16, 64);
      fl |= __shfl_xor(fl, 32, 64);
      if (g == 0 && q >= 0) {
        if (fl & 1) a.alive_next[q] = 1;
        if (fl & 2) a.unsat_next[q] = 1;
        if ((fl & 4) && a.band_next != nullptr) a.band_next[q] = 1;
      }
    }
  }
}

// --------------------------------------------------------------------------
// SIMT gather SpMM: thread = sample, work item = (tile of 256 samples, block of
// 32 outputs); each output sums its CSC inputs read straight from global memory.
// --------------------------------------------------------------------------
template <typename TIn, typename TOut>
__global__ __launch_bounds__(256) void spmm_simt_kernel(StepArgs a, DevLayer w) {
  const int n = a.counts[a.layer];
  const int ntiles = (n + kTileS - 1) / kTileS;
  const int64_t nitems = static_cast<int64_t>(ntiles) * w.nblk;
  const TIn* __restrict__ X = static_cast<const TIn*>(a.x);
  TOut* __restrict__ Y = static_cast<TOut*>(a.y);
  for (int64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int t = static_cast<int>(item / w.nblk);
    const int c = static_cast<int>(item - static_cast<int64_t>(t) * w.nblk);
    const int p = t * kTileS + static_cast<int>(threadIdx.x);
    const int q = p < n ? a.qmap[p] : -1;
    if (q < 0) continue;
    const TIn* __restrict__ xcol = X + p;
    int fl = 0;
    for (int o = 0; o < 32; ++o) {
      const int out = w.out_idx[c * 32 + o];
      if (out < 0) continue;
      const int k0 = w.optr[c * 32 + o], k1 = w.optr[c * 32 + o + 1];
      float acc = 0.0f;
      for (int k = k0; k < k1; ++k) acc += static_cast<float>(xcol[static_cast<int64_t>(w.orow[k]) * a.ldx]) * w.oval[k];
      const float v = acc + a.bias;
      fl |= (v > 0.0f) ? 1 : 0;
      fl |= (v < 32.0f) ? 2 : 0;
      fl |= (fabsf(v) < a.band || fabsf(v - 32.0f) < a.band) ? 4 : 0;
      const float y = fminf(fmaxf(v, 0.0f), 32.0f);
      Y[static_cast<int64_t>(out) * a.ldy + q] = static_cast<TOut>(y);
    }
    if (fl & 1) a.alive_next[q] = 1;
    if (fl & 2) a.unsat_next[q] = 1;
    if ((fl & 4) && a.band_next != nullptr) a.band_next[q] = 1;
  }
}

__global__ __launch_bounds__(256) void gather_rows_kernel(const _Float16* __restrict__ src, int64_t lds,
                                                          const int32_t* __restrict__ ids,
                                                          const int32_t* __restrict__ count,
                                                          _Float16* __restrict__ dst, int64_t ldd, int N) {
  const int f = *count;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f) return;
  const int64_t id = ids[i];
  for (int nn = blockIdx.y; nn < N; nn += gridDim.y) dst[static_cast<int64_t>(nn) * ldd + i] = src[static_cast<int64_t>(nn) * lds + id];
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
static int fill_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  return static_cast<int>(g < 1 ? 1 : g);
}

hipError_t launch_gen_input(_Float16* y0, int64_t ld, int N, int64_t S, uint64_t seed, hipStream_t st) {
  if (S <= 0) return hipSuccess;
  dim3 grid(static_cast<unsigned>((S + 255) / 256), static_cast<unsigned>((N + kGenNeuronsPerThread - 1) / kGenNeuronsPerThread));
  hipLaunchKernelGGL(gen_input_kernel, grid, dim3(256), 0, st, y0, ld, N, S, seed);
  return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u8_kernel, dim3(fill_grid(n)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(fill_grid(n)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_iota_i32(int32_t* p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(iota_i32_kernel, dim3(fill_grid(n)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}
hipError_t launch_set_i32(int32_t* p, int32_t v, hipStream_t st) {
  hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(1), 0, st, p, v);
  return hipGetLastError();
}
hipError_t launch_scan(const ScanArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_spmm_mfma(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  hipLaunchKernelGGL(spmm_mfma_kernel, dim3(grid), dim3(256), 0, st, a, w);
  return hipGetLastError();
}
hipError_t launch_spmm_simt(const StepArgs& a, const DevLayer& w, Dtype tin, Dtype tout, int grid, hipStream_t st) {
  if (tin == Dtype::F16 && tout == Dtype::F16)
    hipLaunchKernelGGL((spmm_simt_kernel<_Float16, _Float16>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F16 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_kernel<_Float16, float>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F32 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_kernel<float, float>), dim3(grid), dim3(256), 0, st, a, w);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_gather_rows(const _Float16* src, int64_t lds, const int32_t* ids, const int32_t* count,
                              int64_t max_rows, _Float16* dst, int64_t ldd, int N, hipStream_t st) {
  if (max_rows <= 0) return hipSuccess;
  dim3 grid(static_cast<unsigned>((max_rows + 255) / 256), static_cast<unsigned>(N < 1024 ? N : 1024));
  hipLaunchKernelGGL(gather_rows_kernel, grid, dim3(256), 0, st, src, lds, ids, count, dst, ldd, N);
  return hipGetLastError();
}

int spmm_mfma_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_mfma_kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
  return nb;
}
int spmm_simt_blocks_per_cu(Dtype tin, Dtype tout) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (tin == Dtype::F16 && tout == Dtype::F16)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<_Float16, _Float16>, 256, 0);
  else if (tin == Dtype::F16 && tout == Dtype::F32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<_Float16, float>, 256, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<float, float>, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

}  // namespace cospdnn
