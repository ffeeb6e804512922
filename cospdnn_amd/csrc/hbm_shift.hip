// HBM store-alignment probe: the cost of writing 32-row x 64-sample FP16 tiles (the SpMM's
// steady-layer shape: 16384 rows, 9984 samples) to output positions shifted by s samples, so
// that every 128-byte line of an output row is written partly by two waves, against the
// aligned copy.  Shifts by whole 16-byte chunks (s = 8, 16, 24) and by 64 bytes (s = 32),
// nontemporal and plain stores, and per-sample 2-byte stores.  Checks the chunked shifted
// stores against the per-sample ones, then prints microseconds and GB/s per variant.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

constexpr int kLd = 80;  // LDS row stride (halves)

template <int H>  // halves H..H+7 of the 16 halves A|B
__device__ __forceinline__ uint4 funnel(uint4 A, uint4 B) {
  const uint32_t w[8] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w};
  uint32_t o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (H % 2 == 0) o[k] = w[H / 2 + k];
    else o[k] = __builtin_amdgcn_alignbyte(w[(H + 1) / 2 + k], w[(H - 1) / 2 + k], 2);
  }
  return make_uint4(o[0], o[1], o[2], o[3]);
}
__device__ __forceinline__ uint4 funnel_dyn(int h, uint4 A, uint4 B) {
  switch (h) {
    case 1: return funnel<1>(A, B);
    case 2: return funnel<2>(A, B);
    case 3: return funnel<3>(A, B);
    case 4: return funnel<4>(A, B);
    case 5: return funnel<5>(A, B);
    case 6: return funnel<6>(A, B);
    case 7: return funnel<7>(A, B);
    default: return B;
  }
}

// MODE 0 aligned nt copy; 1 per-sample plain 2-byte stores at q0 = 64t + s; 2 shifted
// 16-byte chunks (nt) + one 2-byte store per lane; 3 the same with plain chunk stores
template <int MODE>
__global__ __launch_bounds__(256) void shiftcopy(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ldx,
                                                 int64_t ldy, int ntiles, int nblk, int s) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[4][32 * kLd];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r8 = lane >> 3, ci = lane & 7, c8 = ci * 8;
  uint16_t* tl = lds[wave];
  const int nitems = ntiles * nblk;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += gridDim.x * 4) {
    const int c = item / ntiles, t = item - c * ntiles;
    u4v v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v[k] = __builtin_nontemporal_load(reinterpret_cast<const u4v*>(x + (int64_t)(32 * c + 8 * k + r8) * ldx + t * 64 + c8));
    if constexpr (MODE == 0 || MODE == 4) {
      // s: a multiple of 8 (whole 16-byte chunks); 0 = aligned
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        u4v* dst = reinterpret_cast<u4v*>(y + (int64_t)(32 * c + 8 * k + r8) * ldy + t * 64 + s + c8);
        if constexpr (MODE == 0) __builtin_nontemporal_store(v[k], dst); else *dst = v[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) *reinterpret_cast<u4v*>(tl + (8 * k + r8) * kLd + c8) = v[k];
      const int q0 = t * 64 + s;
      if constexpr (MODE == 1) {
#pragma unroll 8
        for (int r = 0; r < 32; ++r) y[(int64_t)(32 * c + r) * ldy + q0 + lane] = tl[r * kLd + lane];
      } else {
        const int h = 8 - (s & 7);
        const int sm = ci < h ? ci : 56 + ci;  // this lane's sample of the two partial chunks
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int row = 8 * k + r8;
          uint16_t* yr = y + (int64_t)(32 * c + row) * ldy;
          if (ci >= 1) {
            const uint4 A = *reinterpret_cast<const uint4*>(tl + row * kLd + c8 - 8);
            const uint4 B = *reinterpret_cast<const uint4*>(tl + row * kLd + c8);
            const uint4 o = funnel_dyn(h, A, B);
            u4v* dst = reinterpret_cast<u4v*>(yr + (q0 & ~7) + c8);
            if constexpr (MODE == 2) __builtin_nontemporal_store(u4v{o.x, o.y, o.z, o.w}, dst);
            else *dst = u4v{o.x, o.y, o.z, o.w};
          }
          yr[q0 + sm] = tl[row * kLd + sm];
        }
      }
    }
  }
}

int main() {
  const int N = 16384, S = 9984, ntiles = S / 64, nblk = N / 32;
  const int64_t ldx = 10048, ldy = 10048 + 64;
  uint16_t *a, *b, *ref;
  const size_t bx = (size_t)N * ldx * 2, by = (size_t)N * ldy * 2;
  CK(hipMalloc(&a, bx)); CK(hipMalloc(&b, by)); CK(hipMalloc(&ref, by));
  {
    uint16_t* h = (uint16_t*)malloc(bx);
    for (size_t i = 0; i < bx / 2; ++i) h[i] = (uint16_t)(i * 2654435761u >> 7);
    CK(hipMemcpy(a, h, bx, hipMemcpyHostToDevice));
    free(h);
  }
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  auto run = [&](const char* name, auto launch) -> int {
    float best = 1e30f;
    for (int it = 0; it < 6; ++it) {
      CK(hipEventRecord(e0, 0)); launch(); CK(hipGetLastError()); CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (it && ms < best) best = ms;
    }
    printf("%-34s %8.1f us  %7.0f GB/s\n", name, best * 1e3, 2.0 * N * S * 2 / (best * 1e-3) / 1e9);
    fflush(stdout);
    return 0;
  };
  // correctness of the shifted variants against the per-sample one (s = 3 and 4)
  uint16_t* h1 = (uint16_t*)malloc(by);
  uint16_t* h2 = (uint16_t*)malloc(by);
  for (int s : {3, 4, 7, 1}) {
    CK(hipMemset(ref, 0, by)); CK(hipMemset(b, 0, by));
    hipLaunchKernelGGL(shiftcopy<1>, dim3(1024), dim3(256), 0, 0, a, ref, ldx, ldy, ntiles, nblk, s);
    hipLaunchKernelGGL(shiftcopy<3>, dim3(1024), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s);
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(h1, ref, by, hipMemcpyDeviceToHost)); CK(hipMemcpy(h2, b, by, hipMemcpyDeviceToHost));
    size_t bad = 0;
    for (size_t i = 0; i < by / 2; ++i) bad += h1[i] != h2[i];
    printf("s=%d mismatches %zu\n", s, bad);
  }
  for (int grid : {1024}) {
    char nm[64];
    for (int s : {8, 16, 24, 32}) {
      snprintf(nm, 64, "16B-shift nt s=%d g=%d", s, grid);
      run(nm, [&] { hipLaunchKernelGGL(shiftcopy<0>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s); });
      snprintf(nm, 64, "16B-shift plain s=%d g=%d", s, grid);
      run(nm, [&] { hipLaunchKernelGGL(shiftcopy<4>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s); });
    }
    snprintf(nm, 64, "aligned plain g=%d", grid);
    run(nm, [&] { hipLaunchKernelGGL(shiftcopy<4>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, 0); });
  }
  for (int grid : {1024, 2048}) {
    char nm[64];
    snprintf(nm, 64, "aligned nt g=%d", grid);
    run(nm, [&] { hipLaunchKernelGGL(shiftcopy<0>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, 0); });
    for (int s : {3, 4}) {
      snprintf(nm, 64, "per-sample s=%d g=%d", s, grid);
      run(nm, [&] { hipLaunchKernelGGL(shiftcopy<1>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s); });
      snprintf(nm, 64, "chunks nt s=%d g=%d", s, grid);
      run(nm, [&] { hipLaunchKernelGGL(shiftcopy<2>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s); });
      snprintf(nm, 64, "chunks plain s=%d g=%d", s, grid);
      run(nm, [&] { hipLaunchKernelGGL(shiftcopy<3>, dim3(grid), dim3(256), 0, 0, a, b, ldx, ldy, ntiles, nblk, s); });
    }
  }
  return 0;
}
