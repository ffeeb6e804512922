// HBM access-pattern probe: read, write and copy bandwidth of plain streams, and of copies
// shaped like the SpMM's activation traffic (32 rows x 64 samples of FP16 per wave step,
// rows one activation row apart) with two lane mappings: one lane per 64 bytes of a row,
// and 8 lanes x 16 bytes per row (every instruction moves whole 128-byte segments).
// Also: the steady layer's shape with sparse and dense row strides, copies through row
// tables (banded gather, permuted and consecutive rows), and sample blocks sized for the
// memory-side cache.  Prints GB/s per pattern; used to choose the SpMM's mapping, row
// stride and row order.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <utility>
#include <vector>
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

__global__ void rd(const uint4* __restrict__ a, size_t n, uint32_t* out) {
  uint32_t s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint4 v = a[i]; s ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (s == 0x12345678u) out[0] = s;
}
__global__ void rd_nt(const uint4* __restrict__ a, size_t n, uint32_t* out) {
  uint32_t s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    u4v v = __builtin_nontemporal_load(reinterpret_cast<const u4v*>(a) + i); s ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (s == 0x12345678u) out[0] = s;
}
__global__ void wr(uint4* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = make_uint4((uint32_t)i, 1, 2, 3);
}
__global__ void wr_nt(uint4* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(u4v{(uint32_t)i, 1u, 2u, 3u}, reinterpret_cast<u4v*>(a) + i);
}
__global__ void cp(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}
__global__ void cp_nt(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(reinterpret_cast<const u4v*>(a) + i), reinterpret_cast<u4v*>(b) + i);
}
// unrolled copy: each thread moves 4 x 16 B per iteration (more bytes in flight)
__global__ void cp4(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i + 3 * stride < n; i += 4 * stride) {
    uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
    b[i] = v0; b[i + stride] = v1; b[i + 2 * stride] = v2; b[i + 3 * stride] = v3;
  }
}
// the SpMM's pattern: a wave reads 32 rows x 128 B (rows ld apart) and writes 32 rows x 128 B
// work item = (tile t of 64 samples, block c); rows of block c = 32c..32c+31 (identity gather)
__global__ void tilecopy(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ld, int ntiles, int nblk, int order) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lrow = lane >> 1, lcol = (lane & 1) * 32;
  const int nitems = ntiles * nblk;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += gridDim.x * 4) {
    int t, c;
    if (order) { c = item / ntiles; t = item - c * ntiles; } else { t = item / nblk; c = item - t * nblk; }
    const uint4* src = reinterpret_cast<const uint4*>(x + (int64_t)(32 * c + lrow) * ld + t * 64 + lcol);
    uint4* dst = reinterpret_cast<uint4*>(y + (int64_t)(32 * c + lrow) * ld + t * 64 + lcol);
    uint4 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
    dst[0] = v0; dst[1] = v1; dst[2] = v2; dst[3] = v3;
  }
}

// coalesced mapping: load instruction k covers rows 8k..8k+7, lane -> row 8k + lane/8, 16 B at (lane%8)*16
template <bool NT>
__global__ void tilecopy_co(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ld, int ntiles, int nblk, int order) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r8 = lane >> 3, c8 = (lane & 7) * 8;
  const int nitems = ntiles * nblk;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += gridDim.x * 4) {
    int t, c;
    if (order) { c = item / ntiles; t = item - c * ntiles; } else { t = item / nblk; c = item - t * nblk; }
    u4v v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u4v* src = reinterpret_cast<const u4v*>(x + (int64_t)(32 * c + 8 * k + r8) * ld + t * 64 + c8);
      v[k] = NT ? __builtin_nontemporal_load(src) : *src;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u4v* dst = reinterpret_cast<u4v*>(y + (int64_t)(32 * c + 8 * k + r8) * ld + t * 64 + c8);
      if (NT) __builtin_nontemporal_store(v[k], dst); else *dst = v[k];
    }
  }
}
// the coalesced nontemporal copy with two items per loop trip (8 loads in flight per lane)
__global__ void tilecopy_co2(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ld, int ntiles, int nblk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r8 = lane >> 3, c8 = (lane & 7) * 8;
  const int nitems = ntiles * nblk;
  const int stride = gridDim.x * 4;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += 2 * stride) {
    const int item2 = item + stride < nitems ? item + stride : item;
    u4v v[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int it = h ? item2 : item;
      const int c = it / ntiles, t = it - c * ntiles;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        v[4 * h + k] = __builtin_nontemporal_load(reinterpret_cast<const u4v*>(x + (int64_t)(32 * c + 8 * k + r8) * ld + t * 64 + c8));
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int it = h ? item2 : item;
      const int c = it / ntiles, t = it - c * ntiles;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        __builtin_nontemporal_store(v[4 * h + k], reinterpret_cast<u4v*>(y + (int64_t)(32 * c + 8 * k + r8) * ld + t * 64 + c8));
    }
  }
}
// the coalesced nontemporal copy through row tables: block c reads rows rin[32c..32c+31]
// and writes rows rout[32c..32c+31] (block-column-major items)
template <bool NT>
__global__ void tilecopy_tab(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ld, int ntiles, int nblk,
                             const int* __restrict__ rin, const int* __restrict__ rout) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r8 = lane >> 3, c8 = (lane & 7) * 8;
  const int nitems = ntiles * nblk;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += gridDim.x * 4) {
    const int c = item / ntiles, t = item - c * ntiles;
    u4v v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const u4v* src = reinterpret_cast<const u4v*>(x + (int64_t)rin[32 * c + 8 * k + r8] * ld + t * 64 + c8);
      v[k] = NT ? __builtin_nontemporal_load(src) : *src;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u4v* dst = reinterpret_cast<u4v*>(y + (int64_t)rout[32 * c + 8 * k + r8] * ld + t * 64 + c8);
      if (NT) __builtin_nontemporal_store(v[k], dst); else *dst = v[k];
    }
  }
}
// 128 samples per wave-step: 16 rows x 256 B per instruction pair
__global__ void tilecopy_w128(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int64_t ld, int ntiles2, int nblk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r4 = lane >> 4, c16 = (lane & 15) * 8;
  const int nitems = ntiles2 * nblk;
  for (int item = blockIdx.x * 4 + wave; item < nitems; item += gridDim.x * 4) {
    const int c = item / ntiles2, t = item - c * ntiles2;
    u4v v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      v[k] = *reinterpret_cast<const u4v*>(x + (int64_t)(32 * c + 4 * k + r4) * ld + t * 128 + c16);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      *reinterpret_cast<u4v*>(y + (int64_t)(32 * c + 4 * k + r4) * ld + t * 128 + c16) = v[k];
  }
}

int main() {
  const size_t bytes = size_t{2} << 30, n = bytes / 16;
  uint4 *a, *b; uint32_t* o;
  CK(hipMalloc(&a, bytes)); CK(hipMalloc(&b, bytes)); CK(hipMalloc(&o, 64));
  CK(hipMemset(a, 1, bytes)); CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  auto run = [&](const char* name, double moved, auto launch) -> int {
    float best = 1e30f;
    for (int it = 0; it < 4; ++it) {
      CK(hipEventRecord(e0, 0)); launch(); CK(hipGetLastError()); CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (it && ms < best) best = ms;
    }
    printf("%-28s %8.3f ms  %7.0f GB/s\n", name, best, moved / (best * 1e-3) / 1e9);
    return 0;
  };
  for (int grid : {1024, 4096, 16384}) {
    char nm[64];
    snprintf(nm, 64, "read g=%d", grid); run(nm, bytes, [&] { hipLaunchKernelGGL(rd, dim3(grid), dim3(256), 0, 0, a, n, o); });
    snprintf(nm, 64, "read_nt g=%d", grid); run(nm, bytes, [&] { hipLaunchKernelGGL(rd_nt, dim3(grid), dim3(256), 0, 0, a, n, o); });
    snprintf(nm, 64, "write g=%d", grid); run(nm, bytes, [&] { hipLaunchKernelGGL(wr, dim3(grid), dim3(256), 0, 0, b, n); });
    snprintf(nm, 64, "write_nt g=%d", grid); run(nm, bytes, [&] { hipLaunchKernelGGL(wr_nt, dim3(grid), dim3(256), 0, 0, b, n); });
    snprintf(nm, 64, "copy g=%d", grid); run(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(cp, dim3(grid), dim3(256), 0, 0, a, b, n); });
    snprintf(nm, 64, "copy_nt g=%d", grid); run(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(cp_nt, dim3(grid), dim3(256), 0, 0, a, b, n); });
    snprintf(nm, 64, "copy4 g=%d", grid); run(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(cp4, dim3(grid), dim3(256), 0, 0, a, b, n); });
  }
  // tile copy shaped like the SpMM at N=16384 rows, 9984 samples (ld 10240) and 60032 (ld 60160)
  for (int S : {9984, 60032}) {
    const int64_t ld = (S + 255) / 256 * 256;
    const int N = 16384;
    if ((size_t)N * ld * 2 > bytes) { printf("skip\n"); continue; }
    const int ntiles = S / 64, nblk = N / 32;
    for (int order : {0, 1}) for (int grid : {512, 1024, 2048}) {
      char nm[64]; snprintf(nm, 64, "tilecopy S=%d o=%d g=%d", S, order, grid);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk, order); });
    }
  }
  for (int S : {9984, 60032}) {
    const int64_t ld = (S + 255) / 256 * 256;
    const int N = 16384;
    const int ntiles = S / 64, nblk = N / 32;
    for (int order : {0, 1}) for (int grid : {512, 1024, 2048}) {
      char nm[64]; snprintf(nm, 64, "tilecopy_co S=%d o=%d g=%d", S, order, grid);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy_co<false>, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk, order); });
      snprintf(nm, 64, "tilecopy_co_nt S=%d o=%d g=%d", S, order, grid);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy_co<true>, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk, order); });
    }
    for (int grid : {512, 1024}) {
      char nm[64]; snprintf(nm, 64, "tilecopy_w128 S=%d g=%d", S, grid);
      run(nm, 2.0 * N * (S / 128 * 128) * 2, [&] { hipLaunchKernelGGL(tilecopy_w128, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, S / 128, nblk); });
    }
  }
  // the steady-state layer shape: 9984 live rows in buffers laid out for 60000 (ld 60160),
  // against the same rows stored compactly (ld 10240); one and two items per loop trip
  for (int64_t ld : {int64_t{10240}, int64_t{60160}}) {
    const int N = 16384, S = 9984, ntiles = S / 64, nblk = N / 32;
    for (int grid : {1024, 2048}) {
      char nm[80]; snprintf(nm, 80, "steady co_nt ld=%lld g=%d", (long long)ld, grid);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy_co<true>, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk, 1); });
      snprintf(nm, 80, "steady co2_nt ld=%lld g=%d", (long long)ld, grid);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy_co2, dim3(grid), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk); });
    }
  }
  // row tables of a banded layer (block c gathers rows c + 512 k, k < 32) against
  // consecutive rows, and the banded gather through a bit-reversed row placement
  {
    const int N = 16384, S = 9984, ntiles = S / 64, nblk = N / 32;
    const int64_t ld = 10048;  // act_ld(9953): odd multiple of 64
    std::vector<int> seq(N), band(N), perm(N), bandp(N);
    for (int i = 0; i < N; ++i) seq[i] = i;
    for (int c = 0; c < nblk; ++c)
      for (int k = 0; k < 32; ++k) band[32 * c + k] = c + nblk * k;
    for (int i = 0; i < N; ++i) {
      int r = 0;
      for (int b = 0; b < 14; ++b) r |= ((i >> b) & 1) << (13 - b);
      perm[i] = r;
    }
    for (int i = 0; i < N; ++i) bandp[i] = perm[band[i]];
    // outputs: a fixed pseudo-random permutation (the banded generator's output order)
    std::vector<int> rnd(N);
    for (int i = 0; i < N; ++i) rnd[i] = i;
    uint64_t st = 12345;
    for (int i = N - 1; i > 0; --i) {
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      const int j = (int)((st >> 33) % (uint64_t)(i + 1));
      std::swap(rnd[i], rnd[j]);
    }
    int *d_seq, *d_band, *d_bandp, *d_rnd;
    CK(hipMalloc(&d_seq, N * 4)); CK(hipMalloc(&d_band, N * 4)); CK(hipMalloc(&d_bandp, N * 4)); CK(hipMalloc(&d_rnd, N * 4));
    CK(hipMemcpy(d_seq, seq.data(), N * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_band, band.data(), N * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_bandp, bandp.data(), N * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_rnd, rnd.data(), N * 4, hipMemcpyHostToDevice));
    struct { const char* name; const int* in; const int* out; } cases[] = {
        {"rows seq -> seq", d_seq, d_seq}, {"rows seq -> rnd", d_seq, d_rnd},
        {"rows band -> rnd", d_band, d_rnd}, {"rows bitrev(band) -> rnd", d_bandp, d_rnd},
        {"rows band -> seq", d_band, d_seq}, {"rows rnd -> seq", d_rnd, d_seq}, {"rows rnd -> rnd", d_rnd, d_rnd}};
    for (auto& cs : cases) {
      char nm[80]; snprintf(nm, 80, "%s ld=%lld", cs.name, (long long)ld);
      run(nm, 2.0 * N * S * 2, [&] { hipLaunchKernelGGL(tilecopy_tab<true>, dim3(1024), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ld, ntiles, nblk, cs.in, cs.out); });
    }
    // sample blocks small enough for the 256 MB memory-side cache: one block ping-ponged
    // between two buffers (rnd -> seq rows), plain and nontemporal accesses
    for (int Sb : {1216, 2496, 4992, 9984}) {
      const int64_t ldb = 64 * (((Sb + 63) / 64) | 1);
      const int nt = Sb / 64;
      for (int ntm = 0; ntm < 2; ++ntm) {
        char nm[80]; snprintf(nm, 80, "block S=%d %s (x2 buf %.0f MB)", Sb, ntm ? "nt" : "plain", 2.0 * N * ldb * 2 / 1e6);
        run(nm, 2.0 * N * Sb * 2, [&] {
          if (ntm) {
            hipLaunchKernelGGL(tilecopy_tab<true>, dim3(1024), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ldb, nt, nblk, d_rnd, d_seq);
          } else {
            hipLaunchKernelGGL(tilecopy_tab<false>, dim3(1024), dim3(256), 0, 0, (const uint16_t*)a, (uint16_t*)b, ldb, nt, nblk, d_rnd, d_seq);
          }
        });
      }
    }
  }
  return 0;
}
