// gfx950 kernels of the sparse-DNN inference path.
//
//   spmm_mfma_kernel   FP16 gathered-BCSC SpMM on v_mfma_f32_16x16x32_f16 with a
//                      fused bias + clamp(0,32) + liveness / guard flags + compacted
//                      store epilogue (replaces cospdnn_single, kernels/kernel.cu:240-379,
//                      and check_active_f_col / matrix_reshape, kernel.cu:129-172).
//   spmm_mfma32_*      FP32 MFMA SpMM (v_mfma_f32_16x16x4_f32) for the FP32 guard pass.
//   spmm_simt_kernel   plain SIMT per-output gather SpMM (FP32 baseline path, the FP64
//                      pass, and the FP16 path for layers without a compact tile image).
//   spmm_simt_rows_kernel  the same arithmetic spread over outputs for passes over few rows.
//   scan_kernel        live flags -> prefix sum -> position map, id compaction, guard
//                      bookkeeping, saturated-row retirement (replaces the host scan of
//                      cospdnn.cu:249-265).
#include <hip/hip_runtime.h>

#include "gen.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace cospdnn {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef __fp16 fp16x4_t __attribute__((__vector_size__(4 * sizeof(__fp16))));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr_t;
// Weight-side arrays are never written while a kernel runs: reading them through the
// constant address space lets wave-uniform indices become scalar loads.
typedef const __attribute__((address_space(4))) int32_t* cint_ptr_t;

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
// scan: work-group g owns positions [4096g, 4096g+4096); thread t of its 256
// owns 16 consecutive positions and moves them with 16-byte loads and stores.
// The flags are 0/1 bytes, so a popcount of (word & 0x01010101) counts four
// of them.  A group reserves its slice of the compacted layout with one
// atomicAdd on counts[layer+1] (zeroed when the pass starts), so the compacted
// order follows the order in which groups finish, not the sample order; the
// category ids are sorted on the host.  Thread offsets inside the group come
// from a wave64 shuffle scan and a 4-entry LDS scan over the waves.
// --------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanThreads * kScanPer;

__device__ __forceinline__ int count_flags16(uint4 v) {
  return __popc(v.x & 0x01010101u) + __popc(v.y & 0x01010101u) + __popc(v.z & 0x01010101u) +
         __popc(v.w & 0x01010101u);
}
__device__ __forceinline__ bool flag_at(const uint4& v, int i) {
  const uint32_t w = i < 4 ? v.x : (i < 8 ? v.y : (i < 12 ? v.z : v.w));
  return ((w >> (8 * (i & 3))) & 0xFFu) != 0;
}

__global__ __launch_bounds__(kScanThreads) void scan_kernel(ScanArgs a) {
  __shared__ int wsum[kScanThreads / 64];
  __shared__ int base;
  const int n = a.counts[a.layer];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (blockIdx.x == 0 && tid == 0 && a.work != nullptr) *a.work = 0;
  const int pc = blockIdx.x * kScanTile + tid * kScanPer;  // multiple of 16: aligned, inside the ld-sized arrays
  if (blockIdx.x * kScanTile >= n) return;                  // uniform per group
  const uint4 zero = {0u, 0u, 0u, 0u};
  uint4 va = zero, vu = zero, vb = zero, vr = zero;
  int cnt = 0;
  if (pc < n) {
    va = *reinterpret_cast<const uint4*>(a.alive + pc);
    vu = *reinterpret_cast<const uint4*>(a.unsat + pc);
    if (a.band != nullptr) vb = *reinterpret_cast<const uint4*>(a.band + pc);
    if (a.retire_sat) {
      // flags are 0/1 bytes: retired = alive & !unsat, kept = alive & unsat
      vr = make_uint4(va.x & ~vu.x & 0x01010101u, va.y & ~vu.y & 0x01010101u, va.z & ~vu.z & 0x01010101u,
                      va.w & ~vu.w & 0x01010101u);
      va = make_uint4(va.x & vu.x, va.y & vu.y, va.z & vu.z, va.w & vu.w);
    }
    if (pc + kScanPer <= n) {
      cnt = count_flags16(va);
    } else {
      for (int i = 0; i < kScanPer; ++i) cnt += (pc + i < n && flag_at(va, i)) ? 1 : 0;
    }
  }
  int x = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (tid == 0) {
    int sum = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
      const int t = wsum[i];
      wsum[i] = sum;
      sum += t;
    }
    base = atomicAdd(a.counts + a.layer + 1, sum);
  }
  __syncthreads();
  if (pc >= n) return;
  int q = base + wsum[wv] + x - cnt;
  int qm[kScanPer];
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    const int p = pc + i;
    const bool al = p < n && flag_at(va, i);
    qm[i] = -1;
    if (al) {
      qm[i] = q;
      a.ids_out[q] = a.ids_in[p];
      ++q;
    }
    if (p < n && flag_at(vr, i)) a.ret_ids[atomicAdd(a.ret_count, 1)] = a.ids_in[p];
    const bool fl = p < n && (flag_at(vb, i) || (a.flag_undecided && al && flag_at(vu, i)));
    if (fl) {
      const int32_t id = a.ids_in[p];
      if (a.flag_mark[id] == 0) {
        a.flag_mark[id] = 1;
        a.flag_ids[atomicAdd(a.flag_count, 1)] = id;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kScanPer; i += 4)
    *reinterpret_cast<int4*>(a.qmap + pc + i) = make_int4(qm[i], qm[i + 1], qm[i + 2], qm[i + 3]);
  *reinterpret_cast<uint4*>(a.clear0 + pc) = zero;
  *reinterpret_cast<uint4*>(a.clear1 + pc) = zero;
  *reinterpret_cast<uint4*>(a.clear2 + pc) = zero;
}

// --------------------------------------------------------------------------
// MFMA SpMM.  Every wave works alone (no work-group barriers): a work item is
// (64 samples, block column c of 32 outputs), items are strided over all waves
// of the grid, and the K chunks of an item are consecutive steps.
// Per step the wave's 32 gathered input rows x 64 samples (4 KB) arrive in
// registers, prefetched one step ahead (lane L loads row L/2, samples
// 32(L%2)..+31 with four 16-byte loads; the row ids come two steps ahead), are
// written to the wave's own LDS slice, and each lane reads its B fragment
// X[8g..8g+7][sample] with two ds_read_b64_tr_b16.  D[out][sample] =
// W^T[out][k] * X[k][sample] on v_mfma_f32_16x16x32_f16 (2 M-tiles x 4 N-tiles).
// Epilogue: the pre-activations are staged in LDS as FP32 [32][64]; lane L then
// owns sample L: it clamps, stores its 32 outputs (each store instruction covers
// 64 consecutive target positions of one output row) and sets its flags.
// --------------------------------------------------------------------------
constexpr int kWaveSamples = 64;
constexpr int kWavesPerGroup = 4;

struct StepPos {
  int item, t, c, ch, ch_end;
  bool valid;
};

__device__ __forceinline__ void step_first(StepPos& s, int item, int nitems, int nblk, cint_ptr_t kptr, bool k1) {
  s.valid = item < nitems;
  s.item = item;
  s.t = item / nblk;
  s.c = item - s.t * nblk;
  if (!s.valid) {
    s.ch = 0;
    s.ch_end = 1;
  } else if (k1) {
    s.ch = s.c;
    s.ch_end = s.c + 1;
  } else {
    s.ch = kptr[s.c];
    s.ch_end = kptr[s.c + 1];
  }
}

__device__ __forceinline__ void step_next(StepPos& s, int stride, int nitems, int nblk, cint_ptr_t kptr, bool k1) {
  if (!s.valid) return;
  if (s.ch + 1 < s.ch_end) {
    ++s.ch;
    return;
  }
  step_first(s, s.item + stride, nitems, nblk, kptr, k1);
}


__global__ __launch_bounds__(256, 4) void spmm_mfma_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) float wlds[kWavesPerGroup][32 * kWaveSamples];  // 8 KB per wave
  const int n = a.counts[a.layer];
  const int nblk = w.nblk;
  const int nitems = (n + kWaveSamples - 1) / kWaveSamples * nblk;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int stride = gridDim.x * kWavesPerGroup;
  const bool k1 = w.one_chunk != 0;
  const cint_ptr_t kptr = (cint_ptr_t)w.kptr;
  const cint_ptr_t out_idx = (cint_ptr_t)w.out_idx;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);

  StepPos cur, nx1, nx2;
  step_first(cur, blockIdx.x * kWavesPerGroup + wave, nitems, nblk, kptr, k1);
  if (!cur.valid) return;  // wave-uniform; nothing below synchronises across waves
  nx1 = cur;
  step_next(nx1, stride, nitems, nblk, kptr, k1);
  nx2 = nx1;
  step_next(nx2, stride, nitems, nblk, kptr, k1);

  _Float16* tile = reinterpret_cast<_Float16*>(wlds[wave]);  // [32][64] FP16 input tile (first 4 KB)
  float* stage = wlds[wave];                                 // [32][64] FP32 pre-activations
  const int lrow = lane >> 1, lcol = (lane & 1) * 32;
  const int j = lane & 15, g = lane >> 4;
  const int32_t* __restrict__ gidx = w.gidx;

  // Everything a step consumes is loaded one step ahead: its tile rows (d0..d3), its A
  // fragments, the target position of this lane's sample; the row ids two steps ahead.
  // After the final step the prefetch re-reads valid (unused) addresses.
  int row_nx1 = gidx[(nx1.valid ? nx1.ch : cur.ch) * kChunk + lrow];
  uint4 d0, d1, d2, d3;
  half8_t a0, a1;
  int qv;
  {
    const uint4* src = reinterpret_cast<const uint4*>(
        X + static_cast<int64_t>(gidx[cur.ch * kChunk + lrow]) * a.ldx + cur.t * kWaveSamples + lcol);
    d0 = src[0];
    d1 = src[1];
    d2 = src[2];
    d3 = src[3];
    a0 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(cur.ch) * 2 + 0) * 512 + lane * 8);
    a1 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(cur.ch) * 2 + 1) * 512 + lane * 8);
    qv = a.qmap[min(cur.t * kWaveSamples + lane, n - 1)];
  }
  float4_t acc[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[m][nt] = float4_t{0.f, 0.f, 0.f, 0.f};

  while (true) {
    const bool last = cur.ch + 1 == cur.ch_end;
    const int p_own = cur.t * kWaveSamples + lane;
    const int q_own = p_own < n ? qv : -1;
    const half8_t ca0 = a0, ca1 = a1;
    // this step's tile into the wave's LDS slice
    {
      uint4* dst = reinterpret_cast<uint4*>(tile + lrow * kWaveSamples + lcol);
      dst[0] = d0;
      dst[1] = d1;
      dst[2] = d2;
      dst[3] = d3;
    }
    // next step's loads
    {
      const int pt = nx1.valid ? nx1.t : cur.t;
      const int pch = nx1.valid ? nx1.ch : cur.ch;
      const uint4* src = reinterpret_cast<const uint4*>(X + static_cast<int64_t>(row_nx1) * a.ldx + pt * kWaveSamples + lcol);
      d0 = src[0];
      d1 = src[1];
      d2 = src[2];
      d3 = src[3];
      a0 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(pch) * 2 + 0) * 512 + lane * 8);
      a1 = *reinterpret_cast<const half8_t*>(w.wt + (static_cast<int64_t>(pch) * 2 + 1) * 512 + lane * 8);
      qv = a.qmap[min(pt * kWaveSamples + lane, n - 1)];
      row_nx1 = gidx[(nx2.valid ? nx2.ch : pch) * kChunk + lrow];
    }
    // B fragments: ds_read_b64_tr_b16 gives lane i of each 16-lane group element i%4 of the
    // 8 bytes read by lanes 4k+i/4, so lane m reads row 8g+4h+m/4 at samples 16nt+4(m%4)..+3
    half8_t bfrag[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const _Float16* p0 = tile + (g * 8 + (j >> 2)) * kWaveSamples + nt * 16 + 4 * (j & 3);
      const fp16x4_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)p0);
      const fp16x4_t hi =
          __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)(p0 + 4 * kWaveSamples));
      bfrag[nt] = __builtin_shufflevector(__builtin_bit_cast(half4_t, lo), __builtin_bit_cast(half4_t, hi), 0, 1, 2, 3,
                                          4, 5, 6, 7);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ca0, bfrag[nt], acc[0][nt], 0, 0, 0);
      acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ca1, bfrag[nt], acc[1][nt], 0, 0, 0);
    }
    if (last) {
      // D[i][j] of a 16x16 tile sits in lane j + 16*(i/4), element i%4: this lane holds
      // outputs 16m + 4g + r of sample 16nt + j
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
          for (int r = 0; r < 4; ++r) stage[(m * 16 + g * 4 + r) * kWaveSamples + nt * 16 + j] = acc[m][nt][r] + a.bias;
      int fl = 0;
      if (q_own >= 0) {
        const int32_t cbase = cur.c * 32;
#pragma unroll 8
        for (int o = 0; o < 32; ++o) {
          const int out = out_idx[cbase + o];
          const float v = stage[o * kWaveSamples + lane];
          if (out >= 0) {
            fl |= (v > 0.0f) ? 1 : 0;
            fl |= (v < 32.0f) ? 2 : 0;
            fl |= (fabsf(v) < a.band || fabsf(v - 32.0f) < a.band) ? 4 : 0;
            Y[static_cast<int64_t>(out) * a.ldy + q_own] = static_cast<_Float16>(fminf(fmaxf(v, 0.0f), 32.0f));
          }
        }
        if (fl & 1) a.alive_next[q_own] = 1;
        if (fl & 2) a.unsat_next[q_own] = 1;
        if ((fl & 4) && a.band_next != nullptr) a.band_next[q_own] = 1;
      }
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[m][nt] = float4_t{0.f, 0.f, 0.f, 0.f};
    }
    if (!nx1.valid) break;
    cur = nx1;
    nx1 = nx2;
    step_next(nx2, stride, nitems, nblk, kptr, k1);
  }
}

// --------------------------------------------------------------------------
// SIMT gather SpMM: thread = sample, work item = (tile of 256 samples, block of
// 32 outputs); each output sums its CSC inputs read straight from global memory.
// --------------------------------------------------------------------------
// Accumulation of one product: FP32 lets the compiler fuse; FP64 keeps the oracle's
// separate multiply and add so FP64 passes are bit-identical to it.
template <typename TAcc>
__device__ __forceinline__ TAcc madd(TAcc acc, TAcc x, float w);
template <>
__device__ __forceinline__ float madd<float>(float acc, float x, float w) { return acc + x * w; }
template <>
__device__ __forceinline__ double madd<double>(double acc, double x, float w) {
  return __dadd_rn(acc, __dmul_rn(x, static_cast<double>(w)));
}

template <typename TIn, typename TOut>
__global__ __launch_bounds__(256) void spmm_simt_kernel(StepArgs a, DevLayer w) {
  typedef typename std::conditional<std::is_same<TOut, double>::value, double, float>::type TAcc;
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
      TAcc acc = 0;
      int k = k0;
      // eight independent gathers in flight per thread, summed in CSC order
      for (; k + 8 <= k1; k += 8) {
        TIn xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xcol[static_cast<int64_t>(w.orow[k + u]) * a.ldx];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = madd<TAcc>(acc, static_cast<TAcc>(xv[u]), w.oval[k + u]);
      }
      for (; k < k1; ++k) acc = madd<TAcc>(acc, static_cast<TAcc>(xcol[static_cast<int64_t>(w.orow[k]) * a.ldx]), w.oval[k]);
      const TAcc v = std::is_same<TAcc, double>::value ? static_cast<TAcc>(__dadd_rn(acc, static_cast<double>(a.bias)))
                                                       : acc + static_cast<TAcc>(a.bias);
      fl |= (v > 0) ? 1 : 0;
      fl |= (v < 32) ? 2 : 0;
      fl |= (fabs(static_cast<double>(v)) < a.band || fabs(static_cast<double>(v) - 32.0) < a.band) ? 4 : 0;
      const TAcc y = v < 0 ? TAcc(0) : (v > 32 ? TAcc(32) : v);
      Y[static_cast<int64_t>(out) * a.ldy + q] = static_cast<TOut>(y);
    }
    if (fl & 1) a.alive_next[q] = 1;
    if (fl & 2) a.unsat_next[q] = 1;
    if ((fl & 4) && a.band_next != nullptr) a.band_next[q] = 1;
  }
}

// Few-row variant for the guard passes: a work item is (8 samples, block column c of 32
// outputs) and thread t computes output t/8 of sample t%8, so a pass over a handful of
// rows still spreads each layer's gathers over the grid (the per-sample kernel above
// would leave 248 of 256 threads idle and run 32 outputs x K gathers in series).
// Same arithmetic and CSC summation order as spmm_simt_kernel.  Several threads set the
// same flag byte of a sample; they all write 1.
template <typename TIn, typename TOut>
__global__ __launch_bounds__(256) void spmm_simt_rows_kernel(StepArgs a, DevLayer w) {
  typedef typename std::conditional<std::is_same<TOut, double>::value, double, float>::type TAcc;
  const int n = a.counts[a.layer];
  const int ntiles = (n + kFewRowsTile - 1) / kFewRowsTile;
  const int64_t nitems = static_cast<int64_t>(ntiles) * w.nblk;
  const TIn* __restrict__ X = static_cast<const TIn*>(a.x);
  TOut* __restrict__ Y = static_cast<TOut*>(a.y);
  const int o = static_cast<int>(threadIdx.x) / kFewRowsTile;
  const int sp = static_cast<int>(threadIdx.x) % kFewRowsTile;
  for (int64_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int t = static_cast<int>(item / w.nblk);
    const int c = static_cast<int>(item - static_cast<int64_t>(t) * w.nblk);
    const int p = t * kFewRowsTile + sp;
    const int q = p < n ? a.qmap[p] : -1;
    if (q < 0) continue;
    const int out = w.out_idx[c * 32 + o];
    if (out < 0) continue;
    const TIn* __restrict__ xcol = X + p;
    const int k0 = w.optr[c * 32 + o], k1 = w.optr[c * 32 + o + 1];
    TAcc acc = 0;
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
      TIn xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = xcol[static_cast<int64_t>(w.orow[k + u]) * a.ldx];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = madd<TAcc>(acc, static_cast<TAcc>(xv[u]), w.oval[k + u]);
    }
    for (; k < k1; ++k) acc = madd<TAcc>(acc, static_cast<TAcc>(xcol[static_cast<int64_t>(w.orow[k]) * a.ldx]), w.oval[k]);
    const TAcc v = std::is_same<TAcc, double>::value ? static_cast<TAcc>(__dadd_rn(acc, static_cast<double>(a.bias)))
                                                     : acc + static_cast<TAcc>(a.bias);
    const TAcc y = v < 0 ? TAcc(0) : (v > 32 ? TAcc(32) : v);
    Y[static_cast<int64_t>(out) * a.ldy + q] = static_cast<TOut>(y);
    if (v > 0) a.alive_next[q] = 1;
    if (v < 32) a.unsat_next[q] = 1;
    if (a.band_next != nullptr && (fabs(static_cast<double>(v)) < a.band || fabs(static_cast<double>(v) - 32.0) < a.band))
      a.band_next[q] = 1;
  }
}

// --------------------------------------------------------------------------
// FP32 MFMA SpMM (v_mfma_f32_16x16x4f32), same work decomposition as the FP16
// kernel; used for the FP32 guard recompute.  TIn = _Float16 reads the input
// store in layer 0.  LDS tile: 32 gathered rows x 256 samples of TIn.
// --------------------------------------------------------------------------
template <typename TIn>
__device__ __forceinline__ void spmm_mfma32_body(const StepArgs& a, const DevLayer& w, TIn* xs) {
  constexpr int kRowsPerLoad = 64 * 16 / (kTileS * static_cast<int>(sizeof(TIn)));  // rows per wave instruction
  constexpr int kLoads = kChunk / 4 / kRowsPerLoad;                                  // instructions per wave
  const int n = a.counts[a.layer];
  const int ntiles = (n + kTileS - 1) / kTileS;
  const int64_t nitems = static_cast<int64_t>(ntiles) * w.nblk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const TIn* __restrict__ X = static_cast<const TIn*>(a.x);
  float* __restrict__ Y = static_cast<float*>(a.y);
  constexpr int kLanesPerRow = kTileS * static_cast<int>(sizeof(TIn)) / 16;
  constexpr int kElemsPerLane = 16 / static_cast<int>(sizeof(TIn));

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
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kLoads; ++i) {
        const int r = wave * (kChunk / 4) + i * kRowsPerLoad;  // wave-uniform first row
        const int row = w.gidx[ch * kChunk + r + lane / kLanesPerRow];
        const TIn* src = X + static_cast<int64_t>(row) * a.ldx + s0 + (lane % kLanesPerRow) * kElemsPerLane;
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(xs + r * kTileS), 16, 0, 0);
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      const float* __restrict__ wa = w.wt32 + static_cast<int64_t>(ch) * 1024;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const float a0 = wa[(0 * 8 + kk) * 64 + lane];
        const float a1 = wa[(1 * 8 + kk) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          const float b = static_cast<float>(xs[(4 * kk + g) * kTileS + wave * 64 + nt * 16 + j]);
          acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b, acc[0][nt], 0, 0, 0);
          acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, acc[1][nt], 0, 0, 0);
        }
      }
    }
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
            if (q >= 0) Y[static_cast<int64_t>(outs[m][r]) * a.ldy + q] = fminf(fmaxf(v, 0.0f), 32.0f);
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

// amdgcn builtins cannot sit in a __global__ template (the host pass rejects the
// instantiation), so the two input types get plain kernels around the body.
__global__ __launch_bounds__(256) void spmm_mfma32_f16in_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) _Float16 xs[kChunk * kTileS];
  spmm_mfma32_body<_Float16>(a, w, xs);
}
__global__ __launch_bounds__(256) void spmm_mfma32_f32in_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) float xs[kChunk * kTileS];
  spmm_mfma32_body<float>(a, w, xs);
}

// alive flags straight from the input (only for networks without layers)
__global__ void input_alive_kernel(const _Float16* __restrict__ store, int64_t ld, int N, int64_t S,
                                   uint8_t* __restrict__ alive) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  int any = 0;
  for (int nn = 0; nn < N; ++nn) any |= static_cast<float>(store[static_cast<int64_t>(nn) * ld + s]) != 0.0f;
  alive[s] = static_cast<uint8_t>(any);
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
hipError_t launch_scan(const ScanArgs& a, int64_t max_rows, hipStream_t st) {
  const int64_t groups = std::max<int64_t>(1, (max_rows + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(scan_kernel, dim3(static_cast<unsigned>(groups)), dim3(kScanThreads), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_spmm_mfma(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  hipLaunchKernelGGL(spmm_mfma_kernel, dim3(grid), dim3(256), 0, st, a, w);
  return hipGetLastError();
}
hipError_t launch_spmm_simt_rows(const StepArgs& a, const DevLayer& w, Dtype tin, Dtype tout, int grid, hipStream_t st) {
  if (tin == Dtype::F16 && tout == Dtype::F16)
    hipLaunchKernelGGL((spmm_simt_rows_kernel<_Float16, _Float16>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F16 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_rows_kernel<_Float16, float>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F32 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_rows_kernel<float, float>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F16 && tout == Dtype::F64)
    hipLaunchKernelGGL((spmm_simt_rows_kernel<_Float16, double>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F64 && tout == Dtype::F64)
    hipLaunchKernelGGL((spmm_simt_rows_kernel<double, double>), dim3(grid), dim3(256), 0, st, a, w);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_spmm_simt(const StepArgs& a, const DevLayer& w, Dtype tin, Dtype tout, int grid, hipStream_t st) {
  if (tin == Dtype::F16 && tout == Dtype::F16)
    hipLaunchKernelGGL((spmm_simt_kernel<_Float16, _Float16>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F16 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_kernel<_Float16, float>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F32 && tout == Dtype::F32)
    hipLaunchKernelGGL((spmm_simt_kernel<float, float>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F16 && tout == Dtype::F64)
    hipLaunchKernelGGL((spmm_simt_kernel<_Float16, double>), dim3(grid), dim3(256), 0, st, a, w);
  else if (tin == Dtype::F64 && tout == Dtype::F64)
    hipLaunchKernelGGL((spmm_simt_kernel<double, double>), dim3(grid), dim3(256), 0, st, a, w);
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

hipError_t launch_spmm_mfma32(const StepArgs& a, const DevLayer& w, Dtype tin, int grid, hipStream_t st) {
  if (tin == Dtype::F16)
    hipLaunchKernelGGL(spmm_mfma32_f16in_kernel, dim3(grid), dim3(256), 0, st, a, w);
  else
    hipLaunchKernelGGL(spmm_mfma32_f32in_kernel, dim3(grid), dim3(256), 0, st, a, w);
  return hipGetLastError();
}
hipError_t launch_input_alive(const _Float16* store, int64_t ld, int N, int64_t S, uint8_t* alive, hipStream_t st) {
  if (S <= 0) return hipSuccess;
  hipLaunchKernelGGL(input_alive_kernel, dim3(static_cast<unsigned>((S + 255) / 256)), dim3(256), 0, st, store, ld, N, S,
                     alive);
  return hipGetLastError();
}
int spmm_mfma32_blocks_per_cu(Dtype tin) {
  int nb = 0;
  hipError_t e = tin == Dtype::F16
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_mfma32_f16in_kernel, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_mfma32_f32in_kernel, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

hipError_t configure_kernels() { return hipSuccess; }

int spmm_mfma_resident_per_cu() {
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(spmm_mfma_kernel)) != hipSuccess) return 1;
  // arch + accumulation VGPRs, allocated in granules of 8
  const int vgprs = ((attr.numRegs > 0 ? attr.numRegs : 256) + 7) / 8 * 8;
  const int waves_per_simd = std::max(1, std::min(8, 512 / vgprs));
  const int by_vgpr = waves_per_simd;  // a 256-thread group puts one wave on each of the 4 SIMDs
  const int lds = static_cast<int>(attr.sharedSizeBytes) + 512;
  const int by_lds = std::max(1, (160 * 1024) / lds);
  return std::max(1, std::min(by_vgpr, by_lds));
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
  else if (tin == Dtype::F32 && tout == Dtype::F32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<float, float>, 256, 0);
  else if (tin == Dtype::F16)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<_Float16, double>, 256, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<double, double>, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

int spmm_simt_rows_blocks_per_cu(Dtype tin, Dtype tout) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (tin == Dtype::F16 && tout == Dtype::F16)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_rows_kernel<_Float16, _Float16>, 256, 0);
  else if (tin == Dtype::F16 && tout == Dtype::F32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_rows_kernel<_Float16, float>, 256, 0);
  else if (tin == Dtype::F32 && tout == Dtype::F32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_rows_kernel<float, float>, 256, 0);
  else if (tin == Dtype::F16)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_rows_kernel<_Float16, double>, 256, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_rows_kernel<double, double>, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

}  // namespace cospdnn
