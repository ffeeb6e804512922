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

// One work-group per copy job (see launch_copy_jobs).
__global__ __launch_bounds__(256) void copy_jobs_kernel(CopyJobs jobs, int32_t* __restrict__ dst) {
  const CopyJob& jb = jobs.job[blockIdx.x];
  const int32_t n = jb.len ? min(*jb.len, jb.cap) : jb.cap;
  if (threadIdx.x == 0 && jb.count_at >= 0) dst[jb.count_at] = n;
  for (int32_t i = threadIdx.x; i < n; i += blockDim.x) dst[jb.off + i] = jb.src[i];
}

// *dst = min(*src, cap); *overflow = 1 when *src > cap (one thread)
__global__ void clamp_count_kernel(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int32_t cap,
                                   int32_t* __restrict__ overflow) {
  const int32_t v = *src;
  *dst = v < cap ? v : cap;
  if (v > cap) *overflow = 1;
}

// dst[i] = outer[inner[i]] for i < min(*count, cap), *dst_count = that number;
// *overflow = 1 when *count > cap.  inner holds positions in the list outer.
__global__ void compose_ids_kernel(const int32_t* __restrict__ outer, const int32_t* __restrict__ inner,
                                   const int32_t* __restrict__ count, int32_t cap, int32_t* __restrict__ dst,
                                   int32_t* __restrict__ dst_count, int32_t* __restrict__ overflow) {
  const int32_t v = *count;
  const int32_t n = v < cap ? v : cap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *dst_count = n;
    if (v > cap) *overflow = 1;
  }
  if (i < n) dst[i] = outer[inner[i]];
}

// --------------------------------------------------------------------------
// scan: work-group g owns positions [4096g, 4096g+4096); thread t of its 256
// owns 16 consecutive positions and moves them with 16-byte loads and stores.
// The flags are 0/1 bytes, so a popcount of (word & 0x01010101) counts four
// of them.  The compacted order is the sample order: group g starts at the
// number of kept rows in positions [0, 4096g), which it counts itself from the
// flag bytes (4096g bytes, L2-resident), so no group waits for another and the
// layout is the same on every run.  When no row dies the layout is the identity,
// which keeps the SpMM's 64-sample tiles contiguous and aligned.  Thread offsets
// inside the group come from a wave64 shuffle scan and a 4-entry LDS scan.
// --------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanThreads * kScanPer;

__device__ __forceinline__ int count_flags16(uint4 v) {
  return __popc(v.x & 0x01010101u) + __popc(v.y & 0x01010101u) + __popc(v.z & 0x01010101u) +
         __popc(v.w & 0x01010101u);
}
// the flag bytes of a 16-byte word at positions below lim (lim - its first position)
__device__ __forceinline__ int count_flags16_lim(uint4 v, int lim) {
  if (lim >= kScanPer) return count_flags16(v);
  auto m = [&](int k) -> uint32_t {
    const int valid = min(max(lim - 4 * k, 0), 4);
    return valid >= 4 ? 0x01010101u : ((1u << (8 * valid)) - 1u) & 0x01010101u;
  };
  return __popc(v.x & m(0)) + __popc(v.y & m(1)) + __popc(v.z & m(2)) + __popc(v.w & m(3));
}
__device__ __forceinline__ bool flag_at(const uint4& v, int i) {
  const uint32_t w = i < 4 ? v.x : (i < 8 ? v.y : (i < 12 ? v.z : v.w));
  return ((w >> (8 * (i & 3))) & 0xFFu) != 0;
}

// Rows a scan keeps, four flag bytes at a time: alive, minus retired saturated rows
// (retire_sat: alive & !unsat), minus flagged undecided rows (drop_flagged at the
// checkpoint: alive & unsat).
__device__ __forceinline__ uint32_t flagged_word(const ScanArgs& a, uint32_t al, uint32_t un) {
  return a.flag_undecided ? (a.flag_all_alive ? al : (al & un)) & 0x01010101u : 0u;
}
__device__ __forceinline__ uint32_t kept_word(const ScanArgs& a, uint32_t al, uint32_t un) {
  uint32_t k = al;
  if (a.retire_sat) k &= un;
  if (a.drop_flagged) k &= ~flagged_word(a, al, un);
  return k & 0x01010101u;
}
__device__ __forceinline__ uint4 kept_flags(const ScanArgs& a, int pos) {
  const uint4 v = *reinterpret_cast<const uint4*>(a.alive + pos);
  uint4 u = make_uint4(0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u);
  if (a.retire_sat || (a.drop_flagged && a.flag_undecided && !a.flag_all_alive))
    u = *reinterpret_cast<const uint4*>(a.unsat + pos);
  return make_uint4(kept_word(a, v.x, u.x), kept_word(a, v.y, u.y), kept_word(a, v.z, u.z), kept_word(a, v.w, u.w));
}

__global__ __launch_bounds__(kScanThreads) void scan_kernel(ScanArgs a) {
  __shared__ int wsum[kScanThreads / 64];
  __shared__ int wpre[kScanThreads / 64];
  __shared__ int wpost[kScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int g0 = blockIdx.x * kScanTile;
  const int pc = g0 + tid * kScanPer;  // multiple of 16: aligned, inside the ld-sized arrays
  const uint4 zero = {0u, 0u, 0u, 0u};
  // Every load is issued before the row count is known (positions below max_rows are
  // inside the arrays), so the kernel waits for memory about once.
  const bool mine = pc < a.max_rows;
  uint4 v0 = zero, vu = zero, vb = zero;
  int4 id4[kScanPer / 4];
  if (mine) {
    v0 = *reinterpret_cast<const uint4*>(a.alive + pc);
    vu = *reinterpret_cast<const uint4*>(a.unsat + pc);
    if (a.band != nullptr) vb = *reinterpret_cast<const uint4*>(a.band + pc);
#pragma unroll
    for (int k = 0; k < kScanPer / 4; ++k) id4[k] = *reinterpret_cast<const int4*>(a.ids_in + pc + 4 * k);
  }
  // kept rows before this group, [0, g0): four 16-flag words per thread in flight
  int before = 0;
  for (int base = 0; base < g0; base += 4 * kScanTile) {
    uint4 w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int pos = base + u * kScanTile + tid * kScanPer;
      w[u] = pos < g0 ? kept_flags(a, pos) : zero;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) before += count_flags16(w[u]);
  }
  const int n = a.counts[a.layer];
  if (blockIdx.x == 0 && tid == 0 && a.work != nullptr) *a.work = 0;
  if (g0 >= n) return;  // uniform per group; positions [0, g0) are all < n
  // With keep_ppm every group also counts the kept rows after it, [g0 + tile, n): whether
  // the layout is kept depends on the total.
  int after = 0;
  if (a.keep_ppm > 0) {
    for (int base = g0 + kScanTile; base < n; base += 4 * kScanTile) {
      uint4 w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int pos = base + u * kScanTile + tid * kScanPer;
        w[u] = pos < n ? kept_flags(a, pos) : zero;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) after += count_flags16_lim(w[u], n - (base + u * kScanTile + tid * kScanPer));
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    before += __shfl_xor(before, d, 64);
    after += __shfl_xor(after, d, 64);
  }
  if (lane == 0) {
    wpre[wv] = before;
    wpost[wv] = after;
  }

  uint4 va = zero, vr = zero, vf = zero;
  int cnt = 0;
  if (pc < n) {
    // flags are 0/1 bytes.  retired = alive & !unsat; flagged (checkpoint) = alive & unsat
    if (a.retire_sat)
      vr = make_uint4(v0.x & ~vu.x & 0x01010101u, v0.y & ~vu.y & 0x01010101u, v0.z & ~vu.z & 0x01010101u,
                      v0.w & ~vu.w & 0x01010101u);
    vf = make_uint4(flagged_word(a, v0.x, vu.x), flagged_word(a, v0.y, vu.y), flagged_word(a, v0.z, vu.z),
                    flagged_word(a, v0.w, vu.w));
    if (a.flag_dead)
      vf = make_uint4(vf.x | (~v0.x & 0x01010101u), vf.y | (~v0.y & 0x01010101u), vf.z | (~v0.z & 0x01010101u),
                      vf.w | (~v0.w & 0x01010101u));
    va = make_uint4(kept_word(a, v0.x, vu.x), kept_word(a, v0.y, vu.y), kept_word(a, v0.z, vu.z), kept_word(a, v0.w, vu.w));
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
  int base = 0, wbase = 0, total = 0, post = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    base += wpre[i];
    post += wpost[i];
    wbase += i < wv ? wsum[i] : 0;
    total += wsum[i];
  }
  const int removed = n - (base + total + post);
  const bool keep = a.keep_ppm > 0 && static_cast<int64_t>(removed) * 1000000 <= static_cast<int64_t>(a.keep_ppm) * n;
  // the last group writes the row count of the next layer and the kept-row total
  if (tid == 0 && g0 + kScanTile >= n) {
    a.counts[a.layer + 1] = keep ? n : base + total;
    if (a.kept != nullptr) a.kept[a.layer + 1] = base + total;
  }
  if (pc >= n) return;
  int q = base + wbase + x - cnt;
  int qm[kScanPer];
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    const int p = pc + i;
    const int32_t id = (i & 3) == 0 ? id4[i >> 2].x : ((i & 3) == 1 ? id4[i >> 2].y : ((i & 3) == 2 ? id4[i >> 2].z : id4[i >> 2].w));
    const bool al = p < n && flag_at(va, i);
    qm[i] = -1;
    if (keep) {
      // kept layout: every position maps to itself, removed rows stay as holes
      if (p < n) {
        qm[i] = p;
        a.ids_out[p] = id;
      }
    } else if (al) {
      qm[i] = q;
      a.ids_out[q] = id;
      if (a.qsrc != nullptr) a.qsrc[q] = p;
      ++q;
    }
    if (p < n && flag_at(vr, i)) a.ret_ids[atomicAdd(a.ret_count, 1)] = id;
    const bool fl = p < n && (flag_at(vb, i) || flag_at(vf, i));
    if (fl && a.flag_mark[id] == 0) {
      a.flag_mark[id] = 1;
      a.flag_ids[atomicAdd(a.flag_count, 1)] = id;
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
// 32(L%2)..+31 with four 16-byte loads; the row ids come two steps ahead), and
// are written to the wave's own LDS slice.  The product is formed transposed,
// D'[sample][out] = X^T[sample][k] * W[k][out], on v_mfma_f32_16x16x32_f16
// (4 sample tiles x 2 output tiles): the X fragment of sample tile t holds the
// samples 16(i/4) + 4t + i%4 (i < 16), read with two ds_read_b64_tr_b16 per tile,
// so that the accumulators of lane j + 16g end up holding output j (and 16+j)
// for the 16 consecutive samples 16g..16g+15.
// Epilogue, when the 64 samples keep 64 consecutive positions starting at a
// multiple of 8 (every tile of a layer in which no row died): bias, clamp and
// FP16 conversion in registers, two 16-byte stores per output row and lane, and
// the liveness bits OR-reduced over the 16 lanes of a row group.  Otherwise the
// pre-activations are staged in LDS as FP32 [32][68] and lane L, owning sample L,
// stores its 32 outputs one by one.
// Gather mode (a layer after a compacting scan, with the output-to-input position map
// qsrc): a work item is 64 consecutive OUTPUT positions, so every tile's stores land on
// whole 128-byte lines (a compacted tile of the other mode straddles lines that another
// wave writes too, which costs about 30% of the store bandwidth).  Lane L gathers its
// output position's input column, 32 rows of 2 bytes (one load per row, all lanes on the
// same row), packs them into 16 dwords and writes them as sample L's 64-byte row of the
// LDS slice ([64 samples][32 rows], 16-byte chunks XOR-swizzled); the X fragments are
// then plain 16-byte LDS reads.
// --------------------------------------------------------------------------
constexpr int kWaveSamples = 64;
constexpr int kWavesPerGroup = 4;
constexpr int kStageLd = kWaveSamples + 4;  // FP32 staging row stride (floats): 16 rows cover all 64 banks
constexpr int kOutLd = kWaveSamples + 8;    // FP16 output staging row stride (halves, 144 B)

// leading dimensions of a layer's input and output buffers
__device__ __forceinline__ int64_t x_ld(const StepArgs& a) {
  return a.layer == 0 ? a.ld_in : act_ld(__builtin_amdgcn_readfirstlane(a.counts[a.layer]));
}
__device__ __forceinline__ int64_t y_ld(const StepArgs& a) {
  return act_ld(__builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]));
}

struct StepPos {
  int item, t, c, ch, ch_end;
  bool valid;
};

// Item order.  0 sample-tile-major: consecutive items are the block columns of one
// 64-sample tile.  1 block-column-major: consecutive items are consecutive tiles of one
// block column, so the waves of a group and neighbouring groups read the same gathered
// rows at adjacent addresses and write adjacent runs of the same output rows.
// 2 XCD-partitioned tile-major: the work-groups of XCD x (blockIdx % 8 under the
// round-robin dispatch) take the tiles t = x (mod 8), tile after tile, so the input rows of a
// tile stay in that XCD's L2 while every block column gathers from them (random layers
// re-read each input row about K times).  3 XCD-partitioned block-column-major: XCD x takes
// the block columns [x nblk / 8, (x + 1) nblk / 8) in block-column-major order, so the
// rows that neighbouring block columns of a banded layer share are gathered through the
// same L2.  4 XCD-partitioned tile-major: XCD x takes the block columns of order 3, tile after
// tile, so its share of the tile image stays in its L2 while the tiles stream through (random
// layers whose tile image is larger than one L2).  Orders 2 to 4 need a grid of at least 8
// work-groups.
struct ItemOrder {
  int mode, nblk, ntiles, part, nparts, c0;
};

__device__ __forceinline__ void step_first(StepPos& s, int item, int nitems, const ItemOrder& io, cint_ptr_t kptr,
                                           bool k1) {
  // wave-uniform by construction; readfirstlane keeps the position arithmetic scalar
  item = __builtin_amdgcn_readfirstlane(item);
  s.valid = item < nitems;
  s.item = item;
  // (c, t) computed into locals and stored once: stores to different fields on the two
  // branches get merged into one store with a selected address, which keeps s in memory
  int c, t;
  if (io.mode == 1 || io.mode == 3) {
    const int cc = item / io.ntiles;
    t = item - cc * io.ntiles;
    c = io.c0 + cc;
  } else {
    const int it = item / io.nblk;
    c = io.c0 + item - it * io.nblk;
    t = io.mode == 2 ? io.part + io.nparts * it : it;
  }
  s.c = c;
  s.t = t;
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

__device__ __forceinline__ void step_next(StepPos& s, int stride, int nitems, const ItemOrder& io, cint_ptr_t kptr,
                                          bool k1) {
  if (!s.valid) return;
  if (s.ch + 1 < s.ch_end) {
    ++s.ch;
    return;
  }
  step_first(s, s.item + stride, nitems, io, kptr, k1);
}

// OR over the 16 lanes of each DPP row (row_ror by 1, 2, 4, 8).
__device__ __forceinline__ uint32_t row_or16(uint32_t x) {
  x |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x121, 0xF, 0xF, false));
  x |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x122, 0xF, 0xF, false));
  x |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x124, 0xF, 0xF, false));
  x |= static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x128, 0xF, 0xF, false));
  return x;
}

// Per-sample flag bits of the epilogue's fast path.  acc[t][m][r] is output 16m + j of
// sample 16g + 4t + r; bit 4t+r of the result (after the OR over the 16 lanes of the
// row group) says whether any of the block's outputs of that sample is > 0 (kind 0),
// < 32 (kind 1) or within `band` of 0 or 32 (kind 2).  ok0/ok1: the lane's outputs exist.
template <int kKind, bool kFull>
__device__ __forceinline__ uint32_t flag_bits(const float4_t (&acc)[4][2], bool ok0, bool ok1, float band) {
  uint32_t bits = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v0 = acc[t][0][r], v1 = acc[t][1][r];
      bool hit;
      if (kKind == 0) {
        hit = kFull ? fmaxf(v0, v1) > 0.0f : ((ok0 && v0 > 0.0f) || (ok1 && v1 > 0.0f));
      } else if (kKind == 1) {
        hit = kFull ? fminf(v0, v1) < 32.0f : ((ok0 && v0 < 32.0f) || (ok1 && v1 < 32.0f));
      } else {
        hit = (ok0 && (fabsf(v0) < band || fabsf(v0 - 32.0f) < band)) ||
              (ok1 && (fabsf(v1) < band || fabsf(v1 - 32.0f) < band));
      }
      bits |= hit ? (1u << (4 * t + r)) : 0u;
    }
  return row_or16(bits);
}


typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Streaming 16-byte global accesses (activations are touched once per layer).
__device__ __forceinline__ uint4 load_stream(const _Float16* p) {
  const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(_Float16* p, uint4 v) {
  __builtin_nontemporal_store(u32x4_t{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4_t*>(p));
}

// 16-byte chunk k (rows 8k..8k+7) of sample s in the gather-mode LDS slice: swizzled so that
// the 8 lanes of each LDS cycle hit different banks, both when lane s writes its sample and
// when the X-fragment reads take samples 16(i/4) + 4t + i%4
__device__ __forceinline__ int gather_chunk(int s, int k) { return s * 4 + (k ^ (((s >> 2) ^ (s >> 4)) & 3)); }

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// The 32 input rows of chunk ch (row ids in lanes 0..31 of grow) at input position src:
// dword k of d holds rows 2k (low half) and 2k+1 (high half).  Each load has every lane on
// the same row (base address scalar) at nearby positions.
__device__ __forceinline__ void gather_rows(uint32_t (&d)[16], const _Float16* __restrict__ X, int64_t ldx, int grow,
                                            int src) {
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r0 = __builtin_amdgcn_readlane(grow, 2 * k);
    const int r1 = __builtin_amdgcn_readlane(grow, 2 * k + 1);
    half2_t v;
    v.x = X[static_cast<int64_t>(r0) * ldx + static_cast<uint32_t>(src)];
    v.y = X[static_cast<int64_t>(r1) * ldx + static_cast<uint32_t>(src)];
    d[k] = __builtin_bit_cast(uint32_t, v);
  }
}

// 32 input bits (samples 32h..32h+31 of one row, bit i = sample 32h+i) as 32 FP16 0/1
// values: dword k holds samples 2k (low half) and 2k+1.
__device__ __forceinline__ void expand_bits(uint32_t wbits, uint4 (&o)[4]) {
  uint32_t d[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t b = (wbits >> (2 * k)) & 3u;
    d[k] = (b & 1u) * 0x3C00u | (b >> 1) * 0x3C000000u;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = make_uint4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
}

// The chunk's two W fragments from its mask-image word (the lane's 16 bits, see format.hpp):
// dword k of fragment m holds elements 2k (low half) and 2k+1 (high half), each the FP16 value
// v where its bit is set and 0 elsewhere.  The odd elements' bits move up by 8 so that bit
// q = 4m + k holds element 2k and bit 16 + q element 2k + 1; then dword q is
// (sp & (0x10001 << q)) * (v >> q): one AND and one 24-bit multiply, exact because a mask
// image exists only for values v with their low 7 bits zero (v >> q * 2^q = v for q < 8), and
// x * (v >> q) with x = (lo + hi * 2^16) 2^q cannot carry between the halves (v < 2^16).
__device__ __forceinline__ void expand_mask(uint32_t bits, uint32_t v, half8_t& w0, half8_t& w1) {
  const uint32_t sp = (bits & 0xFFu) | ((bits & 0xFF00u) << 8);
  uint32_t d[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) d[q] = __umul24(sp & (0x10001u << q), v >> q);
  w0 = __builtin_bit_cast(half8_t, u32x4_t{d[0], d[1], d[2], d[3]});
  w1 = __builtin_bit_cast(half8_t, u32x4_t{d[4], d[5], d[6], d[7]});
}

// W fragments of chunk ch, as loaded (MASK: the lane's mask word in wm; otherwise w0, w1)
template <bool MASK>
__device__ __forceinline__ void load_w(const DevLayer& w, int64_t ch, int lane, half8_t& w0, half8_t& w1,
                                       uint32_t& wm) {
  if constexpr (MASK) {
    wm = w.wmask[ch * 64 + lane];
  } else {
    w0 = *reinterpret_cast<const half8_t*>(w.wt + (ch * 2 + 0) * 512 + lane * 8);
    w1 = *reinterpret_cast<const half8_t*>(w.wt + (ch * 2 + 1) * 512 + lane * 8);
  }
}

// GATHER: tiles of output positions after a compacting scan.  BITS: layer 0 of an identity
// layout with a binary input, read from its bit-packed neuron-major copy: a step's 32 rows x
// 64 samples are 256 bytes instead of 4 KB, expanded to FP16 in registers.
// MASK: the W fragments come from the layer's mask image (expanded in registers).
template <bool GATHER, bool BITS, bool MASK>
__device__ __forceinline__ void mfma_body(const StepArgs& a, const DevLayer& w, float (*wlds)[32 * kStageLd], int n,
                                          int n_out) {
  // ident: no row left the pass in the scan before this layer, every position maps to itself
  const bool ident = n_out == n;
  const int nblk = w.nblk;
  // tiles of input positions, or of output positions in gather mode
  const int ntiles = ((GATHER ? n_out : n) + kWaveSamples - 1) / kWaveSamples;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  ItemOrder io{a.order, nblk, ntiles, 0, 1, 0};
  int nitems = ntiles * nblk, first = blockIdx.x * kWavesPerGroup + wave, stride = gridDim.x * kWavesPerGroup;
  if (a.order == 2) {
    // this XCD's groups and tiles
    io.nparts = kXcds;
    io.part = blockIdx.x % kXcds;
    const int groups = (static_cast<int>(gridDim.x) - io.part + kXcds - 1) / kXcds;
    nitems = (ntiles - io.part + kXcds - 1) / kXcds * nblk;
    first = (blockIdx.x / kXcds) * kWavesPerGroup + wave;
    stride = groups * kWavesPerGroup;
  } else if (a.order == 3) {
    // this XCD's block columns
    io.nparts = kXcds;
    io.part = blockIdx.x % kXcds;
    const int groups = (static_cast<int>(gridDim.x) - io.part + kXcds - 1) / kXcds;
    io.c0 = io.part * nblk / kXcds;
    nitems = ((io.part + 1) * nblk / kXcds - io.c0) * ntiles;
    first = (blockIdx.x / kXcds) * kWavesPerGroup + wave;
    stride = groups * kWavesPerGroup;
  } else if (a.order == 4) {
    // this XCD's block columns, tile-major: io.nblk becomes the XCD's share
    io.nparts = kXcds;
    io.part = blockIdx.x % kXcds;
    const int groups = (static_cast<int>(gridDim.x) - io.part + kXcds - 1) / kXcds;
    io.c0 = io.part * nblk / kXcds;
    io.nblk = (io.part + 1) * nblk / kXcds - io.c0;
    if (io.nblk == 0) return;  // fewer block columns than XCDs (wave-uniform)
    nitems = io.nblk * ntiles;
    first = (blockIdx.x / kXcds) * kWavesPerGroup + wave;
    stride = groups * kWavesPerGroup;
  }
  const bool k1 = w.one_chunk != 0;
  const cint_ptr_t kptr = (cint_ptr_t)w.kptr;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const int64_t ldx = x_ld(a), ldy = y_ld(a);

  StepPos cur, nx1, nx2;
  step_first(cur, first, nitems, io, kptr, k1);
  if (!cur.valid) return;  // wave-uniform; nothing below synchronises across waves
  nx1 = cur;
  step_next(nx1, stride, nitems, io, kptr, k1);
  nx2 = nx1;
  step_next(nx2, stride, nitems, io, kptr, k1);

  _Float16* tile = reinterpret_cast<_Float16*>(wlds[wave]);  // [32][64] FP16 input tile (bytes 0..4095)
  uint4* gtile = reinterpret_cast<uint4*>(wlds[wave]);        // gather mode: [64][4] 16-byte chunks (bytes 0..4095)
  _Float16* ostage = tile + 32 * kWaveSamples;               // [32][kOutLd] FP16 outputs (bytes 4096..8703)
  float* stage = wlds[wave];                                 // [32][kStageLd] FP32 pre-activations (slow path)
  // global <-> registers: access k (0..3) of a lane covers row 8k + r8, samples c8..c8+7
  // (16 B), so each load / store instruction moves 8 whole 128-byte row segments
  const int r8 = lane >> 3, c8 = (lane & 7) * 8;
  const int j = lane & 15, g = lane >> 4;
  const int32_t* __restrict__ gidx = w.gidx;

  // Everything a step consumes is loaded one step ahead: its tile rows (d0..d3), its W
  // fragments, the target position of this lane's sample, the output ids; the row ids
  // two steps ahead.  After the final step the prefetch re-reads valid (unused) addresses.
  int rn0, rn1, rn2, rn3;  // row ids of step nx1
  uint4 d0, d1, d2, d3;
  half8_t w0, w1;
  uint32_t wm = 0;  // MASK: the lane's mask word of the next step's chunk
  const uint32_t wval = w.wmask_val;
  int qv;
  // gather mode: this step's 32 rows of the lane's sample (gd), the row ids of step nx1 (in
  // lanes 0..31 of grow1) and the input position of the lane's sample in nx1's tile
  uint32_t gd[16];
  int grow1 = 0, gsrc1 = 0;
  const int32_t* __restrict__ qsrc = a.qsrc;
  // bits mode: lane L covers row L/2 of the chunk, samples 32(L%2)..+31 (one word); the word
  // of this step and the row id of step nx1
  uint32_t bw = 0;
  int brow1 = 0;
  const int bh = lane & 1;
  if constexpr (BITS) {
    brow1 = gidx[(nx1.valid ? nx1.ch : cur.ch) * kChunk + (lane >> 1)];
    const int brow = gidx[cur.ch * kChunk + (lane >> 1)];
    bw = a.xbits[(static_cast<int64_t>(cur.t) * a.xbits_rows + brow) * 2 + bh];
  } else if constexpr (GATHER) {
    const int pch = nx1.valid ? nx1.ch : cur.ch;
    const int pt = nx1.valid ? nx1.t : cur.t;
    grow1 = gidx[pch * kChunk + (lane & 31)];
    gsrc1 = qsrc[min(pt * kWaveSamples + lane, n_out - 1)];
    const int grow = gidx[cur.ch * kChunk + (lane & 31)];
    const int src = qsrc[min(cur.t * kWaveSamples + lane, n_out - 1)];
    gather_rows(gd, X, ldx, grow, src);
  } else {
    {
      const int ch = (nx1.valid ? nx1.ch : cur.ch) * kChunk + r8;
      rn0 = gidx[ch];
      rn1 = gidx[ch + 8];
      rn2 = gidx[ch + 16];
      rn3 = gidx[ch + 24];
    }
    const int ch = cur.ch * kChunk + r8;
    const _Float16* base = X + cur.t * kWaveSamples + c8;
    d0 = load_stream(base + static_cast<int64_t>(gidx[ch]) * ldx);
    d1 = load_stream(base + static_cast<int64_t>(gidx[ch + 8]) * ldx);
    d2 = load_stream(base + static_cast<int64_t>(gidx[ch + 16]) * ldx);
    d3 = load_stream(base + static_cast<int64_t>(gidx[ch + 24]) * ldx);
    qv = ident ? cur.t * kWaveSamples + lane : a.qmap[min(cur.t * kWaveSamples + lane, n - 1)];
  }
  load_w<MASK>(w, cur.ch, lane, w0, w1, wm);
  // accumulators start at the bias: D' + b comes out of the MFMA chain
  const float4_t bias4 = float4_t{a.bias, a.bias, a.bias, a.bias};
  const bool full = w.full_blocks != 0;
  const bool want_unsat = a.unsat_next != nullptr;
  const bool want_band = a.band_next != nullptr;
  float4_t acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int m = 0; m < 2; ++m) acc[t][m] = bias4;

  // The stores of a fast-path epilogue wait in the wave's output staging area (LDS) and
  // go out at the start of the next step, next to that step's prefetch: a VMEM wait with
  // loads and stores pending waits for everything, so stores issued at the end of a step
  // would put a whole store latency in front of every step.  Pending: the tile's first
  // sample position (wave-uniform; lane i's sample is that + i), its block column (the
  // output rows follow from it) and the lane's flag bits.
  bool pend = false;
  int pq0 = 0, pcb = 0;
  uint32_t pfl = 0;
  auto flush = [&]() {
    if (!pend) return;
    const _Float16* orow = ostage + r8 * kOutLd;
    const uint4 v0 = *reinterpret_cast<const uint4*>(orow + c8);
    const uint4 v1 = *reinterpret_cast<const uint4*>(orow + 8 * kOutLd + c8);
    const uint4 v2 = *reinterpret_cast<const uint4*>(orow + 16 * kOutLd + c8);
    const uint4 v3 = *reinterpret_cast<const uint4*>(orow + 24 * kOutLd + c8);
    const int s0 = pcb * 32 + r8;  // output slots of the lane's four rows: s0, +8, +16, +24
    const int po0 = s0 < w.n_out ? s0 : -1;
    const int po1 = s0 + 8 < w.n_out ? s0 + 8 : -1;
    const int po2 = s0 + 16 < w.n_out ? s0 + 16 : -1;
    const int po3 = s0 + 24 < w.n_out ? s0 + 24 : -1;
    _Float16* ybase = Y + pq0 + c8;
    if (full || po0 >= 0) store_stream(ybase + static_cast<int64_t>(po0) * ldy, v0);
    if (full || po1 >= 0) store_stream(ybase + static_cast<int64_t>(po1) * ldy, v1);
    if (full || po2 >= 0) store_stream(ybase + static_cast<int64_t>(po2) * ldy, v2);
    if (full || po3 >= 0) store_stream(ybase + static_cast<int64_t>(po3) * ldy, v3);
    if (pfl & 1u) a.alive_next[pq0 + lane] = 1;
    if (pfl & 2u) a.unsat_next[pq0 + lane] = 1;
    if (pfl & 4u) a.band_next[pq0 + lane] = 1;
    pend = false;
  };

  while (true) {
    const bool last = cur.ch + 1 == cur.ch_end;
    const int p_own = cur.t * kWaveSamples + lane;
    // the output position of the lane's sample (-1: none)
    const int q_own = GATHER ? (p_own < n_out ? p_own : -1) : (p_own < n ? (BITS ? p_own : qv) : -1);
    half8_t cw0 = w0, cw1 = w1;
    const uint32_t cwm = wm;
    // output rows j and 16 + j of the block column (slot = row, -1 for padding)
    const int cov0 = cur.c * 32 + j < w.n_out ? cur.c * 32 + j : -1;
    const int cov1 = cur.c * 32 + 16 + j < w.n_out ? cur.c * 32 + 16 + j : -1;
    const int cur_c = cur.c;
    const int pt = nx1.valid ? nx1.t : cur.t;
    const int pch = nx1.valid ? nx1.ch : cur.ch;
    if constexpr (BITS) {
      // this step's bits as FP16 into the tile (row L/2, samples 32(L%2)..+31), then the
      // next step's word
      uint4 o[4];
      expand_bits(bw, o);
      uint4* dst = reinterpret_cast<uint4*>(tile + (lane >> 1) * kWaveSamples + 32 * bh);
      dst[0] = o[0];
      dst[1] = o[1];
      dst[2] = o[2];
      dst[3] = o[3];
      bw = a.xbits[(static_cast<int64_t>(pt) * a.xbits_rows + brow1) * 2 + bh];
    } else if constexpr (GATHER) {
      // this step's rows of the lane's sample into the wave's LDS slice, then the next
      // step's gathers (row ids and input positions loaded a step earlier)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        gtile[gather_chunk(lane, k)] = make_uint4(gd[4 * k], gd[4 * k + 1], gd[4 * k + 2], gd[4 * k + 3]);
      gather_rows(gd, X, ldx, grow1, gsrc1);
    } else {
      // this step's tile into the wave's LDS slice (waits for the loads issued last step)
      {
        uint4* dst = reinterpret_cast<uint4*>(tile + r8 * kWaveSamples + c8);
        dst[0 * 8 * kWaveSamples / 8] = d0;
        dst[1 * 8 * kWaveSamples / 8] = d1;
        dst[2 * 8 * kWaveSamples / 8] = d2;
        dst[3 * 8 * kWaveSamples / 8] = d3;
      }
      // Next step's tile rows, last step's stores, then the rest of the next step's loads.
      // Every global access of a step is issued here, at its start, so that the wait at the
      // top of the next step finds them done; the tile loads go first as they need this
      // step's row ids (no load may be waited for after a store: the wait would cover it).
      const _Float16* base = X + pt * kWaveSamples + c8;
      d0 = load_stream(base + static_cast<int64_t>(rn0) * ldx);
      d1 = load_stream(base + static_cast<int64_t>(rn1) * ldx);
      d2 = load_stream(base + static_cast<int64_t>(rn2) * ldx);
      d3 = load_stream(base + static_cast<int64_t>(rn3) * ldx);
    }
    __builtin_amdgcn_sched_barrier(0);
    flush();
    __builtin_amdgcn_sched_barrier(0);
    {
      // W fragments of the next step (gather mode: after this step's MFMAs, so that the
      // gathered rows in flight and the W registers in use do not both need copies)
      if constexpr (!GATHER) load_w<MASK>(w, pch, lane, w0, w1, wm);
      if constexpr (BITS) {
        brow1 = gidx[(nx2.valid ? nx2.ch : pch) * kChunk + (lane >> 1)];
      } else if constexpr (GATHER) {
        // row ids and input positions of step nx2
        grow1 = gidx[(nx2.valid ? nx2.ch : pch) * kChunk + (lane & 31)];
        gsrc1 = qsrc[min((nx2.valid ? nx2.t : pt) * kWaveSamples + lane, n_out - 1)];
      } else {
        qv = ident ? pt * kWaveSamples + lane : a.qmap[min(pt * kWaveSamples + lane, n - 1)];
        const int ch2 = (nx2.valid ? nx2.ch : pch) * kChunk + r8;
        rn0 = gidx[ch2];
        rn1 = gidx[ch2 + 8];
        rn2 = gidx[ch2 + 16];
        rn3 = gidx[ch2 + 24];
      }
    }
    // X fragments.  Lane i of each 16-lane group g needs rows 8g..8g+7 of sample
    // 16(i/4)+4t+i%4.  Gather mode: one 16-byte read of that sample's chunk g.  Otherwise
    // ds_read_b64_tr_b16, which gives lane i element i%4 of the 8 bytes read by lanes
    // 4k+i/4 (k < 4): lane m reads row 8g+4h+m/4 at samples 16(m%4)+4t..+3.
    half8_t xfrag[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if constexpr (GATHER) {
        const int smp = 16 * (j >> 2) + 4 * t + (j & 3);
        xfrag[t] = __builtin_bit_cast(half8_t, gtile[gather_chunk(smp, g)]);
      } else {
        const _Float16* p0 = tile + (g * 8 + (j >> 2)) * kWaveSamples + 16 * (j & 3) + 4 * t;
        const fp16x4_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)p0);
        const fp16x4_t hi = __builtin_amdgcn_ds_read_tr16_b64_v4f16(
            (__attribute__((address_space(3))) fp16x4_t*)(p0 + 4 * kWaveSamples));
        xfrag[t] = __builtin_shufflevector(__builtin_bit_cast(half4_t, lo), __builtin_bit_cast(half4_t, hi), 0, 1, 2,
                                           3, 4, 5, 6, 7);
      }
    }
    if constexpr (MASK) expand_mask(cwm, wval, cw0, cw1);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xfrag[t], cw0, acc[t][0], 0, 0, 0);
      acc[t][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xfrag[t], cw1, acc[t][1], 0, 0, 0);
    }
    if constexpr (GATHER) load_w<MASK>(w, pch, lane, w0, w1, wm);
    if (last) {
      // D'[i][jj] of a 16x16 tile sits in lane jj + 16*(i/4), element i%4: acc[t][m][r] of
      // lane j + 16g is output 16m + j of sample 16g + 4t + r.
      const int q0 = __builtin_amdgcn_readfirstlane(q_own);
      // Gather mode: the tile's outputs are the 64 positions from 64t, always aligned; in the
      // last tile the positions past the row count get stores (inside the buffer's leading
      // dimension, used by no later layer) but no flags.
      const bool fast =
          GATHER || __builtin_amdgcn_ballot_w64(!(q0 >= 0 && (q0 & 7) == 0 && q_own == q0 + lane)) == 0;
      if (fast) {
        // FP16 outputs clamped with med3, transposed through LDS ([32][kOutLd]) so that the
        // stores run 8 lanes x 16 B along each output row
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          uint32_t pk[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int t = e / 2, r = 2 * (e % 2);
            const _Float16 h0 = static_cast<_Float16>(__builtin_amdgcn_fmed3f(acc[t][m][r], 0.0f, 32.0f));
            const _Float16 h1 = static_cast<_Float16>(__builtin_amdgcn_fmed3f(acc[t][m][r + 1], 0.0f, 32.0f));
            pk[e] = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h0)) |
                    (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h1)) << 16);
          }
          uint4* o = reinterpret_cast<uint4*>(ostage + (16 * m + j) * kOutLd + 16 * g);
          o[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
          o[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        }
        // liveness bits of the lane's 16 samples (bit 4t+r), OR-reduced over the 16 lanes of
        // the row group: lane j + 16g owns sample 16g + j
        const bool ok0 = full || cov0 >= 0, ok1 = full || cov1 >= 0;
        const uint32_t ba = full ? flag_bits<0, true>(acc, true, true, 0.0f) : flag_bits<0, false>(acc, ok0, ok1, 0.0f);
        uint32_t fl = (ba >> j) & 1u;
        if (want_unsat) fl |= ((flag_bits<1, false>(acc, ok0, ok1, 0.0f) >> j) & 1u) << 1;
        if (want_band) fl |= ((flag_bits<2, false>(acc, ok0, ok1, a.band) >> j) & 1u) << 2;
        if (GATHER && q_own < 0) fl = 0;
        pend = true;
        pq0 = q0;
        pcb = cur_c;
        pfl = fl;
      } else {
        // rows died in this tile: stage the pre-activations, then each lane stores the 32
        // outputs of its own sample
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            *reinterpret_cast<float4_t*>(stage + (16 * m + j) * kStageLd + 16 * g + 4 * t) = acc[t][m];
          }
        if (q_own >= 0) {
          uint32_t fl = 0;
          const int32_t cbase = cur_c * 32;
#pragma unroll 8
          for (int o = 0; o < 32; ++o) {
            const int out = cbase + o < w.n_out ? cbase + o : -1;  // slot = row
            const float v = stage[o * kStageLd + lane];
            if (out >= 0) {
              fl |= v > 0.0f ? 1u : 0u;
              fl |= v < 32.0f ? 2u : 0u;
              fl |= (fabsf(v) < a.band || fabsf(v - 32.0f) < a.band) ? 4u : 0u;
              Y[static_cast<int64_t>(out) * ldy + q_own] = static_cast<_Float16>(__builtin_amdgcn_fmed3f(v, 0.0f, 32.0f));
            }
          }
          if (fl & 1u) a.alive_next[q_own] = 1;
          if (want_unsat && (fl & 2u)) a.unsat_next[q_own] = 1;
          if (want_band && (fl & 4u)) a.band_next[q_own] = 1;
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < 2; ++m) acc[t][m] = bias4;
    }
    if (!nx1.valid) break;
    cur = nx1;
    nx1 = nx2;
    step_next(nx2, stride, nitems, io, kptr, k1);
  }
  flush();
}

__global__ __launch_bounds__(256, 4) void spmm_mfma_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) float wlds[kWavesPerGroup][32 * kStageLd];  // 8.5 KB per wave
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  if (n_out != n && a.qsrc != nullptr)
    mfma_body<true, false, false>(a, w, wlds, n, n_out);
  else if (n_out == n && a.xbits != nullptr)
    mfma_body<false, true, false>(a, w, wlds, n, n_out);
  else
    mfma_body<false, false, false>(a, w, wlds, n, n_out);
}

// The same kernel on a layer's mask image (the launcher picks it when the layer has one).
__global__ __launch_bounds__(256, 4) void spmm_mfma_mask_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) float wlds[kWavesPerGroup][32 * kStageLd];
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  if (n_out != n && a.qsrc != nullptr)
    mfma_body<true, false, true>(a, w, wlds, n, n_out);
  else if (n_out == n && a.xbits != nullptr)
    mfma_body<false, true, true>(a, w, wlds, n, n_out);
  else
    mfma_body<false, false, true>(a, w, wlds, n, n_out);
}

// --------------------------------------------------------------------------
// LDS-resident MFMA SpMM (plan order 5) for layers with at most kLdsRows input rows (N <= 1024):
// a work-group holds a whole 64-sample tile of the input, every row, in LDS (128 KB), loaded
// once with 16-byte loads, and its 8 waves run all block columns of that tile from it: the X
// fragments are ds_read_b64_tr_b16 at the gathered rows' LDS addresses, so a random layer's
// 20-32-fold re-reads of each input row are LDS reads instead of L2 gathers.  Same operand
// layout, MFMA chain (bias-initialised, K chunks in order) and epilogue arithmetic as the
// plain mode of spmm_mfma_kernel; tiles of input positions, stores compacted through qmap.
// Epilogue: the fast path (64 consecutive aligned positions: layers where no row died) goes
// through a per-wave FP16 staging area as whole 128-byte row segments; otherwise each lane
// stores its 32 values per sample to that sample's position.
// --------------------------------------------------------------------------
constexpr int kLdsRows = kLdsRowsMax;
constexpr int kLdsWaves = 8;  // 2 per SIMD; one work-group per CU (the tile takes up to 128 KB)
constexpr int kLdsAhead = 4;  // K chunks whose W fragments and rows are in flight per wave
// LDS row stride of the tile (halves): 136 bytes, so that the gathered rows of one fragment
// read, whose starts are otherwise all on the same bank, spread over the banks
constexpr int kLdsLd = kWaveSamples + 4;

// W fragments (MASK: the lane's mask word) and fragment-read rows of one K chunk
struct LdsChunk {
  half8_t w0, w1;
  uint32_t wm;
  int rlo, rhi;
};
template <bool MASK>
__device__ __forceinline__ void lds_chunk_load(LdsChunk& k, const DevLayer& w, int ch, int lane, int j, int g) {
  k.rlo = w.gidx[ch * kChunk + g * 8 + (j >> 2)];
  k.rhi = w.gidx[ch * kChunk + g * 8 + 4 + (j >> 2)];
  load_w<MASK>(w, ch, lane, k.w0, k.w1, k.wm);
}
// the chunk's X fragments from the LDS tile, then its 8 MFMAs
template <bool MASK>
__device__ __forceinline__ void lds_chunk_mfma(float4_t (&acc)[4][2], const LdsChunk& k, const _Float16* lds_x, int j,
                                               uint32_t wval) {
  half8_t w0 = k.w0, w1 = k.w1;
  if constexpr (MASK) expand_mask(k.wm, wval, w0, w1);
  half8_t xfrag[4];
  // 32-bit LDS offsets (row ids < 2^16, so one 24-bit multiply each)
  typedef __attribute__((address_space(3))) const _Float16 lds_half_t;
  lds_half_t* xl = (lds_half_t*)lds_x;
  const uint32_t olo = __umul24(static_cast<uint32_t>(k.rlo), kLdsLd) + 16 * (j & 3);
  const uint32_t ohi = __umul24(static_cast<uint32_t>(k.rhi), kLdsLd) + 16 * (j & 3);
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    lds_half_t* plo = xl + olo + 4 * tt;
    lds_half_t* phi = xl + ohi + 4 * tt;
    const fp16x4_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)plo);
    const fp16x4_t hi = __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)phi);
    xfrag[tt] = __builtin_shufflevector(__builtin_bit_cast(half4_t, lo), __builtin_bit_cast(half4_t, hi), 0, 1, 2, 3,
                                        4, 5, 6, 7);
  }
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    acc[tt][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xfrag[tt], w0, acc[tt][0], 0, 0, 0);
    acc[tt][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xfrag[tt], w1, acc[tt][1], 0, 0, 0);
  }
}

// Epilogue of the LDS-staged MFMA kernels for block column c of a 64-sample tile:
// acc[tt][m][r] of lane j + 16g is output 16m + j of sample 16g + 4tt + r.  Bias is in acc.
// Outputs go through the wave's staging area ostage ([16][kOutLd]).  fast: the tile's 64
// samples go to 64 consecutive positions from q0, a multiple of 8 (8 lanes x 16 B per output
// row); otherwise one store per output row, each lane's value to its sample's position (q_own,
// -1: none).  Then the liveness / unsaturated / guard-band flags of the lane's sample.
__device__ __forceinline__ void lds_epilogue(const StepArgs& a, const DevLayer& w, const float4_t (&acc)[4][2], int c,
                                             int q_own, int q0, bool fast, _Float16* ostage,
                                             _Float16* __restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, g = lane >> 4;
  const int r8 = lane >> 3, c8 = (lane & 7) * 8;
  const bool full = w.full_blocks != 0;
  const bool want_unsat = a.unsat_next != nullptr;
  const bool want_band = a.band_next != nullptr;
  // epilogue: acc[tt][m][r] of lane j + 16g is output 16m + j of sample 16g + 4tt + r
  const int cov0 = c * 32 + j < w.n_out ? c * 32 + j : -1;
  const int cov1 = c * 32 + 16 + j < w.n_out ? c * 32 + 16 + j : -1;
  const bool ok0 = full || cov0 >= 0, ok1 = full || cov1 >= 0;
  const uint32_t ba = full ? flag_bits<0, true>(acc, true, true, 0.0f) : flag_bits<0, false>(acc, ok0, ok1, 0.0f);
  uint32_t fl = (ba >> j) & 1u;
  if (want_unsat) fl |= ((flag_bits<1, false>(acc, ok0, ok1, 0.0f) >> j) & 1u) << 1;
  if (want_band) fl |= ((flag_bits<2, false>(acc, ok0, ok1, a.band) >> j) & 1u) << 2;
  // FP16 outputs through the wave's staging area ([16][kOutLd]: row = output 16m + j of the
  // block column, column = the tile's sample), 16 rows at a time
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    uint32_t pk[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int tt = e / 2, r = 2 * (e % 2);
      const _Float16 h0 = static_cast<_Float16>(__builtin_amdgcn_fmed3f(acc[tt][m][r], 0.0f, 32.0f));
      const _Float16 h1 = static_cast<_Float16>(__builtin_amdgcn_fmed3f(acc[tt][m][r + 1], 0.0f, 32.0f));
      pk[e] = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h0)) |
              (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h1)) << 16);
    }
    uint4* o = reinterpret_cast<uint4*>(ostage + j * kOutLd + 16 * g);
    o[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    o[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (fast) {
      // 64 consecutive positions from q0: 8 lanes x 16 B per output row
      _Float16* ybase = Y + q0 + c8;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int slot = c * 32 + 16 * m + r8 + 8 * k;
        const uint4 v = *reinterpret_cast<const uint4*>(ostage + (r8 + 8 * k) * kOutLd + c8);
        if (full || slot < w.n_out) store_stream(ybase + static_cast<int64_t>(slot) * ldy, v);
      }
    } else if (q_own >= 0) {
      // compacted positions: one output row per store, lane L its sample's value at q_own
      // (positions increase with the lane, so a store covers one or two 128-byte lines)
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) {
        const int slot = c * 32 + 16 * m + rr;
        if (full || slot < w.n_out) Y[static_cast<int64_t>(slot) * ldy + q_own] = ostage[rr * kOutLd + lane];
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  if (q_own >= 0) {
    if (fl & 1u) a.alive_next[q_own] = 1;
    if (fl & 2u) a.unsat_next[q_own] = 1;
    if (fl & 4u) a.band_next[q_own] = 1;
  }
}

template <bool MASK>
__device__ __forceinline__ void mfma_lds_body(const StepArgs& a, const DevLayer& w, _Float16* lds_x);

__global__ __launch_bounds__(kLdsWaves * 64, 1) void spmm_mfma_lds_kernel(StepArgs a, DevLayer w) {
  // [x_rows][kLdsLd] input tile, then per wave [16][kOutLd] FP16 output staging
  extern __shared__ __attribute__((aligned(16))) _Float16 lds_x[];
  mfma_lds_body<false>(a, w, lds_x);
}

// The same kernel on a layer's mask image.
__global__ __launch_bounds__(kLdsWaves * 64, 1) void spmm_mfma_lds_mask_kernel(StepArgs a, DevLayer w) {
  extern __shared__ __attribute__((aligned(16))) _Float16 lds_x[];
  mfma_lds_body<true>(a, w, lds_x);
}

template <bool MASK>
__device__ __forceinline__ void mfma_lds_body(const StepArgs& a, const DevLayer& w, _Float16* lds_x) {
  constexpr int kThreads = kLdsWaves * 64;
  const uint32_t wval = w.wmask_val;
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  const bool ident = n_out == n;
  const int x_rows = a.x_rows;  // <= kLdsRows, checked by the launcher
  const int ntiles = (n + kWaveSamples - 1) / kWaveSamples;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int j = lane & 15, g = lane >> 4;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
  const int nblk = w.nblk;
  const float4_t bias4 = float4_t{a.bias, a.bias, a.bias, a.bias};
  const cint_ptr_t kptr = (cint_ptr_t)w.kptr;
  // staging areas from a 16-byte boundary (x_rows may be odd)
  _Float16* ostage = lds_x + (static_cast<int64_t>(x_rows) * kLdsLd + 7) / 8 * 8 + wave * 16 * kOutLd;

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // the tile: x_rows rows x 64 samples; thread tid moves the 16-byte pieces tid, tid + 512,
    // ... (piece i: row i / 8, samples 8 (i % 8)..+7); columns past n lie inside the buffer's
    // leading dimension (a multiple of 64 at least as large as n)
    __syncthreads();  // the previous tile's fragment reads are done
    {
      const _Float16* src = X + static_cast<int64_t>(t) * kWaveSamples;
      const int pieces = x_rows * 8;
      for (int i0 = 0; i0 < pieces; i0 += kThreads * 8) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = i0 + u * kThreads + tid;
          if (i < pieces) v[u] = load_stream(src + static_cast<int64_t>(i >> 3) * ldx + (i & 7) * 8);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = i0 + u * kThreads + tid;
          if (i < pieces) {
            // two 8-byte writes: a row starts on an 8-byte boundary
            uint2* d = reinterpret_cast<uint2*>(lds_x + (i >> 3) * kLdsLd + (i & 7) * 8);
            d[0] = make_uint2(v[u].x, v[u].y);
            d[1] = make_uint2(v[u].z, v[u].w);
          }
        }
      }
    }
    __syncthreads();
    // this lane's sample (lane L owns sample 64t + L) and its output position
    const int p_own = t * kWaveSamples + lane;
    const int q_own = p_own < n ? (ident ? p_own : a.qmap[p_own]) : -1;
    const int q0 = __builtin_amdgcn_readfirstlane(q_own);
    const bool fast = __builtin_amdgcn_ballot_w64(!(q0 >= 0 && (q0 & 7) == 0 && q_own == q0 + lane)) == 0;
    for (int c = wave; c < nblk; c += kLdsWaves) {
      float4_t acc[4][2];
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
#pragma unroll
        for (int m = 0; m < 2; ++m) acc[tt][m] = bias4;
      const int ch0 = kptr[c], ch1 = kptr[c + 1];
      // K chunks in order, kLdsAhead register sets loaded that many chunks ahead (the W
      // fragments come from L2; no register that a load in flight writes is copied); they
      // are loaded only when the block column has any chunk (all-empty inputs: acc = bias)
      LdsChunk kc[kLdsAhead];
      if (ch1 > ch0) {
#pragma unroll
        for (int u = 0; u < kLdsAhead; ++u) lds_chunk_load<MASK>(kc[u], w, min(ch0 + u, ch1 - 1), lane, j, g);
      }
      for (int ch = ch0; ch < ch1; ch += kLdsAhead) {
        // the reloads are unconditional (clamped to the last chunk), so every pass issues the
        // same loads and the wait before each chunk covers only that chunk's set
#pragma unroll
        for (int u = 0; u < kLdsAhead; ++u) {
          if (ch + u < ch1) lds_chunk_mfma<MASK>(acc, kc[u], lds_x, j, wval);
          lds_chunk_load<MASK>(kc[u], w, min(ch + u + kLdsAhead, ch1 - 1), lane, j, g);
        }
      }
      lds_epilogue(a, w, acc, c, q_own, q0, fast, ostage, Y, ldy);
    }
  }
}

// --------------------------------------------------------------------------
// Band-streaming MFMA SpMM (plan order 6) for random layers whose 64-sample input tile does
// not fit LDS.  Gathering a block column's rows from L2 re-reads every input row once per
// block column that uses it (~28 times at N = 4096); here a work item is (64-sample tile,
// group of 32 block columns = 1024 outputs): the item's 8 waves keep the accumulators of four
// block columns each in registers and stream the tile's input rows through LDS in bands of
// kBandRows rows (two LDS buffers), so each input row moves from L2 once per group.  The band
// image splits every block column's K chunks at band boundaries (band-relative row ids), so
// a chunk reads one band.  Per step (item, band): the current band's chunks run from LDS (X
// fragments by ds_read_b64_tr_b16 at the rows' LDS addresses, same MFMA chain as the other
// kernels), the next band's rows (loaded into registers a step earlier) are written to the
// other buffer, one barrier.  The W fragments and row ids of a step's first chunks are loaded
// at the end of the step before.  Tiles are partitioned over the XCDs (blockIdx % 8), and the
// groups of a tile follow each other, so a tile's rows are read from HBM once into its XCD's
// L2.  Epilogue and compacted stores as in the LDS-resident kernel.
// --------------------------------------------------------------------------
constexpr int kBandWaves = 8;  // 2 per SIMD: one work-group per CU
constexpr int kBandBcPerWave = 4;
constexpr int kBandThreads = kBandWaves * 64;
constexpr int kBandGroup = kBandWaves * kBandBcPerWave;            // block columns per item
constexpr int kBandPieces = kBandRows * 8 / kBandThreads;          // 16-byte pieces per thread and band
static_assert(kBandPieces * kBandThreads == kBandRows * 8, "whole 16-byte pieces per thread");
constexpr size_t kBandBufHalves = static_cast<size_t>(kBandRows) * kLdsLd;

size_t spmm_mfma_band_bytes() {
  return (2 * kBandBufHalves + static_cast<size_t>(kBandWaves) * 16 * kOutLd) * sizeof(_Float16);
}

// Rows [b kBandRows, (b + 1) kBandRows) of tile t into registers: piece k of thread tid is row
// b kBandRows + k kBandThreads / 8 + tid / 8, samples 64 t + 8 (tid % 8)..+7 (each load
// instruction moves 8 whole 128-byte row segments); rows past x_rows are zero.
__device__ __forceinline__ void band_load(uint4 (&v)[kBandPieces], const _Float16* __restrict__ X, int64_t ldx, int t,
                                          int b, int x_rows, int tid) {
  const _Float16* src = X + static_cast<int64_t>(t) * kWaveSamples + (tid & 7) * 8;
#pragma unroll
  for (int k = 0; k < kBandPieces; ++k) {
    const int row = b * kBandRows + k * (kBandThreads / 8) + (tid >> 3);
    v[k] = row < x_rows ? load_stream(src + static_cast<int64_t>(row) * ldx) : make_uint4(0u, 0u, 0u, 0u);
  }
}
__device__ __forceinline__ void band_store(const uint4 (&v)[kBandPieces], _Float16* buf, int tid) {
#pragma unroll
  for (int k = 0; k < kBandPieces; ++k) {
    // two 8-byte writes: rows start on 8-byte boundaries (136-byte stride)
    uint2* d = reinterpret_cast<uint2*>(buf + (k * (kBandThreads / 8) + (tid >> 3)) * kLdsLd + (tid & 7) * 8);
    d[0] = make_uint2(v[k].x, v[k].y);
    d[1] = make_uint2(v[k].z, v[k].w);
  }
}

template <bool MASK>
struct BandCfg {
  static constexpr int kAhead = MASK ? 3 : 1;  // K chunks in flight per block column
};

// Loads of the first kAhead chunks of [lo, hi) (clamped; unconditional, so the number of loads
// in flight is the same on every path).
template <bool MASK>
__device__ __forceinline__ void band_prefetch(LdsChunk (&kc)[BandCfg<MASK>::kAhead], const DevLayer& wb, int lo, int hi,
                                              int lane, int j, int g) {
#pragma unroll
  for (int u = 0; u < BandCfg<MASK>::kAhead; ++u)
    lds_chunk_load<MASK>(kc[u], wb, hi > lo ? min(lo + u, hi - 1) : 0, lane, j, g);
}

// The chunks [lo, hi) of one block column from the band in buf (kc holds the first kAhead).
template <bool MASK>
__device__ __forceinline__ void band_chunks(float4_t (&acc)[4][2], LdsChunk (&kc)[BandCfg<MASK>::kAhead],
                                            const DevLayer& wb, int lo, int hi, const _Float16* buf, int lane, int j,
                                            int g, uint32_t wval) {
  constexpr int A = BandCfg<MASK>::kAhead;
  for (int ch = lo; ch < hi; ch += A) {
#pragma unroll
    for (int u = 0; u < A; ++u) {
      if (ch + u < hi) lds_chunk_mfma<MASK>(acc, kc[u], buf, j, wval);
      if (ch + u + A < hi) lds_chunk_load<MASK>(kc[u], wb, ch + u + A, lane, j, g);
    }
  }
}

template <bool MASK>
__device__ __forceinline__ void mfma_band_body(const StepArgs& a, const DevLayer& w, _Float16* lds) {
  constexpr int A = BandCfg<MASK>::kAhead;
  const uint32_t wval = w.wmask_val;
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  const bool ident = n_out == n;
  const int x_rows = a.x_rows;
  const int nb = w.bands;
  const int ntiles = (n + kWaveSamples - 1) / kWaveSamples;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int j = lane & 15, g = lane >> 4;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
  const int nblk = w.nblk;
  const float4_t bias4 = float4_t{a.bias, a.bias, a.bias, a.bias};
  DevLayer wb = w;  // the band image in the fields the chunk loader reads
  wb.gidx = w.bgidx;
  wb.wt = w.bwt;
  wb.wmask = w.bmask;
  const cint_ptr_t bptr = (cint_ptr_t)w.bptr;
  _Float16* buf0 = lds;
  _Float16* buf1 = lds + kBandBufHalves;
  _Float16* ostage = lds + 2 * kBandBufHalves + wave * 16 * kOutLd;
  const int ngroups = (nblk + kBandGroup - 1) / kBandGroup;
  // this XCD's tiles t = xcd (mod 8), each with its groups one after another; this
  // work-group's items are wl, wl + wgs, ... of the XCD's list
  const int xcd = static_cast<int>(blockIdx.x) % kXcds;
  const int wgs = (static_cast<int>(gridDim.x) - xcd + kXcds - 1) / kXcds;
  const int wl = static_cast<int>(blockIdx.x) / kXcds;
  const int ntx = ntiles > xcd ? (ntiles - xcd + kXcds - 1) / kXcds : 0;
  const int items = ntx * ngroups;
  const int my_items = wl < items ? (items - wl + wgs - 1) / wgs : 0;
  const int steps = my_items * nb;
  if (steps == 0) return;  // uniform per work-group
  auto tile_of = [&](int s) { return xcd + kXcds * ((wl + (s / nb) * wgs) / ngroups); };
  auto group_of = [&](int s) { return (wl + (s / nb) * wgs) % ngroups; };
  // chunk range of the wave's block column q (0 .. kBandBcPerWave - 1) of group G in band b
  auto range = [&](int G, int q, int b, int& lo, int& hi) {
    const int c = G * kBandGroup + wave + kBandWaves * q;
    lo = c < nblk ? bptr[c * (nb + 1) + b] : 0;
    hi = c < nblk ? bptr[c * (nb + 1) + b + 1] : 0;
  };

  // Three-stage pipeline over the steps: band s + 1 travels in registers during step s (loaded
  // at the end of step s - 1) and goes to LDS at the end of step s.  Vector loads complete in
  // issue order, so a wait for a chunk load issued inside step s also waits for the band loads
  // issued before it: those were issued a whole step earlier and have landed.
  uint4 v[kBandPieces];
  LdsChunk kc[kBandBcPerWave][A];
  int lo[kBandBcPerWave], hi[kBandBcPerWave];
  // step 0: its band into buffer 0, its first chunks in flight, then step 1's band
  band_load(v, X, ldx, tile_of(0), 0, x_rows, tid);
  band_store(v, buf0, tid);
#pragma unroll
  for (int q = 0; q < kBandBcPerWave; ++q) {
    range(group_of(0), q, 0, lo[q], hi[q]);
    band_prefetch<MASK>(kc[q], wb, lo[q], hi[q], lane, j, g);
  }
  if (steps > 1) band_load(v, X, ldx, tile_of(1), 1 % nb, x_rows, tid);
  __syncthreads();
  float4_t acc[kBandBcPerWave][4][2];
  for (int s = 0; s < steps; ++s) {
    const int b = s % nb;
    const int t = tile_of(s), G = group_of(s);
    if (b == 0) {
#pragma unroll
      for (int q = 0; q < kBandBcPerWave; ++q)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
          for (int m = 0; m < 2; ++m) acc[q][tt][m] = bias4;
    }
    const _Float16* buf = (s & 1) ? buf1 : buf0;
#pragma unroll
    for (int q = 0; q < kBandBcPerWave; ++q) band_chunks<MASK>(acc[q], kc[q], wb, lo[q], hi[q], buf, lane, j, g, wval);
    if (b == nb - 1) {
      // this lane's sample (lane L owns sample 64t + L) and its output position
      const int p_own = t * kWaveSamples + lane;
      const int q_own = p_own < n ? (ident ? p_own : a.qmap[p_own]) : -1;
      const int q0 = __builtin_amdgcn_readfirstlane(q_own);
      const bool fast = __builtin_amdgcn_ballot_w64(!(q0 >= 0 && (q0 & 7) == 0 && q_own == q0 + lane)) == 0;
#pragma unroll
      for (int q = 0; q < kBandBcPerWave; ++q) {
        const int c = G * kBandGroup + wave + kBandWaves * q;
        if (c < nblk) lds_epilogue(a, w, acc[q], c, q_own, q0, fast, ostage, Y, ldy);
      }
    }
    if (s + 1 < steps) {
      // step s + 1: its first chunks in flight, its band (in v) to the buffer step s - 1 read,
      // then step s + 2's band into v
      const int bn = (s + 1) % nb, Gn = group_of(s + 1);
#pragma unroll
      for (int q = 0; q < kBandBcPerWave; ++q) {
        range(Gn, q, bn, lo[q], hi[q]);
        band_prefetch<MASK>(kc[q], wb, lo[q], hi[q], lane, j, g);
      }
      band_store(v, (s & 1) ? buf0 : buf1, tid);
      if (s + 2 < steps) band_load(v, X, ldx, tile_of(s + 2), (s + 2) % nb, x_rows, tid);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBandThreads, 1) void spmm_mfma_band_kernel(StepArgs a, DevLayer w) {
  extern __shared__ __attribute__((aligned(16))) _Float16 lds_band[];
  mfma_band_body<false>(a, w, lds_band);
}

__global__ __launch_bounds__(kBandThreads, 1) void spmm_mfma_band_mask_kernel(StepArgs a, DevLayer w) {
  extern __shared__ __attribute__((aligned(16))) _Float16 lds_band[];
  mfma_band_body<true>(a, w, lds_band);
}

// --------------------------------------------------------------------------
// Sample-slab SpMM (FP16 pass, plan kind SLAB) for layers whose MFMA images waste most of
// their work: on a random layer a block column's 32 outputs gather ~32 chunks of 32 rows
// with about one nonzero per row, so the MFMA kernels multiply ~32 zeros per useful product
// and re-read every input row once per block column (or per band).  Here a work-group holds
// a slab of T consecutive input positions in LDS, all input rows of them (T = slab_samples:
// 64 up to 1023 rows, 16 at 4096, 4 at 16384, 1 at 65536), and computes every output of
// those samples from it.  Each lane takes slab_lane_slots(T) output slots and min(T, 16) of the
// slab's samples: it reads the slots' input rows from the slab image (16-bit row ids, 8 per
// 16-byte load), gathers each row's samples from LDS (one read per plane of up to 4 samples)
// and accumulates them in FP32 with v_fma_mix_f32, one instruction per sample and nonzero: the
// layer's MACs and no more.  Every input value moves from memory once per slab, the flags of a
// sample are the OR over all of its outputs (reduced in the work-group), and the epilogue fuses
// bias, clamp and FP16 store.  In the sample-blocked layout (kLayoutSlab) a slab is one
// contiguous block, copied straight into LDS (global_load_lds), and a wave's stores cover
// contiguous runs; the neuron-major layout and the bit-packed layer-0 input are read (and
// neuron-major written) too, so the kernel can sit next to the MFMA kernels in a plan.
// Stores go to the samples' compacted positions (qmap); a slab whose samples keep T
// consecutive positions from a multiple of T stores whole vectors, any other one value by
// value.
// --------------------------------------------------------------------------
constexpr int kSlabThreads = 512;  // 8 waves, 2 per SIMD (256 VGPRs for deep pipelines); one work-group per CU
constexpr int kSlabWaves = kSlabThreads / 64;

template <int T>
struct SlabShape {
  static constexpr int kP = T < 4 ? T : 4;       // samples per plane
  static constexpr int kSpl = T < 16 ? T : 16;   // samples per lane
  static constexpr int kLpo = T / kSpl;          // lanes per output slot
  static constexpr int kOpw = 64 / kLpo;         // output slots per wave at a time
  static constexpr int kReads = kSpl / kP;       // LDS reads per row id and lane (one per plane)
  static constexpr int kPlanes = T / kP;
  static constexpr int kM = slab_lane_slots(T);  // output slots per lane
  static constexpr int kBatch = slab_batch_slots(T);  // slots of a wave's batch (kOpw * kM)
  // LDS bytes per plane (compile-time, so that plane offsets fold into the reads' immediates)
  static constexpr int kPb = slab_plane_cap(T);
};

typedef __attribute__((address_space(3))) _Float16 lds_half_t;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

// plain stores into LDS (vector types of the HIP headers cannot be assigned in address space 3)
__device__ __forceinline__ void lds_put16(lds_half_t* p, uint4 v) {
  *reinterpret_cast<__attribute__((address_space(3))) u32x4_t*>(p) = u32x4_t{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void lds_put8(lds_half_t* p, uint32_t lo, uint32_t hi) {
  *reinterpret_cast<__attribute__((address_space(3))) u32x2_t*>(p) = u32x2_t{lo, hi};
}
__device__ __forceinline__ void lds_put4(lds_half_t* p, uint32_t v) {
  *reinterpret_cast<__attribute__((address_space(3))) uint32_t*>(p) = v;
}
__device__ __forceinline__ void lds_put2(lds_half_t* p, uint16_t v) {
  *reinterpret_cast<__attribute__((address_space(3))) uint16_t*>(p) = v;
}

// P consecutive FP16 values of one plane row from LDS (byte offset off): one 8-byte read for P = 4
template <int P>
__device__ __forceinline__ void slab_read(const lds_half_t* base, uint32_t off, _Float16* x) {
  const __attribute__((address_space(3))) char* p = reinterpret_cast<const __attribute__((address_space(3))) char*>(base) + off;
  if constexpr (P == 4) {
    const half4_t v = *reinterpret_cast<const __attribute__((address_space(3))) half4_t*>(p);
#pragma unroll
    for (int t = 0; t < 4; ++t) x[t] = v[t];
  } else if constexpr (P == 2) {
    const half2_t v = *reinterpret_cast<const __attribute__((address_space(3))) half2_t*>(p);
    x[0] = v[0];
    x[1] = v[1];
  } else {
    x[0] = *reinterpret_cast<const __attribute__((address_space(3))) _Float16*>(p);
  }
}

// LDS byte offset row * 2P of the 16-bit row id in half `hi` of dw (P = 1, 2 or 4 samples per
// plane row): one SDWA shift, which takes the word and scales it in a single instruction (the
// compiler's form is an extract, a shift and a mask)
template <int P>
__device__ __forceinline__ uint32_t slab_row_off(uint32_t dw, int hi) {
  uint32_t o;
  if (hi) {
    if constexpr (P == 1)
      asm("v_lshlrev_b32_sdwa %0, 1, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1" : "=v"(o) : "v"(dw));
    else if constexpr (P == 2)
      asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1" : "=v"(o) : "v"(dw));
    else
      asm("v_lshlrev_b32_sdwa %0, 3, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1" : "=v"(o) : "v"(dw));
  } else {
    if constexpr (P == 1)
      asm("v_lshlrev_b32_sdwa %0, 1, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0" : "=v"(o) : "v"(dw));
    else if constexpr (P == 2)
      asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0" : "=v"(o) : "v"(dw));
    else
      asm("v_lshlrev_b32_sdwa %0, 3, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0" : "=v"(o) : "v"(dw));
  }
  return o;
}

// P consecutive FP16 values of one plane row from LDS byte address addr.  The slab kernels hold no
// kernel-static LDS (checked at set-up), so their dynamic LDS, the slab, starts at address 0 and
// row offsets are addresses as they are (through the dynamic array's symbol every address would
// take an add of its link-time value).
template <int P>
__device__ __forceinline__ void slab_read_at(uint32_t addr, _Float16* x) {
  slab_read<P>((const lds_half_t*)(uintptr_t)0u, addr, x);
}

// P FP16 values (packed in dwords, low half first) to global memory at p
template <int P>
__device__ __forceinline__ void slab_store(_Float16* p, const uint32_t* d) {
  if constexpr (P == 4)
    *reinterpret_cast<uint2*>(p) = make_uint2(d[0], d[1]);
  else if constexpr (P == 2)
    *reinterpret_cast<uint32_t*>(p) = d[0];
  else
    *reinterpret_cast<uint16_t*>(p) = static_cast<uint16_t>(d[0]);
}

// Element index of input position p, row r in a sample-blocked buffer of R rows per plane.
template <int T>
__device__ __forceinline__ int64_t slab_index(int p, int r, int R) {
  constexpr int P = SlabShape<T>::kP;
  return static_cast<int64_t>(p / T) * T * R + static_cast<int64_t>((p % T) / P) * R * P + static_cast<int64_t>(r) * P + p % P;
}

// Slab sl (positions sl*T ..: input positions = output positions, no compaction) into the LDS
// planes straight from memory: the bit-packed layer-0 input or T values of every neuron-major
// row (through registers).  Plane k at half k * PB/2, row r at + r * P.
template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void slab_load_rows(const StepArgs& a, lds_half_t* lds, int sl, int x_rows, int64_t ldx,
                                               int tid) {
  constexpr int P = SlabShape<T>::kP, PH = SlabShape<T>::kPb / 2;
  const int p0 = sl * T;
  if (a.xbits != nullptr) {
    // 64-sample tile t, row r: words (t * xbits_rows + r) * 2 + h, bit i of word h = sample 64t + 32h + i
    const int t = p0 >> 6, o = p0 & 63;
    for (int r = tid; r < x_rows; r += NT) {
      const uint32_t* wp = a.xbits + (static_cast<int64_t>(t) * a.xbits_rows + r) * 2;
      uint32_t bits[2];
      if constexpr (T == 64) {
        bits[0] = wp[0];
        bits[1] = wp[1];
      } else {
        bits[0] = wp[o >> 5] >> (o & 31);  // the slab's T bits (T divides 32 and o)
        bits[1] = 0;
      }
#pragma unroll
      for (int k = 0; k < T / P; ++k) {
        // plane k: samples kP .. kP+P-1, dword j holds samples kP+2j (low half) and kP+2j+1
        uint32_t d[2] = {0u, 0u};
#pragma unroll
        for (int j = 0; j < (P + 1) / 2; ++j) {
          const int sp = k * P + 2 * j;
          const uint32_t bb = (bits[sp >> 5] >> (sp & 31)) & 3u;
          d[j] = (bb & 1u) * 0x3C00u | (bb >> 1) * 0x3C000000u;
        }
        lds_half_t* dst = lds + k * PH + r * P;
        if constexpr (P == 4)
          lds_put8(dst, d[0], d[1]);
        else if constexpr (P == 2)
          lds_put4(dst, d[0]);
        else
          lds_put2(dst, static_cast<uint16_t>(d[0]));
      }
    }
  } else {
    const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
    if constexpr (T >= 8) {
      // 16-byte pieces of a row (8 samples = two planes)
      constexpr int ppr = T / 8;
      const int pieces = x_rows * ppr;
      for (int i = tid; i < pieces; i += NT) {
        const int r = i / ppr, part = i % ppr;
        const uint4 v = *reinterpret_cast<const uint4*>(X + static_cast<int64_t>(r) * ldx + p0 + 8 * part);
        lds_put8(lds + (2 * part) * PH + r * 4, v.x, v.y);
        lds_put8(lds + (2 * part + 1) * PH + r * 4, v.z, v.w);
      }
    } else {
      for (int r = tid; r < x_rows; r += NT) {
        const _Float16* src = X + static_cast<int64_t>(r) * ldx + p0;
        if constexpr (T == 4) {
          const uint2 v = *reinterpret_cast<const uint2*>(src);
          lds_put8(lds + r * 4, v.x, v.y);
        } else if constexpr (T == 2) {
          lds_put4(lds + r * 2, *reinterpret_cast<const uint32_t*>(src));
        } else {
          lds_put2(lds + r, *reinterpret_cast<const uint16_t*>(src));
        }
      }
    }
  }
}

// A sample-blocked slab (T / P planes of slab_rows(x_rows) rows, contiguous in HBM) straight into
// the LDS planes, 16 bytes a lane (global_load_lds: the lanes of a wave fill 1 KB, pieces
// k0 = w * 64 + k * NT on, NT threads), no registers held.  The loads are non-temporal (nt): a
// slab is read once per work-group, and kept out of the L2 it leaves the slab image there, which
// every slab of the layer reads again (c2 forward 24.3 -> 23.5 ms; non-temporal epilogue stores,
// partial lines, took that back).  The caller waits for vmcnt(0) before the barrier that
// publishes it.
template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void slab_load_blocked(const StepArgs& a, lds_half_t* lds, int sl, int x_rows, int wave,
                                                  int lane) {
  constexpr int P = SlabShape<T>::kP, PH = SlabShape<T>::kPb / 2;
  const int R = slab_rows(x_rows);
  const int pp = R * P / 8;  // 16-byte pieces per plane (R is a multiple of 8)
  const _Float16* src = static_cast<const _Float16*>(a.x) + static_cast<int64_t>(sl) * T * R;
#pragma unroll 1
  for (int k = 0; k < T / P; ++k) {
    const _Float16* sp = src + static_cast<int64_t>(k) * R * P;
    lds_half_t* dp = lds + k * PH;
    for (int k0 = wave * 64; k0 < pp; k0 += NT) {
      const int i = k0 + lane;
      if (i < pp) __builtin_amdgcn_global_load_lds(sp + 8 * i, (lds_ptr_t)(dp + 8 * k0), 16, 0, 2);
    }
  }
}

// The slab of output positions q0 = sl*T .. after a compacting scan: sample t is input position
// psrc[t] (-1: past the live rows, loaded as zeros), gathered value by value from the input
// (sample-blocked, neuron-major or bit-packed).  Thread (t, r): t = tid % T, rows r = tid / T +
// k * (NT / T) (NT threads); neighbouring threads read neighbouring positions of one row.
template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void slab_load_gather(const StepArgs& a, lds_half_t* lds, const int* psrc, int x_rows,
                                                 int64_t ldx, bool x_slab, int tid) {
  constexpr int P = SlabShape<T>::kP, PH = SlabShape<T>::kPb / 2;
  constexpr int kRowsPer = NT / T;
  const int t = tid % T;
  const int p = psrc[t];
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  const int R = slab_rows(x_rows);
  lds_half_t* dst = lds + (t / P) * PH + t % P;
  // 32 loads in flight per thread: the gather is latency-bound (2-byte loads)
  constexpr int kU = 32;
  for (int r0 = tid / T; r0 < x_rows; r0 += kU * kRowsPer) {
    uint16_t v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int r = r0 + u * kRowsPer;
      v[u] = 0;
      if (p >= 0 && r < x_rows) {
        if (a.xbits != nullptr) {
          const uint32_t wd = a.xbits[(static_cast<int64_t>(p >> 6) * a.xbits_rows + r) * 2 + ((p >> 5) & 1)];
          v[u] = ((wd >> (p & 31)) & 1u) ? 0x3C00u : 0u;
        } else if (x_slab) {
          v[u] = __builtin_bit_cast(uint16_t, X[slab_index<T>(p, r, R)]);
        } else {
          v[u] = __builtin_bit_cast(uint16_t, X[static_cast<int64_t>(r) * ldx + p]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int r = r0 + u * kRowsPer;
      if (r < x_rows) lds_put2(dst + r * P, v[u]);
    }
  }
}

// The slab of output positions q0 = sl*T .. after a compacting scan, from a sample-blocked input:
// sample t is input position p = psrc[t], whose values are every P-th of its source plane (slab
// p / T, plane (p % T) / P, lane p % P).  The plane is read in coalesced 16-byte pieces (8 / P
// rows of P samples: neighbouring threads, neighbouring pieces), kU samples' pieces in flight
// per thread, and the sample's values picked out into its LDS plane.  Rows past x_rows (up to
// slab_rows) take whatever the input's pad rows hold; the slab image never reads them.  The
// loads are non-temporal, as in slab_load_blocked (c2 SpMM 21.7 -> 21.5 ms).
template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void slab_load_gather_blocked(const StepArgs& a, lds_half_t* lds, const int* psrc,
                                                         int x_rows, int tid) {
  constexpr int P = SlabShape<T>::kP, PH = SlabShape<T>::kPb / 2;
  constexpr int kRpp = 8 / P;              // rows per 16-byte piece
  constexpr int kU = T < 8 ? T : 8;        // samples in flight per thread
  const int R = slab_rows(x_rows);
  const int pieces = R / kRpp;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
#pragma unroll 1
  for (int t0 = 0; t0 < T; t0 += kU) {
    int src[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) src[u] = psrc[t0 + u];
#pragma unroll 1
    for (int c = tid; c < pieces; c += NT) {
      uint4 v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int p = src[u];
        v[u] = make_uint4(0u, 0u, 0u, 0u);
        if (p >= 0) {
          const u32x4_t x4 = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(
              X + static_cast<int64_t>(p / T) * T * R + static_cast<int64_t>((p % T) / P) * R * P + static_cast<int64_t>(c) * 8));
          v[u] = make_uint4(x4[0], x4[1], x4[2], x4[3]);
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int t = t0 + u, j = src[u] >= 0 ? src[u] % P : 0;
        const uint32_t d[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
        lds_half_t* dst = lds + (t / P) * PH + t % P + c * kRpp * P;
#pragma unroll
        for (int rr = 0; rr < kRpp; ++rr) {
          const int k = rr * P + j;  // half k of the piece: row c * kRpp + rr, lane j
          const uint32_t dw = (k >> 1) == 0 ? d[0] : (k >> 1) == 1 ? d[1] : (k >> 1) == 2 ? d[2] : d[3];
          lds_put2(dst + rr * P, static_cast<uint16_t>(dw >> (16 * (k & 1))));
        }
      }
    }
  }
}

// OR of the lanes' flag bits per slab sample: lane (slot, group g) holds bit b for sample
// g * SPL + b; the result's bit s is set when any lane's bit for sample s is.
template <int T>
__device__ __forceinline__ uint64_t slab_wave_bits(uint32_t f) {
  constexpr int SPL = SlabShape<T>::kSpl, LPO = SlabShape<T>::kLpo;
  // lanes l with l % LPO == 0
  constexpr uint64_t kLanes0 = LPO == 1 ? ~0ull : (LPO == 2 ? 0x5555555555555555ull : 0x1111111111111111ull);
  uint64_t out = 0;
#pragma unroll
  for (int b = 0; b < SPL; ++b) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(((f >> b) & 1u) != 0u);
#pragma unroll
    for (int g = 0; g < LPO; ++g) out |= ((m & (kLanes0 << g)) != 0ull ? 1ull : 0ull) << (g * SPL + b);
  }
  return out;
}

// A wave's position in the slab image: batch b (b >= bend: past the end), octet o of its nk,
// and the lane's entry offset of octet 0 (elements of sid / sval).
struct SlabCur {
  int b, o, nk;
  int64_t e;
};
template <int T>
__device__ __forceinline__ void slab_cur_batch(SlabCur& c, int b, int bend, int oi, cint_ptr_t sgoff) {
  constexpr int B = SlabShape<T>::kBatch, M = SlabShape<T>::kM;
  c.b = b;
  c.o = 0;
  c.nk = 0;
  c.e = 0;
  if (b < bend) {
    const int k0 = sgoff[b];
    c.nk = sgoff[b + 1] - k0;
    c.e = (static_cast<int64_t>(k0) * B + oi * M) * 8;
  }
}
template <int T>
__device__ __forceinline__ void slab_cur_next(SlabCur& c, int bend, int oi, cint_ptr_t sgoff, int bstride) {
  if (c.b >= bend) return;
  if (c.o + 1 < c.nk) {
    ++c.o;
    return;
  }
  slab_cur_batch<T>(c, c.b + bstride, bend, oi, sgoff);
}
// One octet of a lane's M slots: 8 row ids (and values given, 8 FP16 values) of each.
template <int M>
struct SlabOct {
  uint4 id[M], val[M];
};
// Unconditional (a cursor past the end reads octet 0 of the image, never used), so that every
// step issues the same loads.
template <int T, bool UNI>
__device__ __forceinline__ SlabOct<SlabShape<T>::kM> slab_oct_load(const DevLayer& w, const SlabCur& c) {
  constexpr int B = SlabShape<T>::kBatch, M = SlabShape<T>::kM;
  SlabOct<M> k;
  const int64_t at = c.e + static_cast<int64_t>(c.o) * (B * 8);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    k.id[m] = *reinterpret_cast<const uint4*>(w.sid + at + 8 * m);
    if constexpr (!UNI)
      k.val[m] = *reinterpret_cast<const uint4*>(w.sval + at + 8 * m);
    else
      k.val[m] = make_uint4(0u, 0u, 0u, 0u);
  }
  return k;
}

template <int T, bool UNI, bool BAND, int NT = kSlabThreads>
__device__ __forceinline__ void slab_body(const StepArgs& a, const DevLayer& w, lds_half_t* lds) {
  using Sh = SlabShape<T>;
  constexpr int P = Sh::kP, SPL = Sh::kSpl, LPO = Sh::kLpo, OPW = Sh::kOpw, NR = Sh::kReads, PB = Sh::kPb, M = Sh::kM;
  constexpr int ND = (SPL + 1) / 2;       // dwords of a lane's SPL FP16 outputs
  // bookkeeping behind the planes (kernel-static LDS would move the slab off address 0, and the
  // row offsets would then need an add each): input position of each slab sample (-1: none),
  // and per slab parity and flag kind the T flag bits
  int* const psrc = reinterpret_cast<int*>((_Float16*)lds + kSlabMaxBytes / 2);
  uint32_t(*const sflag)[3][2] = reinterpret_cast<uint32_t(*)[3][2]>(psrc + 64);
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  const bool ident = n_out == n;
  const int x_rows = a.x_rows;
  const int nout = w.n_out;
  const int ry = slab_rows(nout);            // rows per plane of a sample-blocked output slab
  // Slabs are output positions; without compaction they are the input positions too.  The work
  // is nsl * nb units (slab, batch of output slots) in slab-major order, split into one contiguous
  // range per work-group, balanced to a unit: a work-group loads each slab of its range once,
  // and its waves take the range's batches of that slab in turn (lo + wave, lo + wave + 8, ..).
  // With fewer slabs than work-groups several work-groups share a slab; a slab at a range
  // boundary is shared by two (flag bytes are only ever set to 1, so work-groups sharing a
  // sample write them without atomics).
  const int nsl = (n_out + T - 1) / T;
  const int nb = (nout + Sh::kBatch - 1) / Sh::kBatch;
  const int64_t units = static_cast<int64_t>(nsl) * nb;
  const int64_t u0 = units * blockIdx.x / gridDim.x, u1 = units * (blockIdx.x + 1) / gridDim.x;
  const int sl_first = static_cast<int>(u0 / nb);
  const int sl_last = u1 > u0 ? static_cast<int>((u1 - 1) / nb) : sl_first - 1;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int oi = lane / LPO, g = lane % LPO;
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const bool y_slab = a.y_layout == kLayoutSlab;
  const bool x_slab = a.x_layout == kLayoutSlab && a.layer > 0 && a.xbits == nullptr;
  const cint_ptr_t sgoff = (cint_ptr_t)w.sgoff;
  // BAND: the guard-band flags are wanted (a.band_next set); a kernel of its own, since the
  // band test per output is as costly as the rest of the epilogue
  const bool want_unsat = a.unsat_next != nullptr;
  constexpr bool want_band = BAND;
  // FP16 1.0 that the compiler cannot fold (in the uniform case sval is null), so that the
  // uniform case keeps the f16 x f16 + f32 form (v_fma_mix_f32) instead of a conversion and
  // an add
  const _Float16 one_h = w.sval == nullptr ? static_cast<_Float16>(1.0f) : static_cast<_Float16>(2.0f);
  const float bias = a.bias, wuni = w.svalue;
  // LDS byte offset of the lane's first plane (its samples are planes g * NR ..)
  const uint32_t lane_base = static_cast<uint32_t>(g * NR * PB);
  // the zero row of every plane (pads of shorter lists)
  const int zr = slab_rows(x_rows);
  for (int i = tid; i < T; i += NT) lds[(i / P) * (PB / 2) + zr * P + i % P] = static_cast<_Float16>(0.0f);
  auto load_slab = [&](int sl) {
    if (ident) {
      if (x_slab)
        slab_load_blocked<T, NT>(a, lds, sl, x_rows, wave, lane);
      else
        slab_load_rows<T, NT>(a, lds, sl, x_rows, ldx, tid);
    } else if (T == 1 && x_slab) {
      // one sample per slab: the sample's rows are one contiguous block of the sample-blocked
      // input (slab psrc[0] of it), loaded whole instead of value by value
      slab_load_blocked<T, NT>(a, lds, __builtin_amdgcn_readfirstlane(psrc[0]), x_rows, wave, lane);
    } else if (x_slab) {
      slab_load_gather_blocked<T, NT>(a, lds, psrc, x_rows, tid);
    } else {
      slab_load_gather<T, NT>(a, lds, psrc, x_rows, ldx, x_slab, tid);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  };
  auto set_src = [&](int sl) {
    if (tid < T) {
      const int q = sl * T + tid;
      psrc[tid] = q < n_out ? (ident ? q : a.qsrc[q]) : -1;
    }
  };
  set_src(sl_first);
  __syncthreads();  // psrc of the first slab
  if (sl_first <= sl_last) load_slab(sl_first);
  for (int sl = sl_first, it = 0; sl <= sl_last; ++sl, ++it) {
    const int par = it & 1;
    if (tid < 6) sflag[par][tid >> 1][tid & 1] = 0u;
    __syncthreads();  // the slab and cleared flag words are in place
    const int q0 = sl * T;
    // the range's batches of this slab, [blo, bhi); the wave's: b0, b0 + bstride, ..
    const int64_t ub = static_cast<int64_t>(sl) * nb;
    const int blo = static_cast<int>((u0 > ub ? u0 : ub) - ub);
    const int bhi = static_cast<int>((u1 < ub + nb ? u1 : ub + nb) - ub);
    const int b0 = blo + wave, bstride = NT / 64;
    // per lane: packed FP16 max (alive: > 0) and min (unsaturated: < 32) of its samples over
    // its slots, reduced to flag bits once per slab
    uint32_t fmx[ND], fmn[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      fmx[j] = 0u;           // two FP16 zeros
      fmn[j] = 0x50005000u;  // two FP16 32.0
    }
    uint32_t fb = 0;
    // The wave's steps, one octet (8 nonzeros of each of the lane's M slots, 128 FMAs a lane) per
    // step: batches b = wave, wave + NT / 64, ... (every batch has at least one octet),
    // octets in order.  Row ids (and values) sit in a ring of kSlabAhead entries: step k
    // processes entry k % kSlabAhead, then reloads it for step k + kSlabAhead; the loop is
    // unrolled over the ring (no register copies that would wait early).  The epilogue's stores
    // make the next waits drain every load (the compiler's counter model), once per batch.
    constexpr int kSlabAhead = 3;
    SlabCur cur, ahead;
    slab_cur_batch<T>(cur, b0, bhi, oi, sgoff);
    ahead = cur;
    SlabOct<M> ring[kSlabAhead];
#pragma unroll
    for (int u = 0; u < kSlabAhead; ++u) {
      ring[u] = slab_oct_load<T, UNI>(w, ahead);
      slab_cur_next<T>(ahead, bhi, oi, sgoff, bstride);
    }
    float acc[M][SPL];
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
      for (int t = 0; t < SPL; ++t) acc[m][t] = 0.0f;
    auto step = [&](SlabOct<M>& slot) {
      // per slot its 8 rows' LDS reads, then the FMAs as they land; the reads of slot m + 1 are
      // issued before the FMAs of slot m, so that a wave has reads in flight while it computes
      _Float16 x[M][8][SPL];
      auto reads = [&](int m) {
        const uint32_t dw[4] = {slot.id[m].x, slot.id[m].y, slot.id[m].z, slot.id[m].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint32_t off = (LPO == 1 ? 0u : lane_base) + slab_row_off<P>(dw[e >> 1], e & 1);
          // plane offsets fold into the reads' immediate offsets (below 64 KB)
#pragma unroll
          for (int q = 0; q < NR; ++q) slab_read_at<P>(off + static_cast<uint32_t>(q * PB), &x[m][e][q * P]);
        }
      };
      reads(0);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (m + 1 < M) reads(m + 1);
        const uint32_t wv[4] = {slot.val[m].x, slot.val[m].y, slot.val[m].z, slot.val[m].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const _Float16 wh =
              UNI ? one_h : __builtin_bit_cast(_Float16, static_cast<uint16_t>(wv[e >> 1] >> (16 * (e & 1))));
#pragma unroll
          for (int t = 0; t < SPL; ++t)
            acc[m][t] = __builtin_fmaf(static_cast<float>(x[m][e][t]), static_cast<float>(wh), acc[m][t]);
        }
      }
      if (cur.o + 1 == cur.nk) {
        // the batch's last octet: epilogue of the lane's M slots (bias, clamp, FP16, aligned store)
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int s = cur.b * Sh::kBatch + m * OPW + oi;
          if (s < nout) {
            uint32_t d[ND];
#pragma unroll
            for (int j = 0; j < ND; ++j) d[j] = 0u;
#pragma unroll
            for (int t = 0; t < SPL; ++t) {
              const float z = UNI ? __builtin_fmaf(acc[m][t], wuni, bias) : acc[m][t] + bias;
              if (want_band) fb |= (fabsf(z) < a.band || fabsf(z - 32.0f) < a.band ? 1u : 0u) << t;
              const _Float16 yv = static_cast<_Float16>(__builtin_amdgcn_fmed3f(z, 0.0f, 32.0f));
              d[t >> 1] |= static_cast<uint32_t>(__builtin_bit_cast(uint16_t, yv)) << (16 * (t & 1));
            }
#pragma unroll
            for (int j = 0; j < ND; ++j) {
              const half2_t v = __builtin_bit_cast(half2_t, d[j]);
              fmx[j] = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(half2_t, fmx[j]), v));
              fmn[j] = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(half2_t, fmn[j]), v));
            }
            if (y_slab) {
              // slab q0 / T of the output, planes g * NR .., row s
              _Float16* yb = Y + static_cast<int64_t>(q0) * ry + static_cast<int64_t>(g * NR) * ry * P +
                             static_cast<int64_t>(s) * P;
#pragma unroll
              for (int q = 0; q < NR; ++q) slab_store<P>(yb + static_cast<int64_t>(q) * ry * P, d + q * ((P + 1) / 2));
            } else {
              _Float16* yp = Y + static_cast<int64_t>(s) * ldy + q0 + g * SPL;
#pragma unroll
              for (int q = 0; q < NR; ++q) slab_store<P>(yp + q * P, d + q * ((P + 1) / 2));
            }
          }
#pragma unroll
          for (int t = 0; t < SPL; ++t) acc[m][t] = 0.0f;
        }
        // Drain the stores here, once per batch: the compiler's wait counting treats loads and
        // stores in flight together as unordered, and with a store possibly pending every later
        // wait for a row-id load would drain everything, the loads just issued included.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      }
      slot = slab_oct_load<T, UNI>(w, ahead);
      slab_cur_next<T>(ahead, bhi, oi, sgoff, bstride);
      slab_cur_next<T>(cur, bhi, oi, sgoff, bstride);
    };
    // the wave's step count (octets over its batches), so that the main loop runs whole rounds
    // of the ring with no exit test between steps (the compiler's waits then count exactly)
    int nsteps = 0;
    for (int b = b0; b < bhi; b += bstride) nsteps += sgoff[b + 1] - sgoff[b];
    int k = 0;
    for (; k + kSlabAhead <= nsteps; k += kSlabAhead) {
#pragma unroll
      for (int u = 0; u < kSlabAhead; ++u) step(ring[u]);
    }
#pragma unroll
    for (int u = 0; u < kSlabAhead - 1; ++u)
      if (k + u < nsteps) step(ring[u]);
    // flags of the slab's samples: per-sample bits from the lane's max / min, OR over the wave,
    // then over the work-group in LDS
    uint32_t fa = 0, fu = 0;
#pragma unroll
    for (int t = 0; t < SPL; ++t) {
      const uint32_t hx = (fmx[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
      const uint32_t hn = (fmn[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
      fa |= (hx != 0u ? 1u : 0u) << t;        // max > 0 (values are in [0, 32])
      fu |= (hn < 0x5000u ? 1u : 0u) << t;    // min < 32
    }
    const uint64_t wa = slab_wave_bits<T>(fa);
    const uint64_t wu = want_unsat ? slab_wave_bits<T>(fu) : 0ull;
    const uint64_t wb = want_band ? slab_wave_bits<T>(fb) : 0ull;
    if (lane == 0) {
      if (wa != 0ull) {
        atomicOr(&sflag[par][0][0], static_cast<uint32_t>(wa));
        if (T > 32) atomicOr(&sflag[par][0][1], static_cast<uint32_t>(wa >> 32));
      }
      if (wu != 0ull) {
        atomicOr(&sflag[par][1][0], static_cast<uint32_t>(wu));
        if (T > 32) atomicOr(&sflag[par][1][1], static_cast<uint32_t>(wu >> 32));
      }
      if (wb != 0ull) {
        atomicOr(&sflag[par][2][0], static_cast<uint32_t>(wb));
        if (T > 32) atomicOr(&sflag[par][2][1], static_cast<uint32_t>(wb >> 32));
      }
    }
    __syncthreads();  // every read of this slab and every flag word are done
    if (tid < T) {
      const int q = q0 + tid;
      if (q < n_out) {
        if ((sflag[par][0][tid >> 5] >> (tid & 31)) & 1u) a.alive_next[q] = 1;
        if (want_unsat && ((sflag[par][1][tid >> 5] >> (tid & 31)) & 1u)) a.unsat_next[q] = 1;
        if (want_band && ((sflag[par][2][tid >> 5] >> (tid & 31)) & 1u)) a.band_next[q] = 1;
      }
    }
    if (sl < sl_last) {
      if (!ident) {
        set_src(sl + 1);
        __syncthreads();  // psrc of the next slab (gathered loads read it)
      }
      load_slab(sl + 1);
    }
  }
}

#define SLAB_KERNEL(T_)                                                                                 \
  __global__ __launch_bounds__(kSlabThreads, 1) void spmm_slab_kernel_##T_(StepArgs a, DevLayer w) {     \
    extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];                                 \
    slab_body<T_, false, false>(a, w, (lds_half_t*)slab_lds);                                           \
  }                                                                                                     \
  __global__ __launch_bounds__(kSlabThreads, 1) void spmm_slab_uni_kernel_##T_(StepArgs a, DevLayer w) { \
    extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];                                 \
    slab_body<T_, true, false>(a, w, (lds_half_t*)slab_lds);                                            \
  }                                                                                                     \
  __global__ __launch_bounds__(kSlabThreads, 1) void spmm_slab_band_kernel_##T_(StepArgs a, DevLayer w) { \
    extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];                                 \
    if (w.sval == nullptr)                                                                              \
      slab_body<T_, true, true>(a, w, (lds_half_t*)slab_lds);                                           \
    else                                                                                                \
      slab_body<T_, false, true>(a, w, (lds_half_t*)slab_lds);                                          \
  }
SLAB_KERNEL(64)
SLAB_KERNEL(32)
SLAB_KERNEL(16)
SLAB_KERNEL(8)
SLAB_KERNEL(4)
SLAB_KERNEL(2)
SLAB_KERNEL(1)
// One sample per slab with 12 waves (3 per SIMD at 168 VGPRs; same slab image): faster alone
// (one more wave per SIMD to overlap LDS reads with FMAs), but its work-group then takes its
// CU's whole register file and cannot share the CU with kernels running beside the pass.
constexpr int kSlabWideThreads = 768;
__global__ __launch_bounds__(kSlabWideThreads, 1) void spmm_slab_uni_wide_kernel_1(StepArgs a, DevLayer w) {
  extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];
  slab_body<1, true, false, kSlabWideThreads>(a, w, (lds_half_t*)slab_lds);
}
#undef SLAB_KERNEL

// --------------------------------------------------------------------------
// Layer 0 of a sample-slab plan on the bit-packed binary input with one weight value w: an
// output is y = clamp(w c + b, 0, 32), c the number of its input rows set in the sample.  A
// work-group holds the input bits of W = 16 P samples of every row in LDS, 2 P bytes a row at
// byte row * 2 P (the slab kernel's plane layout for the layer's slab size T, P = slab_plane(T),
// so the slab image's entry order spreads the LDS banks here too), and a lane counts for each
// of its output slots the W samples' set bits over the slot's rows in bit-sliced form: the 8
// rows of an octet reduce to a 4-bit count through full adders (v_xor3 and a majority), which
// ripples into 6 count planes (lists of at most 63 rows).  The epilogue turns the planes into
// per-sample counts and computes w c + b in FP32 as the slab kernel does (its sum of 1.0 over
// the set rows is the same count), so the outputs are the slab kernel's bit for bit.  One LDS
// read serves W samples where the slab kernel reads one per sample (at T = 1) to four.
// --------------------------------------------------------------------------
// flag bits of an FP16 value h in [0, 32] by arithmetic (as compares, 16 to 64 samples' lane
// masks would live in scalar registers at once): alive 1 when h != 0, unsaturated 1 below 0x5000
__device__ __forceinline__ uint32_t bits_alive(uint32_t h) {
  uint32_t r;
  asm("v_min_u32 %0, %1, 1" : "=v"(r) : "v"(h));
  return r;
}
__device__ __forceinline__ uint32_t bits_unsat(uint32_t h) {
  uint32_t r;
  asm("v_add_u32 %0, 0xffffb000, %1\n\tv_lshrrev_b32 %0, 31, %0" : "=v"(r) : "v"(h));
  return r;
}
__device__ __forceinline__ void bits_fa(uint32_t a, uint32_t b, uint32_t c, uint32_t& s, uint32_t& cy) {
  s = a ^ b ^ c;
  cy = (a & b) | (c & (a ^ b));
}

// W input bits (2 P bytes) of row r for the slab of positions q0.. into LDS, ident layout: the
// bit-packed input holds, per 64-sample tile t and row r, 64 bits at byte (t * rows + r) * 8
// (bit i = sample 64 t + i); gathered: bit by bit through psrc.
template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void bits_load(const StepArgs& a, lds_half_t* lds, int q0, const int* psrc, bool ident,
                                          int x_rows, int tid) {
  constexpr int P = SlabShape<T>::kP, W = 16 * P;
  const int64_t xr = a.xbits_rows;
  if (ident) {
    const int t = q0 >> 6, k = (q0 & 63) / W;
    const char* base = reinterpret_cast<const char*>(a.xbits) + static_cast<int64_t>(t) * xr * 8 + k * 2 * P;
#pragma unroll 4
    for (int r = tid; r < x_rows; r += NT) {
      const char* src = base + static_cast<int64_t>(r) * 8;
      if constexpr (P == 4) {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        lds_put8(lds + r * 4, v.x, v.y);
      } else if constexpr (P == 2) {
        lds_put4(lds + r * 2, *reinterpret_cast<const uint32_t*>(src));
      } else {
        lds_put2(lds + r, *reinterpret_cast<const uint16_t*>(src));
      }
    }
  } else {
    for (int r = tid; r < x_rows; r += NT) {
      uint32_t wd[2] = {0u, 0u};
      for (int i = 0; i < W; ++i) {
        const int p = psrc[i];
        if (p < 0) continue;
        const uint32_t word = a.xbits[(static_cast<int64_t>(p >> 6) * xr + r) * 2 + ((p >> 5) & 1)];
        wd[i >> 5] |= ((word >> (p & 31)) & 1u) << (i & 31);
      }
      if constexpr (P == 4)
        lds_put8(lds + r * 4, wd[0], wd[1]);
      else if constexpr (P == 2)
        lds_put4(lds + r * 2, wd[0]);
      else
        lds_put2(lds + r, static_cast<uint16_t>(wd[0]));
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int T, int NT = kSlabThreads>
__device__ __forceinline__ void bits_body(const StepArgs& a, const DevLayer& w, lds_half_t* lds) {
  using Sh = SlabShape<T>;
  static_assert(Sh::kLpo == 1, "one lane per output slot");
  constexpr int P = Sh::kP, M = Sh::kM, W = 16 * P;
  constexpr int NW = W > 32 ? 2 : 1;  // 32-bit words of a row's bits
  int* const psrc = reinterpret_cast<int*>((_Float16*)lds + kSlabMaxBytes / 2);
  uint32_t* const sflag = reinterpret_cast<uint32_t*>(psrc + 64);  // [parity][alive, unsat][word]
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int n_out = __builtin_amdgcn_readfirstlane(a.counts[a.layer + 1]);
  const bool ident = n_out == n;
  const int x_rows = a.x_rows;
  const int nout = w.n_out;
  const int ry = slab_rows(nout);
  const int nsl = (n_out + W - 1) / W;
  const int nb = (nout + Sh::kBatch - 1) / Sh::kBatch;
  const int64_t units = static_cast<int64_t>(nsl) * nb;
  const int64_t u0 = units * blockIdx.x / gridDim.x, u1 = units * (blockIdx.x + 1) / gridDim.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t ldy = y_ld(a);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const bool y_slab = a.y_layout == kLayoutSlab;
  const cint_ptr_t sgoff = (cint_ptr_t)w.sgoff;
  const bool want_unsat = a.unsat_next != nullptr;
  const float bias = a.bias, wuni = w.svalue;
  const int zr = slab_rows(x_rows);
  if (tid < NW) lds_put4(lds + zr * P + 2 * tid, 0u);  // the zero row (pads of shorter lists; 4 bytes at P = 1)
  auto set_src = [&](int sl) {
    if (tid < W) {
      const int q = sl * W + tid;
      psrc[tid] = q < n_out ? (ident ? q : a.qsrc[q]) : -1;
    }
  };
  int sl = static_cast<int>(u0 / nb);
  const int sl_last = u1 > u0 ? static_cast<int>((u1 - 1) / nb) : sl - 1;
  set_src(sl);
  __syncthreads();  // psrc of the first slab, the zero row
  if (sl <= sl_last) bits_load<T, NT>(a, lds, sl * W, psrc, ident, x_rows, tid);
  for (int it = 0; sl <= sl_last; ++sl, ++it) {
    const int par = it & 1;
    if (tid < 4) sflag[par * 4 + tid] = 0u;
    __syncthreads();  // the slab and cleared flag words are in place
    const int q0 = sl * W;
    const int64_t ub = static_cast<int64_t>(sl) * nb;
    const int blo = static_cast<int>((u0 > ub ? u0 : ub) - ub);
    const int bhi = static_cast<int>((u1 < ub + nb ? u1 : ub + nb) - ub);
    const int b0 = blo + wave, bstride = NT / 64;
    uint32_t fa[NW], fu[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) fa[j] = fu[j] = 0u;
    SlabCur cur, ahead;
    slab_cur_batch<T>(cur, b0, bhi, lane, sgoff);
    ahead = cur;
    uint32_t cnt[M][6][NW];
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int j = 0; j < NW; ++j) cnt[m][k][j] = 0u;
    auto step = [&](const SlabOct<M>& slot) {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const uint32_t dw[4] = {slot.id[m].x, slot.id[m].y, slot.id[m].z, slot.id[m].w};
        uint32_t x[8][NW];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint32_t off = slab_row_off<P>(dw[e >> 1], e & 1);
          const __attribute__((address_space(3))) char* pr = (const __attribute__((address_space(3))) char*)(uintptr_t)off;
          if constexpr (P == 4) {
            const u32x2_t v = *reinterpret_cast<const __attribute__((address_space(3))) u32x2_t*>(pr);
            x[e][0] = v[0];
            x[e][NW - 1] = v[1];
          } else if constexpr (P == 2) {
            x[e][0] = *reinterpret_cast<const __attribute__((address_space(3))) uint32_t*>(pr);
          } else {
            x[e][0] = *reinterpret_cast<const __attribute__((address_space(3))) uint16_t*>(pr);
          }
        }
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          // 8 one-bit rows -> ones, twos, fours, eights
          uint32_t s0, c0, s1, c1, ones, c3, t0, c4, twos, c5;
          bits_fa(x[0][j], x[1][j], x[2][j], s0, c0);
          bits_fa(x[3][j], x[4][j], x[5][j], s1, c1);
          const uint32_t s2 = x[6][j] ^ x[7][j], c2 = x[6][j] & x[7][j];
          bits_fa(s0, s1, s2, ones, c3);
          bits_fa(c0, c1, c2, t0, c4);
          twos = t0 ^ c3;
          c5 = t0 & c3;
          const uint32_t fours = c4 ^ c5, eights = c4 & c5;
          // ripple into the 6 count planes
          uint32_t k = cnt[m][0][j] & ones;
          cnt[m][0][j] ^= ones;
          uint32_t sum, cy;
          bits_fa(cnt[m][1][j], twos, k, sum, cy);
          cnt[m][1][j] = sum;
          bits_fa(cnt[m][2][j], fours, cy, sum, k);
          cnt[m][2][j] = sum;
          bits_fa(cnt[m][3][j], eights, k, sum, cy);
          cnt[m][3][j] = sum;
          k = cnt[m][4][j] & cy;
          cnt[m][4][j] ^= cy;
          cnt[m][5][j] ^= k;
        }
      }
      if (cur.o + 1 == cur.nk) {
        // the batch's last octet: per slot and sample the count, w c + b, clamp, FP16, stores.
        // Samples go in groups of 8 through a loop that is not unrolled (code size, below).
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int s = cur.b * Sh::kBatch + m * 64 + lane;
          if (s < nout) {
            const bool full = q0 + W <= n_out;
            // element offset of the group's first sample; per group + 8 ry (sample-blocked: a
            // group of 8 holds 8 / P planes-groups of P samples, P ry apart) or + 8 (neuron-major)
            int64_t off = y_slab ? slab_index<T>(q0, s, ry) : static_cast<int64_t>(s) * ldy + q0;
            const int64_t gstep = y_slab ? 8 * static_cast<int64_t>(ry) : 8;
#pragma unroll 1
            for (int g = 0; g < W / 8; ++g) {
              const int j = (8 * g) >> 5, b0 = (8 * g) & 31;
              uint32_t pl[6];
#pragma unroll
              for (int k = 0; k < 6; ++k) pl[k] = j ? cnt[m][k][NW - 1] : cnt[m][k][0];
              // the group's byte of each plane as the rows of an 8x8 bit matrix (v_perm_b32 picks
              // byte b of two words), transposed by three delta swaps: byte i of the result is
              // sample i's count (bit k from plane k)
              const uint32_t b = static_cast<uint32_t>(b0 >> 3);
              const uint32_t sel_lo = b | ((4u + b) << 8) | 0x0C0C0000u, sel_hi = 0x0C0Cu | (b << 16) | ((4u + b) << 24);
              const uint32_t lo = __builtin_amdgcn_perm(pl[1], pl[0], sel_lo) | __builtin_amdgcn_perm(pl[3], pl[2], sel_hi);
              const uint32_t hi = __builtin_amdgcn_perm(pl[5], pl[4], sel_lo);
              uint64_t x = (static_cast<uint64_t>(hi) << 32) | lo;
              uint64_t tt = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
              x = x ^ tt ^ (tt << 7);
              tt = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
              x = x ^ tt ^ (tt << 14);
              tt = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
              x = x ^ tt ^ (tt << 28);
              const uint32_t cw[2] = {static_cast<uint32_t>(x), static_cast<uint32_t>(x >> 32)};
              uint32_t d[4] = {0u, 0u, 0u, 0u}, am = 0u, um = 0u;
#pragma unroll
              for (int i = 0; i < 8; ++i) {
                const float c = static_cast<float>((cw[i >> 2] >> (8 * (i & 3))) & 0xFFu);  // v_cvt_f32_ubyte
                const float z = __builtin_fmaf(c, wuni, bias);
                const _Float16 yv = static_cast<_Float16>(__builtin_amdgcn_fmed3f(z, 0.0f, 32.0f));
                const uint32_t h = __builtin_bit_cast(uint16_t, yv);
                // flags from the FP16 value, as the slab kernel takes them (y in [0, 32]): alive
                // when its bits are nonzero, unsaturated below 0x5000 (32.0); integer ops only
                am |= bits_alive(h) << i;
                um |= bits_unsat(h) << i;
                d[i >> 1] |= h << (16 * (i & 1));
              }
              if (j) {
                fa[NW - 1] |= am << b0;
                fu[NW - 1] |= um << b0;
              } else {
                fa[0] |= am << b0;
                fu[0] |= um << b0;
              }
              if (full) {
                if (y_slab) {
#pragma unroll
                  for (int h = 0; h < 8 / P; ++h) {
                    _Float16* yg = Y + off + static_cast<int64_t>(h) * P * ry;
                    if constexpr (P == 1)
                      *reinterpret_cast<uint16_t*>(yg) = static_cast<uint16_t>(d[h >> 1] >> (16 * (h & 1)));
                    else
                      slab_store<P>(yg, d + h * (P / 2));
                  }
                } else {
                  *reinterpret_cast<uint4*>(Y + off) = make_uint4(d[0], d[1], d[2], d[3]);
                }
              } else {
                // the partial last slab: value by value
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                  const int t = 8 * g + i;
                  if (q0 + t < n_out) {
                    const int64_t at = y_slab ? slab_index<T>(q0 + t, s, ry) : static_cast<int64_t>(s) * ldy + q0 + t;
                    *reinterpret_cast<uint16_t*>(Y + at) = static_cast<uint16_t>(d[i >> 1] >> (16 * (i & 1)));
                  }
                }
              }
              off += gstep;
            }
          }
#pragma unroll
          for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int jj = 0; jj < NW; ++jj) cnt[m][k][jj] = 0u;
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      }
      slab_cur_next<T>(cur, bhi, lane, sgoff, bstride);
    };
    int nsteps = 0;
    for (int b = b0; b < bhi; b += bstride) nsteps += sgoff[b + 1] - sgoff[b];
    if constexpr (W <= 16) {
      // 16 samples a slab (~200 VALU instructions a step): three octets in flight, in unrolled
      // rounds of the ring (the body stays small enough to copy)
      constexpr int kAhead = 3;
      SlabOct<M> ring[kAhead];
#pragma unroll
      for (int u = 0; u < kAhead; ++u) {
        ring[u] = slab_oct_load<T, true>(w, ahead);
        slab_cur_next<T>(ahead, bhi, lane, sgoff, bstride);
      }
      int k = 0;
      for (; k + kAhead <= nsteps; k += kAhead) {
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
          step(ring[u]);
          ring[u] = slab_oct_load<T, true>(w, ahead);
          slab_cur_next<T>(ahead, bhi, lane, sgoff, bstride);
        }
      }
#pragma unroll
      for (int u = 0; u < kAhead - 1; ++u)
        if (k + u < nsteps) step(ring[u]);
    } else {
      // 32 or 64 samples a slab: one octet in flight (a step's ~300-600 VALU instructions hide
      // its load) and a single copy of the body: the ring's copies of the epilogue overflowed
      // the instruction cache (the kernel ran 6x slower at 64 samples)
      SlabOct<M> nxt = slab_oct_load<T, true>(w, ahead);
      slab_cur_next<T>(ahead, bhi, lane, sgoff, bstride);
#pragma unroll 1
      for (int k = 0; k < nsteps; ++k) {
        const SlabOct<M> slot = nxt;
        nxt = slab_oct_load<T, true>(w, ahead);
        slab_cur_next<T>(ahead, bhi, lane, sgoff, bstride);
        step(slot);
      }
    }
    // flags: OR over the wave's lanes, then over the work-group
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      uint32_t va = fa[j], vu = fu[j];
#pragma unroll
      for (int dl = 32; dl >= 1; dl >>= 1) {
        va |= __shfl_xor(va, dl);
        vu |= __shfl_xor(vu, dl);
      }
      if (lane == 0) {
        if (va != 0u) atomicOr(&sflag[par * 4 + j], va);
        if (want_unsat && vu != 0u) atomicOr(&sflag[par * 4 + 2 + j], vu);
      }
    }
    __syncthreads();  // every read of this slab and every flag word are done
    if (tid < W) {
      const int q = q0 + tid;
      if (q < n_out) {
        if ((sflag[par * 4 + (tid >> 5)] >> (tid & 31)) & 1u) a.alive_next[q] = 1;
        if (want_unsat && ((sflag[par * 4 + 2 + (tid >> 5)] >> (tid & 31)) & 1u)) a.unsat_next[q] = 1;
      }
    }
    if (sl < sl_last) {
      if (!ident) {
        set_src(sl + 1);
        __syncthreads();  // psrc of the next slab (gathered loads read it)
      }
      bits_load<T, NT>(a, lds, (sl + 1) * W, psrc, ident, x_rows, tid);
    }
  }
}

// The wide variants (bits_wide_threads(T): 768 at T = 1, 1024 above, within the registers each
// needs) overlap LDS reads and adds with more waves, but take whole CUs: for layers with nothing
// running beside the pass (a.order = kSlabWideOrder).
#define BITS_KERNEL(T_)                                                                                 \
  __global__ __launch_bounds__(kSlabThreads, 1) void spmm_bits_kernel_##T_(StepArgs a, DevLayer w) {     \
    extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];                                 \
    bits_body<T_>(a, w, (lds_half_t*)slab_lds);                                                         \
  }                                                                                                     \
  __global__ __launch_bounds__(bits_wide_threads(T_), 1) void spmm_bits_wide_kernel_##T_(StepArgs a,    \
                                                                                     DevLayer w) {     \
    extern __shared__ __attribute__((aligned(16))) _Float16 slab_lds[];                                 \
    bits_body<T_, bits_wide_threads(T_)>(a, w, (lds_half_t*)slab_lds);                                  \
  }
BITS_KERNEL(16)
BITS_KERNEL(8)
BITS_KERNEL(4)
BITS_KERNEL(2)
BITS_KERNEL(1)
#undef BITS_KERNEL

// --------------------------------------------------------------------------
// Layout conversion between the neuron-major and the sample-blocked activation layouts, for a
// sample-slab layer next to a layer of another kernel: with few samples per slab (T <= 4) the
// slab kernel would read (or write) 2 to 8 bytes per row and slab in the neuron-major layout,
// one scattered access each; here a 64-row x 64-position tile moves through LDS, read and
// written as whole runs (128-byte row segments, and 64-row runs of P-sample pieces).
// --------------------------------------------------------------------------
constexpr int kConvLd = 64 + 8;  // LDS tile row stride (halves)

template <int T>
__device__ __forceinline__ void convert_tile(const ConvArgs& c, int n, int64_t ld, int r0, int p0,
                                             _Float16 (*tile)[kConvLd]) {
  constexpr int P = T < 4 ? T : 4;
  const int R = slab_rows(c.rows);
  const int tid = threadIdx.x;
  const _Float16* __restrict__ src = c.src;
  _Float16* __restrict__ dst = c.dst;
  if (c.to_slab) {
    // neuron-major rows (16-byte pieces) -> tile
    for (int i = tid; i < 64 * 8; i += 256) {
      const int r = i >> 3, j = i & 7;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r0 + r < c.rows) v = *reinterpret_cast<const uint4*>(src + static_cast<int64_t>(r0 + r) * ld + p0 + 8 * j);
      *reinterpret_cast<uint4*>(&tile[r][8 * j]) = v;
    }
    __syncthreads();
    // tile -> P-sample pieces of the blocked layout, consecutive threads on consecutive rows
    for (int i = tid; i < 64 * (64 / P); i += 256) {
      const int r = i & 63, j = i >> 6;
      if (r0 + r >= c.rows) continue;
      const int p = p0 + j * P;
      _Float16* d = dst + static_cast<int64_t>(p / T) * T * R + static_cast<int64_t>((p % T) / P) * R * P +
                    static_cast<int64_t>(r0 + r) * P;
      if constexpr (P == 4) {
        *reinterpret_cast<uint2*>(d) = *reinterpret_cast<const uint2*>(&tile[r][j * P]);
      } else if constexpr (P == 2) {
        *reinterpret_cast<uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(&tile[r][j * P]);
      } else {
        *d = tile[r][j];
      }
    }
  } else {
    for (int i = tid; i < 64 * (64 / P); i += 256) {
      const int r = i & 63, j = i >> 6;
      const int p = p0 + j * P;
      const _Float16* q = src + static_cast<int64_t>(p / T) * T * R + static_cast<int64_t>((p % T) / P) * R * P +
                          static_cast<int64_t>(r0 + r) * P;
      if (r0 + r >= c.rows) continue;
      if constexpr (P == 4) {
        *reinterpret_cast<uint2*>(&tile[r][j * P]) = *reinterpret_cast<const uint2*>(q);
      } else if constexpr (P == 2) {
        *reinterpret_cast<uint32_t*>(&tile[r][j * P]) = *reinterpret_cast<const uint32_t*>(q);
      } else {
        tile[r][j] = *q;
      }
    }
    __syncthreads();
    for (int i = tid; i < 64 * 8; i += 256) {
      const int r = i >> 3, j = i & 7;
      if (r0 + r < c.rows)
        *reinterpret_cast<uint4*>(dst + static_cast<int64_t>(r0 + r) * ld + p0 + 8 * j) =
            *reinterpret_cast<const uint4*>(&tile[r][8 * j]);
    }
  }
}

template <int T>
__global__ __launch_bounds__(256) void convert_layout_kernel(ConvArgs c) {
  __shared__ __attribute__((aligned(16))) _Float16 tile[64][kConvLd];
  const int n = __builtin_amdgcn_readfirstlane(c.counts[c.count_idx]);
  const int64_t ld = act_ld(n);
  const int ptiles = (n + 63) / 64, rtiles = (c.rows + 63) / 64;
  for (int t = blockIdx.x; t < ptiles * rtiles; t += gridDim.x) {
    __syncthreads();  // the previous tile's reads of `tile` are done
    convert_tile<T>(c, n, ld, (t % rtiles) * 64, (t / rtiles) * 64, tile);
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
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
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
        for (int u = 0; u < 8; ++u) xv[u] = xcol[static_cast<int64_t>(w.orow[k + u]) * ldx];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = madd<TAcc>(acc, static_cast<TAcc>(xv[u]), w.oval[k + u]);
      }
      for (; k < k1; ++k) acc = madd<TAcc>(acc, static_cast<TAcc>(xcol[static_cast<int64_t>(w.orow[k]) * ldx]), w.oval[k]);
      const TAcc v = std::is_same<TAcc, double>::value ? static_cast<TAcc>(__dadd_rn(acc, static_cast<double>(a.bias)))
                                                       : acc + static_cast<TAcc>(a.bias);
      fl |= (v > 0) ? 1 : 0;
      fl |= (v < 32) ? 2 : 0;
      fl |= (fabs(static_cast<double>(v)) < a.band || fabs(static_cast<double>(v) - 32.0) < a.band) ? 4 : 0;
      const TAcc y = v < 0 ? TAcc(0) : (v > 32 ? TAcc(32) : v);
      Y[static_cast<int64_t>(out) * ldy + q] = static_cast<TOut>(y);
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
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
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
      for (int u = 0; u < 8; ++u) xv[u] = xcol[static_cast<int64_t>(w.orow[k + u]) * ldx];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = madd<TAcc>(acc, static_cast<TAcc>(xv[u]), w.oval[k + u]);
    }
    for (; k < k1; ++k) acc = madd<TAcc>(acc, static_cast<TAcc>(xcol[static_cast<int64_t>(w.orow[k]) * ldx]), w.oval[k]);
    const TAcc v = std::is_same<TAcc, double>::value ? static_cast<TAcc>(__dadd_rn(acc, static_cast<double>(a.bias)))
                                                     : acc + static_cast<TAcc>(a.bias);
    const TAcc y = v < 0 ? TAcc(0) : (v > 32 ? TAcc(32) : v);
    Y[static_cast<int64_t>(out) * ldy + q] = static_cast<TOut>(y);
    if (v > 0) a.alive_next[q] = 1;
    if (v < 32) a.unsat_next[q] = 1;
    if (a.band_next != nullptr && (fabs(static_cast<double>(v)) < a.band || fabs(static_cast<double>(v) - 32.0) < a.band))
      a.band_next[q] = 1;
  }
}

// Row-gather FP16 kernel for the FP16 pass (plan kind ROWS), first meant for layers whose
// live rows fill a handful of 64-sample groups.  A wave takes one 64-sample group at a time
// (lane = sample) and a strided set of its output slots, two at a time, so every
// gather is one 128-byte row segment with all lanes on one row, the CSC lists are read
// once per sample group and no tile image is read at all.  The MFMA kernel instead re-reads its padded tile image (up to 64 times the
// nonzeros) once per 64-sample tile.  Up to 32 row ids and values of each of the two slots
// come in with one vector load (lanes 0..31 and 32..63), and all 64 gathers are issued
// before any is summed (scalar row bases from readlane); the sums run in CSC order with
// FP32 accumulation, like spmm_simt_kernel<_Float16, _Float16>.  The flags of the wave's
// samples are OR-ed over its slots and written once per group.  Work split 0 needs a grid of
// at least ceil(n / 64) waves, split 1 one of at least 8 work-groups.
constexpr int kRowsBatch = 32;  // rows of one slot loaded before they are summed

// Sample group g of the few-row kernel: the wave computes output slots s0, s0 + stride, ...
// (two at a time) for the 64 samples of the group and writes their flags.
__device__ __forceinline__ void rows_group(const StepArgs& a, const DevLayer& w, const _Float16* __restrict__ X,
                                           _Float16* __restrict__ Y, int64_t ldx, int64_t ldy, int n, int g, int lane,
                                           int s0, int stride) {
  const int p = g * 64 + lane;
  const int q = p < n ? a.qmap[p] : -1;
  const _Float16* __restrict__ xcol = X + (p < n ? p : 0);
  const cint_ptr_t out_idx = (cint_ptr_t)w.out_idx;
  const cint_ptr_t optr = (cint_ptr_t)w.optr;
  const int nslots = w.nblk * 32;
  int fl = 0;
  for (int sa = s0; sa < nslots; sa += 2 * stride) {
    const int sb = sa + stride;  // the pair's second slot (none past the last)
    const int out_a = out_idx[sa];
    const int out_b = sb < nslots ? out_idx[sb] : -1;
    int ka = out_a >= 0 ? optr[sa] : 0, ka1 = out_a >= 0 ? optr[sa + 1] : 0;
    int kb = out_b >= 0 ? optr[sb] : 0, kb1 = out_b >= 0 ? optr[sb + 1] : 0;
    float acc_a = 0.0f, acc_b = 0.0f;
    while (ka < ka1 || kb < kb1) {  // wave-uniform
      const int ma = min(kRowsBatch, ka1 - ka), mb = min(kRowsBatch, kb1 - kb);
      // this batch's row ids and values in lanes 0..31 (slot a) and 32..63 (slot b); row 0
      // where a list has ended, so that every gather below is issued unconditionally
      const int lk = lane & 31, lm = lane < 32 ? ma : mb;
      const int rowv = lk < lm ? w.orow[(lane < 32 ? ka : kb) + lk] : 0;
      const float valv = lk < lm ? w.oval[(lane < 32 ? ka : kb) + lk] : 0.0f;
      _Float16 xa[kRowsBatch], xb[kRowsBatch];
#pragma unroll
      for (int u = 0; u < kRowsBatch; ++u)
        xa[u] = xcol[static_cast<int64_t>(__builtin_amdgcn_readlane(rowv, u)) * ldx];
#pragma unroll
      for (int u = 0; u < kRowsBatch; ++u)
        xb[u] = xcol[static_cast<int64_t>(__builtin_amdgcn_readlane(rowv, 32 + u)) * ldx];
#pragma unroll
      for (int u = 0; u < kRowsBatch; ++u)
        if (u < ma)
          acc_a = madd<float>(acc_a, static_cast<float>(xa[u]),
                              __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, valv), u)));
#pragma unroll
      for (int u = 0; u < kRowsBatch; ++u)
        if (u < mb)
          acc_b = madd<float>(acc_b, static_cast<float>(xb[u]),
                              __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, valv), 32 + u)));
      ka += ma > 0 ? ma : 0;
      kb += mb > 0 ? mb : 0;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int out = h == 0 ? out_a : out_b;
      if (out < 0) continue;
      const float vv = (h == 0 ? acc_a : acc_b) + a.bias;
      fl |= (vv > 0.0f) ? 1 : 0;
      fl |= (vv < 32.0f) ? 2 : 0;
      fl |= (fabsf(vv) < a.band || fabsf(vv - 32.0f) < a.band) ? 4 : 0;
      const float y = vv < 0.0f ? 0.0f : (vv > 32.0f ? 32.0f : vv);
      if (q >= 0) Y[static_cast<int64_t>(out) * ldy + q] = static_cast<_Float16>(y);
    }
  }
  if (q < 0) return;
  if (fl & 1) a.alive_next[q] = 1;
  if ((fl & 2) && a.unsat_next != nullptr) a.unsat_next[q] = 1;
  if ((fl & 4) && a.band_next != nullptr) a.band_next[q] = 1;
}

// Work split (a.order): 0, wave v takes group v % ng and the slots v / ng + k * (waves per
// group), every group at once (few groups); 1, the groups are partitioned over the XCDs
// (blockIdx % 8) and all waves of an XCD stream through its groups one after another, the
// slots of a group strided over them, so the rows that the group's 32-fold gathers re-read
// stay in that XCD's L2 (a 64-sample group is N x 128 bytes).
__global__ __launch_bounds__(256) void spmm_wave_rows_kernel(StepArgs a, DevLayer w) {
  const int n = __builtin_amdgcn_readfirstlane(a.counts[a.layer]);
  const int ng = (n + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  if (ng == 0) return;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
  if (a.order == 1) {
    // the launch has at least kXcds work-groups
    const int x = static_cast<int>(blockIdx.x) % kXcds;
    const int nwx = (static_cast<int>(gridDim.x) - x + kXcds - 1) / kXcds * 4;  // waves on XCD x
    const int wv = static_cast<int>(blockIdx.x) / kXcds * 4 + wave;
    for (int g = x; g < ng; g += kXcds) rows_group(a, w, X, Y, ldx, ldy, n, g, lane, wv, nwx);
    return;
  }
  const int v = static_cast<int>(blockIdx.x) * 4 + wave;
  const int per = static_cast<int>(gridDim.x) * 4 / ng;  // >= 1: the launch covers ceil(S_run / 64) >= ng waves
  if (v >= per * ng) return;
  rows_group(a, w, X, Y, ldx, ldy, n, v % ng, lane, v / ng, per);
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
  const int64_t ldx = x_ld(a), ldy = y_ld(a);
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
        const TIn* src = X + static_cast<int64_t>(row) * ldx + s0 + (lane % kLanesPerRow) * kElemsPerLane;
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
            if (q >= 0) Y[static_cast<int64_t>(outs[m][r]) * ldy + q] = fminf(fmaxf(v, 0.0f), 32.0f);
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

// thread = (neuron row, word): bit i of word w of row n = sample 32w + i of the store is nonzero
__global__ __launch_bounds__(256) void pack_cols_kernel(const _Float16* __restrict__ store, int64_t ld, int N,
                                                        uint32_t* __restrict__ bits) {
  const int64_t wpr = ld / 32;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= wpr * N) return;
  const int64_t nrow = i / wpr, wd = i - nrow * wpr;
  const uint4* src = reinterpret_cast<const uint4*>(store + nrow * ld + 32 * wd);
  // tile-major: the two words of 64-sample tile t of every row are contiguous, so one tile's
  // bits (N x 8 bytes) occupy few cache lines
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = src[q];
    const uint32_t h[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if ((h[e] & 0x7FFFu) != 0u) out |= 1u << (8 * q + 2 * e);
      if ((h[e] & 0x7FFF0000u) != 0u) out |= 1u << (8 * q + 2 * e + 1);
    }
  }
  bits[((wd >> 1) * N + nrow) * 2 + (wd & 1)] = out;
}

__global__ __launch_bounds__(256) void pack_rows_kernel(const _Float16* __restrict__ store, int64_t ld, int N,
                                                        int64_t S, uint32_t* __restrict__ packed, int words) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= S) return;
  for (int wd = blockIdx.y; wd < words; wd += gridDim.y) {
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
      const int nn = 32 * wd + b;
      if (nn < N && static_cast<float>(store[static_cast<int64_t>(nn) * ld + s]) != 0.0f) bits |= 1u << b;
    }
    packed[s * words + wd] = bits;
  }
}

// thread = selected row (consecutive rows write consecutive positions of each neuron row)
__global__ __launch_bounds__(256) void gather_packed_kernel(const uint32_t* __restrict__ packed, int words,
                                                            const int32_t* __restrict__ ids,
                                                            const int32_t* __restrict__ count,
                                                            _Float16* __restrict__ dst, int64_t ldd, int N) {
  const int f = *count;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= f) return;
  const int64_t id = ids[r];
  for (int wd = blockIdx.y; wd < words; wd += gridDim.y) {
    const uint32_t bits = packed[id * words + wd];
    const int nb = min(32, N - 32 * wd);
    for (int b = 0; b < nb; ++b)
      dst[static_cast<int64_t>(32 * wd + b) * ldd + r] = (bits >> b) & 1u ? static_cast<_Float16>(1.0f) : static_cast<_Float16>(0.0f);
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
hipError_t launch_clamp_count(const int32_t* src, int32_t* dst, int32_t cap, int32_t* overflow, hipStream_t st) {
  hipLaunchKernelGGL(clamp_count_kernel, dim3(1), dim3(1), 0, st, src, dst, cap, overflow);
  return hipGetLastError();
}
hipError_t launch_compose_ids(const int32_t* outer, const int32_t* inner, const int32_t* count, int32_t cap,
                              int32_t* dst, int32_t* dst_count, int32_t* overflow, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>(std::max<int32_t>(1, (cap + 255) / 256));
  hipLaunchKernelGGL(compose_ids_kernel, dim3(grid), dim3(256), 0, st, outer, inner, count, cap, dst, dst_count, overflow);
  return hipGetLastError();
}
hipError_t launch_copy_jobs(const CopyJobs& jobs, int32_t* dst, hipStream_t st) {
  if (jobs.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_jobs_kernel, dim3(jobs.n), dim3(256), 0, st, jobs, dst);
  return hipGetLastError();
}
hipError_t launch_set_i32(int32_t* p, int32_t v, hipStream_t st) {
  hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(1), 0, st, p, v);
  return hipGetLastError();
}
hipError_t launch_scan(const ScanArgs& a, int64_t max_rows, hipStream_t st) {
  const int64_t groups = std::max<int64_t>(1, (max_rows + kScanTile - 1) / kScanTile);
  ScanArgs b = a;
  b.max_rows = static_cast<int>(max_rows);
  hipLaunchKernelGGL(scan_kernel, dim3(static_cast<unsigned>(groups)), dim3(kScanThreads), 0, st, b);
  return hipGetLastError();
}
hipError_t launch_spmm_mfma(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  if (w.wmask != nullptr)
    hipLaunchKernelGGL(spmm_mfma_mask_kernel, dim3(grid), dim3(256), 0, st, a, w);
  else
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
int spmm_mfma_lds_bytes(int x_rows) {
  return ((x_rows * kLdsLd + 7) / 8 * 8 + kLdsWaves * 16 * kOutLd) * static_cast<int>(sizeof(_Float16));
}
hipError_t launch_spmm_mfma_lds(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  if (a.x_rows < 1 || a.x_rows > kLdsRows || w.kptr == nullptr) return hipErrorInvalidValue;
  if (w.wmask != nullptr)
    hipLaunchKernelGGL(spmm_mfma_lds_mask_kernel, dim3(grid), dim3(kLdsWaves * 64), spmm_mfma_lds_bytes(a.x_rows), st,
                       a, w);
  else
    hipLaunchKernelGGL(spmm_mfma_lds_kernel, dim3(grid), dim3(kLdsWaves * 64), spmm_mfma_lds_bytes(a.x_rows), st, a, w);
  return hipGetLastError();
}

hipError_t launch_spmm_mfma_band(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  if (w.bptr == nullptr || w.bands < 1 || grid < kXcds) return hipErrorInvalidValue;
  if (w.bmask != nullptr)
    hipLaunchKernelGGL(spmm_mfma_band_mask_kernel, dim3(grid), dim3(kBandThreads), spmm_mfma_band_bytes(), st, a, w);
  else
    hipLaunchKernelGGL(spmm_mfma_band_kernel, dim3(grid), dim3(kBandThreads), spmm_mfma_band_bytes(), st, a, w);
  return hipGetLastError();
}
hipError_t launch_convert_layout(const ConvArgs& c, int grid, hipStream_t st) {
  switch (c.t) {
    case 4: hipLaunchKernelGGL(convert_layout_kernel<4>, dim3(grid), dim3(256), 0, st, c); break;
    case 2: hipLaunchKernelGGL(convert_layout_kernel<2>, dim3(grid), dim3(256), 0, st, c); break;
    case 1: hipLaunchKernelGGL(convert_layout_kernel<1>, dim3(grid), dim3(256), 0, st, c); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_spmm_slab(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  const int t = slab_samples(a.x_rows);
  if (t == 0 || w.sid == nullptr || w.sgoff == nullptr || grid < 1) return hipErrorInvalidValue;
  const bool uni = w.sval == nullptr;
  constexpr size_t bytes = kSlabMaxBytes + kSlabMiscBytes;
#define SLAB_LAUNCH(T_)                                                                                       \
  case T_:                                                                                                    \
    if (a.band_next != nullptr)                                                                               \
      hipLaunchKernelGGL(spmm_slab_band_kernel_##T_, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);        \
    else if (uni)                                                                                             \
      hipLaunchKernelGGL(spmm_slab_uni_kernel_##T_, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);         \
    else                                                                                                      \
      hipLaunchKernelGGL(spmm_slab_kernel_##T_, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);             \
    break;
  switch (t) {
    SLAB_LAUNCH(64)
    SLAB_LAUNCH(32)
    SLAB_LAUNCH(16)
    SLAB_LAUNCH(8)
    SLAB_LAUNCH(4)
    SLAB_LAUNCH(2)
    case 1:
      if (a.band_next != nullptr)
        hipLaunchKernelGGL(spmm_slab_band_kernel_1, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);
      else if (uni && a.order == kSlabWideOrder)
        hipLaunchKernelGGL(spmm_slab_uni_wide_kernel_1, dim3(grid), dim3(kSlabWideThreads), bytes, st, a, w);
      else if (uni)
        hipLaunchKernelGGL(spmm_slab_uni_kernel_1, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);
      else
        hipLaunchKernelGGL(spmm_slab_kernel_1, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef SLAB_LAUNCH
  return hipGetLastError();
}
hipError_t launch_spmm_bits(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  const int t = slab_samples(a.x_rows);
  if (t < 1 || t > 16 || w.sid == nullptr || w.sgoff == nullptr || w.sval != nullptr || a.xbits == nullptr || grid < 1 ||
      (a.y_layout == kLayoutSlab && slab_samples(w.n_out) != t))
    return hipErrorInvalidValue;
  constexpr size_t bytes = kSlabMaxBytes + kSlabMiscBytes;
  const bool wide = a.order == kSlabWideOrder;
#define BITS_LAUNCH(T_)                                                                                       \
  case T_:                                                                                                    \
    if (wide)                                                                                                 \
      hipLaunchKernelGGL(spmm_bits_wide_kernel_##T_, dim3(grid), dim3(bits_wide_threads(T_)), bytes, st, a, w); \
    else                                                                                                      \
      hipLaunchKernelGGL(spmm_bits_kernel_##T_, dim3(grid), dim3(kSlabThreads), bytes, st, a, w);            \
    break;
  switch (t) {
    BITS_LAUNCH(16)
    BITS_LAUNCH(8)
    BITS_LAUNCH(4)
    BITS_LAUNCH(2)
    BITS_LAUNCH(1)
    default: return hipErrorInvalidValue;
  }
#undef BITS_LAUNCH
  return hipGetLastError();
}
hipError_t launch_spmm_wave_rows(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st) {
  hipLaunchKernelGGL(spmm_wave_rows_kernel, dim3(grid), dim3(256), 0, st, a, w);
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

hipError_t launch_pack_rows(const _Float16* store, int64_t ld, int N, int64_t S, uint32_t* packed, int words,
                            hipStream_t st) {
  if (S <= 0 || words <= 0) return hipSuccess;
  dim3 grid(static_cast<unsigned>((S + 255) / 256), static_cast<unsigned>(words < 256 ? words : 256));
  hipLaunchKernelGGL(pack_rows_kernel, grid, dim3(256), 0, st, store, ld, N, S, packed, words);
  return hipGetLastError();
}
hipError_t launch_pack_cols(const _Float16* store, int64_t ld, int N, uint32_t* bits, hipStream_t st) {
  const int64_t total = ld / 32 * N;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_cols_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, store, ld, N,
                     bits);
  return hipGetLastError();
}
hipError_t launch_gather_packed(const uint32_t* packed, int words, const int32_t* ids, const int32_t* count,
                                int64_t max_rows, _Float16* dst, int64_t ldd, int N, hipStream_t st) {
  if (max_rows <= 0 || words <= 0) return hipSuccess;
  dim3 grid(static_cast<unsigned>((max_rows + 255) / 256), static_cast<unsigned>(words < 512 ? words : 512));
  hipLaunchKernelGGL(gather_packed_kernel, grid, dim3(256), 0, st, packed, words, ids, count, dst, ldd, N);
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

hipError_t configure_kernels() {
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(spmm_mfma_lds_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, spmm_mfma_lds_bytes(kLdsRows));
  if (e != hipSuccess) return e;
  const hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(spmm_mfma_lds_mask_kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, spmm_mfma_lds_bytes(kLdsRows));
  if (e2 != hipSuccess) return e2;
  const int band_bytes = static_cast<int>(spmm_mfma_band_bytes());
  const hipError_t e3 = hipFuncSetAttribute(reinterpret_cast<const void*>(spmm_mfma_band_kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, band_bytes);
  if (e3 != hipSuccess) return e3;
  const hipError_t e4 = hipFuncSetAttribute(reinterpret_cast<const void*>(spmm_mfma_band_mask_kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, band_bytes);
  if (e4 != hipSuccess) return e4;
#define SLAB_KERNELS(T_)                                                                      \
  reinterpret_cast<const void*>(spmm_slab_kernel_##T_), reinterpret_cast<const void*>(spmm_slab_uni_kernel_##T_), \
      reinterpret_cast<const void*>(spmm_slab_band_kernel_##T_)
  const void* slab_kernels[] = {SLAB_KERNELS(64), SLAB_KERNELS(32), SLAB_KERNELS(16), SLAB_KERNELS(8),
                                SLAB_KERNELS(4),  SLAB_KERNELS(2),  SLAB_KERNELS(1),
                                reinterpret_cast<const void*>(spmm_slab_uni_wide_kernel_1),
                                reinterpret_cast<const void*>(spmm_bits_kernel_16),
                                reinterpret_cast<const void*>(spmm_bits_kernel_8),
                                reinterpret_cast<const void*>(spmm_bits_kernel_4),
                                reinterpret_cast<const void*>(spmm_bits_kernel_2),
                                reinterpret_cast<const void*>(spmm_bits_kernel_1),
                                reinterpret_cast<const void*>(spmm_bits_wide_kernel_16),
                                reinterpret_cast<const void*>(spmm_bits_wide_kernel_8),
                                reinterpret_cast<const void*>(spmm_bits_wide_kernel_4),
                                reinterpret_cast<const void*>(spmm_bits_wide_kernel_2),
                                reinterpret_cast<const void*>(spmm_bits_wide_kernel_1)};
#undef SLAB_KERNELS
  for (const void* k : slab_kernels) {
    const hipError_t e5 =
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kSlabMaxBytes + kSlabMiscBytes);
    if (e5 != hipSuccess) return e5;
    // their LDS reads address the slab from LDS address 0: no kernel-static LDS may precede it
    hipFuncAttributes attr;
    const hipError_t e6 = hipFuncGetAttributes(&attr, k);
    if (e6 != hipSuccess) return e6;
    if (attr.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
  }
  return hipSuccess;
}

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
