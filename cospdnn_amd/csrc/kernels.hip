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

#include <algorithm>

namespace cospdnn {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef __fp16 fp16x4_t __attribute__((__vector_size__(4 * sizeof(__fp16))));
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
  uint4 va = zero, vu = zero, vb = zero;
  int cnt = 0;
  if (pc < n) {
    va = *reinterpret_cast<const uint4*>(a.alive + pc);
    vu = *reinterpret_cast<const uint4*>(a.unsat + pc);
    if (a.band != nullptr) vb = *reinterpret_cast<const uint4*>(a.band + pc);
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
// MFMA SpMM.  Work item = (sample tile t of 256, block column c of 32 outputs),
// handed out by an atomic counter; a work-group walks the K chunks of its
// items as one stream of steps.  Four waves; wave w owns samples
// [64w, 64w+64) of the tile as four 16-sample N-tiles and both 16-output
// M-tiles: D[out][sample] = W^T[out][k] * X[k][sample].
// Per step the 32 gathered input rows x 256 samples (16 KB) arrive in LDS by
// global_load_lds_dwordx4 (lane i writes bytes 16i..16i+15 of the wave's LDS
// window).  The LDS tile is double-buffered: the next step's copy is issued
// before the current step's MFMAs, after the loads whose values this step
// still needs (A fragments, qmap, output ids), so the compiler's vmcnt waits
// for those never wait on the prefetch.  Each lane then gathers its B fragment
// X[8g..8g+7][sample] with eight 16-bit LDS reads.
// --------------------------------------------------------------------------
// Weight-side arrays are never written while a kernel runs: reading them through the
// constant address space lets wave-uniform indices become scalar loads.
typedef const __attribute__((address_space(4))) int32_t* cint_ptr_t;

__device__ __forceinline__ void mfma_issue_tile(const _Float16* __restrict__ X, int64_t ldx, cint_ptr_t gidx,
                                                int ch, int s0, int wave, int lane, _Float16* buf) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = wave * 8 + i * 2;  // wave-uniform first row of the pair
    const int r0 = gidx[ch * kChunk + r];
    const int r1 = gidx[ch * kChunk + r + 1];
    const int row = lane < 32 ? r0 : r1;
    const _Float16* src = X + static_cast<int64_t>(row) * ldx + s0 + (lane & 31) * 8;
    __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(buf + r * kTileS), 16, 0, 0);
  }
}

__device__ __forceinline__ void mfma_issue_weights(const _Float16* __restrict__ wt, int ch, int wave, int lane,
                                                   _Float16* buf) {
  // the chunk's 2 KB A-operand image: waves 0 and 1 copy 1 KB each
  if (wave < 2) {
    const _Float16* src = wt + static_cast<int64_t>(ch) * 1024 + wave * 512 + lane * 8;
    __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(buf + wave * 512), 16, 0, 0);
  }
}

__global__ __launch_bounds__(256) void spmm_mfma_kernel(StepArgs a, DevLayer w) {
  __shared__ __attribute__((aligned(16))) _Float16 xs[2][kChunk * kTileS];  // activation tiles
  __shared__ __attribute__((aligned(16))) _Float16 ws[2][2 * 64 * 8];       // weight tiles
  __shared__ int queue_slot[2];
  const int n = a.counts[a.layer];
  const int nblk = w.nblk;
  const int nitems = (n + kTileS - 1) / kTileS * nblk;
  // Items come from a counter, two ahead: the current item and the one after it are known,
  // and the atomic that fetches the third lands before the step that needs it.
  if (threadIdx.x == 0) {
    queue_slot[0] = atomicAdd(a.work, 1);
    queue_slot[1] = atomicAdd(a.work, 1);
  }
  __syncthreads();
  int item = __builtin_amdgcn_readfirstlane(queue_slot[0]);
  int item_after = __builtin_amdgcn_readfirstlane(queue_slot[1]);
  if (item >= nitems) return;  // uniform
  int fetched = 0;             // thread 0: result of the atomic issued in the previous step
  bool fetch_pending = false;
  int qsel = 0;                // the two queue slots alternate, so a slot is rewritten only
                               // after a barrier that follows every wave's read of it
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: keeps gidx loads scalar
  const int j = lane & 15, g = lane >> 4;
  const _Float16* __restrict__ X = static_cast<const _Float16*>(a.x);
  _Float16* __restrict__ Y = static_cast<_Float16*>(a.y);
  const cint_ptr_t kptr = (cint_ptr_t)w.kptr;
  const cint_ptr_t gidx = (cint_ptr_t)w.gidx;

  // epilogue indices of an item: target positions of its samples, ids of its outputs
  auto load_item_ids = [&](int t_, int c_, int (&q_)[4], int (&o_)[8]) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int p = t_ * kTileS + wave * 64 + nt * 16 + j;
      const int qv = a.qmap[min(p, n - 1)];
      q_[nt] = p < n ? qv : -1;
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) o_[m * 4 + r] = w.out_idx[c_ * 32 + m * 16 + g * 4 + r];
  };

  int t = item / nblk, c = item - t * nblk;
  int ch = kptr[c], ch_end = kptr[c + 1];
  mfma_issue_tile(X, a.ldx, gidx, ch, t * kTileS, wave, lane, xs[0]);
  mfma_issue_weights(w.wt, ch, wave, lane, ws[0]);
  int q[4], outs[8];
  load_item_ids(t, c, q, outs);
  int buf = 0;
  float4_t acc[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[m][nt] = float4_t{0.f, 0.f, 0.f, 0.f};

  // Software pipeline, one step per K chunk.  At the top of a step every copy and load
  // issued by the previous step has landed (one s_waitcnt the compiler can see); all LDS
  // reads of the step come before the next step's copies into the other buffers, and no
  // register loaded after those copies is used in this step.
  while (true) {
    __builtin_amdgcn_s_waitcnt(0);
    if (fetch_pending && tid == 0) queue_slot[qsel] = fetched;
    __syncthreads();
    if (fetch_pending) {
      item_after = __builtin_amdgcn_readfirstlane(queue_slot[qsel]);
      fetch_pending = false;
      qsel ^= 1;
    }
    const _Float16* xb = xs[buf];
    const _Float16* wb = ws[buf];
    const half8_t a0 = *reinterpret_cast<const half8_t*>(wb + (0 * 64 + lane) * 8);
    const half8_t a1 = *reinterpret_cast<const half8_t*>(wb + (1 * 64 + lane) * 8);
    // B fragment of N-tile nt: lane (j, g) needs X[8g+e][16nt+j], e < 8.  ds_read_b64_tr_b16
    // reads 4 halves per lane and, inside each group of 16 lanes, hands lane i element i%4 of
    // lanes 4k+i/4 (k < 4).  So lane m of the group reads row 8g+4h+m/4 at samples
    // 16nt+4(m%4)..+3, and two reads (h = 0, 1) give the eight K values of its own sample.
    half8_t bfrag[4];
    const int tr_row = g * 8 + (j >> 2);
    const int tr_col = wave * 64 + 4 * (j & 3);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const _Float16* p0 = xb + tr_row * kTileS + tr_col + nt * 16;
      const fp16x4_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)p0);
      const fp16x4_t hi =
          __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4_t*)(p0 + 4 * kTileS));
      const half4_t l4 = __builtin_bit_cast(half4_t, lo);
      const half4_t h4 = __builtin_bit_cast(half4_t, hi);
      bfrag[nt] = __builtin_shufflevector(l4, h4, 0, 1, 2, 3, 4, 5, 6, 7);
    }
    // next step: the following chunk of this item, or the first chunk of this group's next item
    const bool last = ch + 1 == ch_end;
    int n_item = item, n_t = t, n_c = c, n_ch = ch + 1, n_end = ch_end;
    bool more = true;
    if (last) {
      n_item = item_after;
      more = n_item < nitems;
      if (more) {
        n_t = n_item / nblk;
        n_c = n_item - n_t * nblk;
        n_ch = kptr[n_c];
        n_end = kptr[n_c + 1];
      }
    }
    // always issued (after the final step it re-reads this step's chunk, which is in bounds)
    const int p_t = more ? n_t : t, p_c = more ? n_c : c, p_ch = more ? n_ch : ch;
    mfma_issue_tile(X, a.ldx, gidx, p_ch, p_t * kTileS, wave, lane, xs[buf ^ 1]);
    mfma_issue_weights(w.wt, p_ch, wave, lane, ws[buf ^ 1]);
    int q_n[4], outs_n[8];
    load_item_ids(p_t, p_c, q_n, outs_n);
    if (last && more) {
      // the item after the next one; its index is read at the top of the next step
      if (tid == 0) fetched = atomicAdd(a.work, 1);
      fetch_pending = true;
    }

#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bfrag[nt], acc[0][nt], 0, 0, 0);
      acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bfrag[nt], acc[1][nt], 0, 0, 0);
    }

    if (last) {
      // Epilogue. D[i][j] of a 16x16 tile sits in lane j + 16*(i/4), element i%4:
      // this lane holds outputs 16m + 4g + r of sample 16nt + j.
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        int fl = 0;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float v = acc[m][nt][r] + a.bias;
            const int out = outs[m * 4 + r];
            if (out >= 0) {
              fl |= (v > 0.0f) ? 1 : 0;
              fl |= (v < 32.0f) ? 2 : 0;
              fl |= (fabsf(v) < a.band || fabsf(v - 32.0f) < a.band) ? 4 : 0;
              const float y = fminf(fmaxf(v, 0.0f), 32.0f);
              if (q[nt] >= 0) Y[static_cast<int64_t>(out) * a.ldy + q[nt]] = static_cast<_Float16>(y);
            }
          }
        }
        fl |= __shfl_xor(fl, 16, 64);
        fl |= __shfl_xor(fl, 32, 64);
        if (g == 0 && q[nt] >= 0) {
          if (fl & 1) a.alive_next[q[nt]] = 1;
          if (fl & 2) a.unsat_next[q[nt]] = 1;
          if ((fl & 4) && a.band_next != nullptr) a.band_next[q[nt]] = 1;
        }
      }
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[m][nt] = float4_t{0.f, 0.f, 0.f, 0.f};
    }
    if (!more) break;
    item = n_item;
    t = n_t;
    c = n_c;
    ch = n_ch;
    ch_end = n_end;
    buf ^= 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) q[i] = q_n[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) outs[i] = outs_n[i];
  }
  __builtin_amdgcn_s_waitcnt(0);  // the final re-read copy lands before the group exits
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
      int k = k0;
      // eight independent gathers in flight per thread, summed in CSC order
      for (; k + 8 <= k1; k += 8) {
        TIn xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = xcol[static_cast<int64_t>(w.orow[k + u]) * a.ldx];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += static_cast<float>(xv[u]) * w.oval[k + u];
      }
      for (; k < k1; ++k) acc += static_cast<float>(xcol[static_cast<int64_t>(w.orow[k]) * a.ldx]) * w.oval[k];
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

// --------------------------------------------------------------------------
// FP64 replay: one work-group carries one sample through all layers in FP64,
// each output summing its inputs in CSC order with separate multiply and add
// (no contraction), i.e. the operation sequence of the FP64 oracle, so the
// activations and the category of a replayed row are bit-identical to it.
// The layer's input vector sits in LDS when it fits (N <= 16384); the gathers
// of an output are issued in groups of eight ahead of the ordered additions.
// --------------------------------------------------------------------------
constexpr int kReplayThreads = 1024;
constexpr int kReplayLdsDoubles = 16384;

template <bool kInLds>
__device__ __forceinline__ void replay_layer(const DevLayer& w, const double* __restrict__ yin, double* __restrict__ yout,
                                             double b) {
  const int slots = w.nblk * 32;
  for (int slot = threadIdx.x; slot < slots; slot += kReplayThreads) {
    const int out = w.out_idx[slot];
    if (out < 0) continue;
    const int k1 = w.optr[slot + 1];
    double acc = 0.0;
    int k = w.optr[slot];
    for (; k + 8 <= k1; k += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = yin[w.orow[k + u]];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = __dadd_rn(acc, __dmul_rn(x[u], static_cast<double>(w.oval[k + u])));
    }
    for (; k < k1; ++k) acc = __dadd_rn(acc, __dmul_rn(yin[w.orow[k]], static_cast<double>(w.oval[k])));
    const double z = __dadd_rn(acc, b);
    yout[out] = z < 0.0 ? 0.0 : (z > 32.0 ? 32.0 : z);
  }
}

__global__ __launch_bounds__(kReplayThreads) void replay_fp64_kernel(ReplayArgs a) {
  extern __shared__ double ylds[];
  const int r = blockIdx.x;
  if (r >= a.rows) return;
  const int N = a.N;
  const int32_t local = a.row_ids[r];
  const int64_t id = a.map != nullptr ? a.map[local] : local;
  double* ya = a.scratch + static_cast<int64_t>(r) * 2 * N;
  double* yb = ya + N;
  const bool in_lds = N <= kReplayLdsDoubles;
  double* yin = in_lds ? ylds : ya;
  for (int nn = threadIdx.x; nn < N; nn += kReplayThreads) yin[nn] = static_cast<double>(a.store[static_cast<int64_t>(nn) * a.ld + id]);
  __syncthreads();
  const double b = static_cast<double>(a.bias);
  for (int l = 0; l < a.L; ++l) {
    const DevLayer w = a.layers[l];
    if (in_lds) {
      replay_layer<true>(w, yin, yb, b);
      __syncthreads();
      for (int nn = threadIdx.x; nn < N; nn += kReplayThreads) yin[nn] = yb[nn];
      __syncthreads();
    } else {
      replay_layer<false>(w, yin, yb, b);
      __syncthreads();
      double* tmp = yin;
      yin = yb;
      yb = tmp;
    }
  }
  int any = 0;
  for (int nn = threadIdx.x; nn < N; nn += kReplayThreads) any |= yin[nn] > 0.0 ? 1 : 0;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) a.alive[r] = static_cast<uint8_t>(any);
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

hipError_t launch_spmm_mfma32(const StepArgs& a, const DevLayer& w, Dtype tin, int grid, hipStream_t st) {
  if (tin == Dtype::F16)
    hipLaunchKernelGGL(spmm_mfma32_f16in_kernel, dim3(grid), dim3(256), 0, st, a, w);
  else
    hipLaunchKernelGGL(spmm_mfma32_f32in_kernel, dim3(grid), dim3(256), 0, st, a, w);
  return hipGetLastError();
}
hipError_t launch_replay_fp64(const ReplayArgs& a, hipStream_t st) {
  if (a.rows <= 0) return hipSuccess;
  const size_t lds = a.N <= kReplayLdsDoubles ? sizeof(double) * static_cast<size_t>(a.N) : 0;
  hipLaunchKernelGGL(replay_fp64_kernel, dim3(a.rows), dim3(kReplayThreads), lds, st, a);
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
  // the FP64 replay keeps up to 128 KB of its row in LDS
  return hipFuncSetAttribute(reinterpret_cast<const void*>(replay_fp64_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(sizeof(double) * kReplayLdsDoubles));
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
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, spmm_simt_kernel<float, float>, 256, 0);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

}  // namespace cospdnn
