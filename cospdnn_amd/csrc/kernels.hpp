// Device kernels of the sparse-DNN engine (gfx950) and their launchers.
//
// Data layout in HBM (one network):
//   activations  [N][ld] neuron-major, samples contiguous, FP16 or FP32
//                (the reference's A[N][Bp] orientation, cospdnn.cu:38-44)
//   flags        three uint8 arrays per layer side (alive / unsaturated /
//                guard-band), indexed by sample position, double-buffered
//   ids          original sample id of every position, double-buffered
//   qmap         target position of every input position (-1: dropped)
//   counts       counts[l] = rows in the input layout of layer l
//                counts[l+1] = samples alive entering layer l (= rows of
//                layer l's output).  counts[0] = S.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cospdnn {

constexpr int kTileS = 256;  // samples per work tile (both SpMM kernels)
constexpr int kBandRows = 512;  // activation rows per band of a layer's band image
constexpr int kXcds = 8;     // MI355X: 8 XCDs, work-groups dispatched to them round-robin

struct DevLayer {
  int nblk = 0;
  // [nblk*32]: the activation row of each output slot.  Rows are stored in block order, so
  // out_idx[s] = s for s < n_out and -1 for the padding slots after them.
  const int32_t* out_idx = nullptr;
  int n_out = 0;
  const int32_t* optr = nullptr;     // [nblk*32+1]
  const int32_t* orow = nullptr;
  const float* oval = nullptr;
  const int32_t* kptr = nullptr;     // [nblk+1] (MFMA image; null if absent)
  const int32_t* gidx = nullptr;
  const _Float16* wt = nullptr;
  // mask image (null: none or not used): per chunk and lane a 16-bit word, every set bit the
  // FP16 value wmask_val (see format.hpp); the MFMA kernels then read it instead of wt
  const uint16_t* wmask = nullptr;
  uint32_t wmask_val = 0;
  // band image (null: none): input rows in bands of kBandRows, chunks per (block column, band)
  // at bptr[c * (bands + 1) + b], band-relative row ids bgidx, fragments bwt or mask words bmask
  int bands = 0;
  int band_nchunk = 0;
  const int32_t* bptr = nullptr;
  const int32_t* bgidx = nullptr;
  const _Float16* bwt = nullptr;
  const uint16_t* bmask = nullptr;
  const float* wt32 = nullptr;       // FP32 tile image (MFMA FP32 kernel)
  int one_chunk = 0;                 // every block column has exactly one K chunk (kptr[c] = c)
  int full_blocks = 0;               // every out_idx slot holds an output (no padding slots)
  // slab image (null: none; the sample-slab kernel), for T = slab_samples(input rows): the output
  // slots in batches of B = slab_batch_slots(T) (a wave's unit of work: lane oi of the wave takes
  // slots oi, oi + B/M, .. of a batch, M = slab_lane_slots(T)); per batch the slots' input rows
  // (16-bit activation rows of the layer's input) in octets of 8: batch b's octets are
  // sgoff[b] .. sgoff[b + 1] - 1, and within octet o the lane's M slots follow each other, slot
  // i = m * B/M + oi of the batch at sid[((o * B) + oi * M + m) * 8 ..].  A slot with fewer rows
  // than its batch's octets hold is padded with the zero row (row slab_rows(x_rows), all zero
  // in the kernel's LDS) at value 0.  sval: FP16 values in the same layout, or null when every
  // weight equals svalue (then pads point at the zero row).  Slots n_out .. repeat slot n_out-1.
  const int32_t* sgoff = nullptr;
  const uint16_t* sid = nullptr;
  const _Float16* sval = nullptr;
  float svalue = 0.0f;
};

// LDS of the sample-slab kernel's slab (at most this many bytes), followed by kSlabMiscBytes of
// per-slab bookkeeping (the slab therefore starts at LDS address 0).
constexpr int kSlabMaxBytes = 131072 + 256;
constexpr int kSlabMiscBytes = 512;
// Slab layout.  A slab holds T consecutive sample positions of every row, in planes of
// P = min(T, 4) samples: plane k holds samples kP..kP+P-1 of all rows, P values a row, so that
// a row's values for one plane are one 8-byte (P = 4) LDS read.  In HBM (the sample-blocked
// activation layout) plane k of slab b holds slab_rows(rows) rows; in LDS every plane has one
// more, the zero row slab_rows(x_rows) (all zero: pads of shorter weight lists), and starts on
// a 16-byte boundary.
__host__ __device__ constexpr int slab_plane(int t) { return t < 4 ? t : 4; }
__host__ __device__ constexpr int slab_rows(int rows) { return (rows + 7) / 8 * 8; }
__host__ __device__ constexpr int slab_lds_plane_bytes(int x_rows, int t) {
  return ((slab_rows(x_rows) + 1) * slab_plane(t) * 2 + 15) / 16 * 16;
}
// LDS bytes per plane the kernel reserves for T samples (a compile-time constant per T, so that
// plane offsets fold into the LDS reads' immediate offsets)
__host__ __device__ constexpr int slab_plane_cap(int t) { return kSlabMaxBytes / (t / slab_plane(t)) / 16 * 16; }
// A lane of the sample-slab kernel computes slab_lane_samples(T) samples of each of
// slab_lane_slots(T) output slots (16 FP32 sums a lane and 128 FMAs per 8-row step down to T = 4;
// 4 slots below); slab_batch_slots(T) slots make a wave's batch.
__host__ __device__ constexpr int slab_lane_samples(int t) { return t < 16 ? t : 16; }
__host__ __device__ constexpr int slab_lane_slots(int t) { return 16 / slab_lane_samples(t) < 4 ? 16 / slab_lane_samples(t) : 4; }
__host__ __device__ constexpr int slab_batch_slots(int t) { return 64 / (t / slab_lane_samples(t)) * slab_lane_slots(t); }
// Samples per slab for a layer with x_rows input rows: the largest power of two up to 64 whose
// planes (with the zero row) fit slab_plane_cap; 0 when none does (no slab kernel for the layer).
__host__ __device__ inline int slab_samples(int x_rows) {
  if (x_rows < 1 || x_rows > 65536) return 0;
  for (int t = 64; t >= 1; t >>= 1)
    if (slab_lds_plane_bytes(x_rows, t) <= slab_plane_cap(t)) return t;
  return 0;
}
// Activation layouts (sample-slab kernel): neuron-major [rows][ld] (position p of row r at
// r * ld + p), or sample-blocked by the slab size T (position p of row r in slab p / T, plane
// (p % T) / P, at ((p / T) * T * R + ((p % T) / P) * R * P) + r * P + p % P with R = slab_rows(rows),
// so that one slab is one contiguous block of planes).  Buffers hold slab_rows(N) * ld values.
constexpr int kLayoutRows = 0;
constexpr int kLayoutSlab = 1;

// Activation buffers are compacted layer by layer: the buffer a layer >= 1 reads holds
// counts[layer] rows of samples at leading dimension act_ld(counts[layer]), the row count
// rounded up to an odd multiple of 64 samples (whole 128-byte FP16 segments; an odd multiple
// keeps the neurons' rows from falling on the same HBM channels).  Only the layer-0 input
// store has a leading dimension of its own.  act_ld(rows) <= round_up(rows, 256) + 64.
__host__ __device__ inline int64_t act_ld(int64_t rows) {
  const int64_t m = ((rows > 1 ? rows : 1) + 63) / 64;
  return 64 * (m | 1);
}

struct StepArgs {
  const void* x = nullptr;
  int64_t ld_in = 0;                 // leading dimension of x at layer 0 (the input store)
  void* y = nullptr;
  float bias = 0.0f;
  const int32_t* counts = nullptr;
  int layer = 0;                     // n = counts[layer]
  const int32_t* qmap = nullptr;
  // MFMA FP16 kernel: input position of each output position (compacted layers), or null
  // (then a compacted layer's tiles are input positions and its stores are compacted)
  const int32_t* qsrc = nullptr;
  // MFMA FP16 kernel, layer 0 with a binary input: its bit-packed copy, tile-major (for 64-sample
  // tile t and row n the two words (t * xbits_rows + n) * 2 + h, bit i of word h = sample
  // 64 t + 32 h + i), or null
  const uint32_t* xbits = nullptr;
  int64_t xbits_rows = 0;
  uint8_t* alive_next = nullptr;
  uint8_t* unsat_next = nullptr;     // MFMA FP16 kernel: may be null (flag not needed after this layer)
  uint8_t* band_next = nullptr;      // null: guard band off for this layer
  float band = 0.0f;
  int32_t* work = nullptr;           // work-item counter of the MFMA kernel (zeroed by the scan)
  int order = 1;                     // MFMA kernel item order (0..4); row-gather kernel work split
  int x_rows = 0;                    // LDS-resident MFMA and sample-slab kernels: rows of the input buffer
  int x_layout = kLayoutRows;        // sample-slab kernel: layout of x (layers >= 1; layer 0 reads
  int y_layout = kLayoutRows;        // the input store or its bit-packed copy) and of y
};

struct ScanArgs {
  int32_t* counts = nullptr;
  int layer = 0;                     // reads counts[layer], writes counts[layer+1]
  int max_rows = 0;                  // upper bound of counts[layer] (set by launch_scan)
  const uint8_t* alive = nullptr;
  const uint8_t* unsat = nullptr;
  const uint8_t* band = nullptr;     // may be null
  uint8_t* clear0 = nullptr;         // next-side flag arrays to zero on [0, n)
  uint8_t* clear1 = nullptr;
  uint8_t* clear2 = nullptr;
  const int32_t* ids_in = nullptr;
  int32_t* ids_out = nullptr;
  int32_t* qmap = nullptr;
  int32_t* qsrc = nullptr;           // compacting scans: qsrc[q] = input position of output q (may be null)
  int flag_undecided = 0;            // append alive & unsaturated rows
  int32_t* flag_ids = nullptr;
  int32_t* flag_count = nullptr;
  uint8_t* flag_mark = nullptr;      // per original id
  int32_t* work = nullptr;           // reset to 0 for the next SpMM's work queue (may be null)
  // Retirement of saturated rows (guard passes only): a live row whose input to this layer
  // is all 32 leaves the pass as a category (its id appended to ret_ids) when all-32 is a
  // fixed point of every remaining layer.
  int retire_sat = 0;
  int32_t* ret_ids = nullptr;
  int32_t* ret_count = nullptr;
  // With flag_undecided: the flagged rows also leave the pass (a later pass decides them).
  int drop_flagged = 0;
  // With flag_undecided: flag every live row, saturated ones too (all-32 is not a fixed
  // point of every remaining layer, so saturation decides nothing yet).
  int flag_all_alive = 0;
  // Flag the rows that died in the previous layer (deaths close to the last layer are
  // rechecked: a later pass at higher precision may keep such a row alive to the end).
  int flag_dead = 0;
  // Keep the layout instead of compacting when at most keep_ppm parts per million of the
  // rows entering the layer are removed (0: always compact).  Kept layout: qmap and ids are
  // the identity, counts[layer+1] = counts[layer], and the removed rows stay in place as
  // holes.  Only for scans that remove dead rows alone: a dead row is all-zero and stays
  // so (bias <= 0), so a hole never sets a flag; the next compacting scan drops it.
  int keep_ppm = 0;
  // kept[layer+1] = rows kept (without holes); may be null
  int32_t* kept = nullptr;
};

enum class Dtype : int { F16 = 0, F32 = 1, F64 = 2 };

// Launchers (all stream-ordered, no host synchronisation).
hipError_t launch_gen_input(_Float16* y0, int64_t ld, int N, int64_t S, uint64_t seed, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
hipError_t launch_iota_i32(int32_t* p, int64_t n, hipStream_t st);
hipError_t launch_set_i32(int32_t* p, int32_t v, hipStream_t st);
// Device-side row counts of passes enqueued before their size is known:
// *dst = min(*src, cap), *overflow = 1 when *src > cap.
hipError_t launch_clamp_count(const int32_t* src, int32_t* dst, int32_t cap, int32_t* overflow, hipStream_t st);
// dst[i] = outer[inner[i]] for i < min(*count, cap) (grid covers cap), *dst_count = that number.
hipError_t launch_compose_ids(const int32_t* outer, const int32_t* inner, const int32_t* count, int32_t cap,
                              int32_t* dst, int32_t* dst_count, int32_t* overflow, hipStream_t st);
// max_rows: upper bound of counts[layer] (the grid covers it).  counts[layer+1] must be 0
// beforehand: no group writes it when counts[layer] is 0.
hipError_t launch_scan(const ScanArgs& a, int64_t max_rows, hipStream_t st);
hipError_t launch_spmm_mfma(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
hipError_t launch_spmm_simt(const StepArgs& a, const DevLayer& w, Dtype tin, Dtype tout, int grid, hipStream_t st);
// Few-row variant (work item = 8 samples x 32 outputs, one thread per output and sample).
constexpr int kFewRowsTile = 8;
hipError_t launch_spmm_simt_rows(const StepArgs& a, const DevLayer& w, Dtype tin, Dtype tout, int grid, hipStream_t st);
// FP16 pass, few live rows: wave = (64-sample group, strided output slots); grid * 4 must be at
// least ceil(counts[layer] / 64).
hipError_t launch_spmm_wave_rows(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
// LDS-resident MFMA kernel (a.x_rows <= kLdsRowsMax input rows, tile image present); any grid.
constexpr int kLdsRowsMax = 1024;
hipError_t launch_spmm_mfma_lds(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
// Band-streaming MFMA kernel (band image present, a.x_rows input rows; w.bmask used when set,
// w.bwt otherwise); one work-group of 512 threads per CU, any grid of at least 8 work-groups.
hipError_t launch_spmm_mfma_band(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
// Sample-slab kernel (FP16 pass; slab image present, slab_samples(a.x_rows) > 0, 512 threads,
// one work-group per CU); any grid (the work is split into one balanced range per work-group).
hipError_t launch_spmm_slab(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
// a.order for launch_spmm_slab: at one sample per slab with one weight value and no guard-band
// flags, the 768-thread variant (faster alone; its work-group fills its CU's registers, so it
// suits layers with nothing running beside the pass).  Any other value: the 512-thread kernel.
constexpr int kSlabWideOrder = 7;
// threads of the bit-sliced layer-0 kernel's wide variant (launch_spmm_bits with a.order =
// kSlabWideOrder): 768 at T = 1 (149 VGPRs at 512 threads), 1024 above (at most 122)
__host__ __device__ constexpr int bits_wide_threads(int t) { return t == 1 ? 768 : 1024; }
// Layer 0 of a slab plan on the bit-packed binary input (a.xbits) with one weight value (slab image
// without values, slab_samples(a.x_rows) <= 16, lists of at most 63 rows): per output the count
// of set input rows by bit-sliced adders, y written in a.y_layout (sample-blocked by the same slab
// size, or neuron-major); 512 threads, one work-group per CU, any grid.
hipError_t launch_spmm_bits(const StepArgs& a, const DevLayer& w, int grid, hipStream_t st);
// Activations of `rows` rows at positions [0, counts[count_idx]) (leading dimension act_ld of that
// count in the neuron-major layout) converted between the neuron-major layout and the
// sample-blocked layout of slab size t (4, 2 or 1); to_slab: neuron-major -> blocked.
struct ConvArgs {
  const _Float16* src = nullptr;
  _Float16* dst = nullptr;
  int rows = 0;
  const int32_t* counts = nullptr;
  int count_idx = 0;
  int t = 0;
  int to_slab = 1;
};
hipError_t launch_convert_layout(const ConvArgs& c, int grid, hipStream_t st);
hipError_t launch_spmm_mfma32(const StepArgs& a, const DevLayer& w, Dtype tin, int grid, hipStream_t st);
hipError_t launch_input_alive(const _Float16* store, int64_t ld, int N, int64_t S, uint8_t* alive, hipStream_t st);
hipError_t launch_gather_rows(const _Float16* src, int64_t lds, const int32_t* ids, const int32_t* count,
                              int64_t max_rows, _Float16* dst, int64_t ldd, int N, hipStream_t st);
// Binary inputs: bit-packed sample-major copy [S][words] (words = ceil(N/32)) of the
// neuron-major FP16 store, and the gather of selected samples from it.
// bits[(t * N + n) * 2 + h] (tile-major, t < ld/64): bit i = store[n][64t + 32h + i] != 0
hipError_t launch_pack_cols(const _Float16* store, int64_t ld, int N, uint32_t* bits, hipStream_t st);
hipError_t launch_pack_rows(const _Float16* store, int64_t ld, int N, int64_t S, uint32_t* packed, int words,
                            hipStream_t st);
hipError_t launch_gather_packed(const uint32_t* packed, int words, const int32_t* ids, const int32_t* count,
                                int64_t max_rows, _Float16* dst, int64_t ldd, int N, hipStream_t st);

// Result read-back in one launch: job j copies min(*len, cap) (len == nullptr: cap) int32
// from src to dst + off and stores that count at dst[count_at] (count_at < 0: nowhere).
// dst is pinned host memory; the host reads it after one synchronisation.
constexpr int kMaxCopyJobs = 16;
struct CopyJob {
  const int32_t* src;
  const int32_t* len;
  int32_t cap;
  int64_t off;
  int64_t count_at;
};
struct CopyJobs {
  CopyJob job[kMaxCopyJobs];
  int n = 0;
};
hipError_t launch_copy_jobs(const CopyJobs& jobs, int32_t* dst, hipStream_t st);

// Kernel attributes set once per process (dynamic LDS limits).
hipError_t configure_kernels();

// Resident work-groups per CU for each SpMM kernel (occupancy query).
int spmm_mfma_blocks_per_cu();
// Work-groups of the MFMA kernel that fit one CU by its VGPR and LDS use (gfx950: 512 VGPRs per
// lane and SIMD, 160 KB LDS per CU).
int spmm_mfma_resident_per_cu();
int spmm_mfma32_blocks_per_cu(Dtype tin);
int spmm_simt_blocks_per_cu(Dtype tin, Dtype tout);
int spmm_simt_rows_blocks_per_cu(Dtype tin, Dtype tout);

}  // namespace cospdnn
