// Weight formats of the engine.
//
// A layer arrives as CSC (column j = output neuron j, rows = input neurons;
// the same orientation as the reference's COO "T" files, utils/matrix.h:282-289).
// The builder groups output columns into block columns of kBlockOut outputs,
// putting columns with identical input sets next to each other (the
// similarity grouping the reference expects to be done offline, README.md:9),
// and emits two device images:
//
//  * SIMT image: the CSC re-ordered block by block (optr/orow/oval) plus the
//    output id of every block slot (out_idx, -1 for padding).
//  * MFMA image ("gathered BCSC"): per block column the sorted union of its
//    input rows, padded to chunks of kChunkK, as a gather list (gidx), and the
//    dense FP16 tile of every chunk stored in the A-operand order of
//    v_mfma_f32_16x16x32_f16 with M = outputs and K = gathered inputs.
//    This generalises the reference's fixed "two 16-row blocks per block
//    column" BCSC (kernels/kernel.cu:310,316,346) to any number of chunks.
#pragma once
#include <cstdint>
#include <vector>

#include "gen.hpp"

namespace cospdnn {

constexpr int kBlockOut = 32;  // outputs per block column
constexpr int kChunkK = 32;    // gathered inputs per K chunk (MFMA K)

struct BlockedLayer {
  int N = 0;
  int nblk = 0;
  int64_t nnz = 0;
  std::vector<int32_t> out_idx;  // nblk*32
  // SIMT image
  std::vector<int32_t> optr;  // nblk*32+1
  std::vector<int32_t> orow;
  std::vector<float> oval;
  // MFMA image
  bool has_tiles = false;
  std::vector<int32_t> kptr;   // nblk+1, offsets in chunks
  std::vector<int32_t> gidx;   // nchunk*32
  std::vector<uint16_t> wt;    // nchunk*1024 FP16 bit patterns (v_mfma_f32_16x16x32_f16 A order)
  std::vector<float> wt32;     // nchunk*1024 FP32 values (v_mfma_f32_16x16x4f32 A order)
  bool has_tiles32 = false;    // wt32 built (tile_waste <= kMaxWasteFp32Tiles)
  int64_t nchunk = 0;
  double tile_waste = 0.0;     // MFMA MACs per useful MAC (nchunk*32*32 / nnz)
  bool fp16_exact = true;      // every weight survives FP16 round trip
  float colsum_min = 0.0f;     // min over columns of the column sum
  bool nonneg = true;          // every weight >= 0
};

// FP32 tile images are built only up to this waste (they serve the FP32 guard pass).
constexpr double kMaxWasteFp32Tiles = 4.0;

// Builds both images. Tiles are built only when tile_waste <= max_waste.
// Returns false (and leaves `out` unspecified) if the CSC is malformed.
bool build_layer(const CscLayer& w, int N, double max_waste, BlockedLayer& out);

// True when an all-32 input row gives an all-32 output row whatever the summation order
// in FP32 or FP64: every weight is nonnegative and, for every column j (empty ones
// included), 32 * colsum_j + bias >= 32 after the worst rounding loss of a K_j-term FP32
// sum (products by 32 are exact), with a factor 2 of slack.
bool all32_fixed_point(const CscLayer& w, int N, float bias);

// Reference BCSC (values [out][in] per 16x16 block, utils/matrix.h:1245;
// block start rows; block-column pointers) + fuse list (row map applied to
// block rows, kernels/kernel.cu:310) -> CSC.  Zero entries are dropped.
bool csc_from_reference_bcsc(const std::vector<float>& values, const std::vector<int>& row_idx,
                             const std::vector<int>& ptr, const std::vector<int>& fuse, int N,
                             int blocksize_c, int blocksize_r, CscLayer& out);

uint16_t float_to_half_bits(float f);  // round to nearest even
float half_bits_to_float(uint16_t h);

}  // namespace cospdnn
