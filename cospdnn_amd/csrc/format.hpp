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
#include "kernels.hpp"

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
  // Mask image: when every nonzero of the tile image is one FP16 value (wmask_val) whose low 7
  // bits are zero, the image
  // is a bit per element: per chunk and lane one 16-bit word, bit 4m + e/2 for the even
  // elements e of fragment m (m = 0, 1), bit 8 + 4m + e/2 for the odd ones (128 bytes a
  // chunk instead of 2 KB); the kernels expand it to the same fragments in registers.
  bool has_mask = false;
  uint16_t wmask_val = 0;
  std::vector<uint16_t> wmask;  // nchunk*64
  // Band image (input rows in bands of kBandRows activation rows): per block column the
  // sorted union of its input rows, split at band boundaries, each band's part padded to
  // chunks of 32, so that every chunk reads one band; band-relative row ids, fragments (and
  // mask words) in the same layout as the tile image.  bptr[c * (bands + 1) + b] = first chunk
  // of block column c in band b.
  bool has_band = false;
  int bands = 0;
  int64_t band_nchunk = 0;
  std::vector<int32_t> bptr;
  std::vector<int32_t> bgidx;
  std::vector<uint16_t> bwt;
  std::vector<uint16_t> bmask;  // when has_mask
  // Slab image (sample-slab kernel, layout in kernels.hpp DevLayer::sid): the SIMT image's lists
  // per batch of slab_batch_slots(T) output slots in octets of 8 rows, interleaved by lane, as
  // 16-bit rows (and FP16 values unless every weight is slab_value).
  bool has_slab = false;
  bool slab_uniform = false;
  float slab_value = 0.0f;
  std::vector<int32_t> sgoff;  // batches+1, offsets in octets
  std::vector<uint16_t> sid;   // sgoff[batches] * slab_batch_slots(T) * 8
  std::vector<uint16_t> sval;  // the same size (FP16 bits) unless slab_uniform
  bool has_tiles32 = false;    // wt32 built (tile_waste <= kMaxWasteFp32Tiles)
  int64_t nchunk = 0;
  double tile_waste = 0.0;     // MFMA MACs per useful MAC (nchunk*32*32 / nnz)
  bool fp16_exact = true;      // every weight survives FP16 round trip
  float colsum_min = 0.0f;     // min over columns of the column sum
  bool nonneg = true;          // every weight >= 0
};

// FP32 tile images are built only up to this waste (they serve the FP32 guard pass).
constexpr double kMaxWasteFp32Tiles = 4.0;

// Output order of a layer (sorted_rows[j]: the sorted distinct input rows of output j) for
// block columns of `group` outputs.  overlap = false: outputs with identical input sets follow
// each other, otherwise natural order.  overlap = true: greedy grouping by shared inputs (each
// group seeded with the lowest-numbered free output, then a free output sharing the most rows
// with the group's union), which also puts identical sets together.
void similarity_order(const std::vector<std::vector<int32_t>>& sorted_rows, int N, int group, bool overlap,
                      std::vector<int32_t>& order);
// Layers up to this many neurons are grouped by overlap (upload cost grows with N times the
// union size); larger ones by identical input sets.
constexpr int kSimilarityMaxN = 16384;

// Builds both images. Tiles are built only when tile_waste <= max_waste.
// Returns false (and leaves `out` unspecified) if the CSC is malformed.
bool build_layer(const CscLayer& w, int N, double max_waste, BlockedLayer& out);

// The band image of a layer from its SIMT image (rows already in the activation-row order of
// the layer's input, in_rows of them).  Fragments follow the values' FP16 bits; mask words are
// built when the tile image has a mask (one value throughout).
void build_band_image(BlockedLayer& bl, int in_rows);

// The slab image of a layer from its SIMT image (rows already in the activation-row order of
// the layer's input, in_rows of them): built when slab_samples(in_rows) > 0 and every weight is
// FP16-exact.  Every batch holds at least one octet; shorter lists are padded with the zero
// row (slab_rows(in_rows)) at value 0, or with row 0 at value 0 past 65528 input rows (no
// 16-bit id for the zero row; the image then carries values even for one weight value).
void build_slab_image(BlockedLayer& bl, int in_rows);

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
