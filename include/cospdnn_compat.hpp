// C++ mirror of the reference's engine entry points, implemented on the C ABI
// (include/cospdnn.h).  Same names, argument meaning and console output as
//   test_graph_challenge_tc     src/cospdnn.cuh:14-21  (src/cospdnn.cu:3-369)
//   test_graph_challenge_tc_co  src/cospdnn.cuh:22-35  (src/cospdnn_double.cu:208-1213)
// so a caller of the reference links against libcospdnn.so instead.
//
// Arguments as in the reference:
//   input      [batch][neuron] input rows (binary in the Graph Challenge data)
//   weight     per layer, the BCSC block values (16x16 blocks, [out][in] inside a block)
//   row_idx    per layer, the first (fused) row of every block
//   ptr        per layer, block-column pointers (neuron/16 + 1 entries)
//   fuse_list  per layer, fused row -> activation row (layer count + 1 rows; the last is unused)
//   gpu_id, gpu_num  sample slice of this call: ceil(batch/gpu_num) rows starting at gpu_id times that
// Behaviour that differs on purpose (each is a reference defect):
//   * every layer uses its own weights (the reference reuses layer L-2's in the last layer,
//     cospdnn.cu:177-208);
//   * a sample is live when any activation is > 0 (the reference tests > 1e-5, kernel.cu:133);
//   * the printed time is the device time of the whole forward (all layers, compaction and the
//     FP32 guard recompute), not only the SpMM kernels.
// Error behaviour matches the reference's Safe_Call: a failure prints to stderr and exits.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

void test_graph_challenge_tc(std::vector<std::vector<float>>& input, std::vector<std::vector<float>>& weight,
                             std::vector<std::vector<int>>& row_idx, std::vector<std::vector<int>>& ptr,
                             std::vector<std::vector<int>>& fuse_list, int batch, int neuron, float bias,
                             int blocksize_c, int blocksize_r, int gpu_id, int gpu_num);

void test_graph_challenge_tc_co(std::vector<std::vector<float>>& input1, std::vector<std::vector<float>>& weight1,
                                std::vector<std::vector<int>>& row_idx1, std::vector<std::vector<int>>& ptr1,
                                std::vector<std::vector<int>>& fuse_list1, int batch1, int neuron1, float bias1,
                                std::vector<std::vector<float>>& input2, std::vector<std::vector<float>>& weight2,
                                std::vector<std::vector<int>>& row_idx2, std::vector<std::vector<int>>& ptr2,
                                std::vector<std::vector<int>>& fuse_list2, int batch2, int neuron2, float bias2,
                                int blocksize_c, int blocksize_r, int gpu_id, int gpu_num);

namespace cospdnn {

// Knobs of the mirror (the reference hard-codes its precision and timing).
struct CompatOptions {
  int precision = 1;          // 1 = FP16-MFMA (default), 0 = FP32 SIMT
  int guard_layer = 12;
  bool use_graph = false;
  std::string categories_path;  // when set, the category ids (1-based, one per line) are written here
};
CompatOptions& compat_options();

struct CompatNetResult {
  std::vector<int64_t> categories;  // 0-based ids of the call's sample slice, ascending
  std::vector<int64_t> live;        // samples alive entering each layer
  double ms_total = 0.0;
  double edges_per_sec = 0.0;
  int64_t flagged = 0;
};

// Same inputs as test_graph_challenge_tc, results returned instead of printed; 0 on success,
// otherwise a cospdnn status code (cospdnn_last_error() has the detail).
int graph_challenge_run(const std::vector<std::vector<float>>& input, const std::vector<std::vector<float>>& weight,
                        const std::vector<std::vector<int>>& row_idx, const std::vector<std::vector<int>>& ptr,
                        const std::vector<std::vector<int>>& fuse_list, int batch, int neuron, float bias,
                        int blocksize_c, int blocksize_r, int gpu_id, int gpu_num, CompatNetResult* out);

// Two networks on one engine in one concurrent cospdnn_run (test_graph_challenge_tc_co's
// inputs as arrays of two); out[k] for network k.
int graph_challenge_run_co(const std::vector<std::vector<float>>* input, const std::vector<std::vector<float>>* weight,
                           const std::vector<std::vector<int>>* row_idx, const std::vector<std::vector<int>>* ptr,
                           const std::vector<std::vector<int>>* fuse_list, const int* batch, const int* neuron,
                           const float* bias, int blocksize_c, int blocksize_r, int gpu_id, int gpu_num,
                           CompatNetResult* out);

// The BCSC arrays test_graph_challenge_tc takes, from a layer in CSC (column j = output j, rows
// in the fused row order of the dataset files): block column c = outputs [c bs_c, (c+1) bs_c),
// one bs_c x bs_r block per row window [w bs_r, (w+1) bs_r) the block column reads, values
// [out][in] inside a block (utils/matrix.h:1245), block start rows, block-column pointers.
// Blocks start at multiples of bs_r (the reference's builder may also start them elsewhere,
// matrix.h:1231-1239; its kernel and this engine read either).
int reference_bcsc_from_csc(const std::vector<int32_t>& colptr, const std::vector<int32_t>& rows,
                            const std::vector<float>& vals, int neuron, int blocksize_c, int blocksize_r,
                            std::vector<float>& values, std::vector<int>& row_idx, std::vector<int>& ptr);

// Reference dataset readers (src/cospdnn.cpp:14-79, utils/matrix.h:270-303).  They return a
// cospdnn status code; `message` receives the text the reference prints on failure.
int read_input_tsv(const std::string& dir, int neuron, int batch, std::vector<std::vector<float>>& input,
                   int64_t* read_num, std::string* message);
int read_fuse_tsv(const std::string& dir, int neuron, int layer, std::vector<std::vector<int>>& fuse,
                  std::string* message);
std::string weight_file_name(const std::string& dir, int neuron, int layer);
// COO file "u v val" (1-based, u = input neuron, v = output neuron) -> CSC by output column
int read_weight_tsv_csc(const std::string& path, int neuron, std::vector<int32_t>& colptr, std::vector<int32_t>& rows,
                        std::vector<float>& vals, std::string* message);

}  // namespace cospdnn
