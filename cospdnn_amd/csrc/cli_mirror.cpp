// Drivers that use the engine only through the C++ mirror of the reference's entry points
// (include/cospdnn_compat.hpp), the way the reference's own main() functions do:
//   graph_challenge_tc.out     <neuron> <batch> <layer> <gpu_num> <gpu_id> <dataset_dir> [--categories=FILE]
//     (src/cospdnn.cpp:80-188: read the input, the fuse lists and every layer's COO file, build
//     the BCSC arrays, call test_graph_challenge_tc)
//   graph_challenge_tc_co.out  <neuron1> <batch1> <layer1> <neuron2> <batch2> <layer2> <gpu_id> <dataset_dir>
//                              [--categories=FILE]
//     (src/cospdnn_double.cpp:81-197: the same for two networks, test_graph_challenge_tc_co)
// The driver kind comes from the MIRROR_CO macro (0: single, 1: concurrent).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "cli_common.hpp"
#include "cospdnn_compat.hpp"

namespace {

struct RefNet {
  std::vector<std::vector<float>> input, weight;
  std::vector<std::vector<int>> row_idx, ptr, fuse;
};

// Reads one network of the dataset directory into the arrays the reference's entry point takes;
// failures print the reference's messages and exit with its codes.
void read_ref_net(const std::string& dir, int neuron, int batch, int layer, RefNet& n) {
  std::string msg;
  int64_t nread = 0;
  int rc = cospdnn::read_input_tsv(dir, neuron, batch, n.input, &nread, &msg);
  if (rc != COSPDNN_OK) {
    std::cout << msg << "\n";
    std::exit(-1);
  }
  std::cout << "Read Input success! read_numeber = " << nread << std::endl;
  rc = cospdnn::read_fuse_tsv(dir, neuron, layer, n.fuse, &msg);
  if (rc != COSPDNN_OK) {
    std::cout << msg << "\n";
    std::exit(rc == COSPDNN_ERR_IO ? -1 : 0);
  }
  n.fuse.resize(layer + 1);  // layer + 1 lists, the last unused (cospdnn.cpp:113)
  n.weight.resize(layer);
  n.row_idx.resize(layer);
  n.ptr.resize(layer);
  for (int l = 0; l < layer; ++l) {
    std::vector<int32_t> colptr, rows;
    std::vector<float> vals;
    rc = cospdnn::read_weight_tsv_csc(cospdnn::weight_file_name(dir, neuron, l), neuron, colptr, rows, vals, &msg);
    if (rc == COSPDNN_OK)
      rc = cospdnn::reference_bcsc_from_csc(colptr, rows, vals, neuron, 16, 16, n.weight[l], n.row_idx[l], n.ptr[l]);
    if (rc != COSPDNN_OK) {
      std::cout << msg << "\n";
      std::exit(-1);
    }
  }
  std::cout << "weight len: " << layer << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
#if MIRROR_CO
  if (argc < 9) {
    std::cout << "Usage: exe neuron1 batch1 layer1 neuron2 batch2 layer2 gpu_id dataset_dir" << std::endl;
    return 0;
  }
  const int n1 = std::atoi(argv[1]), b1 = std::atoi(argv[2]), l1 = std::atoi(argv[3]);
  const int n2 = std::atoi(argv[4]), b2 = std::atoi(argv[5]), l2 = std::atoi(argv[6]);
  const int dev = std::atoi(argv[7]);
  const std::string dir(argv[8]);
  const int first_flag = 9;
#else
  if (argc < 7) {
    std::cout << "Usage: exe neuron batch layer gpu_num gpu_id dataset_dir" << std::endl;
    return 0;
  }
  const int neuron = std::atoi(argv[1]), batch = std::atoi(argv[2]), layer = std::atoi(argv[3]);
  const int gpu_num = std::max(1, std::atoi(argv[4]));
  const int gpu_id = std::atoi(argv[5]);
  const std::string dir(argv[6]);
  const int first_flag = 7;
#endif
  cli::Flags f;
  if (!cli::parse_flags(argc, argv, first_flag, f)) return 2;
  cospdnn::compat_options().categories_path = f.categories;
  cospdnn::compat_options().precision = f.precision;
  std::cout << "[BEGIN]..." << std::endl;
#if MIRROR_CO
  RefNet a, b;
  read_ref_net(dir, n1, b1, l1, a);
  read_ref_net(dir, n2, b2, l2, b);
  if (hipSetDevice(dev) != hipSuccess) return EXIT_FAILURE;
  test_graph_challenge_tc_co(a.input, a.weight, a.row_idx, a.ptr, a.fuse, b1, n1, cli::bias_for(n1), b.input, b.weight,
                             b.row_idx, b.ptr, b.fuse, b2, n2, cli::bias_for(n2), 16, 16, 0, 1);
#else
  RefNet a;
  read_ref_net(dir, neuron, batch, layer, a);
  if (hipSetDevice(gpu_id) != hipSuccess) return EXIT_FAILURE;
  test_graph_challenge_tc(a.input, a.weight, a.row_idx, a.ptr, a.fuse, batch, neuron, cli::bias_for(neuron), 16, 16,
                          gpu_id, gpu_num);
#endif
  std::cout << "[END]..." << std::endl;
  return 0;
}
