// Driver around the reference's own host classes, compiled from the reference
// sources in place by oracle/build_ref.sh into oracle/_ref/ref_driver.
// Test infrastructure only (pins the oracle and the BCSC reader).
//
//   ref_driver bcsc <coo.tsv> <out_prefix>
//       COOMatrix(file, 1, T=true) + BCSRCSCMatrix(coo, 16, 16)
//       (utils/matrix.h:270-303, 1200-1269, 1356-1362); writes
//       <out_prefix>.values (float32), .row_index (int32), .len (int32).
//   ref_driver cpufuse <N> <batch> <bias_int> <L> <weight_prefix> <input.bin> <output.bin>
//       CpuSpmmFuse::run_and_cmp (utils/cpu_spmm_fuse.h:11-95) over the L COO files
//       <weight_prefix><l>.tsv, sample-major float32 input and candidate output;
//       exits 0 when every value agrees within 1e-3, aborts otherwise.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "utils/cpu_spmm_fuse.h"
#include "utils/matrix.h"

using namespace ftxj;

template <typename T>
static void write_vec(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

static std::vector<float> read_floats(const std::string& path, size_t n) {
  std::vector<float> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(float)));
  if (!f) {
    std::fprintf(stderr, "short read %s\n", path.c_str());
    std::exit(2);
  }
  return v;
}

int main(int argc, char** argv) {
  if (argc >= 4 && std::string(argv[1]) == "bcsc") {
    COOMatrix coo(argv[2], 1, true);
    BCSRCSCMatrix b(coo, 16, 16);
    const std::string p = argv[3];
    write_vec(p + ".values", b.bcsc_values_);
    write_vec(p + ".row_index", b.bcsc_row_index_);
    write_vec(p + ".len", b.bcsc_len_);
    return 0;
  }
  if (argc >= 9 && std::string(argv[1]) == "cpufuse") {
    const int N = std::atoi(argv[2]);
    const int batch = std::atoi(argv[3]);
    const int bias = std::atoi(argv[4]);
    const int L = std::atoi(argv[5]);
    std::vector<COOMatrix> weights;
    for (int l = 0; l < L; ++l) weights.emplace_back(std::string(argv[6]) + std::to_string(l) + ".tsv", 1, true);
    std::vector<float> input = read_floats(argv[7], static_cast<size_t>(N) * batch);
    std::vector<float> output = read_floats(argv[8], static_cast<size_t>(N) * batch);
    CpuSpmmFuse::run_and_cmp(weights, input.data(), N, batch, bias, output.data(), L, true, true, true);
    return 0;
  }
  std::fprintf(stderr, "usage: ref_driver bcsc <coo.tsv> <out_prefix> | cpufuse ...\n");
  return 2;
}
