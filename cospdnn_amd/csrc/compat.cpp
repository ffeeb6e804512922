// C++ mirror of the reference entry points and the reference dataset readers,
// on top of the C ABI.  See include/cospdnn_compat.hpp.
#include "cospdnn_compat.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "cospdnn.h"

namespace cospdnn {

CompatOptions& compat_options() {
  static CompatOptions o;
  return o;
}

namespace {

// Whole-file reader with a minimal number tokenizer (the dataset files are large).
struct TextFile {
  std::vector<char> buf;
  size_t pos = 0;
  bool open(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(static_cast<size_t>(n < 0 ? 0 : n) + 1);
    const size_t got = n > 0 ? std::fread(buf.data(), 1, static_cast<size_t>(n), f) : 0;
    buf[got] = '\0';
    buf.resize(got + 1);
    std::fclose(f);
    return true;
  }
  bool next_double(double* v) {
    const char* p = buf.data() + pos;
    while (*p && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (!*p) return false;
    char* end = nullptr;
    *v = std::strtod(p, &end);
    if (end == p) return false;
    pos = static_cast<size_t>(end - buf.data());
    return true;
  }
};

}  // namespace

int read_input_tsv(const std::string& dir, int neuron, int batch, std::vector<std::vector<float>>& input,
                   int64_t* read_num, std::string* message) {
  const std::string name = dir + "/sparse-images-" + std::to_string(neuron) + ".tsv";
  TextFile f;
  if (!f.open(name)) {
    if (message) *message = "FILE:" + name + " does not exists.";
    return COSPDNN_ERR_IO;
  }
  input.assign(batch, std::vector<float>(neuron, 0.0f));
  int64_t cnt = 0;
  double b, n, v;
  while (f.next_double(&b) && f.next_double(&n) && f.next_double(&v)) {
    const int bi = static_cast<int>(b), ni = static_cast<int>(n);
    if (bi <= batch) {
      if (bi < 1 || ni < 1 || ni > neuron) {
        if (message) *message = "input entry out of range in " + name;
        return COSPDNN_ERR_FORMAT;
      }
      ++cnt;
      input[bi - 1][ni - 1] = static_cast<float>(v);
      if (static_cast<float>(v) != 1.0f) std::printf("read input %d, %f\n", bi, v);
    }
  }
  if (read_num) *read_num = cnt;
  return COSPDNN_OK;
}

int read_fuse_tsv(const std::string& dir, int neuron, int layer, std::vector<std::vector<int>>& fuse,
                  std::string* message) {
  const std::string name = dir + "/fuse_matrix/neuron" + std::to_string(neuron) + "/" + std::to_string(neuron) +
                           "_" + std::to_string(layer) + "_fuse.tsv";
  TextFile f;
  if (!f.open(name)) {
    if (message) *message = "FILE:" + name + " does not exists.";
    return COSPDNN_ERR_IO;
  }
  fuse.assign(layer + 1, std::vector<int>(neuron, 0));
  for (int i = 0; i < layer; ++i)
    for (int j = 0; j < neuron; ++j) {
      double v;
      if (!f.next_double(&v)) {
        if (message) *message = " read error";
        return COSPDNN_ERR_FORMAT;
      }
      fuse[i][j] = static_cast<int>(v);
    }
  return COSPDNN_OK;
}

std::string weight_file_name(const std::string& dir, int neuron, int layer) {
  const std::string n = std::to_string(neuron);
  return dir + "/neuron" + n + "/" + n + "-l" + std::to_string(layer + 1) + ".tsv";
}

int read_weight_tsv_csc(const std::string& path, int neuron, std::vector<int32_t>& colptr, std::vector<int32_t>& rows,
                        std::vector<float>& vals, std::string* message) {
  TextFile f;
  if (!f.open(path)) {
    if (message) *message = "File:" + path + " does not exists.";
    return COSPDNN_ERR_IO;
  }
  std::vector<int32_t> us, vs;
  std::vector<float> xs;
  double u, v, x;
  while (f.next_double(&u) && f.next_double(&v) && f.next_double(&x)) {
    const int ui = static_cast<int>(u) - 1, vi = static_cast<int>(v) - 1;
    if (ui < 0 || vi < 0 || ui >= neuron || vi >= neuron) {
      if (message) *message = "weight entry out of range in " + path;
      return COSPDNN_ERR_FORMAT;
    }
    us.push_back(ui);
    vs.push_back(vi);
    xs.push_back(static_cast<float>(x));
  }
  colptr.assign(neuron + 1, 0);
  for (int32_t c : vs) ++colptr[c + 1];
  for (int j = 0; j < neuron; ++j) colptr[j + 1] += colptr[j];
  rows.assign(us.size(), 0);
  vals.assign(us.size(), 0.0f);
  std::vector<int32_t> fill(colptr.begin(), colptr.end() - 1);
  for (size_t i = 0; i < us.size(); ++i) {
    const int32_t d = fill[vs[i]]++;
    rows[d] = us[i];
    vals[d] = xs[i];
  }
  return COSPDNN_OK;
}

int graph_challenge_run(const std::vector<std::vector<float>>& input, const std::vector<std::vector<float>>& weight,
                        const std::vector<std::vector<int>>& row_idx, const std::vector<std::vector<int>>& ptr,
                        const std::vector<std::vector<int>>& fuse_list, int batch, int neuron, float bias,
                        int blocksize_c, int blocksize_r, int gpu_id, int gpu_num, CompatNetResult* out) {
  if (!out || gpu_num < 1 || gpu_id < 0 || neuron <= 0 || batch < 0) return COSPDNN_ERR_ARG;
  const int L = static_cast<int>(weight.size());
  if (row_idx.size() < weight.size() || ptr.size() < weight.size()) return COSPDNN_ERR_ARG;
  // sample slice of this call (cospdnn.cu:23, 36-37)
  const int per = (batch + gpu_num - 1) / gpu_num;
  const int64_t s0 = static_cast<int64_t>(gpu_id) * per;
  const int64_t s1 = std::min<int64_t>(static_cast<int64_t>(gpu_id + 1) * per, static_cast<int64_t>(input.size()));
  *out = CompatNetResult();
  out->live.assign(L + 1, 0);
  if (s1 <= s0) return COSPDNN_OK;
  const int64_t S = s1 - s0;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) device = 0;
  cospdnn_engine* eng = nullptr;
  int rc = cospdnn_engine_create(device, &eng);
  if (rc != COSPDNN_OK) return rc;
  cospdnn_net* net = nullptr;
  rc = cospdnn_net_create(eng, neuron, L, S, bias, &net);
  for (int l = 0; rc == COSPDNN_OK && l < L; ++l) {
    const std::vector<int>* fu = l < static_cast<int>(fuse_list.size()) && !fuse_list[l].empty() ? &fuse_list[l] : nullptr;
    std::vector<int32_t> ri(row_idx[l].begin(), row_idx[l].end()), pt(ptr[l].begin(), ptr[l].end()), fz;
    if (fu) fz.assign(fu->begin(), fu->end());
    rc = cospdnn_net_set_layer_ref_bcsc(net, l, weight[l].data(), static_cast<int64_t>(weight[l].size()), ri.data(),
                                        static_cast<int64_t>(ri.size()), pt.data(), static_cast<int64_t>(pt.size()),
                                        fu ? fz.data() : nullptr, fu ? static_cast<int64_t>(fz.size()) : 0,
                                        blocksize_c, blocksize_r);
  }
  if (rc == COSPDNN_OK) {
    std::vector<int64_t> smp;
    std::vector<int32_t> neu;
    std::vector<float> val;
    for (int64_t s = s0; s < s1; ++s)
      for (int j = 0; j < neuron && j < static_cast<int>(input[s].size()); ++j)
        if (input[s][j] != 0.0f) {
          smp.push_back(s - s0);
          neu.push_back(j);
          val.push_back(input[s][j]);
        }
    rc = cospdnn_net_set_input_coo(net, static_cast<int64_t>(smp.size()), smp.data(), neu.data(), val.data());
  }
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  const CompatOptions& co = compat_options();
  o.precision = co.precision;
  o.guard_layer = co.guard_layer;
  o.use_graph = co.use_graph ? 1 : 0;
  if (rc == COSPDNN_OK) rc = cospdnn_net_upload(net, &o);
  if (rc == COSPDNN_OK) {
    out->categories.assign(S, 0);
    out->live.assign(L + 1, 0);
    cospdnn_result r;
    std::memset(&r, 0, sizeof(r));
    r.categories = out->categories.data();
    r.live = out->live.data();
    rc = cospdnn_run(&net, 1, &o, &r);
    if (rc == COSPDNN_OK) {
      out->categories.resize(r.n_categories);
      out->ms_total = r.ms_total;
      out->edges_per_sec = r.edges_per_sec;
      out->flagged = r.n_flagged;
    }
  }
  cospdnn_net_destroy(net);
  cospdnn_engine_destroy(eng);
  return rc;
}

}  // namespace cospdnn

namespace {

[[noreturn]] void die(int rc) {
  std::fprintf(stderr, "cospdnn error %d (%s): %s\n", rc, cospdnn_strerror(rc), cospdnn_last_error());
  std::exit(EXIT_FAILURE);
}

void write_categories(const std::string& path, const std::vector<int64_t>& cats, int64_t offset) {
  if (path.empty()) return;
  std::ofstream f(path);
  for (int64_t c : cats) f << (c + offset + 1) << "\n";
}

}  // namespace

void test_graph_challenge_tc(std::vector<std::vector<float>>& input, std::vector<std::vector<float>>& weight,
                             std::vector<std::vector<int>>& row_idx, std::vector<std::vector<int>>& ptr,
                             std::vector<std::vector<int>>& fuse_list, int batch, int neuron, float bias,
                             int blocksize_c, int blocksize_r, int gpu_id, int gpu_num) {
  const int per = (batch + gpu_num - 1) / std::max(1, gpu_num);
  std::cout << "initial padding batch " << (per + 255) / 256 * 256 << std::endl;
  cospdnn::CompatNetResult r;
  const int rc = cospdnn::graph_challenge_run(input, weight, row_idx, ptr, fuse_list, batch, neuron, bias, blocksize_c,
                                              blocksize_r, gpu_id, gpu_num, &r);
  if (rc != COSPDNN_OK) die(rc);
  for (size_t l = 0; l + 1 < r.live.size(); ++l) std::cout << "layer " << l << ", feature = " << r.live[l + 1] << std::endl;
  std::cout << "categories = " << r.categories.size() << ", recomputed rows = " << r.flagged << std::endl;
  std::cout << "Kernel Exec Time = " << r.ms_total << "ms" << std::endl;
  write_categories(cospdnn::compat_options().categories_path, r.categories,
                   static_cast<int64_t>(gpu_id) * per);
}

void test_graph_challenge_tc_co(std::vector<std::vector<float>>& input1, std::vector<std::vector<float>>& weight1,
                                std::vector<std::vector<int>>& row_idx1, std::vector<std::vector<int>>& ptr1,
                                std::vector<std::vector<int>>& fuse_list1, int batch1, int neuron1, float bias1,
                                std::vector<std::vector<float>>& input2, std::vector<std::vector<float>>& weight2,
                                std::vector<std::vector<int>>& row_idx2, std::vector<std::vector<int>>& ptr2,
                                std::vector<std::vector<int>>& fuse_list2, int batch2, int neuron2, float bias2,
                                int blocksize_c, int blocksize_r, int gpu_id, int gpu_num) {
  if (weight1.size() != weight2.size()) {
    std::cerr << "Error: Two networks must have the same number of layers!" << std::endl;
    return;
  }
  cospdnn::CompatNetResult r1, r2;
  int rc = cospdnn::graph_challenge_run(input1, weight1, row_idx1, ptr1, fuse_list1, batch1, neuron1, bias1,
                                        blocksize_c, blocksize_r, gpu_id, gpu_num, &r1);
  if (rc != COSPDNN_OK) die(rc);
  rc = cospdnn::graph_challenge_run(input2, weight2, row_idx2, ptr2, fuse_list2, batch2, neuron2, bias2, blocksize_c,
                                    blocksize_r, gpu_id, gpu_num, &r2);
  if (rc != COSPDNN_OK) die(rc);
  std::cout << "categories = " << r1.categories.size() << " and " << r2.categories.size() << std::endl;
  std::cout << "Total Kernel Exec Time = " << (r1.ms_total + r2.ms_total) << "ms" << std::endl;
}
