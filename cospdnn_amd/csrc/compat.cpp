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

namespace {

// A network of the call's sample slice (cospdnn.cu:23, 36-37) on `eng`, uploaded: the
// reference's BCSC + fuse arrays per layer and the input rows as (sample, neuron, value).
// *S_out = 0 (and no network) when the slice is empty.
int build_compat_net(cospdnn_engine* eng, const std::vector<std::vector<float>>& input,
                     const std::vector<std::vector<float>>& weight, const std::vector<std::vector<int>>& row_idx,
                     const std::vector<std::vector<int>>& ptr, const std::vector<std::vector<int>>& fuse_list,
                     int batch, int neuron, float bias, int blocksize_c, int blocksize_r, int gpu_id, int gpu_num,
                     const cospdnn_opts& o, cospdnn_net** out, int64_t* S_out) {
  *out = nullptr;
  *S_out = 0;
  if (gpu_num < 1 || gpu_id < 0 || neuron <= 0 || batch < 0) return COSPDNN_ERR_ARG;
  const int L = static_cast<int>(weight.size());
  if (row_idx.size() < weight.size() || ptr.size() < weight.size()) return COSPDNN_ERR_ARG;
  const int per = (batch + gpu_num - 1) / gpu_num;
  const int64_t s0 = static_cast<int64_t>(gpu_id) * per;
  const int64_t s1 = std::min<int64_t>(static_cast<int64_t>(gpu_id + 1) * per, static_cast<int64_t>(input.size()));
  if (s1 <= s0) return COSPDNN_OK;
  const int64_t S = s1 - s0;
  cospdnn_net* net = nullptr;
  int rc = cospdnn_net_create(eng, neuron, L, S, bias, &net);
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
  if (rc == COSPDNN_OK) rc = cospdnn_net_upload(net, &o);
  if (rc != COSPDNN_OK) {
    cospdnn_net_destroy(net);
    return rc;
  }
  *out = net;
  *S_out = S;
  return COSPDNN_OK;
}

cospdnn_opts compat_run_opts() {
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  const CompatOptions& co = compat_options();
  o.precision = co.precision;
  o.guard_layer = co.guard_layer;
  o.use_graph = co.use_graph ? 1 : 0;
  return o;
}

void fill_result(const cospdnn_result& r, int64_t S, CompatNetResult* out) {
  out->categories.resize(static_cast<size_t>(r.n_categories));
  out->ms_total = r.ms_total;
  out->edges_per_sec = r.edges_per_sec;
  out->flagged = r.n_flagged;
  (void)S;
}

int current_device() {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) device = 0;
  return device;
}

}  // namespace

int graph_challenge_run(const std::vector<std::vector<float>>& input, const std::vector<std::vector<float>>& weight,
                        const std::vector<std::vector<int>>& row_idx, const std::vector<std::vector<int>>& ptr,
                        const std::vector<std::vector<int>>& fuse_list, int batch, int neuron, float bias,
                        int blocksize_c, int blocksize_r, int gpu_id, int gpu_num, CompatNetResult* out) {
  if (!out) return COSPDNN_ERR_ARG;
  *out = CompatNetResult();
  out->live.assign(weight.size() + 1, 0);
  cospdnn_engine* eng = nullptr;
  int rc = cospdnn_engine_create(current_device(), &eng);
  if (rc != COSPDNN_OK) return rc;
  const cospdnn_opts o = compat_run_opts();
  cospdnn_net* net = nullptr;
  int64_t S = 0;
  rc = build_compat_net(eng, input, weight, row_idx, ptr, fuse_list, batch, neuron, bias, blocksize_c, blocksize_r,
                        gpu_id, gpu_num, o, &net, &S);
  if (rc == COSPDNN_OK && net) {
    out->categories.assign(static_cast<size_t>(S), 0);
    cospdnn_result r;
    std::memset(&r, 0, sizeof(r));
    r.categories = out->categories.data();
    r.live = out->live.data();
    rc = cospdnn_run(&net, 1, &o, &r);
    if (rc == COSPDNN_OK) fill_result(r, S, out);
  }
  cospdnn_net_destroy(net);
  cospdnn_engine_destroy(eng);
  return rc;
}

int graph_challenge_run_co(const std::vector<std::vector<float>>* input, const std::vector<std::vector<float>>* weight,
                           const std::vector<std::vector<int>>* row_idx, const std::vector<std::vector<int>>* ptr,
                           const std::vector<std::vector<int>>* fuse_list, const int* batch, const int* neuron,
                           const float* bias, int blocksize_c, int blocksize_r, int gpu_id, int gpu_num,
                           CompatNetResult* out) {
  if (!out) return COSPDNN_ERR_ARG;
  cospdnn_engine* eng = nullptr;
  int rc = cospdnn_engine_create(current_device(), &eng);
  if (rc != COSPDNN_OK) return rc;
  cospdnn_opts o = compat_run_opts();
  o.batch_concurrent = 2;  // both main passes at once, on two side streams
  cospdnn_net* nets[2] = {nullptr, nullptr};
  int64_t S[2] = {0, 0};
  for (int k = 0; k < 2; ++k) {
    out[k] = CompatNetResult();
    out[k].live.assign(weight[k].size() + 1, 0);
    if (rc == COSPDNN_OK)
      rc = build_compat_net(eng, input[k], weight[k], row_idx[k], ptr[k], fuse_list[k], batch[k], neuron[k], bias[k],
                            blocksize_c, blocksize_r, gpu_id, gpu_num, o, &nets[k], &S[k]);
  }
  if (rc == COSPDNN_OK) {
    cospdnn_result r[2];
    std::memset(r, 0, sizeof(r));
    cospdnn_net* run[2];
    int which[2], n = 0;
    for (int k = 0; k < 2; ++k) {
      if (!nets[k]) continue;
      out[k].categories.assign(static_cast<size_t>(S[k]), 0);
      r[n].categories = out[k].categories.data();
      r[n].live = out[k].live.data();
      run[n] = nets[k];
      which[n++] = k;
    }
    if (n > 0) rc = cospdnn_run(run, n, &o, r);
    for (int i = 0; rc == COSPDNN_OK && i < n; ++i) fill_result(r[i], S[which[i]], &out[which[i]]);
  }
  for (cospdnn_net* nt : nets) cospdnn_net_destroy(nt);
  cospdnn_engine_destroy(eng);
  return rc;
}

int reference_bcsc_from_csc(const std::vector<int32_t>& colptr, const std::vector<int32_t>& rows,
                            const std::vector<float>& vals, int neuron, int blocksize_c, int blocksize_r,
                            std::vector<float>& values, std::vector<int>& row_idx, std::vector<int>& ptr) {
  if (neuron <= 0 || blocksize_c <= 0 || blocksize_r <= 0 || static_cast<int>(colptr.size()) != neuron + 1)
    return COSPDNN_ERR_ARG;
  values.clear();
  row_idx.clear();
  const int nbc = (neuron + blocksize_c - 1) / blocksize_c;
  ptr.assign(1, 0);
  std::vector<int> windows;
  for (int c = 0; c < nbc; ++c) {
    const int j0 = c * blocksize_c, j1 = std::min(neuron, j0 + blocksize_c);
    // the row windows [w bs_r, (w + 1) bs_r) that any output of the block column reads
    windows.clear();
    for (int j = j0; j < j1; ++j)
      for (int32_t q = colptr[j]; q < colptr[j + 1]; ++q) {
        if (rows[q] < 0 || rows[q] >= neuron) return COSPDNN_ERR_FORMAT;
        windows.push_back(rows[q] / blocksize_r);
      }
    std::sort(windows.begin(), windows.end());
    windows.erase(std::unique(windows.begin(), windows.end()), windows.end());
    const size_t base = values.size();
    values.resize(base + windows.size() * static_cast<size_t>(blocksize_c) * blocksize_r, 0.0f);
    for (size_t b = 0; b < windows.size(); ++b) row_idx.push_back(windows[b] * blocksize_r);
    for (int j = j0; j < j1; ++j)
      for (int32_t q = colptr[j]; q < colptr[j + 1]; ++q) {
        const int w = rows[q] / blocksize_r;
        const size_t b = static_cast<size_t>(std::lower_bound(windows.begin(), windows.end(), w) - windows.begin());
        // [out][in] inside a block
        values[base + (b * blocksize_c + (j - j0)) * blocksize_r + (rows[q] - w * blocksize_r)] += vals[q];
      }
    ptr.push_back(static_cast<int>(row_idx.size()));
  }
  return COSPDNN_OK;
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
  // both networks on one engine in one concurrent run (the reference co-schedules them in one
  // grid per layer, cospdnn_double.cu:208-1213)
  const std::vector<std::vector<float>> in[2] = {input1, input2};
  const std::vector<std::vector<float>> wt[2] = {weight1, weight2};
  const std::vector<std::vector<int>> ri[2] = {row_idx1, row_idx2};
  const std::vector<std::vector<int>> pt[2] = {ptr1, ptr2};
  const std::vector<std::vector<int>> fu[2] = {fuse_list1, fuse_list2};
  const int batch[2] = {batch1, batch2}, neuron[2] = {neuron1, neuron2};
  const float bias[2] = {bias1, bias2};
  cospdnn::CompatNetResult r[2];
  const int rc = cospdnn::graph_challenge_run_co(in, wt, ri, pt, fu, batch, neuron, bias, blocksize_c, blocksize_r,
                                                  gpu_id, gpu_num, r);
  if (rc != COSPDNN_OK) die(rc);
  std::cout << "categories = " << r[0].categories.size() << " and " << r[1].categories.size() << std::endl;
  // one concurrent run: both results report its time
  std::cout << "Total Kernel Exec Time = " << std::max(r[0].ms_total, r[1].ms_total) << "ms" << std::endl;
  const std::string& path = cospdnn::compat_options().categories_path;
  if (!path.empty()) {
    const int per[2] = {(batch1 + gpu_num - 1) / std::max(1, gpu_num), (batch2 + gpu_num - 1) / std::max(1, gpu_num)};
    std::ofstream f(path);
    for (int k = 0; k < 2; ++k)
      for (int64_t c : r[k].categories) f << (k + 1) << "\t" << (c + static_cast<int64_t>(gpu_id) * per[k] + 1) << "\n";
  }
}
