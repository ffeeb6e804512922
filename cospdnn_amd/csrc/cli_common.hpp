// Shared pieces of the two command-line drivers (cospdnn.out, cospdnn_double.out),
// which keep the reference's positional arguments (src/cospdnn.cpp:80-93,
// src/cospdnn_double.cpp:81-96) and its stdout lines.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "cospdnn.h"
#include "cospdnn_compat.hpp"

namespace cli {

struct Flags {
  int precision = COSPDNN_FP16;
  int guard_layer = 12;
  bool graph = false;
  bool use_fuse = true;
  bool allow_uneven = false;
  bool bias_set = false;
  float bias = 0.0f;
  std::string categories;
};

// Trailing --key=value options (extensions; the positional arguments are the reference's).
inline bool parse_flags(int argc, char** argv, int first, Flags& f) {
  for (int i = first; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--precision=fp32") f.precision = COSPDNN_FP32;
    else if (a == "--precision=fp16") f.precision = COSPDNN_FP16;
    else if (a.rfind("--guard-layer=", 0) == 0) f.guard_layer = std::atoi(a.c_str() + 14);
    else if (a == "--graph") f.graph = true;
    else if (a == "--no-fuse") f.use_fuse = false;
    else if (a == "--allow-uneven") f.allow_uneven = true;
    else if (a.rfind("--bias=", 0) == 0) { f.bias = std::strtof(a.c_str() + 7, nullptr); f.bias_set = true; }
    else if (a.rfind("--categories=", 0) == 0) f.categories = a.substr(13);
    else {
      std::cerr << "unknown option " << a << std::endl;
      return false;
    }
  }
  return true;
}

// Bias by neuron count (src/cospdnn.cpp:97-102); other sizes get 0 like the reference's map.
inline float bias_for(int neuron) {
  std::map<int, float> bias_map = {{65536, -0.45f}, {16384, -0.4f}, {4096, -0.35f}, {1024, -0.3f}};
  auto it = bias_map.find(neuron);
  return it == bias_map.end() ? 0.0f : it->second;
}

struct Dataset {
  int neuron = 0, layer = 0, batch = 0;
  std::vector<std::vector<float>> input;
  std::vector<std::vector<int32_t>> colptr, rows;
  std::vector<std::vector<float>> vals;
};

// Reads input, fuse lists and weights the way the reference main() does; on failure prints the
// reference's message and exits with its code (missing file: -1, short fuse file: 0).
inline void load_dataset(const std::string& dir, int neuron, int batch, int layer, bool use_fuse, Dataset& d) {
  d.neuron = neuron;
  d.layer = layer;
  d.batch = batch;
  std::string msg;
  int64_t nread = 0;
  int rc = cospdnn::read_input_tsv(dir, neuron, batch, d.input, &nread, &msg);
  if (rc != COSPDNN_OK) {
    std::cout << msg << "\n";
    std::exit(-1);
  }
  std::cout << "Read Input success! read_numeber = " << nread << std::endl;
  std::vector<std::vector<int>> fuse;
  if (use_fuse) {
    rc = cospdnn::read_fuse_tsv(dir, neuron, layer, fuse, &msg);
    if (rc == COSPDNN_ERR_IO) {
      std::cout << msg << "\n";
      std::exit(-1);
    }
    if (rc != COSPDNN_OK) {
      std::cout << msg << "\n";
      std::exit(0);
    }
    std::cout << "Read fuse success!" << std::endl;
  }
  d.colptr.resize(layer);
  d.rows.resize(layer);
  d.vals.resize(layer);
  for (int l = 0; l < layer; ++l) {
    const std::string wf = cospdnn::weight_file_name(dir, neuron, l);
    rc = cospdnn::read_weight_tsv_csc(wf, neuron, d.colptr[l], d.rows[l], d.vals[l], &msg);
    if (rc != COSPDNN_OK) {
      std::cout << msg << "\n";
      std::exit(-1);
    }
    if (use_fuse)
      for (auto& r : d.rows[l]) r = fuse[l][r];  // the kernel reads activation row fuse[row] (kernel.cu:310)
  }
  std::cout << "weight len: " << layer << std::endl;
}

// A network for samples [s0, s1) of the dataset on `eng`.
inline int make_net(cospdnn_engine* eng, const Dataset& d, int64_t s0, int64_t s1, float bias, const Flags& f,
                    cospdnn_net** out) {
  int rc = cospdnn_net_create(eng, d.neuron, d.layer, s1 - s0, bias, out);
  for (int l = 0; rc == COSPDNN_OK && l < d.layer; ++l)
    rc = cospdnn_net_set_layer_csc(*out, l, d.colptr[l].data(), d.rows[l].data(), d.vals[l].data());
  if (rc != COSPDNN_OK) return rc;
  std::vector<int64_t> smp;
  std::vector<int32_t> neu;
  std::vector<float> val;
  for (int64_t s = s0; s < s1; ++s)
    for (int j = 0; j < d.neuron; ++j)
      if (d.input[s][j] != 0.0f) {
        smp.push_back(s - s0);
        neu.push_back(j);
        val.push_back(d.input[s][j]);
      }
  rc = cospdnn_net_set_input_coo(*out, static_cast<int64_t>(smp.size()), smp.data(), neu.data(), val.data());
  if (rc != COSPDNN_OK) return rc;
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  (void)f;
  return cospdnn_net_upload(*out, &o);
}

inline cospdnn_opts run_opts(const Flags& f) {
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  o.precision = f.precision;
  o.guard_layer = f.guard_layer;
  o.use_graph = f.graph ? 1 : 0;
  return o;
}

[[noreturn]] inline void die(int rc) {
  std::fprintf(stderr, "cospdnn error %d (%s): %s\n", rc, cospdnn_strerror(rc), cospdnn_last_error());
  std::exit(EXIT_FAILURE);
}

}  // namespace cli
