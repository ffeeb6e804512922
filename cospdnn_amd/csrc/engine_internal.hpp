// Internal types of the engine, shared by engine.cpp (forward passes, C ABI) and
// tuner.cpp (autotuner and cost-model MLP).  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cospdnn.h"
#include "format.hpp"
#include "kernels.hpp"

namespace cospdnn_detail {

using namespace cospdnn;

// Records the detail of a failure for cospdnn_last_error() and returns code.
int fail(int code, const std::string& msg);
const char* last_error_text();


#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(COSPDNN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CALL_TRY(expr)        \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != COSPDNN_OK) return rc_; \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool own = true;  // false: a view into memory the buffer does not own (never freed here)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    own = true;
  }
  void borrow(void* q, size_t n) {
    reset();
    p = q;
    bytes = n;
    own = false;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

int dalloc(DevBuf& b, size_t bytes);

template <typename T>
inline int dupload(DevBuf& b, const std::vector<T>& v) {
  CALL_TRY(dalloc(b, sizeof(T) * v.size()));
  if (!v.empty()) HIP_TRY(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return COSPDNN_OK;
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct LayerStore {
  DevBuf out_idx, optr, orow, oval, kptr, gidx, wt, wt32, wmask;
  DevBuf bptr, bgidx, bwt, bmask;  // band image (order 6)
  DevBuf sgoff, sid, sval;  // slab image (sample-slab kernel)
  DevBuf out_logical;  // N int32: the neuron id of each activation row this layer writes
  DevLayer view;
  bool has_tiles = false;
  bool has_tiles32 = false;  // FP32 tile image present
  double waste = 0.0;
  bool fp16_exact = true;
  bool sat_stable = false;  // an all-32 input row gives an all-32 output row in FP32 and FP64
  int64_t nnz = 0;
  int max_list = 0;  // the longest column (input rows of one output)
  // Streamed MFMA image (the weight budget was exceeded): kptr | gidx | mask words (or FP16
  // fragments) in pinned host memory, copied into one of the network's two device slots one
  // layer ahead of its SpMM (the reference streams every layer so, cospdnn.cu:170-210)
  bool streamed = false;
  bool stream_mask = false;  // the streamed image holds mask words (else FP16 fragments)
  void* host_img = nullptr;  // hipHostMalloc
  size_t img_bytes = 0, off_gidx = 0, off_w = 0;
  LayerStore() = default;
  LayerStore(const LayerStore&) = delete;
  LayerStore& operator=(const LayerStore&) = delete;
  ~LayerStore() {
    if (host_img) (void)hipHostFree(host_img);
  }
};

struct RunBufs {
  int64_t cap = 0, ld = 0;
  int L = 0;
  int64_t gen = 0;  // bumped whenever the buffers are reallocated (captured graphs hold their addresses)
  DevBuf act16[2], act32[2], act64[2];
  DevBuf act16x;  // FP16 networks of at most 4 samples per slab: layout conversions next to slab layers
  DevBuf flags;  // 6 arrays of ld bytes: side s, kind k at (s*3+k)*ld
  DevBuf ids;    // 2 x ld int32
  DevBuf qmap;   // ld int32: output position of each input position (-1: removed)
  DevBuf qsrc;   // ld int32: input position of each output position (compacting FP16 scans)
  DevBuf counts; // L+2 int32
  DevBuf flag_ids, flag_count, flag_mark;
  DevBuf work;   // MFMA work-queue counter
  DevBuf in16;   // recompute input
  DevBuf ret_ids, ret_count;  // rows retired as saturated (guard passes)
  uint8_t* f(int side, int kind) const { return flags.as<uint8_t>() + static_cast<int64_t>(side * 3 + kind) * ld; }
  int32_t* id(int side) const { return ids.as<int32_t>() + static_cast<int64_t>(side) * ld; }
};

int ensure_bufs(RunBufs& rb, int64_t cap, int L, int N, int precision, bool need_in16, hipStream_t st);

// The fixed device array behind cospdnn_run_noalloc (noalloc.hip): its address on `device`,
// looked up once per device (with that device current) and kept, and its size in bytes.
int noalloc_scratch(int device, void** ptr, size_t* bytes);
// Serialises the calls that work in one device's fixed array.
std::mutex& noalloc_mutex(int device);

}  // namespace cospdnn_detail

using cospdnn_detail::DevBuf;
using cospdnn_detail::LayerStore;
using cospdnn_detail::RunBufs;

struct GraphCache {
  hipGraphExec_t exec = nullptr;
  int precision = -1, guard_layer = 0, band_from = 0, use_mfma = 0, mfma32 = 0, order = 0, groups = 0, keep_ppm = 0, gather = 0, bits = 0, wmask = 0, slab = 0, grid_div = 1;
  int64_t plan_version = -1;
  // the device-side guard pipeline captured with the main pass (0: main pass only)
  int dev_guard = 0, guard_layer32 = 0, guard_tail = 0;
  int64_t gen_main = -1, gen_rec = -1, gen_rep = -1;
  float band = 0.0f;
  ~GraphCache() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

// One hipGraph holding the concurrent main passes of several networks (fork/join over
// the engine's side streams), valid for that list of networks and their plans.
struct BatchGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<cospdnn_net*> nets;
  std::vector<int64_t> plan_versions;
  std::vector<int> chain;  // chain (side stream) of each network
  // per network: device-side guard captured, its checkpoint and tail, buffer generations
  std::vector<std::array<int64_t, 6>> guard_keys;
  GraphCache key;  // options of the captured passes (its exec stays null)
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    nets.clear();
    plan_versions.clear();
    chain.clear();
    guard_keys.clear();
  }
  ~BatchGraph() { reset(); }
};

struct cospdnn_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int64_t hbm_bytes = 0;
  int64_t tile_bytes = 0;  // MFMA tile images of this engine's networks (bounded at upload)
  int bpc_mfma = 1, bpc_mfma_resident = 1, bpc_simt16 = 1, bpc_simt1632 = 1, bpc_simt32 = 1, bpc_m32h = 1, bpc_m32f = 1,
      bpc_simt1664 = 1, bpc_simt64 = 1, bpc_rows = 1;
  // concurrent runs of several networks
  std::vector<hipStream_t> side;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  hipEvent_t batch_ev[3] = {nullptr, nullptr, nullptr};
  BatchGraph batch_graph;
};

struct cospdnn_net {
  cospdnn_engine* eng = nullptr;
  int N = 0, L = 0;
  int64_t S = 0;
  float bias = 0.0f;
  // host-side definition
  std::vector<cospdnn::CscLayer> layers;
  std::vector<char> layer_set;
  bool gen_layers = false;
  int gen_kind = 0, gen_K = 32;
  float gen_wval = 0.0625f;
  uint64_t gen_seed = 0;
  int input_mode = 0;  // 0 none, 1 host image, 2 generated
  std::vector<_Float16> input_img;  // [N][ld_store]
  bool input_fp16_exact = true;
  bool input_binary = true;  // every input value is 0 or 1
  uint64_t input_seed = 0;
  // device side
  bool uploaded = false;
  std::vector<std::unique_ptr<LayerStore>> dl;
  DevBuf store;
  int64_t ld_store = 0;
  DevBuf packed;       // binary inputs: [S][words] bit-packed sample-major copy of the store
  int words = 0;
  DevBuf bits;         // binary inputs: bit-packed copy of the store, tile-major ([ld_store/64][N][2] words)
  RunBufs main, rec;
  int64_t nnz_total = 0;
  int64_t tile_bytes = 0;  // this network's MFMA tile images (counted in the engine's too)
  bool weights_fp16_exact = true;
  GraphCache graph;
  std::vector<cospdnn_layer_plan> plan;  // per-layer launch plan of the FP16 pass (empty: options decide)
  int64_t plan_version = 0;
  RunBufs rep;                       // FP64 pass over the rows the FP32 pass leaves undecided
  // device-side guard pipeline: gs = {rows of the FP32 pass, rows of the FP64 pass, overflow},
  // gids = sample ids of the FP64 pass; fork/join events of its side stream
  DevBuf gs, gids;
  int64_t gcap32 = 0, gcap64 = 0;
  // pinned host staging for the results of a pass (hipHostMalloc)
  int32_t* hpin = nullptr;
  size_t hpin_n = 0;
  hipEvent_t gev[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> layer_ev;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double est_main_ms = -1.0;  // last measured main-pass time (chains of a concurrent run)
  // weight streaming: device slots (streamed layer l uses slot l % 2), the copy stream and its
  // events (slot contents copied / slot free again / all copies of a pass done)
  int64_t streamed_layers = 0;
  DevBuf wslot[2];
  hipStream_t copy_st = nullptr;
  hipEvent_t slot_copied[2] = {nullptr, nullptr}, slot_free[2] = {nullptr, nullptr}, copy_done = nullptr;
  ~cospdnn_net() {
    for (hipEvent_t e : {slot_copied[0], slot_copied[1], slot_free[0], slot_free[1], copy_done})
      if (e) (void)hipEventDestroy(e);
    if (copy_st) (void)hipStreamDestroy(copy_st);
    for (hipEvent_t e : layer_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : gev)
      if (e) (void)hipEventDestroy(e);
    if (hpin) (void)hipHostFree(hpin);
  }
};

namespace cospdnn_detail {

// One network's forward with the guard tiers (engine.cpp).
int run_one(cospdnn_net* net, const cospdnn_opts& o, cospdnn_result* res);

}  // namespace cospdnn_detail
