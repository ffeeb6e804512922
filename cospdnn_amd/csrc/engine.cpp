// Host engine and C ABI of the sparse-DNN inference path.
//
// One network = per-layer weight images + an FP16 copy of the input (the
// "store") in HBM.  A forward pass enqueues, per layer l,
//   scan_kernel(l)   live flags of layer l's input -> qmap, ids, counts[l+1]
//   spmm(l)          MFMA (FP16) or SIMT (FP16/FP32/FP64) SpMM with fused epilogue
// and a final scan that yields the categories.  Everything stays on the
// device: no host round trip per layer (the reference syncs and scans on the
// host each layer, cospdnn.cu:240-276).  The FP16 path recomputes the rows
// still undecided at its checkpoint in FP32, and those the FP32 tier leaves
// undecided in FP64: on a side stream forked after the checkpoint scan (one
// network, or each chain of a batch), or after the main pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_internal.hpp"
#include "gen.hpp"

using namespace cospdnn;
using namespace cospdnn_detail;

namespace cospdnn_detail {

namespace {
thread_local std::string g_last_error;
}  // namespace


int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* last_error_text() { return g_last_error.c_str(); }

int dalloc(DevBuf& b, size_t bytes) {
  b.reset();
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(COSPDNN_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return COSPDNN_OK;
}

int ensure_bufs(RunBufs& rb, int64_t cap, int L, int N, int precision, bool need_in16, hipStream_t st) {
  if (rb.cap < cap || rb.L != L) {
    for (int s = 0; s < 2; ++s) {
      rb.act16[s].reset();
      rb.act32[s].reset();
      rb.act64[s].reset();
    }
    rb.act16x.reset();
    rb.in16.reset();
    rb.cap = cap;
    rb.L = L;
    ++rb.gen;
    // room for act_ld(rows) of every rows <= cap, and for the FP32 MFMA kernel's reads up to
    // the next multiple of its 256-sample tile
    rb.ld = round_up(std::max<int64_t>(cap, 1), kTileS) + 64;
    CALL_TRY(dalloc(rb.flags, static_cast<size_t>(6 * rb.ld)));
    CALL_TRY(dalloc(rb.ids, sizeof(int32_t) * 2 * rb.ld));
    CALL_TRY(dalloc(rb.qmap, sizeof(int32_t) * rb.ld));
    CALL_TRY(dalloc(rb.qsrc, sizeof(int32_t) * rb.ld));
    CALL_TRY(dalloc(rb.counts, sizeof(int32_t) * 2 * (L + 2)));  // row counts, then kept-row counts
    CALL_TRY(dalloc(rb.flag_ids, sizeof(int32_t) * rb.ld));
    CALL_TRY(dalloc(rb.flag_count, sizeof(int32_t)));
    CALL_TRY(dalloc(rb.flag_mark, static_cast<size_t>(rb.ld)));
    CALL_TRY(dalloc(rb.work, sizeof(int32_t)));
    CALL_TRY(dalloc(rb.ret_ids, sizeof(int32_t) * rb.ld));
    CALL_TRY(dalloc(rb.ret_count, sizeof(int32_t)));
    HIP_TRY(hipMemsetAsync(rb.flags.p, 0, rb.flags.bytes, st));
  }
  // slab_rows(N) rows: the sample-blocked layout rounds each plane's rows up to a multiple of 8
  const size_t elems = static_cast<size_t>(slab_rows(N)) * rb.ld;
  if (precision == COSPDNN_FP16) {
    for (int s = 0; s < 2; ++s)
      if (!rb.act16[s].p) {
        CALL_TRY(dalloc(rb.act16[s], elems * 2));
        HIP_TRY(hipMemsetAsync(rb.act16[s].p, 0, rb.act16[s].bytes, st));
      }
    // a third buffer for the layout conversions next to slab layers of few samples per slab
    const int t = slab_samples(N);
    if (t > 0 && t <= 4 && !rb.act16x.p) {
      CALL_TRY(dalloc(rb.act16x, elems * 2));
      HIP_TRY(hipMemsetAsync(rb.act16x.p, 0, rb.act16x.bytes, st));
    }
  } else if (precision == COSPDNN_FP32) {
    for (int s = 0; s < 2; ++s)
      if (!rb.act32[s].p) {
        CALL_TRY(dalloc(rb.act32[s], elems * 4));
        HIP_TRY(hipMemsetAsync(rb.act32[s].p, 0, rb.act32[s].bytes, st));
      }
  } else {
    for (int s = 0; s < 2; ++s)
      if (!rb.act64[s].p) {
        CALL_TRY(dalloc(rb.act64[s], elems * 8));
        HIP_TRY(hipMemsetAsync(rb.act64[s].p, 0, rb.act64[s].bytes, st));
      }
  }
  if (need_in16 && !rb.in16.p) {
    CALL_TRY(dalloc(rb.in16, elems * 2));
    HIP_TRY(hipMemsetAsync(rb.in16.p, 0, rb.in16.bytes, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return COSPDNN_OK;
}

}  // namespace cospdnn_detail

namespace {

// Activation rows are stored in each layer's block order: the buffer layer l writes holds
// output slot s of that layer in row s, so the 32 outputs of a block column are 32
// consecutive rows and a wave's stores stream through neighbouring rows (a copy probe of
// the SpMM's traffic runs 10% faster into consecutive rows than into the neurons' own
// rows).  Layer l >= 1 therefore reads its inputs through the previous layer's order;
// the input store keeps neuron order.  prev_order: the previous layer's out_idx.
void remap_inputs(BlockedLayer& bl, const std::vector<int32_t>& prev_order, int N) {
  std::vector<int32_t> row_of(N);
  for (int s = 0; s < N; ++s) row_of[prev_order[s]] = s;
  for (int32_t& r : bl.orow) r = row_of[r];
  for (int32_t& g : bl.gidx) g = row_of[g];
}

// The streamed MFMA image of a layer in pinned host memory: kptr | gidx | mask words (or FP16
// fragments when the layer has no mask), each part on a 256-byte boundary.
int stage_streamed_image(LayerStore& ls, const BlockedLayer& bl) {
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  ls.stream_mask = bl.has_mask;
  ls.off_gidx = al(bl.kptr.size() * sizeof(int32_t));
  ls.off_w = ls.off_gidx + al(bl.gidx.size() * sizeof(int32_t));
  const size_t wbytes = bl.has_mask ? bl.wmask.size() * sizeof(uint16_t) : bl.wt.size() * sizeof(uint16_t);
  ls.img_bytes = ls.off_w + al(wbytes);
  HIP_TRY(hipHostMalloc(&ls.host_img, ls.img_bytes, hipHostMallocDefault));
  char* h = static_cast<char*>(ls.host_img);
  std::memcpy(h, bl.kptr.data(), bl.kptr.size() * sizeof(int32_t));
  std::memcpy(h + ls.off_gidx, bl.gidx.data(), bl.gidx.size() * sizeof(int32_t));
  std::memcpy(h + ls.off_w, bl.has_mask ? static_cast<const void*>(bl.wmask.data()) : static_cast<const void*>(bl.wt.data()),
              wbytes);
  ls.streamed = true;
  return COSPDNN_OK;
}

// Device images of one layer from its host images (built by build_layer, inputs remapped);
// streamed: the MFMA image goes to pinned host memory instead (weight budget exceeded).
int upload_layer(cospdnn_net* net, int l, const BlockedLayer& bl, bool sat_stable, bool streamed) {
  auto ls = std::make_unique<LayerStore>();
  // the first N slots hold a permutation of the neurons, the rest is padding
  std::vector<int32_t> rows(bl.out_idx.size());
  for (size_t s = 0; s < rows.size(); ++s) rows[s] = bl.out_idx[s] >= 0 ? static_cast<int32_t>(s) : -1;
  CALL_TRY(dupload(ls->out_idx, rows));
  CALL_TRY(dupload(ls->out_logical, std::vector<int32_t>(bl.out_idx.begin(), bl.out_idx.begin() + net->N)));
  CALL_TRY(dupload(ls->optr, bl.optr));
  CALL_TRY(dupload(ls->orow, bl.orow));
  CALL_TRY(dupload(ls->oval, bl.oval));
  ls->view.nblk = bl.nblk;
  ls->view.n_out = net->N;
  ls->view.out_idx = ls->out_idx.as<int32_t>();
  ls->view.optr = ls->optr.as<int32_t>();
  ls->view.orow = ls->orow.as<int32_t>();
  ls->view.oval = ls->oval.as<float>();
  ls->has_tiles = bl.has_tiles;
  ls->waste = bl.tile_waste;
  ls->fp16_exact = bl.fp16_exact;
  ls->nnz = bl.nnz;
  ls->sat_stable = sat_stable;
  for (size_t s = 0; s + 1 < bl.optr.size(); ++s) ls->max_list = std::max(ls->max_list, bl.optr[s + 1] - bl.optr[s]);
  if (bl.has_tiles && streamed) {
    CALL_TRY(stage_streamed_image(*ls, bl));
    if (bl.has_mask) ls->view.wmask_val = bl.wmask_val;
    ls->view.one_chunk = bl.nchunk == bl.nblk ? 1 : 0;
    ls->view.full_blocks = std::all_of(bl.out_idx.begin(), bl.out_idx.end(), [](int32_t o) { return o >= 0; }) ? 1 : 0;
    ++net->streamed_layers;
  } else if (bl.has_tiles) {
    CALL_TRY(dupload(ls->kptr, bl.kptr));
    CALL_TRY(dupload(ls->gidx, bl.gidx));
    CALL_TRY(dupload(ls->wt, bl.wt));
    if (bl.has_mask) {
      CALL_TRY(dupload(ls->wmask, bl.wmask));
      ls->view.wmask = ls->wmask.as<uint16_t>();
      ls->view.wmask_val = bl.wmask_val;
    }
    if (bl.has_tiles32) {
      CALL_TRY(dupload(ls->wt32, bl.wt32));
      ls->view.wt32 = ls->wt32.as<float>();
      ls->has_tiles32 = true;
    }
    ls->view.kptr = ls->kptr.as<int32_t>();
    ls->view.gidx = ls->gidx.as<int32_t>();
    ls->view.wt = ls->wt.as<_Float16>();
    if (bl.has_band) {
      // the band image: FP16 fragments only where the layer has no mask (the mask words hold it)
      CALL_TRY(dupload(ls->bptr, bl.bptr));
      CALL_TRY(dupload(ls->bgidx, bl.bgidx));
      ls->view.bands = bl.bands;
      ls->view.band_nchunk = static_cast<int>(bl.band_nchunk);
      ls->view.bptr = ls->bptr.as<int32_t>();
      ls->view.bgidx = ls->bgidx.as<int32_t>();
      if (!bl.bmask.empty()) {
        CALL_TRY(dupload(ls->bmask, bl.bmask));
        ls->view.bmask = ls->bmask.as<uint16_t>();
      } else {
        CALL_TRY(dupload(ls->bwt, bl.bwt));
        ls->view.bwt = ls->bwt.as<_Float16>();
      }
    }
    ls->view.one_chunk = bl.nchunk == bl.nblk ? 1 : 0;
    ls->view.full_blocks = std::all_of(bl.out_idx.begin(), bl.out_idx.end(), [](int32_t o) { return o >= 0; }) ? 1 : 0;
  }
  if (bl.has_slab) {
    CALL_TRY(dupload(ls->sgoff, bl.sgoff));
    CALL_TRY(dupload(ls->sid, bl.sid));
    ls->view.sgoff = ls->sgoff.as<int32_t>();
    ls->view.sid = ls->sid.as<uint16_t>();
    if (!bl.slab_uniform) {
      CALL_TRY(dupload(ls->sval, bl.sval));
      ls->view.sval = ls->sval.as<_Float16>();
    }
    ls->view.svalue = bl.slab_value;
  }
  net->nnz_total += bl.nnz;
  net->weights_fp16_exact = net->weights_fp16_exact && bl.fp16_exact;
  net->dl[l] = std::move(ls);
  return COSPDNN_OK;
}

constexpr int kMfmaGroupsPerCu = 4;  // __launch_bounds__(256, 4): four groups of four waves fit a CU

struct RunCfg {
  int precision;    // COSPDNN_FP16, COSPDNN_FP32 or COSPDNN_FP64 for a layer-major pass
  int guard_layer;  // checkpoint layer: rows not all-0/all-32 entering it are flagged (-1: none)
  float band;
  int band_from;
  int use_mfma;     // FP16 pass: MFMA kernel for tiled layers
  int mfma32;       // FP32 pass: MFMA FP32 kernel for tiled layers
  int order = 1;    // MFMA SpMM item order
  int groups_per_cu = kMfmaGroupsPerCu;
  int retire_sat = 0;  // guard passes: saturated rows leave the pass once all-32 is a fixed point of the rest
  int drop_flagged = 0;  // rows flagged at the checkpoint leave the pass (a later pass decides them)
  // Early stop (passes that are not captured in a graph): after the scan of layer check_from,
  // and every check_every layers after it, read the kept-row count and stop once it is 0.
  int check_from = -1, check_every = 16;
  int tail = 0;  // guarded passes: rows that die in the last `tail` layers are flagged too
  int stop_after_scan = -1;  // >= 0: the pass ends with the scan of this layer (no rows can be left)
  int grid_cap = 0;          // > 0: at most this many work-groups per SpMM launch
  int grid_div = 1;          // > 1: SpMM grids divided by this (at least kXcds work-groups)
  int keep_ppm = 0;          // FP16 pass: scans that only drop dead rows keep the layout when
                             // at most this many ppm of the rows are removed (0: never)
  int gather = 0;            // FP16 pass: MFMA layers after a compacting scan gather by output position
  int bits = 0;              // FP16 pass: an MFMA layer 0 reads the bit-packed copy of a binary input
  int wmask = 0;             // FP16 pass: MFMA layers read the mask image where the layer has one
  int slab = 0;              // FP16 pass: layers with a slab image run the sample-slab kernel
  int shared = 0;            // other passes run beside this one from its first layer on (batch chains)
};

constexpr int64_t kFewRows = 512;  // passes up to this many rows use the few-row SIMT kernel
constexpr double kTileBudget = 0.4;  // MFMA tile images: at most this share of the HBM per engine
// Band images (order 6) for layers whose tile image wastes more than this (random rows) and
// whose input has at most kBandMaxN rows: past that a group of 1024 outputs uses too few of a
// tile's rows for streaming all of them through LDS to beat gathering the used ones.
constexpr double kBandMinWaste = 4.0;
constexpr int kBandMaxN = 16384;

int grid_rows(const cospdnn_engine* e, int bpc, int64_t S_run, int nblk) {
  const int64_t items = (S_run + kFewRowsTile - 1) / kFewRowsTile * static_cast<int64_t>(nblk);
  const int64_t g = std::min<int64_t>(std::max<int64_t>(items, 1), static_cast<int64_t>(bpc) * e->cu_count);
  return static_cast<int>(g);
}

int grid_for(const cospdnn_engine* e, int bpc, int64_t S_run, int nblk) {
  const int64_t items = (S_run + kTileS - 1) / kTileS * static_cast<int64_t>(nblk);
  const int64_t g = std::min<int64_t>(std::max<int64_t>(items, 1), static_cast<int64_t>(bpc) * e->cu_count);
  return static_cast<int>(g);
}

// Layer l's SpMM in a pass with options cfg: the options' choice, overridden by the network's
// per-layer plan (FP16).
struct LayerLaunch {
  bool mfma = false, gather = false, rows = false, slab = false;
  int order = 1, groups = 1;
};

// Rows of the buffer layer l reads: the input store's neurons for layer 0, the previous
// layer's output slots after it.
int input_rows(const cospdnn_net* net, int l) { return l == 0 ? net->N : net->dl[l - 1]->view.n_out; }

LayerLaunch layer_launch(const cospdnn_net* net, const RunCfg& cfg, int l) {
  const bool fp16 = cfg.precision == COSPDNN_FP16;
  LayerLaunch x;
  const LayerStore& ls = *net->dl[l];
  // a streamed image holding mask words serves only passes that read the mask
  const bool tiles = ls.has_tiles && !(ls.streamed && ls.stream_mask && !cfg.wmask);
  // the sample-slab kernel needs the slab image and a slab that fits LDS
  const bool slab_ok = fp16 && ls.view.sid != nullptr && slab_samples(input_rows(net, l)) > 0;
  x.mfma = fp16 && tiles && cfg.use_mfma;
  x.slab = slab_ok && cfg.slab;
  if (x.slab) x.mfma = false;
  x.order = cfg.order;
  x.groups = cfg.groups_per_cu;
  bool compact = false;
  if (fp16 && !net->plan.empty()) {
    const cospdnn_layer_plan& pl = net->plan[l];
    if (pl.kernel != COSPDNN_PLAN_AUTO) x.slab = false;
    if (pl.kernel == COSPDNN_PLAN_SLAB) {
      x.mfma = false;
      x.slab = slab_ok;
    }
    if (pl.kernel == COSPDNN_PLAN_SIMT) x.mfma = false;
    if (pl.kernel == COSPDNN_PLAN_ROWS) {
      x.mfma = false;
      x.rows = true;
      x.order = pl.order;
      x.groups = pl.groups_per_cu;
    }
    if (pl.kernel == COSPDNN_PLAN_MFMA || pl.kernel == COSPDNN_PLAN_MFMA_COMPACT) {
      x.mfma = tiles;
      x.order = pl.order;
      x.groups = pl.groups_per_cu;
      compact = pl.kernel == COSPDNN_PLAN_MFMA_COMPACT;
    }
  }
  // the LDS-resident (order 5) and band (order 6) kernels take tiles of input positions and
  // store compacted
  x.gather = x.mfma && cfg.gather && !compact && x.order != 5 && x.order != 6;
  return x;
}

// The SpMM kernel (COSPDNN_KERNEL_*) that a pass of S_run rows with options cfg launches for
// layer l; enqueue_forward dispatches on it, and the result reports it per layer.
int pick_kernel(const cospdnn_net* net, const RunCfg& cfg, int l, int64_t S_run) {
  const bool fp16 = cfg.precision == COSPDNN_FP16;
  const bool fp64 = cfg.precision == COSPDNN_FP64;
  const bool few = S_run <= kFewRows;
  const LayerLaunch lx = layer_launch(net, cfg, l);
  const LayerStore& ls = *net->dl[l];
  if (lx.slab) {
    // layer 0 on the bit-packed binary input with one weight value: bit-sliced counts (6 count
    // planes: lists of at most 63 rows; one lane per output slot: at most 16 samples per slab;
    // no guard-band flags)
    const bool band0 = cfg.band > 0.0f && cfg.band_from <= 0;
    const int t = slab_samples(net->N);
    if (l == 0 && fp16 && cfg.bits && net->bits.p && ls.view.sid && !ls.view.sval && t >= 1 && t <= 16 &&
        ls.max_list <= 63 && !band0)
      return COSPDNN_KERNEL_SLAB_BITS;
    return COSPDNN_KERNEL_SLAB;
  }
  if (lx.mfma) {
    // the mask image where the pass reads it: resident, or streamed in mask words
    const bool mask = cfg.wmask && (ls.streamed ? ls.stream_mask : ls.view.wmask != nullptr);
    const bool band_ok = ls.view.bptr != nullptr && ((cfg.wmask && ls.view.bmask != nullptr) || ls.view.bwt != nullptr);
    // order 5: the whole input tile in LDS, where the input has at most 1024 rows (the input
    // store's neurons for layer 0, the previous layer's output slots after it)
    const int x_rows = l == 0 ? net->N : net->dl[l - 1]->view.n_out;
    if (lx.order == 6 && band_ok)
      return cfg.wmask && ls.view.bmask != nullptr ? COSPDNN_KERNEL_MFMA_BAND_MASK : COSPDNN_KERNEL_MFMA_BAND;
    if (lx.order == 5 && x_rows <= kLdsRowsMax) return mask ? COSPDNN_KERNEL_MFMA_LDS_MASK : COSPDNN_KERNEL_MFMA_LDS;
    return mask ? COSPDNN_KERNEL_MFMA_MASK : COSPDNN_KERNEL_MFMA;
  }
  if (fp16 && lx.rows) return COSPDNN_KERNEL_WAVE_ROWS;
  if (fp16 || fp64) return few ? COSPDNN_KERNEL_SIMT_ROWS : COSPDNN_KERNEL_SIMT;
  if (ls.has_tiles32 && cfg.mfma32) return COSPDNN_KERNEL_MFMA32;
  return few ? COSPDNN_KERNEL_SIMT_ROWS : COSPDNN_KERNEL_SIMT;
}

// Optional parts of a pass: its row count read from device memory (S_run is then only an
// upper bound, used for grids and buffer clears), and a callback enqueued right after the
// scan of fork_layer (the guard pipeline forks from there).
struct PassHooks {
  const int32_t* rows_dev = nullptr;
  int32_t* overflow = nullptr;  // set to 1 when *rows_dev > S_run (the pass then takes S_run rows)
  int fork_layer = -1;
  std::function<int()> at_fork;
};

// Enqueues one layer-major pass (no synchronisation). *cat_side receives the ids side
// holding the categories after the final scan.
int enqueue_forward(cospdnn_net* net, RunBufs& rb, const _Float16* input, int64_t ld_in, int64_t S_run,
                    const RunCfg& cfg, bool timing, hipStream_t st, int* cat_side, const PassHooks* hooks = nullptr) {
  cospdnn_engine* e = net->eng;
  const int L = net->L;
  HIP_TRY(launch_fill_u8(rb.f(0, 0), 1, S_run, st));
  HIP_TRY(launch_fill_u8(rb.f(0, 1), 1, S_run, st));
  HIP_TRY(launch_fill_u8(rb.f(0, 2), 0, S_run, st));
  if (L == 0) HIP_TRY(launch_input_alive(input, ld_in, net->N, S_run, rb.f(0, 0), st));
  HIP_TRY(launch_iota_i32(rb.id(0), S_run, st));
  HIP_TRY(launch_fill_i32(rb.counts.as<int32_t>() + 1, 0, 2 * L + 3, st));  // row counts and kept-row counts
  if (hooks && hooks->rows_dev)
    HIP_TRY(launch_clamp_count(hooks->rows_dev, rb.counts.as<int32_t>(), static_cast<int32_t>(S_run), hooks->overflow, st));
  else
    HIP_TRY(launch_set_i32(rb.counts.as<int32_t>(), static_cast<int32_t>(S_run), st));
  HIP_TRY(launch_set_i32(rb.flag_count.as<int32_t>(), 0, st));
  HIP_TRY(launch_fill_u8(rb.flag_mark.as<uint8_t>(), 0, S_run, st));
  HIP_TRY(launch_set_i32(rb.ret_count.as<int32_t>(), 0, st));
  const int D = cfg.guard_layer >= 0 ? std::min(cfg.guard_layer, L) : -1;
  const bool fp16 = cfg.precision == COSPDNN_FP16;
  // sat_from[l]: all-32 is a fixed point of layers l..L-1
  std::vector<char> sat_from(L + 1, 1);
  for (int l = L - 1; l >= 0; --l) sat_from[l] = sat_from[l + 1] && net->dl[l]->sat_stable;
  auto capg = [&](int g) {
    if (cfg.grid_div > 1) g = std::max(std::min(g, kXcds), g / cfg.grid_div);
    return cfg.grid_cap > 0 ? std::min(g, cfg.grid_cap) : g;
  };
  // Streamed MFMA images (weight budget exceeded): layer l's image is copied into slot l % 2 on
  // the copy stream, one layer ahead, after the SpMM that last read the slot; the SpMM waits for
  // its copy.  The copy stream forks from st and joins it at the end (a captured graph holds both).
  const bool streaming = fp16 && net->streamed_layers > 0 && net->copy_st != nullptr;
  auto streamed_mfma = [&](int l) { return l < L && net->dl[l]->streamed && layer_launch(net, cfg, l).mfma; };
  auto stream_copy = [&](int l) -> int {
    if (!streamed_mfma(l)) return COSPDNN_OK;
    const LayerStore& ls = *net->dl[l];
    HIP_TRY(hipStreamWaitEvent(net->copy_st, net->slot_free[l & 1], 0));
    HIP_TRY(hipMemcpyAsync(net->wslot[l & 1].p, ls.host_img, ls.img_bytes, hipMemcpyHostToDevice, net->copy_st));
    HIP_TRY(hipEventRecord(net->slot_copied[l & 1], net->copy_st));
    return COSPDNN_OK;
  };
  if (streaming) {
    HIP_TRY(hipEventRecord(net->slot_free[0], st));
    HIP_TRY(hipEventRecord(net->slot_free[1], st));
    HIP_TRY(hipStreamWaitEvent(net->copy_st, net->slot_free[0], 0));
    CALL_TRY(stream_copy(0));
  }
  for (int l = 0; l <= L; ++l) {
    const int cur = l & 1, nxt = cur ^ 1;
    ScanArgs sa;
    sa.counts = rb.counts.as<int32_t>();
    sa.layer = l;
    sa.alive = rb.f(cur, 0);
    sa.unsat = rb.f(cur, 1);
    sa.band = rb.f(cur, 2);
    sa.clear0 = rb.f(nxt, 0);
    sa.clear1 = rb.f(nxt, 1);
    sa.clear2 = rb.f(nxt, 2);
    sa.ids_in = rb.id(cur);
    sa.ids_out = rb.id(nxt);
    sa.qmap = rb.qmap.as<int32_t>();
    // the output-to-input position map, for a layer in gather mode
    const LayerLaunch lx = l < L ? layer_launch(net, cfg, l) : LayerLaunch();
    sa.qsrc = lx.gather || lx.slab ? rb.qsrc.as<int32_t>() : nullptr;
    sa.flag_undecided = (l == D) ? 1 : 0;
    sa.flag_ids = rb.flag_ids.as<int32_t>();
    sa.flag_count = rb.flag_count.as<int32_t>();
    sa.flag_mark = rb.flag_mark.as<uint8_t>();
    sa.work = rb.work.as<int32_t>();
    sa.retire_sat = (cfg.retire_sat && l < L && sat_from[l]) ? 1 : 0;
    sa.ret_ids = rb.ret_ids.as<int32_t>();
    sa.ret_count = rb.ret_count.as<int32_t>();
    sa.drop_flagged = cfg.drop_flagged;
    // at the checkpoint, all-32 decides a row only if it is a fixed point of the rest
    sa.flag_all_alive = (l == D && !sat_from[l]) ? 1 : 0;
    // scan l sees the deaths of layer l-1
    sa.flag_dead = (D >= 0 && cfg.tail > 0 && l >= 1 && l - 1 >= L - cfg.tail) ? 1 : 0;
    sa.kept = rb.counts.as<int32_t>() + (L + 2);
    // A kept layout only where the scan drops nothing but dead rows (all-zero, and with a
    // bias <= 0 they stay so), and never in front of the tail: the first scan that flags
    // deaths must see no holes.
    const bool keep_ok = fp16 && cfg.keep_ppm > 0 && l < L && !sa.flag_undecided && !sa.retire_sat &&
                         cfg.band <= 0.0f && net->bias <= 0.0f && !(D >= 0 && cfg.tail > 0 && l >= L - cfg.tail);
    sa.keep_ppm = keep_ok ? cfg.keep_ppm : 0;
    HIP_TRY(launch_scan(sa, S_run, st));
    if (hooks && hooks->at_fork && l == hooks->fork_layer) CALL_TRY(hooks->at_fork());
    if (l == L || l == cfg.stop_after_scan) break;
    if (cfg.check_from >= 0 && l >= cfg.check_from && (l - cfg.check_from) % std::max(1, cfg.check_every) == 0) {
      int32_t kept = 0;
      HIP_TRY(hipMemcpyAsync(&kept, rb.counts.as<int32_t>() + l + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (kept == 0) break;  // counts[l+2..L+1] stay 0: no live rows, only retired ones
    }

    const LayerStore& ls = *net->dl[l];
    // the layer's images; the mask image only where the pass reads it
    DevLayer view = ls.view;
    if (!(fp16 && cfg.wmask)) view.wmask = view.bmask = nullptr;
    const bool slot_l = streaming && streamed_mfma(l);
    if (streaming) CALL_TRY(stream_copy(l + 1));
    if (slot_l) {
      // this layer's MFMA image is in slot l % 2 once its copy is done
      char* base = static_cast<char*>(net->wslot[l & 1].p);
      view.kptr = reinterpret_cast<const int32_t*>(base);
      view.gidx = reinterpret_cast<const int32_t*>(base + ls.off_gidx);
      if (ls.stream_mask)
        view.wmask = reinterpret_cast<const uint16_t*>(base + ls.off_w);
      else
        view.wt = reinterpret_cast<const _Float16*>(base + ls.off_w);
      HIP_TRY(hipStreamWaitEvent(st, net->slot_copied[l & 1], 0));
    }
    StepArgs a;
    const bool fp64 = cfg.precision == COSPDNN_FP64;
    auto side_ptr = [&](int sd) -> void* { return fp16 ? rb.act16[sd].p : (fp64 ? rb.act64[sd].p : rb.act32[sd].p); };
    a.x = l == 0 ? static_cast<const void*>(input) : side_ptr(cur);
    a.ld_in = ld_in;  // layers >= 1: compacted buffers, act_ld(rows)
    a.y = side_ptr(nxt);
    a.bias = net->bias;
    a.counts = rb.counts.as<int32_t>();
    a.layer = l;
    a.qmap = rb.qmap.as<int32_t>();
    a.qsrc = lx.gather || lx.slab ? rb.qsrc.as<int32_t>() : nullptr;
    if (l == 0 && (lx.mfma || lx.slab) && cfg.bits && net->bits.p && input == net->store.as<_Float16>()) {
      a.xbits = net->bits.as<uint32_t>();
      a.xbits_rows = net->N;
    }
    a.alive_next = rb.f(nxt, 0);
    a.unsat_next = rb.f(nxt, 1);
    const bool band_on = cfg.band > 0.0f && l >= cfg.band_from;
    a.band_next = band_on ? rb.f(nxt, 2) : nullptr;
    a.band = band_on ? cfg.band : 0.0f;
    a.work = rb.work.as<int32_t>();
    if (timing) HIP_TRY(hipEventRecord(net->layer_ev[2 * l], st));
    const int nblk = view.nblk;
    const int kern = pick_kernel(net, cfg, l, S_run);
    const int order_l = lx.order, groups_l = lx.groups;
    // the images each kernel reads (the view built above must agree with the choice)
    const bool mask_k = kern == COSPDNN_KERNEL_MFMA_MASK || kern == COSPDNN_KERNEL_MFMA_LDS_MASK;
    if (lx.mfma && kern != COSPDNN_KERNEL_MFMA_BAND && kern != COSPDNN_KERNEL_MFMA_BAND_MASK &&
        mask_k != (view.wmask != nullptr))
      return fail(COSPDNN_ERR_INTERNAL, "layer " + std::to_string(l) + ": mask image and kernel choice disagree");
    if (lx.mfma) {
      // the unsaturated flags are read only by the checkpoint scan (layer D) and by retirement
      if (!cfg.retire_sat && l + 1 != D) a.unsat_next = nullptr;
      a.x_rows = l == 0 ? net->N : net->dl[l - 1]->view.n_out;
    }
    const Dtype t_in = l == 0 ? Dtype::F16 : (fp16 ? Dtype::F16 : (fp64 ? Dtype::F64 : Dtype::F32));
    const Dtype t_out = fp16 ? Dtype::F16 : (fp64 ? Dtype::F64 : Dtype::F32);
    switch (kern) {
      case COSPDNN_KERNEL_SLAB_BITS:
        if (a.xbits != nullptr) {
          // writes the next layer's layout itself (16 to 64 samples of a row per lane)
          a.x_rows = input_rows(net, l);
          a.x_layout = kLayoutRows;
          a.y_layout = l + 1 < L && pick_kernel(net, cfg, l + 1, S_run) == COSPDNN_KERNEL_SLAB ? kLayoutSlab : kLayoutRows;
          // the wide variant where nothing runs beside the pass (as for the slab kernel below)
          a.order = !cfg.shared && D > 0 && l < D ? kSlabWideOrder : 1;
          HIP_TRY(launch_spmm_bits(a, view, capg(e->cu_count), st));
          break;
        }
        [[fallthrough]];  // (an input other than the network's store has no bit-packed copy)
      case COSPDNN_KERNEL_SLAB: {
        // consecutive slab layers pass the activations sample-blocked by the slab size; the
        // buffers any other kernel reads or writes (and the last layer's output) stay neuron-major
        a.x_rows = input_rows(net, l);
        const int kprev = l >= 1 ? pick_kernel(net, cfg, l - 1, S_run) : -1;
        const bool prev_slab = kprev == COSPDNN_KERNEL_SLAB || kprev == COSPDNN_KERNEL_SLAB_BITS;
        const bool next_slab = l + 1 < L && pick_kernel(net, cfg, l + 1, S_run) == COSPDNN_KERNEL_SLAB;
        a.x_layout = prev_slab ? kLayoutSlab : kLayoutRows;
        a.y_layout = next_slab ? kLayoutSlab : kLayoutRows;
        // With 4 samples per slab or fewer the kernel would read (write) a few bytes per row and
        // slab of a neuron-major buffer: the layout is converted beside it instead, through the
        // third buffer (the input then, when both convert, the output goes to the input's own
        // buffer, free once converted).
        const int t = slab_samples(a.x_rows);
        const bool conv_in = fp16 && l >= 1 && !prev_slab && t <= 4 && rb.act16x.p;
        const bool conv_out = fp16 && !next_slab && t <= 4 && rb.act16x.p;
        const int64_t tiles = ((S_run + 63) / 64) * ((std::max(a.x_rows, view.n_out) + 63) / 64);
        const int cgrid = capg(static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(tiles, 4 * e->cu_count))));
        void* final_y = a.y;
        if (conv_in) {
          ConvArgs c;
          c.src = static_cast<const _Float16*>(a.x);
          c.dst = rb.act16x.as<_Float16>();
          c.rows = a.x_rows;
          c.counts = rb.counts.as<int32_t>();
          c.count_idx = l;
          c.t = t;
          c.to_slab = 1;
          HIP_TRY(launch_convert_layout(c, cgrid, st));
          a.x = rb.act16x.p;
          a.x_layout = kLayoutSlab;
        }
        if (conv_out) {
          a.y = conv_in ? side_ptr(cur) : rb.act16x.p;
          a.y_layout = kLayoutSlab;
        }
        // one work-group per CU; with fewer slabs than CUs the kernel splits each slab's outputs
        // over several work-groups (it knows the live row count, the host only its bound)
        // One sample per slab: the 768-thread kernel on the layers before the checkpoint, where
        // nothing runs beside the pass (the guard tiers fork at the checkpoint; batch chains
        // share the device throughout); its work-groups would not share a CU with them.
        a.order = !cfg.shared && D > 0 && l < D ? kSlabWideOrder : 1;
        HIP_TRY(launch_spmm_slab(a, view, capg(e->cu_count), st));
        if (conv_out) {
          ConvArgs c;
          c.src = static_cast<const _Float16*>(a.y);
          c.dst = static_cast<_Float16*>(final_y);
          c.rows = view.n_out;
          c.counts = rb.counts.as<int32_t>();
          c.count_idx = l + 1;
          c.t = t;
          c.to_slab = 0;
          HIP_TRY(launch_convert_layout(c, cgrid, st));
        }
        break;
      }
      case COSPDNN_KERNEL_MFMA_BAND:
      case COSPDNN_KERNEL_MFMA_BAND_MASK: {
        a.order = 1;
        HIP_TRY(launch_spmm_mfma_band(a, view, std::max(kXcds, capg(e->cu_count)), st));
        break;
      }
      case COSPDNN_KERNEL_MFMA_LDS:
      case COSPDNN_KERNEL_MFMA_LDS_MASK: {
        a.order = 1;
        const int lds_grid =
            capg(static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((S_run + 63) / 64, e->cu_count))));
        HIP_TRY(launch_spmm_mfma_lds(a, view, lds_grid, st));
        break;
      }
      case COSPDNN_KERNEL_MFMA:
      case COSPDNN_KERNEL_MFMA_MASK: {
        // items are strided over the waves of a grid that is resident at once
        const int64_t items = (S_run + 63) / 64 * static_cast<int64_t>(nblk);
        const int64_t resident = std::min(groups_l, e->bpc_mfma_resident);
        int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((items + 3) / 4, resident * e->cu_count)));
        grid = capg(grid);
        if (order_l >= 2) grid = std::max(grid, kXcds);  // every XCD's share needs a work-group
        a.order = order_l >= 5 ? 1 : order_l;
        HIP_TRY(launch_spmm_mfma(a, view, grid, st));
        break;
      }
      case COSPDNN_KERNEL_WAVE_ROWS: {
        // groups per CU, but never below what the work split needs (split 0: one wave per
        // 64-sample group; split 1: one work-group per XCD), whatever the options' grid cap
        const int64_t need = order_l == 1 ? kXcds : ((S_run + 63) / 64 + 3) / 4;
        const int grid = static_cast<int>(std::max<int64_t>(need, capg(groups_l * e->cu_count)));
        a.order = order_l;
        HIP_TRY(launch_spmm_wave_rows(a, view, grid, st));
        break;
      }
      case COSPDNN_KERNEL_MFMA32:
        HIP_TRY(launch_spmm_mfma32(a, view, t_in, capg(grid_for(e, l == 0 ? e->bpc_m32h : e->bpc_m32f, S_run, nblk)), st));
        break;
      case COSPDNN_KERNEL_SIMT_ROWS:
        HIP_TRY(launch_spmm_simt_rows(a, view, t_in, t_out, capg(grid_rows(e, e->bpc_rows, S_run, nblk)), st));
        break;
      case COSPDNN_KERNEL_SIMT: {
        int bpc = e->bpc_simt16;
        if (fp64) bpc = l == 0 ? e->bpc_simt1664 : e->bpc_simt64;
        if (!fp16 && !fp64) bpc = l == 0 ? e->bpc_simt1632 : e->bpc_simt32;
        HIP_TRY(launch_spmm_simt(a, view, t_in, t_out, capg(grid_for(e, bpc, S_run, nblk)), st));
        break;
      }
      default:
        return fail(COSPDNN_ERR_INTERNAL, "no SpMM kernel for layer " + std::to_string(l));
    }
    if (timing) HIP_TRY(hipEventRecord(net->layer_ev[2 * l + 1], st));
    if (slot_l) HIP_TRY(hipEventRecord(net->slot_free[l & 1], st));
  }
  if (streaming) {
    HIP_TRY(hipEventRecord(net->copy_done, net->copy_st));
    HIP_TRY(hipStreamWaitEvent(st, net->copy_done, 0));
  }
  *cat_side = (L & 1) ^ 1;
  return COSPDNN_OK;
}

// Input rows ids[0..*count) of the network into dst ([N][ldd] FP16), stream-ordered.
int gather_input_rows(cospdnn_net* net, const int32_t* ids, const int32_t* count, int64_t max_rows, _Float16* dst,
                      int64_t ldd, hipStream_t st) {
  if (net->packed.p)
    HIP_TRY(launch_gather_packed(net->packed.as<uint32_t>(), net->words, ids, count, max_rows, dst, ldd, net->N, st));
  else
    HIP_TRY(launch_gather_rows(net->store.as<_Float16>(), net->ld_store, ids, count, max_rows, dst, ldd, net->N, st));
  return COSPDNN_OK;
}

// Host view of a finished pass (call after the stream is synchronised).
struct PassOut {
  std::vector<int32_t> counts;  // 2(L+2): row counts (holes included), then kept-row counts
  std::vector<int32_t> cats;    // ids (pass-local) alive after the last layer
  std::vector<int32_t> flags;   // ids (pass-local) flagged at the checkpoint
};

// Pinned host staging of at least n int32 (contents not kept when it grows).
int host_stage(cospdnn_net* net, size_t n) {
  if (net->hpin_n >= n) return COSPDNN_OK;
  if (net->hpin) HIP_TRY(hipHostFree(net->hpin));
  net->hpin = nullptr;
  net->hpin_n = 0;
  const size_t want = std::max<size_t>(n, 4096);
  HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&net->hpin), sizeof(int32_t) * want, hipHostMallocDefault));
  net->hpin_n = want;
  return COSPDNN_OK;
}

// One pass to read back: its buffers, the ids side holding its categories, the output.
struct PassRead {
  const RunBufs* rb;
  int cat_side;
  bool sort_flags;
  PassOut* out;
};

// A fixed-size int32 array read with the counters.
struct ExtraRead {
  const void* dev;
  int n;
  int32_t* host;
};

// Reads several passes and the extra arrays with one copy kernel into pinned host memory
// and one synchronisation: per pass its row counts and kept-row counts, then its live ids,
// retired ids and flagged ids, each list as long as the device-side count says (up to the
// buffer's capacity).  plain: no pinned staging (the allocation-free entry point); the list
// lengths, then the lists, are copied into ordinary host memory, with two synchronisations.
int read_passes(cospdnn_net* net, const std::vector<PassRead>& reqs, const std::vector<ExtraRead>& extras = {},
                bool plain = false) {
  const int L = net->L;
  hipStream_t st = net->eng->stream;
  CopyJobs jobs;
  int64_t off = 0;
  auto add = [&](const int32_t* src, const int32_t* len, int64_t cap, bool counted) -> int64_t {
    if (jobs.n >= kMaxCopyJobs) return -1;
    const int64_t at = off;
    // counted lists: [count][items...]
    jobs.job[jobs.n++] = CopyJob{src, len, static_cast<int32_t>(cap), counted ? at + 1 : at, counted ? at : -1};
    off += cap + (counted ? 1 : 0);
    return at;
  };
  struct Where { int64_t counts, cats, rets, flags; };
  std::vector<Where> where(reqs.size());
  for (size_t r = 0; r < reqs.size(); ++r) {
    const RunBufs& rb = *reqs[r].rb;
    const int32_t* counts = rb.counts.as<int32_t>();
    where[r].counts = add(counts, nullptr, 2 * (L + 2), false);
    where[r].cats = add(rb.id(reqs[r].cat_side), counts + L + 1, rb.ld, true);
    where[r].rets = add(rb.ret_ids.as<int32_t>(), rb.ret_count.as<int32_t>(), rb.ld, true);
    where[r].flags = add(rb.flag_ids.as<int32_t>(), rb.flag_count.as<int32_t>(), rb.ld, true);
  }
  std::vector<int64_t> xat(extras.size());
  for (size_t k = 0; k < extras.size(); ++k) xat[k] = add(static_cast<const int32_t*>(extras[k].dev), nullptr, extras[k].n, false);
  for (const Where& w : where)
    if (w.counts < 0 || w.cats < 0 || w.rets < 0 || w.flags < 0) return fail(COSPDNN_ERR_ARG, "too many result arrays");
  for (int64_t x : xat)
    if (x < 0) return fail(COSPDNN_ERR_ARG, "too many result arrays");
  std::vector<int32_t> hplain;
  int32_t* h = nullptr;
  if (plain) {
    hplain.assign(static_cast<size_t>(off), 0);
    h = hplain.data();
    for (int j = 0; j < jobs.n; ++j)
      if (jobs.job[j].len) HIP_TRY(hipMemcpyAsync(h + jobs.job[j].count_at, jobs.job[j].len, sizeof(int32_t),
                                                  hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int j = 0; j < jobs.n; ++j) {
      const CopyJob& jb = jobs.job[j];
      const int32_t n = jb.len ? std::min(h[jb.count_at], jb.cap) : jb.cap;
      if (jb.len) h[jb.count_at] = n;
      if (n > 0) HIP_TRY(hipMemcpyAsync(h + jb.off, jb.src, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    }
  } else {
    CALL_TRY(host_stage(net, static_cast<size_t>(off)));
    h = net->hpin;
    HIP_TRY(launch_copy_jobs(jobs, h, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  for (size_t r = 0; r < reqs.size(); ++r) {
    PassOut& out = *reqs[r].out;
    const Where& w = where[r];
    out.counts.assign(h + w.counts, h + w.counts + 2 * (L + 2));
    out.cats.assign(h + w.cats + 1, h + w.cats + 1 + h[w.cats]);
    out.cats.insert(out.cats.end(), h + w.rets + 1, h + w.rets + 1 + h[w.rets]);
    std::sort(out.cats.begin(), out.cats.end());  // the compacted order is not the sample order
    out.flags.assign(h + w.flags + 1, h + w.flags + 1 + h[w.flags]);
    if (reqs[r].sort_flags) std::sort(out.flags.begin(), out.flags.end());
  }
  for (size_t k = 0; k < extras.size(); ++k) std::copy(h + xat[k], h + xat[k] + extras[k].n, extras[k].host);
  return COSPDNN_OK;
}

int read_pass(cospdnn_net* net, const RunBufs& rb, int cat_side, PassOut& out, bool sort_flags = true,
              bool plain = false) {
  return read_passes(net, {PassRead{&rb, cat_side, sort_flags, &out}}, {}, plain);
}

// The buffers of a guarded FP16 forward: the main pass, the FP32 tier (rec, up to gcap32
// rows), the FP64 tier (rep, up to gcap64 rows), the tiers' device-side counters gs
// {FP32 rows, FP64 rows, overflow} and the FP64 tier's sample ids gids.  The network's own
// (allocated on demand), or views into the allocation-free entry point's fixed array
// (`fixed`: tiers are bounded by the capacities, results are read without pinned memory).
struct GuardWs {
  RunBufs* main;
  RunBufs* rec;
  RunBufs* rep;
  int32_t* gs;
  int32_t* gids;
  int64_t gcap32, gcap64;
  bool fixed;
};

GuardWs net_ws(cospdnn_net* net) {
  return GuardWs{&net->main, &net->rec, &net->rep, net->gs.as<int32_t>(), net->gids.as<int32_t>(), net->gcap32,
                 net->gcap64, false};
}

// The FP32 tier: the rows undecided at the checkpoint, layer-major in FP32 (MFMA FP32 on
// tiled layers); rows still undecided at guard_layer32 leave it for the FP64 tier.
RunCfg tier32_cfg(const cospdnn_opts& o, int L) {
  RunCfg c;
  c.precision = COSPDNN_FP32;
  c.guard_layer = o.guard_layer32;
  c.band = 0.0f;
  c.band_from = 1 << 30;
  c.use_mfma = 0;
  c.mfma32 = 1;
  c.retire_sat = 1;
  c.drop_flagged = 1;  // the rows handed to the FP64 tier are not carried further here
  c.check_from = o.guard_layer32 >= 0 ? std::min(o.guard_layer32, L) : 16;
  c.check_every = 16;
  c.tail = o.guard_tail;
  return c;
}

// The FP64 tier (bit-identical to the FP64 oracle).
RunCfg tier64_cfg() {
  RunCfg c;
  c.precision = COSPDNN_FP64;
  c.guard_layer = -1;
  c.band = 0.0f;
  c.band_from = 1 << 30;
  c.use_mfma = 0;
  c.mfma32 = 0;
  c.retire_sat = 1;
  c.check_from = 16;
  c.check_every = 16;
  return c;
}

// Layer-major FP64 pass over the given original sample ids (bit-identical to the FP64
// oracle); alive[i] for ids[i].
int fp64_rows(cospdnn_net* net, const std::vector<int32_t>& ids, std::vector<uint8_t>& alive,
              const GuardWs* fixed = nullptr) {
  hipStream_t st = net->eng->stream;
  alive.assign(ids.size(), 0);
  if (ids.empty()) return COSPDNN_OK;
  const int32_t nrows = static_cast<int32_t>(ids.size());
  RunBufs& rb = fixed ? *fixed->rep : net->rep;
  if (fixed) {
    if (nrows > fixed->gcap64)
      return fail(COSPDNN_ERR_ARG, "cospdnn_run_noalloc: " + std::to_string(nrows) +
                                       " rows left for the FP64 tier, its fixed capacity is " +
                                       std::to_string(fixed->gcap64) + " (use cospdnn_run)");
  } else {
    CALL_TRY(ensure_bufs(rb, nrows, net->L, net->N, COSPDNN_FP64, true, st));
  }
  HIP_TRY(hipMemcpyAsync(rb.flag_ids.p, ids.data(), sizeof(int32_t) * nrows, hipMemcpyHostToDevice, st));
  HIP_TRY(launch_set_i32(rb.flag_count.as<int32_t>(), nrows, st));
  CALL_TRY(gather_input_rows(net, rb.flag_ids.as<int32_t>(), rb.flag_count.as<int32_t>(), nrows, rb.in16.as<_Float16>(),
                             rb.ld, st));
  const RunCfg cfg = tier64_cfg();
  int side = 0;
  CALL_TRY(enqueue_forward(net, rb, rb.in16.as<_Float16>(), rb.ld, nrows, cfg, false, st, &side));
  HIP_TRY(hipStreamSynchronize(st));
  PassOut p;
  CALL_TRY(read_pass(net, rb, side, p, true, fixed != nullptr));
  for (int32_t local : p.cats) alive[local] = 1;
  return COSPDNN_OK;
}

constexpr int kMaxSideStreams = 8;

int ensure_side_streams(cospdnn_engine* e, int n) {
  const int want = std::min(std::max(n, 1), kMaxSideStreams);
  while (static_cast<int>(e->side.size()) < want) {
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
    e->side.push_back(s);
  }
  return COSPDNN_OK;
}

// Checks the options against the network and fills cfg (no allocation, nothing enqueued).
int run_config(cospdnn_net* net, const cospdnn_opts& o, RunCfg& cfg) {
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  if (o.precision != COSPDNN_FP16 && o.precision != COSPDNN_FP32 && o.precision != COSPDNN_FP64)
    return fail(COSPDNN_ERR_ARG, "precision");
  if (o.spmm_order < 0 || o.spmm_order > 6) return fail(COSPDNN_ERR_ARG, "spmm_order must be 0..6");
  if (o.mfma_groups_per_cu < 1 || o.mfma_groups_per_cu > 16)
    return fail(COSPDNN_ERR_ARG, "mfma_groups_per_cu must be in 1..16");
  if (o.guard_tail < 0) return fail(COSPDNN_ERR_ARG, "guard_tail must be >= 0");
  if (o.precision == COSPDNN_FP16 && (!net->weights_fp16_exact || !net->input_fp16_exact))
    return fail(COSPDNN_ERR_PRECISION, "FP16 path needs weights and inputs exactly representable in FP16");
  if (o.precision != COSPDNN_FP16 && !net->input_fp16_exact)
    return fail(COSPDNN_ERR_PRECISION, "inputs must be exactly representable in FP16 (binary inputs are)");
  cfg = RunCfg();
  cfg.precision = o.precision;
  cfg.guard_layer = o.precision == COSPDNN_FP16 ? o.guard_layer : (o.precision == COSPDNN_FP32 ? o.guard_layer32 : -1);
  cfg.band = o.precision == COSPDNN_FP16 ? o.guard_band : 0.0f;
  cfg.band_from = o.guard_band_from;
  cfg.use_mfma = o.use_mfma;
  cfg.mfma32 = o.fp32_mfma;
  cfg.order = o.spmm_order;
  cfg.groups_per_cu = o.mfma_groups_per_cu;
  cfg.tail = o.guard_tail;
  cfg.keep_ppm = o.precision == COSPDNN_FP16 ? o.keep_holes_ppm : 0;
  cfg.gather = o.precision == COSPDNN_FP16 ? o.spmm_gather : 0;
  cfg.bits = o.precision == COSPDNN_FP16 ? o.input_bits : 0;
  cfg.wmask = o.precision == COSPDNN_FP16 ? o.weight_mask : 0;
  cfg.slab = o.precision == COSPDNN_FP16 ? o.spmm_slab : 0;
  return COSPDNN_OK;
}

// run_config, then the network's events and main-pass buffers.  Nothing is enqueued.
int run_prepare(cospdnn_net* net, const cospdnn_opts& o, RunCfg& cfg) {
  cospdnn_engine* e = net->eng;
  const int N = net->N, L = net->L;
  CALL_TRY(run_config(net, o, cfg));
  if (o.layer_timing != 0 && net->layer_ev.size() != 2 * static_cast<size_t>(L)) {
    for (hipEvent_t ev : net->layer_ev) (void)hipEventDestroy(ev);
    net->layer_ev.assign(2 * static_cast<size_t>(L), nullptr);
    for (auto& ev : net->layer_ev) HIP_TRY(hipEventCreate(&ev));
  }
  for (auto& ev : net->ev)
    if (!ev) HIP_TRY(hipEventCreate(&ev));
  CALL_TRY(ensure_bufs(net->main, net->S, L, N, cfg.precision, false, e->stream));
  return COSPDNN_OK;
}

// What a captured graph depends on besides the pass options: the device-side guard
// pipeline (if captured) and the allocations it references.
struct GraphExtra {
  int dev_guard = 0, guard_layer32 = 0, guard_tail = 0;
  int64_t gen_main = -1, gen_rec = -1, gen_rep = -1;
};

bool graph_matches(const GraphCache& gc, const RunCfg& cfg, int64_t plan_version, const GraphExtra& x = GraphExtra()) {
  return gc.exec && gc.precision == cfg.precision && gc.guard_layer == cfg.guard_layer && gc.band == cfg.band &&
         gc.band_from == cfg.band_from && gc.use_mfma == cfg.use_mfma && gc.mfma32 == cfg.mfma32 &&
         gc.order == cfg.order && gc.groups == cfg.groups_per_cu && gc.plan_version == plan_version &&
         gc.keep_ppm == cfg.keep_ppm && gc.gather == cfg.gather && gc.bits == cfg.bits && gc.wmask == cfg.wmask &&
         gc.slab == cfg.slab &&
         gc.grid_div == cfg.grid_div &&
         gc.dev_guard == x.dev_guard && gc.guard_layer32 == x.guard_layer32 && gc.guard_tail == x.guard_tail &&
         gc.gen_main == x.gen_main && gc.gen_rec == x.gen_rec && gc.gen_rep == x.gen_rep;
}

void graph_remember(GraphCache& gc, const RunCfg& cfg, int64_t plan_version, const GraphExtra& x = GraphExtra()) {
  gc.dev_guard = x.dev_guard;
  gc.guard_layer32 = x.guard_layer32;
  gc.guard_tail = x.guard_tail;
  gc.gen_main = x.gen_main;
  gc.gen_rec = x.gen_rec;
  gc.gen_rep = x.gen_rep;
  gc.precision = cfg.precision;
  gc.guard_layer = cfg.guard_layer;
  gc.band = cfg.band;
  gc.band_from = cfg.band_from;
  gc.use_mfma = cfg.use_mfma;
  gc.mfma32 = cfg.mfma32;
  gc.order = cfg.order;
  gc.groups = cfg.groups_per_cu;
  gc.keep_ppm = cfg.keep_ppm;
  gc.gather = cfg.gather;
  gc.bits = cfg.bits;
  gc.wmask = cfg.wmask;
  gc.slab = cfg.slab;
  gc.grid_div = cfg.grid_div;
  gc.plan_version = plan_version;
}

// Device-side guard tiers, enqueued on s1 once the checkpoint scan of the main pass has
// listed its undecided rows (main.flag_ids[0, gs[0])): the FP32 tier over those rows, then
// the FP64 tier over the rows the FP32 tier leaves undecided (their sample ids composed
// into gids, count gs[1]).  Both passes take their row counts from device memory and
// never synchronise with the host; more rows than the buffers hold set gs[2] (the host
// then redoes the tiers one after another).
int enqueue_device_guard(cospdnn_net* net, const GuardWs& ws, const cospdnn_opts& o, hipStream_t s1) {
  const int L = net->L;
  int32_t* gs = ws.gs;
  RunBufs& rec = *ws.rec;
  RunBufs& rep = *ws.rep;
  CALL_TRY(gather_input_rows(net, ws.main->flag_ids.as<int32_t>(), gs, ws.gcap32, rec.in16.as<_Float16>(), rec.ld, s1));
  RunCfg c32 = tier32_cfg(o, L);
  c32.check_from = -1;  // no host checks: the pass runs on the device only
  // When all-32 is a fixed point from its checkpoint on, the checkpoint scan retires or
  // drops every row still alive, so the pass ends there.
  const int d32 = c32.guard_layer;
  bool sat_rest = d32 >= 0 && d32 < L;
  for (int l = std::max(d32, 0); sat_rest && l < L; ++l) sat_rest = net->dl[l]->sat_stable;
  if (sat_rest) c32.stop_after_scan = d32;
  // few rows: one work-group per CU leaves the main pass most of the GPU
  c32.grid_cap = net->eng->cu_count;
  PassHooks h32;
  h32.rows_dev = gs;
  h32.overflow = gs + 2;
  int side = 0;
  CALL_TRY(enqueue_forward(net, rec, rec.in16.as<_Float16>(), rec.ld, ws.gcap32, c32, false, s1, &side, &h32));
  HIP_TRY(launch_compose_ids(ws.main->flag_ids.as<int32_t>(), rec.flag_ids.as<int32_t>(), rec.flag_count.as<int32_t>(),
                             static_cast<int32_t>(ws.gcap64), ws.gids, gs + 1, gs + 2, s1));
  CALL_TRY(gather_input_rows(net, ws.gids, gs + 1, ws.gcap64, rep.in16.as<_Float16>(), rep.ld, s1));
  RunCfg c64 = tier64_cfg();
  c64.check_from = -1;
  c64.grid_cap = net->eng->cu_count;
  PassHooks h64;
  h64.rows_dev = gs + 1;
  h64.overflow = gs + 2;
  CALL_TRY(enqueue_forward(net, rep, rep.in16.as<_Float16>(), rep.ld, ws.gcap64, c64, false, s1, &side, &h64));
  return COSPDNN_OK;
}

// The main pass on st with the guard tiers forked off after its checkpoint scan onto s1,
// and joined back at the end.
int enqueue_main_guarded(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, hipStream_t st, int* cat_side,
                         hipStream_t s1) {
  int32_t* gs = net->gs.as<int32_t>();
  HIP_TRY(launch_set_i32(gs + 2, 0, st));
  PassHooks h;
  h.fork_layer = cfg.guard_layer;
  h.at_fork = [&]() -> int {
    // the checkpoint's list before later scans append the rows that die in the tail layers
    HIP_TRY(launch_clamp_count(net->main.flag_count.as<int32_t>(), gs, static_cast<int32_t>(net->gcap32), gs + 2, st));
    HIP_TRY(hipEventRecord(net->gev[0], st));
    HIP_TRY(hipStreamWaitEvent(s1, net->gev[0], 0));
    CALL_TRY(enqueue_device_guard(net, net_ws(net), o, s1));
    HIP_TRY(hipEventRecord(net->gev[1], s1));
    return COSPDNN_OK;
  };
  CALL_TRY(enqueue_forward(net, net->main, net->store.as<_Float16>(), net->ld_store, net->S, cfg, false, st, cat_side, &h));
  HIP_TRY(hipStreamWaitEvent(st, net->gev[1], 0));
  return COSPDNN_OK;
}

GraphExtra guard_extra(const cospdnn_net* net, const cospdnn_opts& o) {
  GraphExtra x;
  x.dev_guard = 1;
  x.guard_layer32 = o.guard_layer32;
  x.guard_tail = o.guard_tail;
  x.gen_main = net->main.gen;
  x.gen_rec = net->rec.gen;
  x.gen_rep = net->rep.gen;
  return x;
}

// The main pass of one network on stream st (with the device-side guard tiers when
// dev_guard): eagerly, or (use_graph, no per-layer events) replayed from the network's
// cached hipGraph.
int enqueue_main(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, hipStream_t st, int* cat_side,
                 bool dev_guard = false) {
  const bool timing = o.layer_timing != 0;
  *cat_side = (net->L & 1) ^ 1;
  auto body = [&](int* side) -> int {
    if (dev_guard) return enqueue_main_guarded(net, o, cfg, st, side, net->eng->side[0]);
    return enqueue_forward(net, net->main, net->store.as<_Float16>(), net->ld_store, net->S, cfg, timing, st, side);
  };
  if (o.use_graph == 0 || timing) return body(cat_side);
  GraphCache& gc = net->graph;
  const GraphExtra x = dev_guard ? guard_extra(net, o) : GraphExtra();
  if (!graph_matches(gc, cfg, net->plan_version, x)) {
    if (gc.exec) {
      (void)hipGraphExecDestroy(gc.exec);
      gc.exec = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int side = 0;
    int rc = body(&side);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &g);
    if (rc != COSPDNN_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(ce);
    hipError_t ie = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(ie);
    graph_remember(gc, cfg, net->plan_version, x);
  }
  HIP_TRY(hipGraphLaunch(gc.exec, st));
  return COSPDNN_OK;
}

// After the main pass has finished (stream synchronised): categories, the FP32 and FP64
// guard tiers on the engine's stream, and the result fields except the times.
int finish_result(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, std::vector<int64_t>& live,
                  const std::vector<char>& alive, const std::vector<char>& first, int64_t n_flagged,
                  int64_t n_replayed, cospdnn_result* res);

int run_finish(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, int cat_side, cospdnn_result* res,
               const GuardWs* fixed = nullptr) {
  cospdnn_engine* e = net->eng;
  hipStream_t st = e->stream;
  const int N = net->N, L = net->L;
  const int64_t S = net->S;
  const bool plain = fixed != nullptr;
  RunBufs& main = fixed ? *fixed->main : net->main;
  RunBufs& rec = fixed ? *fixed->rec : net->rec;
  std::vector<char> alive(S, 0), first(S, 0);
  std::vector<int64_t> live(L + 1, -1);
  int64_t n_flagged = 0, n_replayed = 0;
  PassOut p1;
  CALL_TRY(read_pass(net, main, cat_side, p1, true, plain));
  for (int l = 0; l <= L; ++l) live[l] = p1.counts[L + 2 + l + 1];  // kept rows (holes not counted)
  if (res->positions)
    for (int l = 0; l <= L; ++l) res->positions[l] = p1.counts[l + 1];
  for (int32_t id : p1.cats) alive[id] = 1;
  first = alive;

  std::vector<int32_t> replay_ids;
  if (cfg.precision == COSPDNN_FP16 && !p1.flags.empty()) {
    // tier 2: the flagged rows again, layer-major in FP32 (MFMA FP32 on tiled layers)
    const int32_t nflag = static_cast<int32_t>(p1.flags.size());
    n_flagged = nflag;
    if (fixed) {
      if (nflag > fixed->gcap32)
        return fail(COSPDNN_ERR_ARG, "cospdnn_run_noalloc: " + std::to_string(nflag) +
                                         " rows flagged for the FP32 tier, its fixed capacity is " +
                                         std::to_string(fixed->gcap32) + " (use cospdnn_run)");
    } else {
      CALL_TRY(ensure_bufs(rec, nflag, L, N, COSPDNN_FP32, true, st));
    }
    HIP_TRY(hipMemcpyAsync(main.flag_ids.p, p1.flags.data(), sizeof(int32_t) * nflag, hipMemcpyHostToDevice, st));
    CALL_TRY(gather_input_rows(net, main.flag_ids.as<int32_t>(), main.flag_count.as<int32_t>(), nflag,
                               rec.in16.as<_Float16>(), rec.ld, st));
    const RunCfg rcfg = tier32_cfg(o, L);
    int rside = 0;
    CALL_TRY(enqueue_forward(net, rec, rec.in16.as<_Float16>(), rec.ld, nflag, rcfg, false, st, &rside));
    HIP_TRY(hipStreamSynchronize(st));
    PassOut p2;
    CALL_TRY(read_pass(net, rec, rside, p2, true, plain));
    for (int32_t id : p1.flags) alive[id] = 0;
    for (int32_t local : p2.cats) alive[p1.flags[local]] = 1;
    for (int32_t local : p2.flags) replay_ids.push_back(p1.flags[local]);
  } else if (cfg.precision == COSPDNN_FP32) {
    replay_ids = p1.flags;
  }
  if (!replay_ids.empty()) {
    // tier 3: FP64 replay, bit-identical to the FP64 oracle
    std::sort(replay_ids.begin(), replay_ids.end());
    std::vector<uint8_t> a64;
    CALL_TRY(fp64_rows(net, replay_ids, a64, fixed));
    for (size_t i = 0; i < replay_ids.size(); ++i) alive[replay_ids[i]] = static_cast<char>(a64[i]);
    n_replayed = static_cast<int64_t>(replay_ids.size());
  }
  return finish_result(net, o, cfg, live, alive, first, n_flagged, n_replayed, res);
}

// After a main pass with the device-side guard tiers (stream synchronised): the FP32 tier's
// and the FP64 tier's decisions over the checkpoint's rows, then the FP64 tier for the rows
// listed after the checkpoint (deaths in the tail layers).  An overflow of the tiers'
// buffers falls back to run_finish.
int run_finish_dev_guard(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, int cat_side,
                         cospdnn_result* res) {
  const int L = net->L;
  const int64_t S = net->S;
  // the three passes and the guard counters in two round trips
  const int side = (L & 1) ^ 1;  // the tiers run every layer
  PassOut p1, p2, p3;
  int32_t gsc[3] = {0, 0, 0};
  std::vector<int32_t> gids(net->gcap64);  // sample ids of the FP64 tier (the first gsc[1])
  CALL_TRY(read_passes(net,
                       {PassRead{&net->main, cat_side, false, &p1},  // flags in list order
                        PassRead{&net->rec, side, true, &p2}, PassRead{&net->rep, side, true, &p3}},
                       {ExtraRead{net->gs.p, 3, gsc}, ExtraRead{net->gids.p, static_cast<int>(net->gcap64), gids.data()}}));
  if (gsc[2] != 0) return run_finish(net, o, cfg, cat_side, res);
  std::vector<int64_t> live(L + 1, -1);
  for (int l = 0; l <= L; ++l) live[l] = p1.counts[L + 2 + l + 1];  // kept rows (holes not counted)
  if (res->positions)
    for (int l = 0; l <= L; ++l) res->positions[l] = p1.counts[l + 1];
  std::vector<char> alive(S, 0);
  for (int32_t id : p1.cats) alive[id] = 1;
  const std::vector<char> first = alive;
  const int32_t n32 = gsc[0], n64 = gsc[1];
  // FP32 tier: its pass-local ids are positions in the checkpoint's list
  for (int32_t i = 0; i < n32; ++i) alive[p1.flags[i]] = 0;
  for (int32_t local : p2.cats) alive[p1.flags[local]] = 1;
  // FP64 tier: its pass-local ids are positions in gids
  for (int32_t i = 0; i < n64; ++i) alive[gids[i]] = 0;
  for (int32_t local : p3.cats) alive[gids[local]] = 1;
  // rows listed after the checkpoint: straight to the FP64 tier
  std::vector<int32_t> late(p1.flags.begin() + n32, p1.flags.end());
  if (!late.empty()) {
    std::sort(late.begin(), late.end());
    std::vector<uint8_t> a64;
    CALL_TRY(fp64_rows(net, late, a64));
    for (size_t i = 0; i < late.size(); ++i) alive[late[i]] = static_cast<char>(a64[i]);
  }
  return finish_result(net, o, cfg, live, alive, first, static_cast<int64_t>(p1.flags.size()),
                       static_cast<int64_t>(n64) + static_cast<int64_t>(late.size()), res);
}

// The device-side guard tiers apply to a guarded FP16 pass with its checkpoint inside the
// network, run without per-layer events.
bool use_device_guard(const cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg) {
  return o.guard_concurrent != 0 && cfg.precision == COSPDNN_FP16 && cfg.guard_layer > 0 && cfg.guard_layer < net->L &&
         o.layer_timing == 0;
}

// Buffers, stream and events of the device-side guard tiers (before anything is enqueued:
// allocations synchronise the device).  The FP32 tier holds up to a quarter of the
// samples (at least 4096), the FP64 tier up to kFewRows.
int prepare_device_guard(cospdnn_net* net) {
  cospdnn_engine* e = net->eng;
  CALL_TRY(ensure_side_streams(e, 1));
  for (auto& ev : net->gev)
    if (!ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const int64_t S = net->S;
  net->gcap32 = std::min<int64_t>(S, std::max<int64_t>(4096, S / 4));
  net->gcap64 = std::min<int64_t>(net->gcap32, kFewRows);
  CALL_TRY(ensure_bufs(net->rec, net->gcap32, net->L, net->N, COSPDNN_FP32, true, e->stream));
  CALL_TRY(ensure_bufs(net->rep, net->gcap64, net->L, net->N, COSPDNN_FP64, true, e->stream));
  if (!net->gs.p) CALL_TRY(dalloc(net->gs, 4 * sizeof(int32_t)));
  if (net->gids.bytes < sizeof(int32_t) * static_cast<size_t>(net->gcap64))
    CALL_TRY(dalloc(net->gids, sizeof(int32_t) * static_cast<size_t>(net->gcap64)));
  return COSPDNN_OK;
}

// The result fields (categories, live counts, counters, bytes; the dump of an inner
// layer) from the final per-sample decisions.  live[l]: rows entering layer l in the
// main pass; first: the main pass's own decisions.
int finish_result(cospdnn_net* net, const cospdnn_opts& o, const RunCfg& cfg, std::vector<int64_t>& live,
                  const std::vector<char>& alive, const std::vector<char>& first, int64_t n_flagged,
                  int64_t n_replayed, cospdnn_result* res) {
  hipStream_t st = net->eng->stream;
  const int N = net->N, L = net->L;
  const int64_t S = net->S;
  int64_t ncat = 0, changed = 0;
  for (int64_t s = 0; s < S; ++s) {
    if (alive[s]) {
      if (res->categories) res->categories[ncat] = s;
      ++ncat;
    }
    changed += alive[s] != first[s];
  }
  live[L] = ncat;
  res->n_categories = ncat;
  if (res->live)
    for (int l = 0; l <= L; ++l) res->live[l] = live[l];
  res->n_flagged = n_flagged;
  res->n_replayed = n_replayed;
  res->n_changed = changed;
  res->edges = static_cast<double>(S) * static_cast<double>(net->nnz_total);
  const double ab = o.precision == COSPDNN_FP16 ? 2.0 : (o.precision == COSPDNN_FP32 ? 4.0 : 8.0);
  double bytes = 0.0;
  int64_t mfma_layers = 0;
  for (int l = 0; l < L; ++l) {
    const double lv = live[l] >= 0 ? static_cast<double>(live[l]) : static_cast<double>(S);
    bytes += 2.0 * lv * N * ab + static_cast<double>(net->dl[l]->nnz) * ab;
    const int k = pick_kernel(net, cfg, l, S);
    if (res->layer_kernel) res->layer_kernel[l] = k;
    const bool m = k != COSPDNN_KERNEL_SIMT && k != COSPDNN_KERNEL_SIMT_ROWS && k != COSPDNN_KERNEL_WAVE_ROWS &&
                   k != COSPDNN_KERNEL_SLAB && k != COSPDNN_KERNEL_SLAB_BITS;
    mfma_layers += m ? 1 : 0;
  }
  res->alg_bytes = bytes;
  res->mfma_layers = mfma_layers;
  res->streamed_layers = net->streamed_layers;
  res->ms_spmm = 0.0;
  if (o.layer_timing != 0) {
    double sum = 0.0;
    for (int l = 0; l < L; ++l) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, net->layer_ev[2 * l], net->layer_ev[2 * l + 1]));
      sum += ms;
      if (res->layer_ms) res->layer_ms[l] = ms;
    }
    res->ms_spmm = sum;
  }
  res->dump_rows = 0;
  if (o.dump_layer >= 0 && o.dump_layer < L && res->dump_vals && res->dump_ids && o.precision != COSPDNN_FP64) {
    // The activations of an inner layer are overwritten by later layers, so
    // the forward is re-run truncated after dump_layer (outside the timing).
    const int dl = o.dump_layer;
    const int saveL = net->L;
    net->L = dl + 1;
    int side = 0;
    RunCfg dcfg = cfg;
    dcfg.guard_layer = -1;
    dcfg.keep_ppm = 0;  // the dump lists the live rows of the layer
    int rc = enqueue_forward(net, net->main, net->store.as<_Float16>(), net->ld_store, S, dcfg, false, st, &side);
    net->L = saveL;
    CALL_TRY(rc);
    HIP_TRY(hipStreamSynchronize(st));
    int32_t rows = 0;
    HIP_TRY(hipMemcpy(&rows, net->main.counts.as<int32_t>() + dl + 1, sizeof(int32_t), hipMemcpyDeviceToHost));
    // layer dl wrote rows [0, rows) of activation side (dl&1)^1; their ids sit on the same side
    const int out_side = (dl & 1) ^ 1;
    const int64_t nr = std::min<int64_t>(rows, res->dump_cap_rows);
    std::vector<int32_t> ids(nr);
    if (nr > 0) HIP_TRY(hipMemcpy(ids.data(), net->main.id(out_side), sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
    const int64_t ld = act_ld(rows);  // compacted buffer of `rows` rows
    // activation row k holds neuron neuron_of[k] (the layer's block order)
    std::vector<int32_t> neuron_of(N);
    HIP_TRY(hipMemcpy(neuron_of.data(), net->dl[dl]->out_logical.p, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    if (nr > 0 && cfg.precision == COSPDNN_FP16) {
      std::vector<_Float16> buf(static_cast<size_t>(N) * nr);
      HIP_TRY(hipMemcpy2D(buf.data(), sizeof(_Float16) * nr, net->main.act16[out_side].p, sizeof(_Float16) * ld,
                          sizeof(_Float16) * nr, N, hipMemcpyDeviceToHost));
      for (int64_t r = 0; r < nr; ++r)
        for (int k = 0; k < N; ++k)
          res->dump_vals[r * N + neuron_of[k]] = static_cast<float>(buf[static_cast<size_t>(k) * nr + r]);
    } else if (nr > 0) {
      std::vector<float> buf(static_cast<size_t>(N) * nr);
      HIP_TRY(hipMemcpy2D(buf.data(), sizeof(float) * nr, net->main.act32[out_side].p, sizeof(float) * ld,
                          sizeof(float) * nr, N, hipMemcpyDeviceToHost));
      for (int64_t r = 0; r < nr; ++r)
        for (int k = 0; k < N; ++k) res->dump_vals[r * N + neuron_of[k]] = buf[static_cast<size_t>(k) * nr + r];
    }
    for (int64_t r = 0; r < nr; ++r) res->dump_ids[r] = ids[r];
    res->dump_rows = nr;
  }
  return COSPDNN_OK;
}

void set_times(cospdnn_result* res, float ms_first, float ms_all) {
  res->ms_total = ms_all;
  res->ms_recompute = ms_all - ms_first;
  res->edges_per_sec = ms_all > 0 ? res->edges / (ms_all * 1e-3) : 0.0;
}

}  // namespace

namespace cospdnn_detail {

int run_one(cospdnn_net* net, const cospdnn_opts& o, cospdnn_result* res) {
  hipStream_t st = net->eng->stream;
  RunCfg cfg;
  CALL_TRY(run_prepare(net, o, cfg));
  const bool dev_guard = use_device_guard(net, o, cfg);
  if (dev_guard) CALL_TRY(prepare_device_guard(net));
  HIP_TRY(hipEventRecord(net->ev[0], st));
  int cat_side = 0;
  CALL_TRY(enqueue_main(net, o, cfg, st, &cat_side, dev_guard));
  HIP_TRY(hipEventRecord(net->ev[1], st));
  HIP_TRY(hipStreamSynchronize(st));
  if (dev_guard)
    CALL_TRY(run_finish_dev_guard(net, o, cfg, cat_side, res));
  else
    CALL_TRY(run_finish(net, o, cfg, cat_side, res));
  HIP_TRY(hipEventRecord(net->ev[2], st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms_first = 0.0f, ms_all = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms_first, net->ev[0], net->ev[1]));
  HIP_TRY(hipEventElapsedTime(&ms_all, net->ev[0], net->ev[2]));
  set_times(res, ms_first, ms_all);
  net->est_main_ms = ms_first;
  return COSPDNN_OK;
}

// Views of one pass's buffers (as ensure_bufs would allocate them for cap rows) laid out
// from byte `at` of base, each on a 256-byte boundary; base == nullptr only measures.
// Returns the end of the last view.
size_t carve_bufs(RunBufs& rb, int64_t cap, int L, int N, int precision, bool need_in16, char* base, size_t at) {
  rb.cap = cap;
  rb.L = L;
  rb.ld = round_up(std::max<int64_t>(cap, 1), kTileS) + 64;
  const size_t ld = static_cast<size_t>(rb.ld);
  auto take = [&](DevBuf& b, size_t bytes) {
    at = static_cast<size_t>(round_up(static_cast<int64_t>(at), 256));
    if (base) b.borrow(base + at, bytes);
    at += bytes;
  };
  take(rb.flags, 6 * ld);
  take(rb.ids, sizeof(int32_t) * 2 * ld);
  take(rb.qmap, sizeof(int32_t) * ld);
  take(rb.qsrc, sizeof(int32_t) * ld);
  take(rb.counts, sizeof(int32_t) * 2 * (L + 2));
  take(rb.flag_ids, sizeof(int32_t) * ld);
  take(rb.flag_count, sizeof(int32_t));
  take(rb.flag_mark, ld);
  take(rb.work, sizeof(int32_t));
  take(rb.ret_ids, sizeof(int32_t) * ld);
  take(rb.ret_count, sizeof(int32_t));
  const size_t elems = static_cast<size_t>(slab_rows(N)) * ld;  // as ensure_bufs
  DevBuf* act = precision == COSPDNN_FP16 ? rb.act16 : (precision == COSPDNN_FP32 ? rb.act32 : rb.act64);
  const size_t esz = precision == COSPDNN_FP16 ? 2 : (precision == COSPDNN_FP32 ? 4 : 8);
  take(act[0], elems * esz);
  take(act[1], elems * esz);
  if (need_in16) take(rb.in16, elems * 2);
  return at;
}

// The guarded FP16 forward of one network in the fixed array: the main pass, then the FP32
// and FP64 tiers one after the other (as run_finish does), every buffer a view into the
// array, the results read back without pinned memory.  Calls on one device are serialised.
int run_noalloc(cospdnn_net* net, const cospdnn_opts& o, cospdnn_result* res) {
  cospdnn_engine* e = net->eng;
  hipStream_t st = e->stream;
  if (o.precision != COSPDNN_FP16) return fail(COSPDNN_ERR_ARG, "cospdnn_run_noalloc: precision must be COSPDNN_FP16");
  if (o.layer_timing != 0 || o.dump_layer >= 0)
    return fail(COSPDNN_ERR_ARG, "cospdnn_run_noalloc: no per-layer timing and no layer dump");
  RunCfg cfg;
  CALL_TRY(run_config(net, o, cfg));
  const int N = net->N, L = net->L;
  const int64_t S = net->S;
  void* base = nullptr;
  size_t cap = 0;
  CALL_TRY(noalloc_scratch(e->device, &base, &cap));
  std::lock_guard<std::mutex> lk(noalloc_mutex(e->device));
  struct Views {
    RunBufs main, rec, rep;
  } v;
  const int64_t gcap32 = std::min<int64_t>(S, std::max<int64_t>(4096, S / 4));
  const int64_t gcap64 = std::min<int64_t>(gcap32, kFewRows);
  auto carve = [&](char* b) {
    size_t at = carve_bufs(v.main, S, L, N, COSPDNN_FP16, false, b, 0);
    at = carve_bufs(v.rec, gcap32, L, N, COSPDNN_FP32, true, b, at);
    return carve_bufs(v.rep, gcap64, L, N, COSPDNN_FP64, true, b, at);
  };
  const size_t need = carve(nullptr);
  if (need > cap)
    return fail(COSPDNN_ERR_ARG, "cospdnn_run_noalloc: the network needs " + std::to_string(need) +
                                     " bytes of working memory, the fixed array holds " + std::to_string(cap) +
                                     " (use cospdnn_run)");
  carve(static_cast<char*>(base));
  // the array keeps the last call's contents: everything this call uses starts at zero
  HIP_TRY(hipMemsetAsync(base, 0, need, st));
  const GuardWs ws{&v.main, &v.rec, &v.rep, nullptr, nullptr, gcap32, gcap64, true};
  const auto t0 = std::chrono::steady_clock::now();
  int cat_side = 0;
  CALL_TRY(enqueue_forward(net, v.main, net->store.as<_Float16>(), net->ld_store, S, cfg, false, st, &cat_side));
  HIP_TRY(hipStreamSynchronize(st));
  const auto t1 = std::chrono::steady_clock::now();
  CALL_TRY(run_finish(net, o, cfg, cat_side, res, &ws));
  HIP_TRY(hipStreamSynchronize(st));
  const auto t2 = std::chrono::steady_clock::now();
  const auto ms = [](std::chrono::steady_clock::duration d) {
    return static_cast<float>(std::chrono::duration<double, std::milli>(d).count());
  };
  set_times(res, ms(t1 - t0), ms(t2 - t0));
  return COSPDNN_OK;
}

// Several networks at once: their main passes run in `chains` concurrent chains, each a
// stream of its own forked from (and joined back to) the engine's stream, either eagerly
// or from one hipGraph that holds all the branches; then each network's guard tiers.
// Networks go to chains longest first, each to the chain with the least estimated work
// (last measured main-pass time, else edges), so that short networks run in the shadow
// of long ones without the long ones, which lean on L2, running against each other.
// Every result reports the times of the whole batch.
int run_batch(cospdnn_net* const* nets, int n, const cospdnn_opts& o, cospdnn_result* res) {
  cospdnn_engine* e = nets[0]->eng;
  hipStream_t st = e->stream;
  std::vector<RunCfg> cfg(n);
  for (int i = 0; i < n; ++i) CALL_TRY(run_prepare(nets[i], o, cfg[i]));
  const int chains = std::max(1, std::min({n, o.batch_concurrent, kMaxSideStreams / 2}));
  if (o.batch_cu_share != 0 && chains > 1)
    for (int i = 0; i < n; ++i) cfg[i].grid_div = chains;
  if (chains > 1)
    for (int i = 0; i < n; ++i) cfg[i].shared = 1;
  // chain c runs its networks' main passes on side[c] and their guard tiers on side[chains + c]
  CALL_TRY(ensure_side_streams(e, 2 * chains));
  std::vector<char> dev_guard(n, 0);
  for (int i = 0; i < n; ++i) {
    dev_guard[i] = use_device_guard(nets[i], o, cfg[i]) ? 1 : 0;
    if (dev_guard[i]) CALL_TRY(prepare_device_guard(nets[i]));
  }
  std::vector<std::array<int64_t, 6>> gkeys(n);
  for (int i = 0; i < n; ++i) {
    const GraphExtra x = dev_guard[i] ? guard_extra(nets[i], o) : GraphExtra();
    gkeys[i] = {x.dev_guard, x.guard_layer32, x.guard_tail, x.gen_main, x.gen_rec, x.gen_rep};
  }
  std::vector<int> chain(n, 0);
  {
    std::vector<double> est(n);
    for (int i = 0; i < n; ++i)
      est[i] = nets[i]->est_main_ms > 0.0 ? nets[i]->est_main_ms
                                          : 1e-10 * static_cast<double>(nets[i]->S) * static_cast<double>(nets[i]->nnz_total);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return est[x] > est[y]; });
    std::vector<double> load(chains, 0.0);
    for (int i : order) {
      const int c = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      chain[i] = c;
      load[c] += est[i];
    }
  }
  for (auto* ev : {&e->batch_ev[0], &e->batch_ev[1], &e->batch_ev[2]})
    if (!*ev) HIP_TRY(hipEventCreateWithFlags(ev, hipEventDefault));
  if (static_cast<int>(e->join_ev.size()) < n) {
    const size_t old = e->join_ev.size();
    e->join_ev.resize(n, nullptr);
    for (size_t i = old; i < e->join_ev.size(); ++i) HIP_TRY(hipEventCreateWithFlags(&e->join_ev[i], hipEventDisableTiming));
  }
  if (!e->fork_ev) HIP_TRY(hipEventCreateWithFlags(&e->fork_ev, hipEventDisableTiming));
  // the side holding each network's final ids, also on graph replays (which enqueue nothing)
  std::vector<int> side(n, 0);
  for (int i = 0; i < n; ++i) side[i] = (nets[i]->L & 1) ^ 1;
  const bool graph = o.use_graph != 0 && o.layer_timing == 0;
  auto enqueue_all = [&]() -> int {
    HIP_TRY(hipEventRecord(e->fork_ev, st));
    for (int c = 0; c < chains; ++c) HIP_TRY(hipStreamWaitEvent(e->side[c], e->fork_ev, 0));
    for (int i = 0; i < n; ++i) {
      hipStream_t si = e->side[chain[i]];
      if (dev_guard[i])
        CALL_TRY(enqueue_main_guarded(nets[i], o, cfg[i], si, &side[i], e->side[chains + chain[i]]));
      else
        CALL_TRY(enqueue_forward(nets[i], nets[i]->main, nets[i]->store.as<_Float16>(), nets[i]->ld_store, nets[i]->S,
                                 cfg[i], o.layer_timing != 0, si, &side[i]));
    }
    for (int c = 0; c < chains; ++c) {
      HIP_TRY(hipEventRecord(e->join_ev[c], e->side[c]));
      HIP_TRY(hipStreamWaitEvent(st, e->join_ev[c], 0));
    }
    return COSPDNN_OK;
  };
  HIP_TRY(hipEventRecord(e->batch_ev[0], st));
  if (graph) {
    BatchGraph& bg = e->batch_graph;
    bool hit = bg.exec != nullptr && bg.nets.size() == static_cast<size_t>(n);
    for (int i = 0; hit && i < n; ++i)
      hit = bg.nets[i] == nets[i] && bg.plan_versions[i] == nets[i]->plan_version && bg.chain[i] == chain[i] &&
            bg.guard_keys[i] == gkeys[i] && graph_matches(bg.key, cfg[i], 0);
    if (!hit) {
      bg.reset();
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_all();
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(st, &g);
      if (rc != COSPDNN_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIP_TRY(ce);
      hipError_t ie = hipGraphInstantiate(&bg.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ie);
      bg.nets.assign(nets, nets + n);
      bg.plan_versions.resize(n);
      for (int i = 0; i < n; ++i) bg.plan_versions[i] = nets[i]->plan_version;
      bg.chain = chain;
      bg.guard_keys = gkeys;
      graph_remember(bg.key, cfg[0], 0);
    }
    HIP_TRY(hipGraphLaunch(bg.exec, st));
  } else {
    CALL_TRY(enqueue_all());
  }
  HIP_TRY(hipEventRecord(e->batch_ev[1], st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < n; ++i) {
    if (dev_guard[i])
      CALL_TRY(run_finish_dev_guard(nets[i], o, cfg[i], side[i], &res[i]));
    else
      CALL_TRY(run_finish(nets[i], o, cfg[i], side[i], &res[i]));
  }
  HIP_TRY(hipEventRecord(e->batch_ev[2], st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms_first = 0.0f, ms_all = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms_first, e->batch_ev[0], e->batch_ev[1]));
  HIP_TRY(hipEventElapsedTime(&ms_all, e->batch_ev[0], e->batch_ev[2]));
  for (int i = 0; i < n; ++i) set_times(&res[i], ms_first, ms_all);
  return COSPDNN_OK;
}

}  // namespace cospdnn_detail

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* cospdnn_version(void) { return "cospdnn-mi355x 0.1.0 (gfx950)"; }

const char* cospdnn_strerror(int code) {
  switch (code) {
    case COSPDNN_OK: return "ok";
    case COSPDNN_ERR_ARG: return "invalid argument";
    case COSPDNN_ERR_HIP: return "HIP runtime error";
    case COSPDNN_ERR_OOM: return "out of memory";
    case COSPDNN_ERR_FORMAT: return "malformed weights or input";
    case COSPDNN_ERR_PRECISION: return "values not exact in FP16";
    case COSPDNN_ERR_STATE: return "invalid call order";
    case COSPDNN_ERR_IO: return "file missing or unparsable";
    case COSPDNN_ERR_MISMATCH: return "networks disagree";
    case COSPDNN_ERR_INTERNAL: return "engine invariant broken";
    default: return "unknown error";
  }
}

const char* cospdnn_last_error(void) { return last_error_text(); }

const char* cospdnn_kernel_name(int kernel) {
  switch (kernel) {
    case COSPDNN_KERNEL_SIMT: return "spmm_simt_kernel";
    case COSPDNN_KERNEL_SIMT_ROWS: return "spmm_simt_rows_kernel";
    case COSPDNN_KERNEL_MFMA: return "spmm_mfma_kernel";
    case COSPDNN_KERNEL_MFMA_MASK: return "spmm_mfma_mask_kernel";
    case COSPDNN_KERNEL_MFMA_LDS: return "spmm_mfma_lds_kernel";
    case COSPDNN_KERNEL_MFMA_LDS_MASK: return "spmm_mfma_lds_mask_kernel";
    case COSPDNN_KERNEL_MFMA_BAND: return "spmm_mfma_band_kernel";
    case COSPDNN_KERNEL_MFMA_BAND_MASK: return "spmm_mfma_band_mask_kernel";
    case COSPDNN_KERNEL_WAVE_ROWS: return "spmm_wave_rows_kernel";
    case COSPDNN_KERNEL_MFMA32: return "spmm_mfma32_f32in_kernel";  // layer 0: spmm_mfma32_f16in_kernel
    // spmm_slab_kernel_<T>, or spmm_slab_uni_kernel_<T> for a layer of one weight value
    case COSPDNN_KERNEL_SLAB: return "spmm_slab_kernel";
    case COSPDNN_KERNEL_SLAB_BITS: return "spmm_bits_kernel";
    default: return nullptr;
  }
}

void cospdnn_default_opts(cospdnn_opts* o) {
  if (!o) return;
  o->precision = COSPDNN_FP16;
  o->guard_layer = 12;
  o->guard_layer32 = 20;
  o->fp32_mfma = 0;
  o->guard_band = 0.0f;
  o->guard_band_from = 1 << 30;
  o->use_mfma = 1;
  o->max_tile_waste = 64.0f;
  o->layer_timing = 0;
  o->use_graph = 0;
  o->dump_layer = -1;
  o->spmm_order = 1;
  o->mfma_groups_per_cu = 4;
  o->batch_concurrent = 2;
  o->guard_tail = 8;
  o->guard_concurrent = 1;
  o->keep_holes_ppm = 0;
  o->spmm_gather = 1;
  o->input_bits = 1;
  o->weight_mask = 1;
  o->weight_budget_mb = 0;
  o->batch_cu_share = 0;
  o->spmm_slab = 0;
}

int cospdnn_engine_create(int device, cospdnn_engine** out) {
  if (!out) return fail(COSPDNN_ERR_ARG, "out is null");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(COSPDNN_ERR_ARG, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  auto e = std::make_unique<cospdnn_engine>();
  e->device = device;
  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIP_TRY(hipDeviceGetAttribute(&e->cu_count, hipDeviceAttributeMultiprocessorCount, device));
  size_t free_b = 0, total_b = 0;
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  e->hbm_bytes = static_cast<int64_t>(total_b);
  HIP_TRY(configure_kernels());
  e->bpc_mfma = spmm_mfma_blocks_per_cu();
  e->bpc_mfma_resident = spmm_mfma_resident_per_cu();
  e->bpc_simt16 = spmm_simt_blocks_per_cu(Dtype::F16, Dtype::F16);
  e->bpc_simt1632 = spmm_simt_blocks_per_cu(Dtype::F16, Dtype::F32);
  e->bpc_simt32 = spmm_simt_blocks_per_cu(Dtype::F32, Dtype::F32);
  e->bpc_simt1664 = spmm_simt_blocks_per_cu(Dtype::F16, Dtype::F64);
  e->bpc_simt64 = spmm_simt_blocks_per_cu(Dtype::F64, Dtype::F64);
  e->bpc_m32h = spmm_mfma32_blocks_per_cu(Dtype::F16);
  e->bpc_m32f = spmm_mfma32_blocks_per_cu(Dtype::F32);
  e->bpc_rows = spmm_simt_rows_blocks_per_cu(Dtype::F64, Dtype::F64);
  // the device address of the allocation-free entry point's fixed array on this device
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  CALL_TRY(noalloc_scratch(device, &scratch, &scratch_bytes));
  *out = e.release();
  return COSPDNN_OK;
}

void cospdnn_engine_destroy(cospdnn_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  e->batch_graph.reset();
  for (hipStream_t s : e->side) (void)hipStreamDestroy(s);
  for (hipEvent_t ev : e->join_ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : e->batch_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork_ev) (void)hipEventDestroy(e->fork_ev);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int cospdnn_engine_info(cospdnn_engine* e, int* cu_count, int64_t* hbm_bytes) {
  if (!e) return fail(COSPDNN_ERR_ARG, "engine is null");
  if (cu_count) *cu_count = e->cu_count;
  if (hbm_bytes) *hbm_bytes = e->hbm_bytes;
  return COSPDNN_OK;
}

int cospdnn_net_create(cospdnn_engine* e, int N, int L, int64_t S, float bias, cospdnn_net** out) {
  if (!e || !out) return fail(COSPDNN_ERR_ARG, "null argument");
  *out = nullptr;
  if (N <= 0 || L < 0 || S <= 0 || S > (int64_t{1} << 31) - kTileS) return fail(COSPDNN_ERR_ARG, "bad N, L or S");
  if (!std::isfinite(bias)) return fail(COSPDNN_ERR_ARG, "bias not finite");
  auto net = std::make_unique<cospdnn_net>();
  net->eng = e;
  net->N = N;
  net->L = L;
  net->S = S;
  net->bias = bias;
  net->layers.resize(L);
  net->layer_set.assign(L, 0);
  net->dl.resize(L);
  net->ld_store = round_up(S, kTileS);
  *out = net.release();
  return COSPDNN_OK;
}

void cospdnn_net_destroy(cospdnn_net* net) {
  if (!net) return;
  (void)hipSetDevice(net->eng->device);
  // a batch graph that captured this network's buffers must not be replayed again
  BatchGraph& bg = net->eng->batch_graph;
  if (std::find(bg.nets.begin(), bg.nets.end(), net) != bg.nets.end()) bg.reset();
  net->eng->tile_bytes -= net->tile_bytes;
  delete net;
}

int cospdnn_net_set_layer_csc(cospdnn_net* net, int layer, const int32_t* colptr, const int32_t* rows,
                              const float* vals) {
  if (!net || !colptr || layer < 0 || layer >= net->L) return fail(COSPDNN_ERR_ARG, "bad layer or null colptr");
  if (net->uploaded) return fail(COSPDNN_ERR_STATE, "network already uploaded");
  const int N = net->N;
  if (colptr[0] != 0) return fail(COSPDNN_ERR_FORMAT, "colptr[0] != 0");
  for (int j = 0; j < N; ++j)
    if (colptr[j + 1] < colptr[j]) return fail(COSPDNN_ERR_FORMAT, "colptr not monotone");
  const int64_t nnz = colptr[N];
  if (nnz > 0 && (!rows || !vals)) return fail(COSPDNN_ERR_ARG, "null rows or vals");
  CscLayer& c = net->layers[layer];
  c.colptr.assign(colptr, colptr + N + 1);
  c.rows.assign(rows, rows + nnz);
  c.vals.assign(vals, vals + nnz);
  for (int64_t q = 0; q < nnz; ++q) {
    if (c.rows[q] < 0 || c.rows[q] >= N) return fail(COSPDNN_ERR_FORMAT, "row index out of range");
    if (!std::isfinite(c.vals[q])) return fail(COSPDNN_ERR_FORMAT, "weight not finite");
  }
  net->layer_set[layer] = 1;
  net->gen_layers = false;
  return COSPDNN_OK;
}

int cospdnn_net_set_layer_ref_bcsc(cospdnn_net* net, int layer, const float* values, int64_t n_values,
                                   const int32_t* row_idx, int64_t n_row_idx, const int32_t* ptr, int64_t n_ptr,
                                   const int32_t* fuse, int64_t n_fuse, int blocksize_c, int blocksize_r) {
  if (!net || layer < 0 || layer >= net->L || !values || !row_idx || !ptr) return fail(COSPDNN_ERR_ARG, "bad argument");
  std::vector<float> v(values, values + n_values);
  std::vector<int> ri(row_idx, row_idx + n_row_idx), pt(ptr, ptr + n_ptr);
  std::vector<int> fu;
  if (fuse) fu.assign(fuse, fuse + n_fuse);
  CscLayer c;
  if (!csc_from_reference_bcsc(v, ri, pt, fu, net->N, blocksize_c, blocksize_r, c))
    return fail(COSPDNN_ERR_FORMAT, "malformed reference BCSC");
  return cospdnn_net_set_layer_csc(net, layer, c.colptr.data(), c.rows.data(), c.vals.data());
}

int cospdnn_net_generate_layers(cospdnn_net* net, int kind, int K, float wval, uint64_t seed) {
  if (!net || (kind != COSPDNN_GEN_RANDOM && kind != COSPDNN_GEN_BANDED)) return fail(COSPDNN_ERR_ARG, "bad kind");
  if (K <= 0 || K > net->N || (kind == COSPDNN_GEN_BANDED && net->N % K != 0))
    return fail(COSPDNN_ERR_ARG, "bad K for N");
  if (net->uploaded) return fail(COSPDNN_ERR_STATE, "network already uploaded");
  net->gen_layers = true;
  net->gen_kind = kind;
  net->gen_K = K;
  net->gen_wval = wval;
  net->gen_seed = seed;
  std::fill(net->layer_set.begin(), net->layer_set.end(), 1);
  return COSPDNN_OK;
}

int cospdnn_net_set_input_dense(cospdnn_net* net, const float* y0) {
  if (!net || !y0) return fail(COSPDNN_ERR_ARG, "null argument");
  if (net->uploaded) return fail(COSPDNN_ERR_STATE, "network already uploaded");
  const int N = net->N;
  const int64_t ld = net->ld_store;
  net->input_img.assign(static_cast<size_t>(N) * ld, static_cast<_Float16>(0.0f));
  net->input_fp16_exact = true;
  net->input_binary = true;
  for (int64_t s = 0; s < net->S; ++s)
    for (int nn = 0; nn < N; ++nn) {
      const float v = y0[s * N + nn];
      if (!std::isfinite(v)) return fail(COSPDNN_ERR_FORMAT, "input not finite");
      const _Float16 h = static_cast<_Float16>(v);
      if (static_cast<float>(h) != v) net->input_fp16_exact = false;
      if (v != 0.0f && v != 1.0f) net->input_binary = false;
      net->input_img[static_cast<size_t>(nn) * ld + s] = h;
    }
  net->input_mode = 1;
  return COSPDNN_OK;
}

int cospdnn_net_set_input_coo(cospdnn_net* net, int64_t nnz, const int64_t* sample, const int32_t* neuron,
                              const float* val) {
  if (!net || nnz < 0 || (nnz > 0 && (!sample || !neuron || !val))) return fail(COSPDNN_ERR_ARG, "bad argument");
  if (net->uploaded) return fail(COSPDNN_ERR_STATE, "network already uploaded");
  const int N = net->N;
  const int64_t ld = net->ld_store;
  net->input_img.assign(static_cast<size_t>(N) * ld, static_cast<_Float16>(0.0f));
  net->input_fp16_exact = true;
  net->input_binary = true;
  for (int64_t i = 0; i < nnz; ++i) {
    if (sample[i] < 0 || sample[i] >= net->S || neuron[i] < 0 || neuron[i] >= N)
      return fail(COSPDNN_ERR_FORMAT, "input entry out of range");
    if (!std::isfinite(val[i])) return fail(COSPDNN_ERR_FORMAT, "input not finite");
    const _Float16 h = static_cast<_Float16>(val[i]);
    if (static_cast<float>(h) != val[i]) net->input_fp16_exact = false;
    if (val[i] != 0.0f && val[i] != 1.0f) net->input_binary = false;
    net->input_img[static_cast<size_t>(neuron[i]) * ld + sample[i]] = h;
  }
  net->input_mode = 1;
  return COSPDNN_OK;
}

int cospdnn_net_generate_input(cospdnn_net* net, uint64_t seed) {
  if (!net) return fail(COSPDNN_ERR_ARG, "null net");
  if (net->uploaded) return fail(COSPDNN_ERR_STATE, "network already uploaded");
  net->input_mode = 2;
  net->input_seed = seed;
  net->input_fp16_exact = true;
  net->input_binary = true;
  net->input_img.clear();
  return COSPDNN_OK;
}

int cospdnn_net_upload(cospdnn_net* net, const cospdnn_opts* opts) {
  if (!net) return fail(COSPDNN_ERR_ARG, "null net");
  if (net->uploaded) return COSPDNN_OK;
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  if (opts) o = *opts;
  for (int l = 0; l < net->L; ++l)
    if (!net->layer_set[l]) return fail(COSPDNN_ERR_STATE, "layer " + std::to_string(l) + " not set");
  if (net->input_mode == 0) return fail(COSPDNN_ERR_STATE, "input not set");
  HIP_TRY(hipSetDevice(net->eng->device));
  net->nnz_total = 0;
  net->weights_fp16_exact = true;
  {
    // generate (when synthetic) and build the host images of a batch of layers on host
    // threads, one layer per thread, then upload them in order
    const int L = net->L;
    const double budget = o.weight_budget_mb > 0 ? static_cast<double>(o.weight_budget_mb) * 1048576.0
                                                  : kTileBudget * static_cast<double>(net->eng->hbm_bytes);
    net->streamed_layers = 0;
    const int batch = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<int32_t> carry;  // out_idx of the last layer of the previous batch
    for (int l0 = 0; l0 < L; l0 += batch) {
      const int l1 = std::min(L, l0 + batch);
      std::vector<BlockedLayer> bl(l1 - l0);
      std::vector<char> ok(l1 - l0, 0), stable(l1 - l0, 0), streamed(l1 - l0, 0);
      std::vector<std::thread> th;
      for (int l = l0; l < l1; ++l)
        th.emplace_back([&, l]() {
          CscLayer gen;
          const CscLayer* csc = &net->layers[l];
          if (net->gen_layers) {
            generate_layer(net->gen_seed, net->gen_kind, net->N, net->gen_K, l, net->gen_wval, gen, 1);
            csc = &gen;
          }
          ok[l - l0] = build_layer(*csc, net->N, o.max_tile_waste, bl[l - l0]) ? 1 : 0;
          stable[l - l0] = all32_fixed_point(*csc, net->N, net->bias) ? 1 : 0;
        });
      for (auto& t : th) t.join();
      for (int l = l0; l < l1; ++l) {
        if (!ok[l - l0]) return fail(COSPDNN_ERR_FORMAT, "layer " + std::to_string(l) + ": malformed CSC");
        // tile images of random layers are ~32x their SIMT image: past the weight budget (all
        // of this engine's networks together) a layer's MFMA image is streamed from pinned host
        // memory instead (its FP32 image is dropped: the FP32 tier takes the SIMT kernel there)
        BlockedLayer& b = bl[l - l0];
        if (b.has_tiles) {
          const int64_t tb = static_cast<int64_t>(b.wt.size()) * 2 + static_cast<int64_t>(b.wt32.size()) * 4 +
                             static_cast<int64_t>(b.gidx.size()) * 4 + static_cast<int64_t>(b.wmask.size()) * 2;
          if (static_cast<double>(net->eng->tile_bytes + tb) > budget) {
            streamed[l - l0] = 1;
            b.has_tiles32 = false;
            b.wt32.clear();
          } else {
            net->eng->tile_bytes += tb;
            net->tile_bytes += tb;
          }
        }
      }
      // inputs of layer l through the row order of layer l-1 (rows of the input store: neurons);
      // then the band image (bands are ranges of those rows) of the random layers it serves
      th.clear();
      for (int l = l0; l < l1; ++l)
        th.emplace_back([&, l]() {
          BlockedLayer& b = bl[l - l0];
          if (l >= 1) remap_inputs(b, l > l0 ? bl[l - 1 - l0].out_idx : carry, net->N);
          if (b.has_tiles && !streamed[l - l0] && b.tile_waste > kBandMinWaste && net->N <= kBandMaxN)
            build_band_image(b, net->N);
          build_slab_image(b, net->N);
        });
      for (auto& t : th) t.join();
      for (int l = l0; l < l1; ++l) {
        BlockedLayer& b = bl[l - l0];
        if (!b.has_band) continue;
        const int64_t bb = static_cast<int64_t>(b.bgidx.size()) * 4 + static_cast<int64_t>(b.bptr.size()) * 4 +
                           (b.bmask.empty() ? static_cast<int64_t>(b.bwt.size()) * 2
                                            : static_cast<int64_t>(b.bmask.size()) * 2);
        if (static_cast<double>(net->eng->tile_bytes + bb) > budget) {
          b.has_band = false;
        } else {
          net->eng->tile_bytes += bb;
          net->tile_bytes += bb;
        }
      }
      for (int l = l0; l < l1; ++l) {
        CALL_TRY(upload_layer(net, l, bl[l - l0], stable[l - l0] != 0, streamed[l - l0] != 0));
        if (!net->gen_layers) net->layers[l] = CscLayer();
      }
      carry = std::move(bl[l1 - 1 - l0].out_idx);
    }
  }
  if (net->streamed_layers > 0) {
    // two device slots of the largest streamed image, the copy stream and its events
    size_t most = 0;
    for (const auto& ls : net->dl)
      if (ls->streamed) most = std::max(most, ls->img_bytes);
    for (auto& slot : net->wslot) CALL_TRY(dalloc(slot, most));
    if (!net->copy_st) HIP_TRY(hipStreamCreateWithFlags(&net->copy_st, hipStreamNonBlocking));  // default priority
    for (hipEvent_t* ev : {&net->slot_copied[0], &net->slot_copied[1], &net->slot_free[0], &net->slot_free[1],
                           &net->copy_done})
      if (!*ev) HIP_TRY(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  }
  CALL_TRY(dalloc(net->store, sizeof(_Float16) * static_cast<size_t>(net->N) * net->ld_store));
  if (net->input_mode == 1) {
    HIP_TRY(hipMemcpy(net->store.p, net->input_img.data(), net->store.bytes, hipMemcpyHostToDevice));
    net->input_img.clear();
    net->input_img.shrink_to_fit();
  } else {
    HIP_TRY(hipMemsetAsync(net->store.p, 0, net->store.bytes, net->eng->stream));
    HIP_TRY(launch_gen_input(net->store.as<_Float16>(), net->ld_store, net->N, net->S, net->input_seed, net->eng->stream));
    HIP_TRY(hipStreamSynchronize(net->eng->stream));
  }
  net->packed.reset();
  net->bits.reset();
  net->words = (net->N + 31) / 32;
  if (net->input_binary) {
    CALL_TRY(dalloc(net->bits, sizeof(uint32_t) * static_cast<size_t>(net->N) * static_cast<size_t>(net->ld_store / 32)));
    HIP_TRY(launch_pack_cols(net->store.as<_Float16>(), net->ld_store, net->N, net->bits.as<uint32_t>(), net->eng->stream));
    CALL_TRY(dalloc(net->packed, sizeof(uint32_t) * static_cast<size_t>(net->words) * static_cast<size_t>(net->S)));
    HIP_TRY(launch_pack_rows(net->store.as<_Float16>(), net->ld_store, net->N, net->S, net->packed.as<uint32_t>(),
                             net->words, net->eng->stream));
    HIP_TRY(hipStreamSynchronize(net->eng->stream));
  }
  net->uploaded = true;
  return COSPDNN_OK;
}

int cospdnn_run(cospdnn_net* const* nets, int n_nets, const cospdnn_opts* opts, cospdnn_result* results) {
  if (!nets || n_nets <= 0 || !results) return fail(COSPDNN_ERR_ARG, "bad argument");
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  if (opts) o = *opts;
  for (int i = 0; i < n_nets; ++i) {
    if (!nets[i]) return fail(COSPDNN_ERR_ARG, "null net");
    if (nets[i]->eng != nets[0]->eng) return fail(COSPDNN_ERR_MISMATCH, "networks on different engines");
  }
  HIP_TRY(hipSetDevice(nets[0]->eng->device));
  for (int i = 0; i < n_nets; ++i)
    for (int k = 0; k < i; ++k)
      if (nets[i] == nets[k]) return fail(COSPDNN_ERR_ARG, "a network appears twice in one call");
  if (o.batch_concurrent < 0) return fail(COSPDNN_ERR_ARG, "batch_concurrent must be >= 0");
  if (n_nets > 1 && o.batch_concurrent > 0) return run_batch(nets, n_nets, o, results);
  for (int i = 0; i < n_nets; ++i) CALL_TRY(run_one(nets[i], o, &results[i]));
  return COSPDNN_OK;
}

int cospdnn_run_noalloc(cospdnn_net* net, const cospdnn_opts* opts, cospdnn_result* result) {
  if (!net || !result) return fail(COSPDNN_ERR_ARG, "bad argument");
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  if (opts) o = *opts;
  HIP_TRY(hipSetDevice(net->eng->device));
  return run_noalloc(net, o, result);
}

int cospdnn_similarity_order(int N, const int32_t* colptr, const int32_t* rows, int group, int overlap,
                             int32_t* order) {
  if (N <= 0 || !colptr || !rows || !order || group < 1 || colptr[0] != 0) return fail(COSPDNN_ERR_ARG, "bad argument");
  std::vector<std::vector<int32_t>> sr(N);
  for (int j = 0; j < N; ++j) {
    if (colptr[j + 1] < colptr[j]) return fail(COSPDNN_ERR_FORMAT, "colptr not monotone");
    sr[j].assign(rows + colptr[j], rows + colptr[j + 1]);
    for (int32_t r : sr[j])
      if (r < 0 || r >= N) return fail(COSPDNN_ERR_FORMAT, "row out of range");
    std::sort(sr[j].begin(), sr[j].end());
    sr[j].erase(std::unique(sr[j].begin(), sr[j].end()), sr[j].end());
  }
  std::vector<int32_t> ord;
  similarity_order(sr, N, group, overlap != 0, ord);
  std::copy(ord.begin(), ord.end(), order);
  return COSPDNN_OK;
}

int cospdnn_generate_layer(uint64_t seed, int kind, int N, int K, int layer, float wval, int32_t* colptr,
                           int32_t* rows, float* vals) {
  if (!colptr || !rows || !vals || N <= 0 || K <= 0 || K > N || layer < 0) return fail(COSPDNN_ERR_ARG, "bad argument");
  if (kind == COSPDNN_GEN_BANDED && N % K != 0) return fail(COSPDNN_ERR_ARG, "banded needs K | N");
  CscLayer c;
  generate_layer(seed, kind, N, K, layer, wval, c);
  std::memcpy(colptr, c.colptr.data(), sizeof(int32_t) * c.colptr.size());
  std::memcpy(rows, c.rows.data(), sizeof(int32_t) * c.rows.size());
  std::memcpy(vals, c.vals.data(), sizeof(float) * c.vals.size());
  return COSPDNN_OK;
}

int cospdnn_generate_input_rows(uint64_t seed, int N, int64_t s0, int64_t count, uint8_t* out) {
  if (!out || N <= 0 || s0 < 0 || count < 0) return fail(COSPDNN_ERR_ARG, "bad argument");
  generate_input_rows(seed, N, s0, count, out);
  return COSPDNN_OK;
}

static bool csc_from_ptrs(int N, const int32_t* colptr, const int32_t* rows, const float* vals, CscLayer& c) {
  if (!colptr || N <= 0 || colptr[0] != 0) return false;
  for (int j = 0; j < N; ++j)
    if (colptr[j + 1] < colptr[j]) return false;
  const int64_t nnz = colptr[N];
  if (nnz > 0 && (!rows || !vals)) return false;
  c.colptr.assign(colptr, colptr + N + 1);
  c.rows.assign(rows, rows + nnz);
  c.vals.assign(vals, vals + nnz);
  return true;
}

int cospdnn_layer_stats(int N, const int32_t* colptr, const int32_t* rows, const float* vals, int64_t* nblk,
                        int64_t* nchunk, double* tile_waste, int* fp16_exact) {
  CscLayer c;
  if (!csc_from_ptrs(N, colptr, rows, vals, c)) return fail(COSPDNN_ERR_ARG, "bad CSC");
  BlockedLayer bl;
  if (!build_layer(c, N, 1e30, bl)) return fail(COSPDNN_ERR_FORMAT, "malformed CSC");
  if (nblk) *nblk = bl.nblk;
  if (nchunk) *nchunk = bl.nchunk;
  if (tile_waste) *tile_waste = bl.tile_waste;
  if (fp16_exact) *fp16_exact = bl.fp16_exact ? 1 : 0;
  return COSPDNN_OK;
}

int cospdnn_layer_tiles_decode(int N, const int32_t* colptr, const int32_t* rows, const float* vals, int64_t cap,
                               int32_t* in_ids, int32_t* out_ids, float* values, int64_t* n_out) {
  CscLayer c;
  if (!csc_from_ptrs(N, colptr, rows, vals, c) || !n_out) return fail(COSPDNN_ERR_ARG, "bad argument");
  BlockedLayer bl;
  if (!build_layer(c, N, 1e30, bl)) return fail(COSPDNN_ERR_FORMAT, "malformed CSC");
  int64_t k = 0;
  for (int blk = 0; blk < bl.nblk; ++blk) {
    for (int ch = bl.kptr[blk]; ch < bl.kptr[blk + 1]; ++ch) {
      for (int m = 0; m < 2; ++m)
        for (int lane = 0; lane < 64; ++lane)
          for (int e = 0; e < 8; ++e) {
            const uint16_t h = bl.wt[((static_cast<size_t>(ch) * 2 + m) * 64 + lane) * 8 + e];
            if ((h & 0x7FFFu) == 0) continue;
            const int o = m * 16 + (lane & 15);
            const int kk = 8 * (lane >> 4) + e;
            if (k < cap && in_ids && out_ids && values) {
              in_ids[k] = bl.gidx[static_cast<size_t>(ch) * kChunkK + kk];
              out_ids[k] = bl.out_idx[static_cast<size_t>(blk) * kBlockOut + o];
              values[k] = half_bits_to_float(h);
            }
            ++k;
          }
    }
  }
  *n_out = k;
  return COSPDNN_OK;
}

int cospdnn_layer_image_decode(int N, const int32_t* colptr, const int32_t* rows, const float* vals, int image,
                               int64_t cap, int32_t* in_ids, int32_t* out_ids, float* values, int64_t* n_out) {
  CscLayer c;
  if (!csc_from_ptrs(N, colptr, rows, vals, c) || !n_out || image < 0 || image > 4)
    return fail(COSPDNN_ERR_ARG, "bad argument");
  BlockedLayer bl;
  if (!build_layer(c, N, 1e30, bl)) return fail(COSPDNN_ERR_FORMAT, "malformed CSC");
  if (image == 4) {
    // the slab image as the sample-slab kernel walks it: batch b, octet o, lane oi's slot m at
    // lane position oi * M + m, entry e; pads (the zero row, or value 0) and the slots past the
    // layer's outputs skipped
    build_slab_image(bl, N);
    if (!bl.has_slab) return fail(COSPDNN_ERR_STATE, "the layer has no such image");
    const int T = slab_samples(N), bs = slab_batch_slots(T), mper = slab_lane_slots(T), opw = bs / mper;
    const int zrow = slab_rows(N);
    int64_t k = 0;
    for (size_t b = 0; b + 1 < bl.sgoff.size(); ++b)
      for (int o = bl.sgoff[b]; o < bl.sgoff[b + 1]; ++o)
        for (int lp = 0; lp < bs; ++lp)
          for (int e = 0; e < 8; ++e) {
            const size_t at = (static_cast<size_t>(o) * bs + lp) * 8 + e;
            const int slot = static_cast<int>(b) * bs + (lp % mper) * opw + lp / mper;
            const int id = bl.sid[at];
            const float v = bl.slab_uniform ? bl.slab_value : half_bits_to_float(bl.sval[at]);
            if (slot >= N || id == zrow || (!bl.slab_uniform && v == 0.0f)) continue;
            if (k < cap && in_ids && out_ids && values) {
              in_ids[k] = id;
              out_ids[k] = bl.out_idx[slot];
              values[k] = v;
            }
            ++k;
          }
    *n_out = k;
    return COSPDNN_OK;
  }
  const bool band = image >= 2, mask = image == 1 || image == 3;
  if (band) build_band_image(bl, N);
  if ((band && !bl.has_band) || (mask && !bl.has_mask))
    return fail(COSPDNN_ERR_STATE, "the layer has no such image");
  const std::vector<int32_t>& gidx = band ? bl.bgidx : bl.gidx;
  const std::vector<uint16_t>& wt = band ? bl.bwt : bl.wt;
  const std::vector<uint16_t>& wm = band ? bl.bmask : bl.wmask;
  int64_t k = 0;
  // one chunk: its fragments as the kernels read them (mask words through the kernels'
  // expansion: dword q = (sp & (0x10001 << q)) * (v >> q), sp = the word with its high byte
  // moved up by 8), row ids offset by the chunk's band
  auto chunk = [&](int blk, int64_t ch, int row0) {
    for (int m = 0; m < 2; ++m)
      for (int lane = 0; lane < 64; ++lane) {
        uint16_t frag[8];
        if (mask) {
          const uint32_t bits = wm[static_cast<size_t>(ch) * 64 + lane];
          const uint32_t sp = (bits & 0xFFu) | ((bits & 0xFF00u) << 8);
          for (int kq = 0; kq < 4; ++kq) {
            const int q = 4 * m + kq;
            const uint64_t x = static_cast<uint64_t>(sp & (0x10001u << q)) * (bl.wmask_val >> q);
            const uint32_t d = static_cast<uint32_t>(x);  // the 24-bit multiply's low 32 bits
            frag[2 * kq] = static_cast<uint16_t>(d & 0xFFFFu);
            frag[2 * kq + 1] = static_cast<uint16_t>(d >> 16);
          }
        } else {
          for (int e = 0; e < 8; ++e) frag[e] = wt[((static_cast<size_t>(ch) * 2 + m) * 64 + lane) * 8 + e];
        }
        for (int e = 0; e < 8; ++e) {
          if ((frag[e] & 0x7FFFu) == 0) continue;
          const int o = m * 16 + (lane & 15);
          const int kk = 8 * (lane >> 4) + e;
          if (k < cap && in_ids && out_ids && values) {
            in_ids[k] = row0 + gidx[static_cast<size_t>(ch) * kChunkK + kk];
            out_ids[k] = bl.out_idx[static_cast<size_t>(blk) * kBlockOut + o];
            values[k] = half_bits_to_float(frag[e]);
          }
          ++k;
        }
      }
  };
  for (int blk = 0; blk < bl.nblk; ++blk) {
    if (!band) {
      for (int ch = bl.kptr[blk]; ch < bl.kptr[blk + 1]; ++ch) chunk(blk, ch, 0);
      continue;
    }
    for (int b = 0; b < bl.bands; ++b)
      for (int ch = bl.bptr[static_cast<size_t>(blk) * (bl.bands + 1) + b];
           ch < bl.bptr[static_cast<size_t>(blk) * (bl.bands + 1) + b + 1]; ++ch)
        chunk(blk, ch, b * kBandRows);
  }
  *n_out = k;
  return COSPDNN_OK;
}

int cospdnn_ref_bcsc_to_csc(int N, const float* values, int64_t n_values, const int32_t* row_idx, int64_t n_row_idx,
                            const int32_t* ptr, int64_t n_ptr, const int32_t* fuse, int64_t n_fuse, int blocksize_c,
                            int blocksize_r, int64_t cap, int32_t* colptr, int32_t* rows, float* vals) {
  if (N <= 0 || !values || !row_idx || !ptr || !colptr) return fail(COSPDNN_ERR_ARG, "bad argument");
  std::vector<float> v(values, values + n_values);
  std::vector<int> ri(row_idx, row_idx + n_row_idx), pt(ptr, ptr + n_ptr), fu;
  if (fuse) fu.assign(fuse, fuse + n_fuse);
  CscLayer c;
  if (!csc_from_reference_bcsc(v, ri, pt, fu, N, blocksize_c, blocksize_r, c))
    return fail(COSPDNN_ERR_FORMAT, "malformed reference BCSC");
  if (static_cast<int64_t>(c.rows.size()) > cap) return fail(COSPDNN_ERR_ARG, "capacity too small");
  std::memcpy(colptr, c.colptr.data(), sizeof(int32_t) * (N + 1));
  if (!c.rows.empty()) {
    std::memcpy(rows, c.rows.data(), sizeof(int32_t) * c.rows.size());
    std::memcpy(vals, c.vals.data(), sizeof(float) * c.vals.size());
  }
  return COSPDNN_OK;
}

}  // extern "C"
