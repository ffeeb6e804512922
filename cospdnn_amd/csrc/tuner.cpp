// Autotuner of the FP16 pass: per-layer launch plans, exhaustive search, and the
// cost-model MLP (trained and evaluated by the kernels in mlp.hip).
//
// A candidate is a launch shape for one layer: the SIMT gather kernel, or the MFMA kernel
// with a work order and a grid size.  The time of a layer depends on the candidate and on
// the layer (neurons, fan-in, tile image) and the rows entering it, not on the choices of
// the other layers, so the exhaustive optimum is the per-layer argmin over runs that apply
// one candidate to every layer.  The cost model predicts those per-layer times from
// features; its plan is the per-layer argmin of the predictions, and the exhaustive plan
// is both its fallback and the yardstick of its regret.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "engine_internal.hpp"
#include "mlp.hpp"

using namespace cospdnn;
using namespace cospdnn_detail;

struct cospdnn_mlp {
  cospdnn_engine* eng = nullptr;
  MlpDims dims;
  DevBuf params, grad, m, v;
  DevBuf X, y, work;
  int cap = 0;  // batch rows the device buffers hold
  int step = 0;
  MlpWork w;
};

namespace {

const cospdnn_layer_plan kCandidates[] = {
    {COSPDNN_PLAN_SIMT, 0, 0},         {COSPDNN_PLAN_MFMA, 0, 4},         {COSPDNN_PLAN_MFMA, 1, 2},
    {COSPDNN_PLAN_MFMA, 1, 4},         {COSPDNN_PLAN_MFMA, 2, 2},         {COSPDNN_PLAN_MFMA, 2, 4},
    {COSPDNN_PLAN_MFMA, 3, 4},         {COSPDNN_PLAN_MFMA, 4, 4},         {COSPDNN_PLAN_MFMA_COMPACT, 1, 4},
    {COSPDNN_PLAN_MFMA_COMPACT, 2, 2}, {COSPDNN_PLAN_MFMA_COMPACT, 2, 4}, {COSPDNN_PLAN_MFMA_COMPACT, 3, 4},
    {COSPDNN_PLAN_MFMA_COMPACT, 4, 4}, {COSPDNN_PLAN_MFMA, 5, 1},         {COSPDNN_PLAN_ROWS, 0, 5},
    {COSPDNN_PLAN_ROWS, 1, 5},         {COSPDNN_PLAN_MFMA, 6, 1},         {COSPDNN_PLAN_SLAB, 0, 1},
};
constexpr int kNumCandidates = static_cast<int>(sizeof(kCandidates) / sizeof(kCandidates[0]));

int check_plan(const cospdnn_layer_plan& p) {
  if (p.kernel < COSPDNN_PLAN_AUTO || p.kernel > COSPDNN_PLAN_SLAB) return fail(COSPDNN_ERR_ARG, "plan: kernel");
  const bool mfma = p.kernel == COSPDNN_PLAN_MFMA || p.kernel == COSPDNN_PLAN_MFMA_COMPACT;
  if (mfma && (p.order < 0 || p.order > 6)) return fail(COSPDNN_ERR_ARG, "plan: order");
  if (p.kernel == COSPDNN_PLAN_ROWS && (p.order < 0 || p.order > 1)) return fail(COSPDNN_ERR_ARG, "plan: rows split");
  if ((mfma || p.kernel == COSPDNN_PLAN_ROWS) && (p.groups_per_cu < 1 || p.groups_per_cu > 16))
    return fail(COSPDNN_ERR_ARG, "plan: groups_per_cu");
  return COSPDNN_OK;
}

// rows_prev: rows entering the layer before (== live for layer 0).  Every MFMA tile after
// the first row the scan before this layer removed leaves the fast epilogue (its output
// positions are shifted off 64 consecutive ones).  With k removals at uniform positions
// the expected fraction of positions after the first is k / (k + 1), so that is f[11]:
// one removal already moves about half the tiles, a handful moves nearly all of them.
void layer_features(const cospdnn_net* net, int l, int64_t live, int64_t rows_prev, const cospdnn_layer_plan& c,
                    float* f) {
  const LayerStore& ls = *net->dl[l];
  const double rows = static_cast<double>(std::max<int64_t>(live, 0));
  const bool mfma = (c.kernel == COSPDNN_PLAN_MFMA || c.kernel == COSPDNN_PLAN_MFMA_COMPACT) && ls.has_tiles;
  f[0] = static_cast<float>(std::log2(rows + 1.0) / 16.0);
  f[1] = static_cast<float>(std::log2(static_cast<double>(net->N)) / 16.0);
  f[2] = static_cast<float>(static_cast<double>(ls.nnz) / net->N / 64.0);
  f[3] = ls.has_tiles ? 1.0f : 0.0f;
  f[4] = ls.has_tiles ? static_cast<float>(std::log2(std::max(1.0, ls.waste)) / 6.0) : 0.0f;
  f[5] = ls.view.one_chunk ? 1.0f : 0.0f;
  f[6] = mfma ? 1.0f : 0.0f;
  const bool rows_k = c.kernel == COSPDNN_PLAN_ROWS;
  // the sample-slab kernel where the layer has a slab image (it falls back to SIMT otherwise)
  const int slab_t = slab_samples(l == 0 ? net->N : net->dl[l - 1]->view.n_out);
  const bool slab = c.kernel == COSPDNN_PLAN_SLAB && ls.view.sid != nullptr && slab_t > 0;
  f[7] = mfma || rows_k || slab ? 0.0f : 1.0f;
  f[8] = mfma ? static_cast<float>(c.order) / 5.0f : 0.0f;
  f[9] = mfma ? static_cast<float>(c.groups_per_cu) / 4.0f : 0.0f;
  f[10] = static_cast<float>(std::log2(rows * static_cast<double>(ls.nnz) + 1.0) / 48.0);
  const double removed = static_cast<double>(std::max<int64_t>(rows_prev - live, 0));
  f[11] = static_cast<float>(removed / (removed + 1.0));
  // gather mode allowed after a compacting scan (plain MFMA) or compacted stores; the
  // LDS-resident kernel (order 5 on inputs of at most kLdsRowsMax rows) always stores compacted
  const bool lds_resident = (c.order == 5 && net->N <= kLdsRowsMax) || (c.order == 6 && ls.view.bptr != nullptr);
  f[12] = mfma && c.kernel == COSPDNN_PLAN_MFMA && !lds_resident ? 1.0f : 0.0f;
  // the row-gather kernel and its work split (0: all groups at once, 1: groups per XCD)
  f[13] = rows_k ? 0.5f * static_cast<float>(1 + c.order) : 0.0f;
  // the sample-slab kernel and its slab size (samples per work-group: 64 .. 1)
  f[14] = slab ? static_cast<float>(1.0 + std::log2(static_cast<double>(slab_t))) / 7.0f : 0.0f;
}

int mlp_reserve(cospdnn_mlp* m, int n) {
  if (n <= m->cap) return COSPDNN_OK;
  const int H = m->dims.hidden;
  CALL_TRY(dalloc(m->X, sizeof(float) * static_cast<size_t>(n) * m->dims.in));
  CALL_TRY(dalloc(m->y, sizeof(float) * static_cast<size_t>(n)));
  // h1, h2, d1, d2: [n][H]; pred, dpred, sq: [n]
  CALL_TRY(dalloc(m->work, sizeof(float) * static_cast<size_t>(n) * (4 * static_cast<size_t>(H) + 3)));
  float* base = m->work.as<float>();
  const size_t nh = static_cast<size_t>(n) * H;
  m->w.h1 = base;
  m->w.h2 = base + nh;
  m->w.d1 = base + 2 * nh;
  m->w.d2 = base + 3 * nh;
  m->w.pred = base + 4 * nh;
  m->w.dpred = m->w.pred + n;
  m->w.sq = m->w.dpred + n;
  m->cap = n;
  return COSPDNN_OK;
}

int mlp_upload_batch(cospdnn_mlp* m, const float* X, const float* y, int n) {
  CALL_TRY(mlp_reserve(m, n));
  HIP_TRY(hipMemcpy(m->X.p, X, sizeof(float) * static_cast<size_t>(n) * m->dims.in, hipMemcpyHostToDevice));
  if (y) HIP_TRY(hipMemcpy(m->y.p, y, sizeof(float) * static_cast<size_t>(n), hipMemcpyHostToDevice));
  return COSPDNN_OK;
}

int mlp_loss(cospdnn_mlp* m, int n, float* loss) {
  std::vector<float> sq(n);
  HIP_TRY(hipMemcpy(sq.data(), m->w.sq, sizeof(float) * n, hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float v : sq) s += v;
  *loss = static_cast<float>(s / std::max(1, n));
  return COSPDNN_OK;
}

}  // namespace

extern "C" {

int cospdnn_net_set_plan(cospdnn_net* net, const cospdnn_layer_plan* plan) {
  if (!net) return fail(COSPDNN_ERR_ARG, "null net");
  if (!plan) {
    net->plan.clear();
  } else {
    for (int l = 0; l < net->L; ++l) CALL_TRY(check_plan(plan[l]));
    net->plan.assign(plan, plan + net->L);
  }
  ++net->plan_version;
  return COSPDNN_OK;
}

int cospdnn_num_candidates(void) { return kNumCandidates; }

int cospdnn_candidate(int k, cospdnn_layer_plan* out) {
  if (k < 0 || k >= kNumCandidates || !out) return fail(COSPDNN_ERR_ARG, "candidate index");
  *out = kCandidates[k];
  return COSPDNN_OK;
}

int cospdnn_net_layer_features(cospdnn_net* net, int layer, int64_t live_rows, int64_t rows_prev,
                               const cospdnn_layer_plan* cand, float* feat) {
  if (!net || !cand || !feat) return fail(COSPDNN_ERR_ARG, "null argument");
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  if (layer < 0 || layer >= net->L) return fail(COSPDNN_ERR_ARG, "layer index");
  CALL_TRY(check_plan(*cand));
  layer_features(net, layer, live_rows, rows_prev, *cand, feat);
  return COSPDNN_OK;
}

int cospdnn_tune_exhaustive(cospdnn_net* net, const cospdnn_opts* opts, int reps, float* ms_table, int64_t* live,
                            cospdnn_layer_plan* plan) {
  if (!net || !plan || reps < 1) return fail(COSPDNN_ERR_ARG, "bad argument");
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  const int L = net->L;
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  if (opts) o = *opts;
  o.precision = COSPDNN_FP16;
  o.layer_timing = 1;
  o.use_graph = 0;
  o.dump_layer = -1;
  const std::vector<cospdnn_layer_plan> saved = net->plan;
  const int64_t saved_version = net->plan_version;
  std::vector<float> table(static_cast<size_t>(kNumCandidates) * std::max(L, 1), 0.0f);
  std::vector<int64_t> live0(L + 1, 0);
  std::vector<float> lms(std::max(L, 1));
  int rc = COSPDNN_OK;
  for (int k = 0; k < kNumCandidates && rc == COSPDNN_OK; ++k) {
    net->plan.assign(L, kCandidates[k]);
    ++net->plan_version;
    std::vector<std::vector<float>> runs(L);
    for (int r = 0; r <= reps && rc == COSPDNN_OK; ++r) {  // run 0 warms up
      cospdnn_result res{};
      res.layer_ms = lms.data();
      if (k == 0 && r == 0) res.positions = live0.data();  // rows each layer's kernel works on
      rc = run_one(net, o, &res);
      if (rc != COSPDNN_OK || r == 0) continue;
      for (int l = 0; l < L; ++l) runs[l].push_back(lms[l]);
    }
    for (int l = 0; l < L && rc == COSPDNN_OK; ++l) {
      std::sort(runs[l].begin(), runs[l].end());
      table[static_cast<size_t>(k) * L + l] = runs[l][runs[l].size() / 2];
    }
  }
  net->plan = saved;
  net->plan_version = saved_version + kNumCandidates + 1;
  CALL_TRY(rc);
  for (int l = 0; l < L; ++l) {
    int best = 0;
    for (int k = 1; k < kNumCandidates; ++k) {
      if (kCandidates[k].kernel != COSPDNN_PLAN_SIMT && !net->dl[l]->has_tiles) continue;  // ran as SIMT
      if (table[static_cast<size_t>(k) * L + l] < table[static_cast<size_t>(best) * L + l]) best = k;
    }
    plan[l] = kCandidates[best];
  }
  if (ms_table) std::memcpy(ms_table, table.data(), sizeof(float) * static_cast<size_t>(kNumCandidates) * L);
  if (live) std::memcpy(live, live0.data(), sizeof(int64_t) * (L + 1));
  return COSPDNN_OK;
}

// The compaction threshold: whole FP16 forwards (the options as given, e.g. replayed from a
// hipGraph) per candidate keep_holes_ppm, median of reps runs after one warm-up run.  Keeping
// the layout saves the slow stores of shifted tiles in the layer after a scan that removed
// rows; the holes it leaves cost bandwidth in every later layer until a scan compacts, so
// the best threshold is measured rather than modelled per layer.
int cospdnn_tune_keep(cospdnn_net* net, const cospdnn_opts* opts, const int32_t* cand_ppm, int ncand, int reps,
                      float* ms, int32_t* best_ppm) {
  if (!net || !cand_ppm || ncand < 1 || reps < 1 || !best_ppm) return fail(COSPDNN_ERR_ARG, "bad argument");
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  for (int k = 0; k < ncand; ++k)
    if (cand_ppm[k] < 0 || cand_ppm[k] > 1000000) return fail(COSPDNN_ERR_ARG, "keep ppm out of [0, 1e6]");
  cospdnn_opts o;
  cospdnn_default_opts(&o);
  if (opts) o = *opts;
  o.precision = COSPDNN_FP16;
  o.layer_timing = 0;
  o.dump_layer = -1;
  float best_ms = 0.0f;
  for (int k = 0; k < ncand; ++k) {
    o.keep_holes_ppm = cand_ppm[k];
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
      cospdnn_result res{};
      CALL_TRY(run_one(net, o, &res));
      if (r > 0) t.push_back(static_cast<float>(res.ms_total));
    }
    std::sort(t.begin(), t.end());
    const float med = t[t.size() / 2];
    if (ms) ms[k] = med;
    if (k == 0 || med < best_ms) {
      best_ms = med;
      *best_ppm = cand_ppm[k];
    }
  }
  return COSPDNN_OK;
}

int cospdnn_tune_samples(cospdnn_net* net, const float* ms_table, const int64_t* live, float* X, float* y, int cap,
                         int* n_out) {
  if (!net || !ms_table || !live || !X || !y || !n_out || cap < 0) return fail(COSPDNN_ERR_ARG, "bad argument");
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  const int L = net->L;
  int n = 0;
  for (int k = 0; k < kNumCandidates; ++k)
    for (int l = 0; l < L; ++l) {
      if (kCandidates[k].kernel != COSPDNN_PLAN_SIMT && !net->dl[l]->has_tiles) continue;
      if (n >= cap) return fail(COSPDNN_ERR_ARG, "sample capacity too small");
      layer_features(net, l, live[l], live[l > 0 ? l - 1 : 0], kCandidates[k], X + static_cast<size_t>(n) * COSPDNN_TUNE_FEATURES);
      y[n] = static_cast<float>(std::log(1000.0 * ms_table[static_cast<size_t>(k) * L + l] + 1.0));
      ++n;
    }
  *n_out = n;
  return COSPDNN_OK;
}

int cospdnn_mlp_create(cospdnn_engine* engine, int n_in, int hidden, uint64_t seed, cospdnn_mlp** out) {
  if (!engine || !out || n_in < 1 || hidden < 1 || n_in > 1024 || hidden > 1024) return fail(COSPDNN_ERR_ARG, "bad argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(engine->device));
  auto m = std::make_unique<cospdnn_mlp>();
  m->eng = engine;
  m->dims.in = n_in;
  m->dims.hidden = hidden;
  const int np = m->dims.num_params();
  // He-uniform initialisation from a seeded generator, zero biases
  std::vector<float> p(np, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](float* w, int fan_in, int count) {
    const float a = std::sqrt(6.0f / static_cast<float>(fan_in));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < count; ++i) w[i] = u(rng);
  };
  float* W1 = p.data();
  float* W2 = W1 + n_in * hidden + hidden;
  float* W3 = W2 + hidden * hidden + hidden;
  fill(W1, n_in, n_in * hidden);
  fill(W2, hidden, hidden * hidden);
  fill(W3, hidden, hidden);
  CALL_TRY(dalloc(m->params, sizeof(float) * np));
  CALL_TRY(dalloc(m->grad, sizeof(float) * np));
  CALL_TRY(dalloc(m->m, sizeof(float) * np));
  CALL_TRY(dalloc(m->v, sizeof(float) * np));
  HIP_TRY(hipMemcpy(m->params.p, p.data(), sizeof(float) * np, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->m.p, 0, sizeof(float) * np));
  HIP_TRY(hipMemset(m->v.p, 0, sizeof(float) * np));
  *out = m.release();
  return COSPDNN_OK;
}

void cospdnn_mlp_destroy(cospdnn_mlp* mlp) {
  if (!mlp) return;
  (void)hipSetDevice(mlp->eng->device);
  delete mlp;
}

int cospdnn_mlp_num_params(const cospdnn_mlp* mlp) { return mlp ? mlp->dims.num_params() : 0; }

int cospdnn_mlp_get_params(cospdnn_mlp* mlp, float* params) {
  if (!mlp || !params) return fail(COSPDNN_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(mlp->eng->device));
  HIP_TRY(hipMemcpy(params, mlp->params.p, sizeof(float) * mlp->dims.num_params(), hipMemcpyDeviceToHost));
  return COSPDNN_OK;
}

int cospdnn_mlp_set_params(cospdnn_mlp* mlp, const float* params) {
  if (!mlp || !params) return fail(COSPDNN_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(mlp->eng->device));
  HIP_TRY(hipMemcpy(mlp->params.p, params, sizeof(float) * mlp->dims.num_params(), hipMemcpyHostToDevice));
  return COSPDNN_OK;
}

int cospdnn_mlp_loss_grad(cospdnn_mlp* mlp, const float* X, const float* y, int n, float* loss, float* grad) {
  if (!mlp || !X || !y || n < 1 || !loss || !grad) return fail(COSPDNN_ERR_ARG, "bad argument");
  HIP_TRY(hipSetDevice(mlp->eng->device));
  hipStream_t st = mlp->eng->stream;
  CALL_TRY(mlp_upload_batch(mlp, X, y, n));
  HIP_TRY(mlp_forward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), n, mlp->w, st));
  HIP_TRY(mlp_backward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), mlp->y.as<float>(), n, mlp->w,
                       mlp->grad.as<float>(), st));
  HIP_TRY(hipStreamSynchronize(st));
  CALL_TRY(mlp_loss(mlp, n, loss));
  HIP_TRY(hipMemcpy(grad, mlp->grad.p, sizeof(float) * mlp->dims.num_params(), hipMemcpyDeviceToHost));
  return COSPDNN_OK;
}

int cospdnn_mlp_train(cospdnn_mlp* mlp, const float* X, const float* y, int n, int steps, float lr, float* loss) {
  if (!mlp || !X || !y || n < 1 || steps < 0 || !(lr > 0.0f)) return fail(COSPDNN_ERR_ARG, "bad argument");
  HIP_TRY(hipSetDevice(mlp->eng->device));
  hipStream_t st = mlp->eng->stream;
  CALL_TRY(mlp_upload_batch(mlp, X, y, n));
  const int np = mlp->dims.num_params();
  for (int s = 0; s < steps; ++s) {
    HIP_TRY(mlp_forward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), n, mlp->w, st));
    HIP_TRY(mlp_backward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), mlp->y.as<float>(), n, mlp->w,
                         mlp->grad.as<float>(), st));
    ++mlp->step;
    HIP_TRY(mlp_adam(mlp->params.as<float>(), mlp->grad.as<float>(), mlp->m.as<float>(), mlp->v.as<float>(), np, lr,
                     mlp->step, st));
  }
  if (loss) {
    HIP_TRY(mlp_forward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), n, mlp->w, st));
    HIP_TRY(mlp_backward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), mlp->y.as<float>(), n, mlp->w,
                         mlp->grad.as<float>(), st));
    HIP_TRY(hipStreamSynchronize(st));
    CALL_TRY(mlp_loss(mlp, n, loss));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return COSPDNN_OK;
}

int cospdnn_mlp_predict(cospdnn_mlp* mlp, const float* X, int n, float* out) {
  if (!mlp || !X || n < 1 || !out) return fail(COSPDNN_ERR_ARG, "bad argument");
  HIP_TRY(hipSetDevice(mlp->eng->device));
  hipStream_t st = mlp->eng->stream;
  CALL_TRY(mlp_upload_batch(mlp, X, nullptr, n));
  HIP_TRY(mlp_forward(mlp->dims, mlp->params.as<float>(), mlp->X.as<float>(), n, mlp->w, st));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipMemcpy(out, mlp->w.pred, sizeof(float) * n, hipMemcpyDeviceToHost));
  return COSPDNN_OK;
}

int cospdnn_tune_mlp(cospdnn_net* net, cospdnn_mlp* mlp, const int64_t* live, cospdnn_layer_plan* plan,
                     float* pred_ms) {
  if (!net || !mlp || !live || !plan) return fail(COSPDNN_ERR_ARG, "null argument");
  if (!net->uploaded) return fail(COSPDNN_ERR_STATE, "network not uploaded");
  if (mlp->dims.in != COSPDNN_TUNE_FEATURES) return fail(COSPDNN_ERR_ARG, "MLP input width is not the feature count");
  const int L = net->L;
  std::vector<float> X(static_cast<size_t>(kNumCandidates) * L * COSPDNN_TUNE_FEATURES);
  for (int k = 0; k < kNumCandidates; ++k)
    for (int l = 0; l < L; ++l)
      layer_features(net, l, live[l], live[l > 0 ? l - 1 : 0], kCandidates[k],
                     X.data() + (static_cast<size_t>(k) * L + l) * COSPDNN_TUNE_FEATURES);
  std::vector<float> pred(static_cast<size_t>(kNumCandidates) * L);
  CALL_TRY(cospdnn_mlp_predict(mlp, X.data(), kNumCandidates * L, pred.data()));
  for (int l = 0; l < L; ++l) {
    int best = 0;
    for (int k = 1; k < kNumCandidates; ++k) {
      if (kCandidates[k].kernel != COSPDNN_PLAN_SIMT && !net->dl[l]->has_tiles) continue;
      if (pred[static_cast<size_t>(k) * L + l] < pred[static_cast<size_t>(best) * L + l]) best = k;
    }
    plan[l] = kCandidates[best];
  }
  if (pred_ms)
    for (size_t i = 0; i < pred.size(); ++i) pred_ms[i] = static_cast<float>((std::exp(static_cast<double>(pred[i])) - 1.0) / 1000.0);
  return COSPDNN_OK;
}

}  // extern "C"
