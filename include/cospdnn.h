/*
 * cospdnn.h -- C ABI of the MI355X (gfx950) sparse-DNN inference engine.
 *
 * The engine evaluates Y_{l+1} = min(max(Y_l . W_l + b, 0), 32) for L layers
 * and returns the "categories": the samples (rows of Y) that are still nonzero
 * after the last layer.
 *
 * This ABI replaces the reference's C++ entry points (paths relative to the
 * CoSpDNN source tree):
 *   test_graph_challenge_tc     src/cospdnn.cuh:14-21, src/cospdnn.cu:3-369
 *   test_graph_challenge_tc_co  src/cospdnn.cuh:22-35, src/cospdnn_double.cu:208-1213
 * whose argument sets map onto a cospdnn_net (N = neuron, S = batch, bias,
 * per-layer weights, input) plus cospdnn_run().  A C++ mirror with the
 * reference's exact signatures sits on top of this ABI
 * (cospdnn_amd/csrc/cospdnn_compat.hpp).
 *
 * Conventions: status codes instead of exit() (the reference's Safe_Call,
 * src/gpu_lib/gpu_runtime.h:6-17, terminates the process); the caller owns
 * every host buffer, the library owns every device buffer; one engine per
 * (device, host thread); calls are stream-ordered on the engine's stream.
 */
#ifndef COSPDNN_H
#define COSPDNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  COSPDNN_OK = 0,
  COSPDNN_ERR_ARG = 1,       /* invalid argument (null pointer, bad size, bad layer index) */
  COSPDNN_ERR_HIP = 2,       /* HIP runtime error; detail in cospdnn_last_error() */
  COSPDNN_ERR_OOM = 3,       /* device or host allocation failed */
  COSPDNN_ERR_FORMAT = 4,    /* malformed weights or input */
  COSPDNN_ERR_PRECISION = 5, /* FP16 path asked for values not exact in FP16 */
  COSPDNN_ERR_STATE = 6,     /* call order: layers or input missing, not uploaded */
  COSPDNN_ERR_IO = 7,        /* file missing or unparsable */
  COSPDNN_ERR_MISMATCH = 8,  /* networks of one run disagree where they must agree */
  COSPDNN_ERR_INTERNAL = 9   /* an engine invariant does not hold (a bug); detail in cospdnn_last_error() */
};

enum { COSPDNN_FP32 = 0, COSPDNN_FP16 = 1, COSPDNN_FP64 = 2 };
enum { COSPDNN_GEN_RANDOM = 0, COSPDNN_GEN_BANDED = 1 };

/* The SpMM kernel of a layer, reported per layer in cospdnn_result.layer_kernel (name of the
   device function: cospdnn_kernel_name). */
enum {
  COSPDNN_KERNEL_SIMT = 0,            /* one thread per (sample, output): gathers over the CSC */
  COSPDNN_KERNEL_SIMT_ROWS = 1,       /* the same for passes of at most 512 rows */
  COSPDNN_KERNEL_MFMA = 2,            /* FP16 MFMA over the layer's tile image */
  COSPDNN_KERNEL_MFMA_MASK = 3,       /* the same reading the mask image */
  COSPDNN_KERNEL_MFMA_LDS = 4,        /* FP16 MFMA with the whole 64-sample input tile in LDS (order 5) */
  COSPDNN_KERNEL_MFMA_LDS_MASK = 5,
  COSPDNN_KERNEL_MFMA_BAND = 6,       /* FP16 MFMA streaming 512-row bands of the input tile (order 6) */
  COSPDNN_KERNEL_MFMA_BAND_MASK = 7,
  COSPDNN_KERNEL_WAVE_ROWS = 8,       /* row-gather kernel (COSPDNN_PLAN_ROWS) */
  COSPDNN_KERNEL_MFMA32 = 9,          /* FP32 MFMA over the FP32 tile image */
  COSPDNN_KERNEL_SLAB = 10,           /* FP16 sample-slab kernel: T whole input samples in LDS, every
                                         output of them by FP32 FMAs over the slab image
                                         (spmm_slab[_uni]_kernel_<T>) */
  COSPDNN_KERNEL_SLAB_BITS = 11       /* layer 0 of a slab plan on the bit-packed binary input with one
                                         weight value: 16 to 64 samples' input bits in LDS, per output
                                         the count of set input rows by bit-sliced adders
                                         (spmm_bits_kernel_<T>) */
};

typedef struct cospdnn_engine cospdnn_engine;
typedef struct cospdnn_net cospdnn_net;

typedef struct {
  int precision;         /* COSPDNN_FP16: MFMA with FP16 activations; COSPDNN_FP32: SIMT FP32;
                            COSPDNN_FP64: every row replayed in FP64 (bit-identical to the oracle) */
  int guard_layer;       /* FP16: rows not at an exact fixed point (all-0 / all-32) entering
                            this layer are recomputed in FP32 (default 12; < 0: no recompute) */
  int guard_layer32;     /* FP32 passes (the FP32 path and the FP16 path's recompute): rows not at
                            an exact fixed point entering this layer are replayed in FP64
                            (default 20; < 0: no replay) */
  int fp32_mfma;         /* FP32 path: 1 = MFMA FP32 kernel where a tile image exists, 0 = plain
                            SIMT gather (default 0; the FP16 path's recompute always uses MFMA) */
  float guard_band;      /* FP16: also recompute rows with a pre-activation within this
                            distance of 0 or 32 ... */
  int guard_band_from;   /* ... at any layer >= guard_band_from (default: 1 << 30 = never) */
  int use_mfma;          /* FP16: 1 = MFMA kernel where a tile image exists, 0 = SIMT only */
  float max_tile_waste;  /* build the MFMA tile image of a layer only when its MFMA work is at
                            most this multiple of the useful work (default 64: random layers
                            too, ~32x; the images of an engine's networks take at most 40%
                            of the HBM, later layers keep the SIMT image only) */
  int layer_timing;      /* 1: time every layer's SpMM with events (reference-style sum) */
  int use_graph;         /* 1: capture the layer loop of each network in a hipGraph and replay */
  int dump_layer;        /* >= 0: copy the activations after this layer into the result */
  int spmm_order;        /* MFMA SpMM work order: 0 sample-tile-major, 1 block-column-major
                            (default), 2 sample tiles partitioned over the XCDs (L2 reuse on
                            random layers), 3 block columns partitioned over the XCDs,
                            4 block columns partitioned over the XCDs, sample tiles in turn
                            (each XCD's share of the tile image stays in its L2), 5 a whole
                            64-sample input tile in LDS per work-group (layers with at most
                            1024 input rows; order 1 otherwise), 6 the input tile streamed
                            through LDS in bands of 512 rows while a work-group accumulates 32
                            block columns (random layers of at most 16384 neurons, which have
                            a band image; order 1 otherwise) */
  int mfma_groups_per_cu;/* MFMA SpMM grid: resident work-groups per CU (1..16) */
  int batch_concurrent;  /* several networks in one cospdnn_run: k > 0 = their main passes
                            run in k concurrent chains on side streams (one hipGraph with
                            use_graph), networks assigned longest first; 0 = one after another */
  int guard_tail;        /* guarded passes: rows that die in the last guard_tail layers are
                            recomputed at the next precision too */
  int guard_concurrent;  /* FP16 with a checkpoint, one network, no per-layer timing: 1 (default)
                            = the FP32 and FP64 tiers run on the device on a side stream, forked
                            off after the checkpoint scan, next to the rest of the main pass;
                            0 = after it, one tier after the other */
  int keep_holes_ppm;    /* FP16: a scan that removes only dead rows keeps the layout (the dead
                            rows stay as all-zero holes until a later scan compacts) when it
                            removes at most this many ppm of the rows entering the layer;
                            compacting shifts every later tile off its 128-byte lines, which
                            slows its stores (default 0: always compact) */
  int spmm_gather;       /* FP16 MFMA layers after a compacting scan: 1 (default) = work items are
                            64 output positions whose input columns are gathered, so every store
                            covers whole 128-byte lines; 0 = items are 64 input positions and the
                            surviving samples are stored to their compacted positions */
  int input_bits;        /* FP16 MFMA layer 0 with a binary input: 1 (default) = read the input's
                            bit-packed copy (1 bit per value instead of 16), expanded to
                            FP16 in registers; 0 = read the FP16 input */
  int weight_mask;       /* FP16 MFMA layers whose weights are all one value: 1 (default) = read the
                            weight tiles as a bit mask (128 bytes a K chunk instead of 2 KB),
                            expanded to the same MFMA operands in registers; 0 = the FP16 tiles */
  int batch_cu_share;    /* several networks with batch_concurrent = k > 1: 1 = every SpMM grid of a
                            chain is divided by k (at least 8 work-groups), so the chains share the
                            CUs instead of queueing whole-GPU launches behind each other; 0 (default)
                            = full grids */
  int weight_budget_mb;  /* upload: MFMA weight images resident in HBM, all of the engine's networks
                            together, at most this many MiB (0, the default: 40% of the HBM);
                            the layers past it keep their images in pinned host memory and the
                            FP16 pass streams them into two device slots, one layer ahead of
                            its SpMM, on a copy stream (eager and in captured graphs) */
  int spmm_slab;         /* FP16: 1 = every layer with a slab image (FP16-exact weights, at most
                            65536 input rows) runs the sample-slab kernel: a work-group keeps T
                            whole input samples in LDS (T*rows*2 bytes <= 128 KB: T = 64 up to
                            1023 rows, 16 at 4096, 4 at 16384, 1 at 65536) and computes all of
                            their outputs with one FP32 FMA per nonzero and sample, no tile
                            waste; consecutive slab layers pass activations in a sample-blocked
                            layout.  0 (default) = the MFMA / SIMT choice above; a plan can pick
                            it per layer (COSPDNN_PLAN_SLAB) */
} cospdnn_opts;

typedef struct {
  /* caller-provided buffers (may be NULL) */
  int64_t* categories;   /* capacity S: ascending original sample ids alive after layer L-1 */
  int64_t* live;         /* capacity L+1: samples alive entering layer l; live[L] = #categories */
  float* dump_vals;      /* capacity dump_cap_rows*N: row-major activations after dump_layer */
  int64_t* dump_ids;     /* capacity dump_cap_rows: original sample id of each dumped row */
  int64_t dump_cap_rows;
  float* layer_ms;       /* capacity L: per-layer SpMM time when layer_timing = 1 */
  /* filled by cospdnn_run */
  int64_t n_categories;
  int64_t dump_rows;
  double ms_total;       /* device time of the whole forward: all layers, compaction, recompute */
  double ms_spmm;        /* sum of per-layer SpMM times (layer_timing = 1), else 0 */
  double ms_recompute;   /* part of ms_total spent in the FP32 guard recompute */
  int64_t n_flagged;     /* rows recomputed in FP32 (FP16 path) */
  int64_t n_replayed;    /* rows replayed in FP64 */
  int64_t n_changed;     /* recomputed or replayed rows whose category differs from the first pass */
  double edges;          /* S * sum_l nnz(W_l) */
  double edges_per_sec;  /* edges / ms_total */
  double alg_bytes;      /* sum_l (2 * live_l * N * act_bytes + nnz_l * weight_bytes) */
  int64_t mfma_layers;   /* layers that ran on the MFMA kernel */
  int64_t* positions;    /* caller-provided (may be NULL), capacity L+1: row positions entering
                            layer l in the first pass, live rows plus the holes of a kept
                            layout (keep_holes_ppm) */
  int64_t streamed_layers; /* layers of the network whose MFMA images are streamed from host memory */
  int32_t* layer_kernel;   /* caller-provided (may be NULL), capacity L: the SpMM kernel
                              (COSPDNN_KERNEL_*) the main pass launches for each layer */
} cospdnn_result;

const char* cospdnn_version(void);
const char* cospdnn_strerror(int code);
const char* cospdnn_last_error(void); /* detail of the last failure on this thread */
const char* cospdnn_kernel_name(int kernel); /* device function of a COSPDNN_KERNEL_* id, or NULL */
void cospdnn_default_opts(cospdnn_opts* opts);

/* engine: one device, one stream (default priority) */
int cospdnn_engine_create(int device, cospdnn_engine** out);
void cospdnn_engine_destroy(cospdnn_engine* engine);
int cospdnn_engine_info(cospdnn_engine* engine, int* cu_count, int64_t* hbm_bytes);

/* network definition (host side; nothing touches the device before upload) */
int cospdnn_net_create(cospdnn_engine* engine, int N, int L, int64_t S, float bias, cospdnn_net** out);
void cospdnn_net_destroy(cospdnn_net* net);
/* layer l as CSC: colptr[N+1], rows/vals[colptr[N]]; column j = output neuron j */
int cospdnn_net_set_layer_csc(cospdnn_net* net, int layer, const int32_t* colptr, const int32_t* rows,
                              const float* vals);
/* layer l in the reference's BCSC (values [out][in] per block, block start rows, block-column
   pointers) with its fuse list (may be NULL: identity), as test_graph_challenge_tc receives it */
int cospdnn_net_set_layer_ref_bcsc(cospdnn_net* net, int layer, const float* values, int64_t n_values,
                                   const int32_t* row_idx, int64_t n_row_idx, const int32_t* ptr, int64_t n_ptr,
                                   const int32_t* fuse, int64_t n_fuse, int blocksize_c, int blocksize_r);
/* all layers from the synthetic generator (kind COSPDNN_GEN_*; K nonzeros per column) */
int cospdnn_net_generate_layers(cospdnn_net* net, int kind, int K, float wval, uint64_t seed);
/* input: dense row-major [S][N] floats, or 0-based (sample, neuron, value) triples, or generated */
int cospdnn_net_set_input_dense(cospdnn_net* net, const float* y0);
int cospdnn_net_set_input_coo(cospdnn_net* net, int64_t nnz, const int64_t* sample, const int32_t* neuron,
                              const float* val);
int cospdnn_net_generate_input(cospdnn_net* net, uint64_t seed);
/* builds the device formats and moves weights + input into HBM (not part of any timing) */
int cospdnn_net_upload(cospdnn_net* net, const cospdnn_opts* opts);

/* runs the networks one after another on the engine's stream; results[i] for nets[i] */
int cospdnn_run(cospdnn_net* const* nets, int n_nets, const cospdnn_opts* opts, cospdnn_result* results);

/* Allocation-free variant of cospdnn_run for one uploaded network.  It has no counterpart in
   the reference; it is for callers that cannot allocate memory once they are running (serving
   loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree, hipHostMalloc or
   other device or pinned allocation; host-side bookkeeping uses ordinary heap memory.
   Every working buffer (activations, per-sample flags and position maps, the FP32 and FP64
   guard tiers) is a view into one fixed device array of COSPDNN_NOALLOC_SCRATCH_BYTES that each
   GPU the library runs on holds until the process ends; a network that needs more is refused
   with COSPDNN_ERR_ARG (about 4*N*S bytes for the main pass plus the tiers: N = 1024 fits up to
   some 35000 samples).  Same results as cospdnn_run (precision COSPDNN_FP16 only; the guard
   tiers run after the main pass, one after the other; no per-layer timing, no layer dump,
   use_graph ignored); COSPDNN_ERR_ARG also when more rows need the FP32 tier than
   max(4096, S/4) or the FP64 tier than 512.  Calls share their GPU's array and keep its last
   contents: each call clears what it uses, and calls on one device are serialised inside the
   library (calls overlapping on one GPU would otherwise race).  ms_total is host wall time. */
#define COSPDNN_NOALLOC_SCRATCH_BYTES (256ll << 20)
int cospdnn_run_noalloc(cospdnn_net* net, const cospdnn_opts* opts, cospdnn_result* result);

/* host-only helpers (no device needed) */
int cospdnn_generate_layer(uint64_t seed, int kind, int N, int K, int layer, float wval, int32_t* colptr,
                           int32_t* rows, float* vals);
int cospdnn_generate_input_rows(uint64_t seed, int N, int64_t s0, int64_t count, uint8_t* out);
/* format statistics of one CSC layer: block columns, MFMA chunks, tile waste, FP16 exactness */
int cospdnn_layer_stats(int N, const int32_t* colptr, const int32_t* rows, const float* vals, int64_t* nblk,
                        int64_t* nchunk, double* tile_waste, int* fp16_exact);
/* decodes the MFMA tile image of one CSC layer back to (input, output, value) triples;
   returns the triple count in *n_out (capacity cap); used to pin the tile layout */
int cospdnn_layer_tiles_decode(int N, const int32_t* colptr, const int32_t* rows, const float* vals,
                               int64_t cap, int32_t* in_ids, int32_t* out_ids, float* values, int64_t* n_out);
/* the same for the layer's other weight images: image 0 the FP16 tile image, 1 its mask image
   (expanded with the kernels' arithmetic), 2 the band image, 3 the band image's mask words,
   4 the slab image (16-bit rows in the sample-slab kernel's batch order, pads skipped);
   COSPDNN_ERR_STATE when the layer has no such image (mask: not one FP16 value with its low
   7 bits zero; band: built at upload for every layer, here on request) */
int cospdnn_layer_image_decode(int N, const int32_t* colptr, const int32_t* rows, const float* vals, int image,
                               int64_t cap, int32_t* in_ids, int32_t* out_ids, float* values, int64_t* n_out);
/* Similarity order of one layer (host): the output order that puts outputs reading the same
   inputs into the same group of `group` consecutive outputs; order[i] = the output at position i.
   overlap = 1: greedy by shared inputs (each group seeded with the lowest-numbered free output,
   then a free output sharing the most inputs with the group's union); 0: identical input sets
   only.  The transform the reference expects done offline (README.md:9, the fuse lists read at
   cospdnn.cpp:14-44); cospdnn_amd/transform.py applies it to a dataset directory. */
int cospdnn_similarity_order(int N, const int32_t* colptr, const int32_t* rows, int group, int overlap,
                             int32_t* order);
/* reference BCSC + fuse -> CSC (host); colptr capacity N+1, rows/vals capacity cap */
int cospdnn_ref_bcsc_to_csc(int N, const float* values, int64_t n_values, const int32_t* row_idx,
                            int64_t n_row_idx, const int32_t* ptr, int64_t n_ptr, const int32_t* fuse,
                            int64_t n_fuse, int blocksize_c, int blocksize_r, int64_t cap, int32_t* colptr,
                            int32_t* rows, float* vals);

/* ---- per-layer launch plans and the autotuner ----------------------------------
   No counterpart in the reference, which hard-codes its choices (16x16 blocks,
   cospdnn.cpp:94-95; compaction up to layer 21, cospdnn.cu:240; evaluateConfig's
   fixed split, cospdnn_double.cu:3-30).  A plan fixes, for each layer of the FP16
   pass, the SpMM kernel and its launch shape.  COSPDNN_PLAN_MFMA_COMPACT is the MFMA
   kernel with gather mode off for that layer (after a compacting scan its tiles are input
   positions and the surviving samples are stored compacted, see spmm_gather).
   COSPDNN_PLAN_ROWS is the row-gather kernel (a wave per 64-sample group and output slots,
   CSC lists, no tile image); its order is the work split: 0 all sample groups at once
   (few live rows), 1 sample groups partitioned over the XCDs and streamed (L2 reuse).
   COSPDNN_PLAN_SLAB is the sample-slab kernel (see spmm_slab; order and groups unused); it
   falls back to SIMT on a layer without a slab image. */
enum {
  COSPDNN_PLAN_AUTO = 0,
  COSPDNN_PLAN_MFMA = 1,
  COSPDNN_PLAN_SIMT = 2,
  COSPDNN_PLAN_MFMA_COMPACT = 3,
  COSPDNN_PLAN_ROWS = 4,
  COSPDNN_PLAN_SLAB = 5
};
typedef struct {
  int kernel;        /* COSPDNN_PLAN_*; MFMA falls back to SIMT on a layer without a tile image */
  int order;         /* MFMA work order (as cospdnn_opts.spmm_order); ROWS work split (0, 1) */
  int groups_per_cu; /* MFMA and ROWS grid (as cospdnn_opts.mfma_groups_per_cu) */
} cospdnn_layer_plan;
/* L entries (copied), or NULL to go back to the options' choice for every layer */
int cospdnn_net_set_plan(cospdnn_net* net, const cospdnn_layer_plan* plan);
/* the candidate launch shapes the tuner compares for each layer */
int cospdnn_num_candidates(void);
int cospdnn_candidate(int k, cospdnn_layer_plan* out);
/* feature vector of (layer, candidate, rows entering the layer, rows entering the layer
   before it; pass live_rows again for layer 0) for the cost model */
#define COSPDNN_TUNE_FEATURES 15
int cospdnn_net_layer_features(cospdnn_net* net, int layer, int64_t live_rows, int64_t rows_prev,
                               const cospdnn_layer_plan* cand, float* feat);
/* Exhaustive tuning: runs the FP16 forward once per candidate applied to every layer
   (median of reps runs, per-layer HIP events) and keeps, per layer, the fastest.
   ms_table [ncand][L] (may be NULL), live [L+1] row positions entering each layer (the
   rows its kernel works on; may be NULL), plan [L].  The network's plan is left unchanged. */
int cospdnn_tune_exhaustive(cospdnn_net* net, const cospdnn_opts* opts, int reps, float* ms_table, int64_t* live,
                            cospdnn_layer_plan* plan);
/* Compaction threshold: the FP16 forward (opts as given) once per candidate keep_holes_ppm,
   median ms_total of reps runs; ms [ncand] (may be NULL), *best_ppm = the fastest. */
int cospdnn_tune_keep(cospdnn_net* net, const cospdnn_opts* opts, const int32_t* cand_ppm, int ncand, int reps,
                      float* ms, int32_t* best_ppm);

/* Cost-model MLP: n_in -> hidden -> hidden -> 1, ReLU, FP32, on the engine's device.
   Training (Adam on mean squared error), gradients and inference run in HIP kernels. */
typedef struct cospdnn_mlp cospdnn_mlp;
int cospdnn_mlp_create(cospdnn_engine* engine, int n_in, int hidden, uint64_t seed, cospdnn_mlp** out);
void cospdnn_mlp_destroy(cospdnn_mlp* mlp);
int cospdnn_mlp_num_params(const cospdnn_mlp* mlp);
/* parameters: W1[n_in][hidden], b1, W2[hidden][hidden], b2, W3[hidden], b3 (row-major) */
int cospdnn_mlp_get_params(cospdnn_mlp* mlp, float* params);
int cospdnn_mlp_set_params(cospdnn_mlp* mlp, const float* params);
/* loss = mean((MLP(X) - y)^2) and its gradient w.r.t. the parameters; X [n][n_in] */
int cospdnn_mlp_loss_grad(cospdnn_mlp* mlp, const float* X, const float* y, int n, float* loss, float* grad);
/* steps of full-batch Adam; *loss = final training loss (may be NULL) */
int cospdnn_mlp_train(cospdnn_mlp* mlp, const float* X, const float* y, int n, int steps, float lr, float* loss);
int cospdnn_mlp_predict(cospdnn_mlp* mlp, const float* X, int n, float* out);
/* Training samples from an exhaustive table: one row per (candidate, layer),
   y = ln(1000 * ms + 1).  cap rows of X [cap][COSPDNN_TUNE_FEATURES] and y [cap]. */
int cospdnn_tune_samples(cospdnn_net* net, const float* ms_table, const int64_t* live, float* X, float* y, int cap,
                         int* n_out);
/* Plan from the cost model: per layer the candidate with the smallest predicted time.
   pred_ms [ncand][L] (may be NULL) receives the predictions. */
int cospdnn_tune_mlp(cospdnn_net* net, cospdnn_mlp* mlp, const int64_t* live, cospdnn_layer_plan* plan,
                     float* pred_ms);

#ifdef __cplusplus
}
#endif

#endif /* COSPDNN_H */
