"""MI355X (gfx950) sparse-DNN inference engine: Python binding of libcospdnn.so.

The engine computes Y_{l+1} = min(max(Y_l . W_l + b, 0), 32) over L layers and
returns the categories (samples still nonzero after the last layer).  All
compute runs in the HIP kernels of libcospdnn.so; this package only moves
arguments across the C ABI (include/cospdnn.h).
"""
from ._lib import (  # noqa: F401
    EXPORTS,
    FP16,
    FP32,
    FP64,
    GEN_BANDED,
    GEN_RANDOM,
    LIB_PATH,
    PLAN_AUTO,
    PLAN_MFMA,
    PLAN_SIMT,
    PLAN_MFMA_COMPACT,
    PLAN_ROWS,
    TUNE_FEATURES,
    CospdnnError,
    Engine,
    LayerPlan,
    Mlp,
    Net,
    candidates,
    build,
    default_opts,
    generate_input_rows,
    generate_layer,
    layer_stats,
    layer_tiles_decode,
    lib,
    ref_bcsc_to_csc,
    run,
    run_noalloc,
    similarity_order,
    tune_exhaustive,
    tune_keep,
    KEEP_CANDIDATES,
    tune_samples,
    version,
)
from .configs import CONFIGS, MASTER_SEED, config_nets  # noqa: F401
