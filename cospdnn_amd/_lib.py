"""ctypes binding of libcospdnn.so (the C ABI declared in include/cospdnn.h).

The library is built in-tree (cospdnn_amd/libcospdnn.so) by build().  Loading
fails loudly when it is missing: there is no Python or CPU fallback.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcospdnn.so")

FP32 = 0
FP16 = 1
FP64 = 2
GEN_RANDOM = 0
GEN_BANDED = 1

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_f32p = ctypes.POINTER(ctypes.c_float)
_c_u8p = ctypes.POINTER(ctypes.c_uint8)
_c_f64p = ctypes.POINTER(ctypes.c_double)


class Opts(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("guard_layer", ctypes.c_int),
        ("guard_layer32", ctypes.c_int),
        ("fp32_mfma", ctypes.c_int),
        ("guard_band", ctypes.c_float),
        ("guard_band_from", ctypes.c_int),
        ("use_mfma", ctypes.c_int),
        ("max_tile_waste", ctypes.c_float),
        ("layer_timing", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("dump_layer", ctypes.c_int),
        ("spmm_order", ctypes.c_int),
        ("mfma_groups_per_cu", ctypes.c_int),
        ("batch_concurrent", ctypes.c_int),
        ("guard_tail", ctypes.c_int),
        ("guard_concurrent", ctypes.c_int),
        ("keep_holes_ppm", ctypes.c_int),
        ("spmm_gather", ctypes.c_int),
        ("input_bits", ctypes.c_int),
        ("weight_mask", ctypes.c_int),
        ("batch_cu_share", ctypes.c_int),
        ("weight_budget_mb", ctypes.c_int),
        ("spmm_slab", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("categories", _c_i64p),
        ("live", _c_i64p),
        ("dump_vals", _c_f32p),
        ("dump_ids", _c_i64p),
        ("dump_cap_rows", ctypes.c_int64),
        ("layer_ms", _c_f32p),
        ("n_categories", ctypes.c_int64),
        ("dump_rows", ctypes.c_int64),
        ("ms_total", ctypes.c_double),
        ("ms_spmm", ctypes.c_double),
        ("ms_recompute", ctypes.c_double),
        ("n_flagged", ctypes.c_int64),
        ("n_replayed", ctypes.c_int64),
        ("n_changed", ctypes.c_int64),
        ("edges", ctypes.c_double),
        ("edges_per_sec", ctypes.c_double),
        ("alg_bytes", ctypes.c_double),
        ("mfma_layers", ctypes.c_int64),
        ("positions", _c_i64p),
        ("streamed_layers", ctypes.c_int64),
        ("layer_kernel", _c_i32p),
    ]


PLAN_AUTO, PLAN_MFMA, PLAN_SIMT, PLAN_MFMA_COMPACT, PLAN_ROWS, PLAN_SLAB = 0, 1, 2, 3, 4, 5
# SpMM kernels reported per layer (RunOutput.layer_kernel)
(KERNEL_SIMT, KERNEL_SIMT_ROWS, KERNEL_MFMA, KERNEL_MFMA_MASK, KERNEL_MFMA_LDS, KERNEL_MFMA_LDS_MASK, KERNEL_MFMA_BAND,
 KERNEL_MFMA_BAND_MASK, KERNEL_WAVE_ROWS, KERNEL_MFMA32, KERNEL_SLAB, KERNEL_SLAB_BITS) = range(12)
TUNE_FEATURES = 15


class LayerPlan(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int), ("order", ctypes.c_int), ("groups_per_cu", ctypes.c_int)]

    def as_tuple(self):
        return (self.kernel, self.order, self.groups_per_cu)


# every symbol include/cospdnn.h declares
EXPORTS = [
    "cospdnn_version", "cospdnn_strerror", "cospdnn_last_error", "cospdnn_kernel_name", "cospdnn_default_opts",
    "cospdnn_engine_create", "cospdnn_engine_destroy", "cospdnn_engine_info",
    "cospdnn_net_create", "cospdnn_net_destroy", "cospdnn_net_set_layer_csc", "cospdnn_net_set_layer_ref_bcsc",
    "cospdnn_net_generate_layers", "cospdnn_net_set_input_dense", "cospdnn_net_set_input_coo",
    "cospdnn_net_generate_input", "cospdnn_net_upload", "cospdnn_run", "cospdnn_run_noalloc",
    "cospdnn_generate_layer", "cospdnn_generate_input_rows", "cospdnn_layer_stats", "cospdnn_layer_tiles_decode",
    "cospdnn_layer_image_decode",
    "cospdnn_ref_bcsc_to_csc", "cospdnn_similarity_order",
    "cospdnn_net_set_plan", "cospdnn_num_candidates", "cospdnn_candidate", "cospdnn_net_layer_features",
    "cospdnn_tune_exhaustive", "cospdnn_tune_keep", "cospdnn_mlp_create", "cospdnn_mlp_destroy", "cospdnn_mlp_num_params",
    "cospdnn_mlp_get_params", "cospdnn_mlp_set_params", "cospdnn_mlp_loss_grad", "cospdnn_mlp_train",
    "cospdnn_mlp_predict", "cospdnn_tune_samples", "cospdnn_tune_mlp",
]


class CospdnnError(RuntimeError):
    def __init__(self, code, where, detail):
        super().__init__(f"{where} failed: code {code} ({detail})")
        self.code = code


_lib = None


def build():
    """Compiles libcospdnn.so for gfx950 with the package Makefile (hipcc)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(the engine has no fallback path)")
        L = ctypes.CDLL(LIB_PATH)
        L.cospdnn_version.restype = ctypes.c_char_p
        L.cospdnn_strerror.restype = ctypes.c_char_p
        L.cospdnn_strerror.argtypes = [ctypes.c_int]
        L.cospdnn_last_error.restype = ctypes.c_char_p
        L.cospdnn_kernel_name.restype = ctypes.c_char_p
        L.cospdnn_kernel_name.argtypes = [ctypes.c_int]
        L.cospdnn_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.cospdnn_default_opts.restype = None
        L.cospdnn_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.cospdnn_engine_destroy.argtypes = [ctypes.c_void_p]
        L.cospdnn_engine_destroy.restype = None
        L.cospdnn_engine_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _c_i64p]
        L.cospdnn_net_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                         ctypes.c_float, ctypes.POINTER(ctypes.c_void_p)]
        L.cospdnn_net_destroy.argtypes = [ctypes.c_void_p]
        L.cospdnn_net_destroy.restype = None
        L.cospdnn_net_set_layer_csc.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_i32p, _c_i32p, _c_f32p]
        L.cospdnn_net_set_layer_ref_bcsc.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_f32p, ctypes.c_int64,
                                                     _c_i32p, ctypes.c_int64, _c_i32p, ctypes.c_int64, _c_i32p,
                                                     ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.cospdnn_net_generate_layers.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                  ctypes.c_uint64]
        L.cospdnn_net_set_input_dense.argtypes = [ctypes.c_void_p, _c_f32p]
        L.cospdnn_net_set_input_coo.argtypes = [ctypes.c_void_p, ctypes.c_int64, _c_i64p, _c_i32p, _c_f32p]
        L.cospdnn_net_generate_input.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.cospdnn_net_upload.argtypes = [ctypes.c_void_p, ctypes.POINTER(Opts)]
        L.cospdnn_run.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(Opts),
                                  ctypes.POINTER(Result)]
        L.cospdnn_run_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Opts), ctypes.POINTER(Result)]
        L.cospdnn_similarity_order.argtypes = [ctypes.c_int, _c_i32p, _c_i32p, ctypes.c_int, ctypes.c_int, _c_i32p]
        L.cospdnn_generate_layer.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_float, _c_i32p, _c_i32p, _c_f32p]
        L.cospdnn_generate_input_rows.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                  _c_u8p]
        L.cospdnn_layer_stats.argtypes = [ctypes.c_int, _c_i32p, _c_i32p, _c_f32p, _c_i64p, _c_i64p, _c_f64p,
                                          ctypes.POINTER(ctypes.c_int)]
        L.cospdnn_layer_tiles_decode.argtypes = [ctypes.c_int, _c_i32p, _c_i32p, _c_f32p, ctypes.c_int64, _c_i32p,
                                                 _c_i32p, _c_f32p, _c_i64p]
        L.cospdnn_layer_image_decode.argtypes = [ctypes.c_int, _c_i32p, _c_i32p, _c_f32p, ctypes.c_int, ctypes.c_int64,
                                                 _c_i32p, _c_i32p, _c_f32p, _c_i64p]
        L.cospdnn_ref_bcsc_to_csc.argtypes = [ctypes.c_int, _c_f32p, ctypes.c_int64, _c_i32p, ctypes.c_int64,
                                              _c_i32p, ctypes.c_int64, _c_i32p, ctypes.c_int64, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int64, _c_i32p, _c_i32p, _c_f32p]
        _pl = ctypes.POINTER(LayerPlan)
        L.cospdnn_net_set_plan.argtypes = [ctypes.c_void_p, _pl]
        L.cospdnn_num_candidates.argtypes = []
        L.cospdnn_candidate.argtypes = [ctypes.c_int, _pl]
        L.cospdnn_net_layer_features.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _pl, _c_f32p]
        L.cospdnn_tune_exhaustive.argtypes = [ctypes.c_void_p, ctypes.POINTER(Opts), ctypes.c_int, _c_f32p, _c_i64p, _pl]
        L.cospdnn_tune_keep.argtypes = [ctypes.c_void_p, ctypes.POINTER(Opts), _c_i32p, ctypes.c_int, ctypes.c_int,
                                        _c_f32p, _c_i32p]
        L.cospdnn_mlp_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.cospdnn_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.cospdnn_mlp_destroy.restype = None
        L.cospdnn_mlp_num_params.argtypes = [ctypes.c_void_p]
        L.cospdnn_mlp_get_params.argtypes = [ctypes.c_void_p, _c_f32p]
        L.cospdnn_mlp_set_params.argtypes = [ctypes.c_void_p, _c_f32p]
        L.cospdnn_mlp_loss_grad.argtypes = [ctypes.c_void_p, _c_f32p, _c_f32p, ctypes.c_int, _c_f32p, _c_f32p]
        L.cospdnn_mlp_train.argtypes = [ctypes.c_void_p, _c_f32p, _c_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        _c_f32p]
        L.cospdnn_mlp_predict.argtypes = [ctypes.c_void_p, _c_f32p, ctypes.c_int, _c_f32p]
        L.cospdnn_tune_samples.argtypes = [ctypes.c_void_p, _c_f32p, _c_i64p, _c_f32p, _c_f32p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int)]
        L.cospdnn_tune_mlp.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _c_i64p, _pl, _c_f32p]
        _lib = L
    return _lib


def _check(rc, where):
    if rc != 0:
        L = lib()
        raise CospdnnError(rc, where, f"{L.cospdnn_strerror(rc).decode()}: {L.cospdnn_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(t)


def default_opts(**kw):
    o = Opts()
    lib().cospdnn_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Engine:
    """One device and one stream."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        _check(lib().cospdnn_engine_create(device, ctypes.byref(self.h)), "cospdnn_engine_create")

    def info(self):
        cu = ctypes.c_int()
        hbm = ctypes.c_int64()
        _check(lib().cospdnn_engine_info(self.h, ctypes.byref(cu), ctypes.byref(hbm)), "cospdnn_engine_info")
        return cu.value, hbm.value

    def close(self):
        if self.h:
            lib().cospdnn_engine_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    def __init__(self, engine, N, L, S, bias):
        self.engine = engine
        self.N, self.L, self.S, self.bias = N, L, S, float(bias)
        self.h = ctypes.c_void_p()
        _check(lib().cospdnn_net_create(engine.h, N, L, S, float(bias), ctypes.byref(self.h)), "cospdnn_net_create")

    def set_layer_csc(self, l, colptr, rows, vals):
        cp = np.ascontiguousarray(colptr, dtype=np.int32)
        rw = np.ascontiguousarray(rows, dtype=np.int32)
        vl = np.ascontiguousarray(vals, dtype=np.float32)
        _check(lib().cospdnn_net_set_layer_csc(self.h, l, _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p)),
               "cospdnn_net_set_layer_csc")

    def set_layer_ref_bcsc(self, l, values, row_idx, ptr, fuse=None, blocksize_c=16, blocksize_r=16):
        v = np.ascontiguousarray(values, dtype=np.float32)
        ri = np.ascontiguousarray(row_idx, dtype=np.int32)
        pt = np.ascontiguousarray(ptr, dtype=np.int32)
        fu = None if fuse is None else np.ascontiguousarray(fuse, dtype=np.int32)
        _check(lib().cospdnn_net_set_layer_ref_bcsc(self.h, l, _p(v, _c_f32p), len(v), _p(ri, _c_i32p), len(ri),
                                                    _p(pt, _c_i32p), len(pt),
                                                    None if fu is None else _p(fu, _c_i32p),
                                                    0 if fu is None else len(fu), blocksize_c, blocksize_r),
               "cospdnn_net_set_layer_ref_bcsc")

    def generate_layers(self, kind, seed, K=32, wval=0.0625):
        _check(lib().cospdnn_net_generate_layers(self.h, kind, K, float(wval), seed), "cospdnn_net_generate_layers")

    def set_input_dense(self, y0):
        y = np.ascontiguousarray(y0, dtype=np.float32)
        assert y.shape == (self.S, self.N)
        _check(lib().cospdnn_net_set_input_dense(self.h, _p(y, _c_f32p)), "cospdnn_net_set_input_dense")

    def set_input_coo(self, sample, neuron, val):
        s = np.ascontiguousarray(sample, dtype=np.int64)
        n = np.ascontiguousarray(neuron, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float32)
        _check(lib().cospdnn_net_set_input_coo(self.h, len(s), _p(s, _c_i64p), _p(n, _c_i32p), _p(v, _c_f32p)),
               "cospdnn_net_set_input_coo")

    def generate_input(self, seed):
        _check(lib().cospdnn_net_generate_input(self.h, seed), "cospdnn_net_generate_input")

    def upload(self, opts=None):
        o = opts if opts is not None else default_opts()
        _check(lib().cospdnn_net_upload(self.h, ctypes.byref(o)), "cospdnn_net_upload")

    def set_plan(self, plan):
        """plan: sequence of L (kernel, order, groups_per_cu) tuples or LayerPlan, or None."""
        if plan is None:
            _check(lib().cospdnn_net_set_plan(self.h, None), "cospdnn_net_set_plan")
            return
        arr = (LayerPlan * self.L)()
        for l, p in enumerate(plan):
            t = p.as_tuple() if isinstance(p, LayerPlan) else tuple(p)
            arr[l] = LayerPlan(*t)
        _check(lib().cospdnn_net_set_plan(self.h, arr), "cospdnn_net_set_plan")

    def layer_features(self, layer, live_rows, cand, rows_prev=None):
        """rows_prev: rows entering the layer before (default: live_rows, nothing removed)."""
        f = np.zeros(TUNE_FEATURES, dtype=np.float32)
        c = cand if isinstance(cand, LayerPlan) else LayerPlan(*cand)
        prev = int(live_rows if rows_prev is None else rows_prev)
        _check(lib().cospdnn_net_layer_features(self.h, layer, int(live_rows), prev, ctypes.byref(c), _p(f, _c_f32p)),
               "cospdnn_net_layer_features")
        return f

    def close(self):
        if self.h:
            lib().cospdnn_net_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RunOutput:
    pass


def run(nets, opts=None, dump_cap_rows=0):
    """Runs the networks; returns one RunOutput per network."""
    o = opts if opts is not None else default_opts()
    n = len(nets)
    handles = (ctypes.c_void_p * n)(*[x.h for x in nets])
    return _run_with(nets, dump_cap_rows,
                     lambda res: _check(lib().cospdnn_run(handles, n, ctypes.byref(o), res), "cospdnn_run"))


def run_noalloc(net, opts=None):
    """One network through the allocation-free entry point (cospdnn_run_noalloc)."""
    o = opts if opts is not None else default_opts()
    return _run_with([net], 0, lambda res: _check(lib().cospdnn_run_noalloc(net.h, ctypes.byref(o), res),
                                                  "cospdnn_run_noalloc"))[0]


def _run_with(nets, dump_cap_rows, call):
    n = len(nets)
    res = (Result * n)()
    keep = []
    for i, net in enumerate(nets):
        cats = np.zeros(net.S, dtype=np.int64)
        live = np.zeros(net.L + 1, dtype=np.int64)
        lms = np.zeros(max(net.L, 1), dtype=np.float32)
        res[i].categories = _p(cats, _c_i64p)
        res[i].live = _p(live, _c_i64p)
        pos = np.zeros(net.L + 1, dtype=np.int64)
        res[i].positions = _p(pos, _c_i64p)
        res[i].layer_ms = _p(lms, _c_f32p)
        lk = np.full(max(net.L, 1), -1, dtype=np.int32)
        res[i].layer_kernel = _p(lk, _c_i32p)
        dv = di = None
        if dump_cap_rows > 0:
            dv = np.zeros((dump_cap_rows, net.N), dtype=np.float32)
            di = np.zeros(dump_cap_rows, dtype=np.int64)
            res[i].dump_vals = _p(dv, _c_f32p)
            res[i].dump_ids = _p(di, _c_i64p)
            res[i].dump_cap_rows = dump_cap_rows
        keep.append((cats, live, lms, dv, di, pos, lk))
    call(res)
    outs = []
    for i in range(n):
        cats, live, lms, dv, di, pos, lk = keep[i]
        r = res[i]
        out = RunOutput()
        out.categories = cats[: r.n_categories].copy()
        out.live = live.copy()
        out.positions = pos.copy()
        out.layer_ms = lms[: nets[i].L].copy()
        out.layer_kernel = lk[: nets[i].L].copy()
        for f in ("ms_total", "ms_spmm", "ms_recompute", "n_flagged", "n_replayed", "n_changed", "edges", "edges_per_sec",
                  "alg_bytes", "mfma_layers", "dump_rows", "streamed_layers"):
            setattr(out, f, getattr(r, f))
        if dv is not None:
            out.dump_vals = dv[: r.dump_rows].copy()
            out.dump_ids = di[: r.dump_rows].copy()
        outs.append(out)
    return outs


def version():
    return lib().cospdnn_version().decode()


def kernel_name(kernel):
    """Device function of a per-layer SpMM kernel id (RunOutput.layer_kernel), or None."""
    name = lib().cospdnn_kernel_name(int(kernel))
    return name.decode() if name is not None else None


def generate_layer(seed, kind, N, K, l, wval=0.0625):
    cp = np.zeros(N + 1, dtype=np.int32)
    rw = np.zeros(N * K, dtype=np.int32)
    vl = np.zeros(N * K, dtype=np.float32)
    _check(lib().cospdnn_generate_layer(seed, kind, N, K, l, float(wval), _p(cp, _c_i32p), _p(rw, _c_i32p),
                                        _p(vl, _c_f32p)), "cospdnn_generate_layer")
    return cp, rw, vl


def generate_input_rows(seed, N, s0, count):
    out = np.zeros((count, N), dtype=np.uint8)
    _check(lib().cospdnn_generate_input_rows(seed, N, s0, count, _p(out, _c_u8p)), "cospdnn_generate_input_rows")
    return out


def layer_stats(N, colptr, rows, vals):
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    rw = np.ascontiguousarray(rows, dtype=np.int32)
    vl = np.ascontiguousarray(vals, dtype=np.float32)
    nblk, nch = ctypes.c_int64(), ctypes.c_int64()
    waste = ctypes.c_double()
    ex = ctypes.c_int()
    _check(lib().cospdnn_layer_stats(N, _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p), ctypes.byref(nblk),
                                     ctypes.byref(nch), ctypes.byref(waste), ctypes.byref(ex)), "cospdnn_layer_stats")
    return {"nblk": nblk.value, "nchunk": nch.value, "tile_waste": waste.value, "fp16_exact": bool(ex.value)}


def layer_tiles_decode(N, colptr, rows, vals):
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    rw = np.ascontiguousarray(rows, dtype=np.int32)
    vl = np.ascontiguousarray(vals, dtype=np.float32)
    n = ctypes.c_int64()
    _check(lib().cospdnn_layer_tiles_decode(N, _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p), 0, None, None,
                                            None, ctypes.byref(n)), "cospdnn_layer_tiles_decode")
    ins = np.zeros(n.value, dtype=np.int32)
    outs = np.zeros(n.value, dtype=np.int32)
    v = np.zeros(n.value, dtype=np.float32)
    _check(lib().cospdnn_layer_tiles_decode(N, _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p), n.value,
                                            _p(ins, _c_i32p), _p(outs, _c_i32p), _p(v, _c_f32p), ctypes.byref(n)),
           "cospdnn_layer_tiles_decode")
    return ins, outs, v


IMAGE_TILES, IMAGE_MASK, IMAGE_BAND, IMAGE_BAND_MASK, IMAGE_SLAB = 0, 1, 2, 3, 4


def layer_image_decode(N, colptr, rows, vals, image):
    """(input, output, value) triples of one of a layer's weight images (IMAGE_*), decoded as
    the kernels read them; raises CospdnnError when the layer has no such image."""
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    rw = np.ascontiguousarray(rows, dtype=np.int32)
    vl = np.ascontiguousarray(vals, dtype=np.float32)
    n = ctypes.c_int64()
    args = (N, _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p), int(image))
    _check(lib().cospdnn_layer_image_decode(*args, 0, None, None, None, ctypes.byref(n)), "cospdnn_layer_image_decode")
    ins = np.zeros(n.value, dtype=np.int32)
    outs = np.zeros(n.value, dtype=np.int32)
    v = np.zeros(n.value, dtype=np.float32)
    _check(lib().cospdnn_layer_image_decode(*args, n.value, _p(ins, _c_i32p), _p(outs, _c_i32p), _p(v, _c_f32p),
                                            ctypes.byref(n)), "cospdnn_layer_image_decode")
    return ins, outs, v


def ref_bcsc_to_csc(N, values, row_idx, ptr, fuse=None, blocksize_c=16, blocksize_r=16):
    v = np.ascontiguousarray(values, dtype=np.float32)
    ri = np.ascontiguousarray(row_idx, dtype=np.int32)
    pt = np.ascontiguousarray(ptr, dtype=np.int32)
    fu = None if fuse is None else np.ascontiguousarray(fuse, dtype=np.int32)
    cap = int(np.count_nonzero(v))
    cp = np.zeros(N + 1, dtype=np.int32)
    rw = np.zeros(max(cap, 1), dtype=np.int32)
    vl = np.zeros(max(cap, 1), dtype=np.float32)
    _check(lib().cospdnn_ref_bcsc_to_csc(N, _p(v, _c_f32p), len(v), _p(ri, _c_i32p), len(ri), _p(pt, _c_i32p),
                                         len(pt), None if fu is None else _p(fu, _c_i32p),
                                         0 if fu is None else len(fu), blocksize_c, blocksize_r, cap,
                                         _p(cp, _c_i32p), _p(rw, _c_i32p), _p(vl, _c_f32p)), "cospdnn_ref_bcsc_to_csc")
    nnz = cp[N]
    return cp, rw[:nnz], vl[:nnz]


def similarity_order(N, colptr, rows, group=32, overlap=True):
    """Output order of one layer that groups outputs reading the same inputs (see
    cospdnn_similarity_order); order[i] = the output at position i."""
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    rw = np.ascontiguousarray(rows, dtype=np.int32)
    order = np.zeros(N, dtype=np.int32)
    _check(lib().cospdnn_similarity_order(N, _p(cp, _c_i32p), _p(rw, _c_i32p), group, 1 if overlap else 0,
                                          _p(order, _c_i32p)), "cospdnn_similarity_order")
    return order


def candidates():
    out = []
    for k in range(lib().cospdnn_num_candidates()):
        c = LayerPlan()
        _check(lib().cospdnn_candidate(k, ctypes.byref(c)), "cospdnn_candidate")
        out.append(c.as_tuple())
    return out


def tune_exhaustive(net, opts=None, reps=3):
    """Per-layer exhaustive tuning of the FP16 pass: returns (plan [L tuples], ms table
    [ncand][L], row positions entering each layer [L+1])."""
    o = opts if opts is not None else default_opts()
    nc = lib().cospdnn_num_candidates()
    table = np.zeros((nc, net.L), dtype=np.float32)
    live = np.zeros(net.L + 1, dtype=np.int64)
    plan = (LayerPlan * max(net.L, 1))()
    _check(lib().cospdnn_tune_exhaustive(net.h, ctypes.byref(o), reps, _p(table, _c_f32p), _p(live, _c_i64p), plan),
           "cospdnn_tune_exhaustive")
    return [plan[l].as_tuple() for l in range(net.L)], table, live


KEEP_CANDIDATES = (0, 3000, 10000, 30000, 100000, 300000, 1000000)


def tune_keep(net, opts=None, candidates=KEEP_CANDIDATES, reps=3):
    """Compaction threshold: returns (best keep_holes_ppm, median forward ms per candidate)."""
    o = opts if opts is not None else default_opts()
    cand = np.ascontiguousarray(candidates, dtype=np.int32)
    ms = np.zeros(len(cand), dtype=np.float32)
    best = ctypes.c_int32(0)
    _check(lib().cospdnn_tune_keep(net.h, ctypes.byref(o), _p(cand, _c_i32p), len(cand), reps, _p(ms, _c_f32p),
                                   ctypes.byref(best)), "cospdnn_tune_keep")
    return int(best.value), ms


def tune_samples(net, table, live):
    """Cost-model training rows (features, ln(1000*ms + 1)) from an exhaustive table."""
    nc = table.shape[0]
    cap = nc * net.L
    X = np.zeros((cap, TUNE_FEATURES), dtype=np.float32)
    y = np.zeros(cap, dtype=np.float32)
    n = ctypes.c_int()
    t = np.ascontiguousarray(table, dtype=np.float32)
    lv = np.ascontiguousarray(live, dtype=np.int64)
    _check(lib().cospdnn_tune_samples(net.h, _p(t, _c_f32p), _p(lv, _c_i64p), _p(X, _c_f32p), _p(y, _c_f32p), cap,
                                      ctypes.byref(n)), "cospdnn_tune_samples")
    return X[:n.value], y[:n.value]


class Mlp:
    """Cost-model MLP on an engine's device (n_in -> hidden -> hidden -> 1, ReLU)."""

    def __init__(self, engine, n_in=TUNE_FEATURES, hidden=64, seed=1):
        self.engine = engine
        self.n_in, self.hidden = n_in, hidden
        self.h = ctypes.c_void_p()
        _check(lib().cospdnn_mlp_create(engine.h, n_in, hidden, seed, ctypes.byref(self.h)), "cospdnn_mlp_create")

    @property
    def num_params(self):
        return lib().cospdnn_mlp_num_params(self.h)

    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        _check(lib().cospdnn_mlp_get_params(self.h, _p(p, _c_f32p)), "cospdnn_mlp_get_params")
        return p

    def set_params(self, p):
        q = np.ascontiguousarray(p, dtype=np.float32)
        assert q.size == self.num_params
        _check(lib().cospdnn_mlp_set_params(self.h, _p(q, _c_f32p)), "cospdnn_mlp_set_params")

    def loss_grad(self, X, y):
        Xc = np.ascontiguousarray(X, dtype=np.float32)
        yc = np.ascontiguousarray(y, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        _check(lib().cospdnn_mlp_loss_grad(self.h, _p(Xc, _c_f32p), _p(yc, _c_f32p), len(yc), _p(loss, _c_f32p),
                                           _p(g, _c_f32p)), "cospdnn_mlp_loss_grad")
        return float(loss[0]), g

    def train(self, X, y, steps=2000, lr=3e-3):
        Xc = np.ascontiguousarray(X, dtype=np.float32)
        yc = np.ascontiguousarray(y, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().cospdnn_mlp_train(self.h, _p(Xc, _c_f32p), _p(yc, _c_f32p), len(yc), steps, float(lr),
                                       _p(loss, _c_f32p)), "cospdnn_mlp_train")
        return float(loss[0])

    def predict(self, X):
        Xc = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(Xc.shape[0], dtype=np.float32)
        _check(lib().cospdnn_mlp_predict(self.h, _p(Xc, _c_f32p), Xc.shape[0], _p(out, _c_f32p)), "cospdnn_mlp_predict")
        return out

    def plan(self, net, live):
        """Per-layer argmin of the predicted times: (plan [L tuples], predicted ms [ncand][L])."""
        nc = lib().cospdnn_num_candidates()
        pred = np.zeros((nc, net.L), dtype=np.float32)
        lv = np.ascontiguousarray(live, dtype=np.int64)
        plan = (LayerPlan * max(net.L, 1))()
        _check(lib().cospdnn_tune_mlp(net.h, self.h, _p(lv, _c_i64p), plan, _p(pred, _c_f32p)), "cospdnn_tune_mlp")
        return [plan[l].as_tuple() for l in range(net.L)], pred

    def close(self):
        if self.h:
            lib().cospdnn_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
