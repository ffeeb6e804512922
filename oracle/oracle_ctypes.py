"""ctypes binding of the CPU oracle in oracle.c.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library never
loads it.

    net = OracleNet.generated(N=1024, L=120, bias=-0.3, kind=RANDOM, seed=20261203)
    alive, live = forward_fp64(net, s0=0, S=60000)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

RANDOM = 0
BANDED = 1

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_f32p = ctypes.POINTER(ctypes.c_float)
_c_u8p = ctypes.POINTER(ctypes.c_uint8)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_f64p = ctypes.POINTER(ctypes.c_double)


class _OrcNet(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int),
        ("L", ctypes.c_int),
        ("bias", ctypes.c_float),
        ("gen_kind", ctypes.c_int),
        ("K", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("wval", ctypes.c_float),
        ("colptr", ctypes.POINTER(_c_i32p)),
        ("rows", ctypes.POINTER(_c_i32p)),
        ("vals", ctypes.POINTER(_c_f32p)),
        ("y0_gen", ctypes.c_int),
        ("y0", _c_u8p),
    ]


_lib = None


def build():
    """Compiles liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_forward_fp64.argtypes = [ctypes.POINTER(_OrcNet), ctypes.c_int64, ctypes.c_int64, _c_u8p, _c_i64p,
                                       _c_i64p, ctypes.c_int64, _c_f64p, ctypes.c_int]
        L.orc_forward_fp64.restype = ctypes.c_int
        L.orc_forward_fp32.argtypes = [ctypes.POINTER(_OrcNet), ctypes.c_int64, ctypes.c_int64, _c_u8p, _c_i64p,
                                       ctypes.c_int]
        L.orc_forward_fp32.restype = ctypes.c_int
        L.orc_gen_layer.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, _c_i32p, _c_i32p, _c_f32p]
        L.orc_gen_layer.restype = None
        L.orc_gen_y0_row.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int64, _c_u8p]
        L.orc_gen_y0_row.restype = None
        L.orc_y0_row_thr.argtypes = [ctypes.c_uint64, ctypes.c_int64]
        L.orc_y0_row_thr.restype = ctypes.c_uint32
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class OracleNet:
    """A network for the oracle: generated from a seed, or explicit CSC + dense Y0."""

    def __init__(self):
        self._keep = []
        self.s = _OrcNet()

    @classmethod
    def generated(cls, N, L, bias, kind, seed, K=32, wval=0.0625):
        n = cls()
        n.N, n.L = N, L
        n.s.N, n.s.L, n.s.bias = N, L, float(bias)
        n.s.gen_kind, n.s.K, n.s.seed, n.s.wval = kind, K, seed, float(wval)
        n.s.y0_gen = 1
        return n

    @classmethod
    def explicit(cls, N, bias, layers, y0):
        """layers: list of (colptr int32[N+1], rows int32[nnz], vals float32[nnz]); y0: uint8 [S][N]."""
        n = cls()
        L = len(layers)
        n.N, n.L = N, L
        n.s.N, n.s.L, n.s.bias = N, L, float(bias)
        n.s.gen_kind = -1
        cps = (_c_i32p * max(L, 1))()
        rws = (_c_i32p * max(L, 1))()
        vls = (_c_f32p * max(L, 1))()
        for l, (cp, rw, vl) in enumerate(layers):
            cp = np.ascontiguousarray(cp, dtype=np.int32)
            rw = np.ascontiguousarray(rw, dtype=np.int32)
            vl = np.ascontiguousarray(vl, dtype=np.float32)
            n._keep += [cp, rw, vl]
            cps[l], rws[l], vls[l] = _ptr(cp, _c_i32p), _ptr(rw, _c_i32p), _ptr(vl, _c_f32p)
        n._keep += [cps, rws, vls]
        n.s.colptr = ctypes.cast(cps, ctypes.POINTER(_c_i32p))
        n.s.rows = ctypes.cast(rws, ctypes.POINTER(_c_i32p))
        n.s.vals = ctypes.cast(vls, ctypes.POINTER(_c_f32p))
        y0 = np.ascontiguousarray(y0, dtype=np.uint8)
        assert y0.ndim == 2 and y0.shape[1] == N
        n._keep.append(y0)
        n.s.y0_gen = 0
        n.s.y0 = _ptr(y0, _c_u8p)
        n.S = y0.shape[0]
        return n


def forward_fp64(net, s0, S, final_ids=None, nthreads=0):
    """FP64 oracle over samples [s0, s0+S). Returns (alive uint8[S], live int64[L+1]) and, if
    final_ids (offsets relative to s0, ascending) is given, the final FP64 rows [len][N]."""
    alive = np.zeros(S, dtype=np.uint8)
    live = np.zeros(net.L + 1, dtype=np.int64)
    if final_ids is not None:
        fid = np.ascontiguousarray(final_ids, dtype=np.int64)
        out = np.zeros((len(fid), net.N), dtype=np.float64)
        rc = lib().orc_forward_fp64(ctypes.byref(net.s), s0, S, _ptr(alive, _c_u8p), _ptr(live, _c_i64p),
                                    _ptr(fid, _c_i64p), len(fid), _ptr(out, _c_f64p), nthreads)
        assert rc == 0
        return alive, live, out
    rc = lib().orc_forward_fp64(ctypes.byref(net.s), s0, S, _ptr(alive, _c_u8p), _ptr(live, _c_i64p), None, 0, None,
                                nthreads)
    assert rc == 0
    return alive, live


def forward_fp32(net, s0, S, nthreads=0):
    """Repo-style FP32 CPU forward (layer-major, dead-row pruning). Returns (alive, live)."""
    alive = np.zeros(S, dtype=np.uint8)
    live = np.zeros(net.L + 1, dtype=np.int64)
    rc = lib().orc_forward_fp32(ctypes.byref(net.s), s0, S, _ptr(alive, _c_u8p), _ptr(live, _c_i64p), nthreads)
    assert rc == 0
    return alive, live


def gen_layer(seed, kind, N, K, l, wval=0.0625):
    cp = np.zeros(N + 1, dtype=np.int32)
    rw = np.zeros(N * K, dtype=np.int32)
    vl = np.zeros(N * K, dtype=np.float32)
    lib().orc_gen_layer(seed, kind, N, K, l, wval, _ptr(cp, _c_i32p), _ptr(rw, _c_i32p), _ptr(vl, _c_f32p))
    return cp, rw, vl


def gen_y0_rows(seed, N, s0, count):
    out = np.zeros((count, N), dtype=np.uint8)
    for r in range(count):
        lib().orc_gen_y0_row(seed, N, s0 + r, _ptr(out[r], _c_u8p))
    return out


def y0_row_threshold(seed, s):
    return int(lib().orc_y0_row_thr(seed, s))


def num_threads():
    return int(lib().orc_num_threads())
