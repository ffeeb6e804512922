"""Writes a synthetic network as a reference-format dataset directory.

Layout (README.md:5-9 and src/cospdnn.cpp:14-79 of the reference):
    <dir>/sparse-images-<N>.tsv                      "sample neuron 1" per nonzero, 1-based
    <dir>/neuron<N>/<N>-l<k>.tsv                     "row col value", 1-based, k = 1..L
    <dir>/fuse_matrix/neuron<N>/<N>_<L>_fuse.tsv     L lines of N fused-row -> neuron ids

Banded layers are stored in the fused row order the similarity transform
produces: the 32 inputs read by one group of outputs are consecutive fused rows,
and the fuse list maps them back to neurons.  Random layers use the identity, and so does
every layer with fused=False (the raw network, as cospdnn_amd.transform takes it).
"""
import os

import numpy as np

from . import _lib


def banded_fuse(N, K=32):
    s = N // K
    r = np.arange(N)
    return ((r % K) * s + r // K).astype(np.int64)


def write_dataset(path, N, L, S, kind, seed, K=32, wval=0.0625, chunk=4096, fused=True):
    os.makedirs(os.path.join(path, f"neuron{N}"), exist_ok=True)
    os.makedirs(os.path.join(path, "fuse_matrix", f"neuron{N}"), exist_ok=True)
    with open(os.path.join(path, f"sparse-images-{N}.tsv"), "w") as f:
        for s0 in range(0, S, chunk):
            rows = _lib.generate_input_rows(seed, N, s0, min(chunk, S - s0))
            ss, nn = np.nonzero(rows)
            f.write("".join(f"{s0 + a + 1}\t{b + 1}\t1\n" for a, b in zip(ss.tolist(), nn.tolist())))
    fuse = banded_fuse(N, K) if kind == _lib.GEN_BANDED and fused else np.arange(N, dtype=np.int64)
    inv = np.empty(N, dtype=np.int64)
    inv[fuse] = np.arange(N)
    with open(os.path.join(path, "fuse_matrix", f"neuron{N}", f"{N}_{L}_fuse.tsv"), "w") as ff:
        for l in range(L):
            cp, rw, vl = _lib.generate_layer(seed, kind, N, K, l, wval)
            cols = np.repeat(np.arange(N), np.diff(cp))
            with open(os.path.join(path, f"neuron{N}", f"{N}-l{l + 1}.tsv"), "w") as f:
                f.write("".join(f"{r + 1}\t{c + 1}\t{v:g}\n" for r, c, v in zip(inv[rw].tolist(), cols.tolist(),
                                                                               vl.tolist())))
            ff.write("\t".join(str(x) for x in fuse.tolist()) + "\n")
    return path
