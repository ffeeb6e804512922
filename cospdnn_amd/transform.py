"""Similarity transform of a reference-format dataset directory: the preprocessing the
reference expects done offline (README.md:9) and whose result it reads as fuse lists
(src/cospdnn.cpp:14-44).

Layer by layer, outputs that read the same or overlapping inputs get consecutive neuron ids
(cospdnn_similarity_order, groups of 32), the next layer's inputs are relabelled with them,
and the layer's fuse list orders its input rows by first use in the new output order, so a
group whose members share their inputs reads consecutive fused rows.  For banded networks
(groups of 32 outputs reading the same 32 rows) that is exactly two 16-row blocks per
16-output block column, the layout the reference kernel assumes (src/kernels/kernel.cu:310,
316).  Sample ids are kept and the relabelling is consistent between layers, so every
sample's values and the category set are unchanged.

Reads <src>/sparse-images-<N>.tsv, <src>/neuron<N>/<N>-l<k>.tsv and, when present,
<src>/fuse_matrix/neuron<N>/<N>_<L>_fuse.tsv (rows of the weight files are then fused ids);
writes the same files under <dst>.

    python -m cospdnn_amd.transform <N> <L> <src_dir> <dst_dir>
"""
import os
import shutil
import sys

import numpy as np

from . import _lib

GROUP = 32


def _weight_path(d, N, l):
    return os.path.join(d, f"neuron{N}", f"{N}-l{l + 1}.tsv")


def _fuse_path(d, N, L):
    return os.path.join(d, "fuse_matrix", f"neuron{N}", f"{N}_{L}_fuse.tsv")


def read_layer(path, N):
    """COO "row col value" (1-based) -> (rows, cols, vals) 0-based."""
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if a.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32)
    return a[:, 0].astype(np.int64) - 1, a[:, 1].astype(np.int64) - 1, a[:, 2].astype(np.float32)


def to_csc(rows, cols, vals, N):
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    cp = np.zeros(N + 1, np.int64)
    np.add.at(cp, cols + 1, 1)
    return np.cumsum(cp).astype(np.int32), rows.astype(np.int32), vals


def transform_layer(cp, rows, vals, N, group=GROUP):
    """One layer whose input rows are already in the new labels of the layer before.  Returns
    (new output label of each old output, fuse list, COO of the transformed layer: fused rows,
    new columns, values)."""
    order = _lib.similarity_order(N, cp, rows, group, overlap=True)
    lab = np.empty(N, np.int64)
    lab[order] = np.arange(N)
    # fused rows: every group's union, in the new output order, rows at their first use
    seen = np.zeros(N, bool)
    fuse = []
    for g in range(0, N, group):
        u = np.unique(np.concatenate([rows[cp[j]:cp[j + 1]] for j in order[g:g + group]] + [np.zeros(0, np.int32)]))
        fresh = u[~seen[u]]
        seen[fresh] = True
        fuse.extend(fresh.tolist())
    fuse.extend(np.flatnonzero(~seen).tolist())
    fuse = np.array(fuse, np.int64)
    inv = np.empty(N, np.int64)
    inv[fuse] = np.arange(N)
    cols = np.repeat(np.arange(N), np.diff(cp))
    return lab, fuse, (inv[rows], lab[cols], vals)


def transform_dataset(N, L, src, dst):
    os.makedirs(os.path.join(dst, f"neuron{N}"), exist_ok=True)
    os.makedirs(os.path.join(dst, "fuse_matrix", f"neuron{N}"), exist_ok=True)
    shutil.copyfile(os.path.join(src, f"sparse-images-{N}.tsv"), os.path.join(dst, f"sparse-images-{N}.tsv"))
    src_fuse = None
    if os.path.exists(_fuse_path(src, N, L)):
        src_fuse = np.loadtxt(_fuse_path(src, N, L), dtype=np.int64, ndmin=2)[:L]
    lab_prev = np.arange(N)  # layer 0 reads the input features: no relabelling
    fuses = []
    for l in range(L):
        r, c, v = read_layer(_weight_path(src, N, l), N)
        if src_fuse is not None:
            r = src_fuse[l][r]  # fused row -> input neuron of the source network
        cp, rows, vals = to_csc(lab_prev[r], c, v, N)
        lab, fuse, (fr, nc, nv) = transform_layer(cp, rows, vals, N)
        with open(_weight_path(dst, N, l), "w") as f:
            f.write("".join(f"{a + 1}\t{b + 1}\t{x:g}\n" for a, b, x in zip(fr.tolist(), nc.tolist(), nv.tolist())))
        fuses.append(fuse)
        lab_prev = lab
    with open(_fuse_path(dst, N, L), "w") as f:
        for fuse in fuses:
            f.write("\t".join(str(x) for x in fuse.tolist()) + "\n")
    return dst


def main(argv):
    if len(argv) != 5:
        print(__doc__)
        return 2
    transform_dataset(int(argv[1]), int(argv[2]), argv[3], argv[4])
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
