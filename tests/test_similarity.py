"""The similarity grouping of outputs (cospdnn_similarity_order, also the column order of the
MFMA tile image) against constructions whose best grouping is known."""
import os
import subprocess

import numpy as np

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

REF_DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_driver")


def csc(cols, N):
    cp = np.zeros(N + 1, np.int32)
    cp[1:] = np.cumsum([len(c) for c in cols])
    return cp, np.concatenate(cols).astype(np.int32)


def unions(order, cp, rw, group=32):
    return [len(set(np.concatenate([rw[cp[j]:cp[j + 1]] for j in order[g:g + group]]).tolist()))
            for g in range(0, len(order), group)]


def test_planted_overlapping_groups_are_recovered():
    """32 groups of 32 outputs; every output of group g reads the group's 28 rows plus 4 rows of
    its own (so no two outputs share an input set); outputs scrambled.  Grouping by overlap puts
    each planted group into one block column (union 28 + 32 x 4 at most); grouping by identical
    sets cannot (every set is distinct) and keeps the scrambled order."""
    rng = np.random.default_rng(7)
    N, G = 1024, 32
    perm = rng.permutation(N)  # output j of the layer is planted output perm[j]
    base = [rng.choice(N, 28, replace=False) for _ in range(N // G)]
    cols = []
    for j in range(N):
        g = perm[j] // G
        extra = rng.choice(np.setdiff1d(np.arange(N), base[g]), 4, replace=False)
        cols.append(np.sort(np.concatenate([base[g], extra])))
    cp, rw = csc(cols, N)
    order = cd.similarity_order(N, cp, rw, G, overlap=True)
    assert sorted(order.tolist()) == list(range(N))
    for b in range(N // G):
        assert len({perm[j] // G for j in order[b * G:(b + 1) * G]}) == 1
    assert max(unions(order, cp, rw)) <= 28 + 4 * G
    plain = cd.similarity_order(N, cp, rw, G, overlap=False)
    assert np.array_equal(plain, np.arange(N))
    assert min(unions(plain, cp, rw)) > 28 + 4 * G
    st = cd.layer_stats(N, cp, rw, np.full(len(rw), 0.0625, np.float32))
    assert st["tile_waste"] <= (28 + 4 * G + 31) // 32 * 32 * G / (32 * G)  # 5 chunks a block column


def test_banded_classes_and_identical_sets():
    """Banded layers (output j reads (pi(j) + k N/32) mod N): the 32 outputs of a residue class
    read the same rows and land in one block column, with either grouping."""
    N = 2048
    cp, rw, _ = cd.generate_layer(11, cd.GEN_BANDED, N, 32, 3)
    for overlap in (True, False):
        order = cd.similarity_order(N, cp, rw, 32, overlap=overlap)
        assert max(unions(order, cp, rw)) == 32


def test_random_layer_unions_shrink():
    """Random rows: grouping by overlap gives smaller block-column unions than natural order."""
    N = 4096
    cp, rw, _ = cd.generate_layer(12, cd.GEN_RANDOM, N, 32, 0)
    greedy = np.mean(unions(cd.similarity_order(N, cp, rw, 32, overlap=True), cp, rw))
    natural = np.mean(unions(np.arange(N), cp, rw))
    # natural: 4096 (1 - exp(-0.25)) ~ 906 rows; greedy about 10% fewer
    assert 880 < natural < 930
    assert greedy < 0.94 * natural


def _transformed_layers(dst, N, L):
    from cospdnn_amd import transform as tf

    fuse = np.loadtxt(os.path.join(dst, "fuse_matrix", f"neuron{N}", f"{N}_{L}_fuse.tsv"), dtype=np.int64, ndmin=2)
    out = []
    for l in range(L):
        r, c, v = tf.read_layer(os.path.join(dst, f"neuron{N}", f"{N}-l{l + 1}.tsv"), N)
        out.append((r, c, v, fuse[l]))
    return out


def test_transform_gives_the_fused_layout_and_keeps_categories(tmp_path):
    """A banded network written with natural neuron labels (the outputs reading one set of rows
    are scattered over the layer) becomes, after the transform, one in which each 16-output
    block column reads exactly two aligned 16-row blocks of fused rows (the reference kernel's
    layout); the reference's own coo2bcsc (oracle/_ref, when built) then builds two blocks per
    block column, and the FP64 oracle's categories are unchanged."""
    from cospdnn_amd import transform as tf
    from cospdnn_amd.dataset import write_dataset

    N, L, S, seed, bias = 256, 4, 300, 71, -0.3
    src, dst = str(tmp_path / "raw"), str(tmp_path / "fused")
    write_dataset(src, N, L, S, cd.GEN_BANDED, seed, fused=False)
    tf.transform_dataset(N, L, src, dst)
    layers = _transformed_layers(dst, N, L)
    csc_layers = []
    for r, c, v, fuse in layers:
        for b in range(N // 16):
            sel = (c >= 16 * b) & (c < 16 * b + 16)
            win = np.unique(r[sel] // 16)
            assert len(win) == 2 and win[1] == win[0] + 1 and win[0] % 2 == 0
        cp, rows, vals = tf.to_csc(fuse[r], c, v, N)  # activation rows of the transformed net
        csc_layers.append((cp, rows, vals))
        if os.path.exists(REF_DRIVER):
            with open(tmp_path / "w.tsv", "w") as f:
                f.write("".join(f"{a + 1}\t{b + 1}\t{x}\n" for a, b, x in zip(r.tolist(), c.tolist(), v.tolist())))
            res = subprocess.run([REF_DRIVER, "bcsc", str(tmp_path / "w.tsv"), str(tmp_path / "b")], capture_output=True)
            assert res.returncode == 0
            assert np.all(np.diff(np.fromfile(tmp_path / "b.len", np.int32)) == 2)
    y0 = orc.gen_y0_rows(seed, N, 0, S)
    raw_alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.BANDED, seed), 0, S)
    new_alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, csc_layers, y0), 0, S)
    assert raw_alive.sum() > 0 and np.array_equal(raw_alive, new_alive)
