"""Host-side weight formats: MFMA tile image and the reference BCSC reader."""
import os
import subprocess

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_driver")


def triples(cp, rw, vl):
    d = {}
    for j in range(len(cp) - 1):
        for q in range(cp[j], cp[j + 1]):
            d[(int(rw[q]), j)] = d.get((int(rw[q]), j), 0.0) + float(vl[q])
    return {k: v for k, v in d.items() if v != 0.0}


def decoded(N, cp, rw, vl):
    ins, outs, v = cd.layer_tiles_decode(N, cp, rw, vl)
    d = {}
    for i, o, x in zip(ins, outs, v):
        assert (int(i), int(o)) not in d, "an entry appears twice in the tile image"
        d[(int(i), int(o))] = float(x)
    return d


def mixed_layer(rng, N, maxk):
    cp, rows, vals = [0], [], []
    for _ in range(N):
        k = int(rng.integers(0, maxk))
        rows += sorted(rng.choice(N, k, replace=False).tolist())
        vals += rng.choice([0.5, -0.25, 0.125, 1.0, -2.0], size=k).tolist()
        cp.append(len(rows))
    return np.array(cp, np.int32), np.array(rows, np.int32), np.array(vals, np.float32)


@pytest.mark.parametrize("case", ["random1024", "banded4096", "mixed80", "mixed200"])
def test_tile_image_decodes_to_the_csc(case):
    rng = np.random.default_rng(7)
    if case == "random1024":
        N, (cp, rw, vl) = 1024, cd.generate_layer(1, cd.GEN_RANDOM, 1024, 32, 0)
    elif case == "banded4096":
        N, (cp, rw, vl) = 4096, cd.generate_layer(2, cd.GEN_BANDED, 4096, 32, 5)
    elif case == "mixed80":
        N, (cp, rw, vl) = 80, mixed_layer(rng, 80, 40)
    else:
        N, (cp, rw, vl) = 200, mixed_layer(rng, 200, 70)
    assert decoded(N, cp, rw, vl) == triples(cp, rw, vl)


def decoded_image(N, cp, rw, vl, image):
    ins, outs, v = cd.layer_image_decode(N, cp, rw, vl, image)
    d = {}
    for i, o, x in zip(ins, outs, v):
        assert (int(i), int(o)) not in d, "an entry appears twice in the image"
        d[(int(i), int(o))] = float(x)
    return d


@pytest.mark.parametrize("case", ["random1024", "random4096", "random2000", "banded4096"])
def test_mask_and_band_images_decode_to_the_csc(case):
    """The mask image (through the kernels' two-instruction expansion), the band image (rows
    split at 512-row bands, band-relative ids) and the band image's mask words all decode to
    the layer's nonzeros, each exactly once."""
    N = {"random1024": 1024, "random4096": 4096, "random2000": 2000, "banded4096": 4096}[case]
    kind = cd.GEN_BANDED if case.startswith("banded") else cd.GEN_RANDOM
    cp, rw, vl = cd.generate_layer(4, kind, N, 32, 0)
    want = triples(cp, rw, vl)
    for image in (cd.IMAGE_TILES, cd.IMAGE_MASK, cd.IMAGE_BAND, cd.IMAGE_BAND_MASK):
        assert decoded_image(N, cp, rw, vl, image) == want, image


@pytest.mark.parametrize("case", ["random1024", "random4096", "random16384", "banded4096", "mixed80", "mixed3000",
                                  "uneven_one_value"])
def test_slab_image_decodes_to_the_csc(case):
    """The slab image (16-bit rows in batch order, one value or FP16 values, lists padded with
    the zero row) decodes to the layer's nonzeros, each exactly once, for slab sizes 64 (N <= 1023
    rows ... 1024), 16 (N = 3000, 4096) and 4 (N = 16384); uneven lists of one value pad with
    the zero row and still need no value array."""
    rng = np.random.default_rng(11)
    if case.startswith("random"):
        N = int(case[6:])
        cp, rw, vl = cd.generate_layer(6, cd.GEN_RANDOM, N, 32, 1)
    elif case == "banded4096":
        N, (cp, rw, vl) = 4096, cd.generate_layer(6, cd.GEN_BANDED, 4096, 32, 1)
    elif case == "uneven_one_value":
        N = 1500
        cp, rw, _ = mixed_layer(rng, N, 40)
        vl = np.full(len(rw), 0.0625, np.float32)
    else:
        N = int(case[5:])
        cp, rw, vl = mixed_layer(rng, N, 40)
    assert decoded_image(N, cp, rw, vl, cd.IMAGE_SLAB) == triples(cp, rw, vl)


@pytest.mark.parametrize("val", [0.0625, 0.046875, -0.75, 1.0, 2.0 ** -14])
def test_mask_expansion_is_exact_for_seven_trailing_zero_bits(val):
    """Values whose FP16 bits end in 7 zeros get a mask image, and its expansion
    ((sp & (0x10001 << q)) * (v >> q)) reproduces them exactly; 0.3, whose bits do not, gets
    none (the FP16 tiles serve it)."""
    cp, rw, _ = cd.generate_layer(5, cd.GEN_RANDOM, 256, 16, 0)
    vl = np.full(len(rw), val, np.float32)
    assert decoded_image(256, cp, rw, vl, cd.IMAGE_MASK) == triples(cp, rw, vl)
    with pytest.raises(cd.CospdnnError):
        cd.layer_image_decode(256, cp, rw, np.full(len(rw), 0.3, np.float32), cd.IMAGE_MASK)


def test_tile_waste():
    cp, rw, vl = cd.generate_layer(3, cd.GEN_BANDED, 16384, 32, 0)
    st = cd.layer_stats(16384, cp, rw, vl)
    assert st["tile_waste"] == 1.0 and st["nchunk"] == 512 and st["nblk"] == 512 and st["fp16_exact"]
    cp, rw, vl = cd.generate_layer(3, cd.GEN_RANDOM, 1024, 32, 0)
    st = cd.layer_stats(1024, cp, rw, vl)
    # union of 32 columns x 32 random rows of 1024 ~ 1024 * (1 - e^-1) ~ 647 -> ~21 chunks per block
    # in natural order; grouping by shared rows brings it to ~17.4
    assert 15 < st["tile_waste"] < 19.5
    st = cd.layer_stats(4, np.array([0, 1, 2, 3, 4], np.int32), np.array([0, 1, 2, 3], np.int32),
                        np.array([0.1, 1, 1, 1], np.float32))
    assert not st["fp16_exact"]


def test_reference_bcsc_with_fuse_roundtrip_synthetic():
    """Hand-built reference-style BCSC: two 16-row blocks per block column, rows through a fuse list."""
    rng = np.random.default_rng(3)
    N = 64
    fuse = rng.permutation(N).astype(np.int32)
    values, row_idx, ptr = [], [], [0]
    want = {}
    for bc in range(N // 16):
        for start in (int(rng.integers(0, N - 16)), int(rng.integers(0, N - 16))):
            blk = np.zeros((16, 16), np.float32)  # [out][in]
            for oc in range(16):
                for ir in rng.choice(16, 3, replace=False):
                    v = float(rng.choice([0.0625, 0.5, -1.0]))
                    blk[oc, ir] = v
                    key = (int(fuse[start + ir]), bc * 16 + oc)
                    want[key] = want.get(key, 0.0) + v
            values += blk.ravel().tolist()
            row_idx.append(start)
        ptr.append(len(row_idx))
    cp, rw, vl = cd.ref_bcsc_to_csc(N, np.array(values, np.float32), np.array(row_idx, np.int32),
                                    np.array(ptr, np.int32), fuse)
    got = triples(cp, rw, vl)
    assert got == {k: v for k, v in want.items() if v != 0.0}


def _ref_bcsc(tmp_path, N, cp, file_rows, vl):
    with open(tmp_path / "w.tsv", "w") as f:
        for j in range(N):
            for q in range(cp[j], cp[j + 1]):
                f.write(f"{file_rows[q] + 1}\t{j + 1}\t{vl[q]}\n")
    r = subprocess.run([REF_DRIVER, "bcsc", str(tmp_path / "w.tsv"), str(tmp_path / "b")], capture_output=True)
    assert r.returncode == 0
    values = np.fromfile(tmp_path / "b.values", np.float32)
    row_idx = np.fromfile(tmp_path / "b.row_index", np.int32)
    ptr = np.fromfile(tmp_path / "b.len", np.int32)
    assert len(ptr) == N // 16 + 1
    return values, row_idx, ptr


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref/ref_driver not built")
def test_reference_bcsc_builder_roundtrip_fused(tmp_path):
    """COO file of a fused (banded) layer -> the reference's own coo2bcsc (oracle/_ref) -> our reader == COO.

    The weight file is stored in fused row order: each 32-output group reads 32
    consecutive fused rows, and the fuse list maps fused row -> activation row.
    """
    N = 256
    cp, rw, vl = orc.gen_layer(9, orc.BANDED, N, 32, 0)
    s = N // 32
    # fused row r -> (class r // 32, member r % 32) -> activation row member*s + class;
    # order the outputs so that block column c holds the 32 outputs of class c
    fuse = np.array([(r % 32) * s + r // 32 for r in range(N)], np.int32)
    inv = np.empty(N, np.int32)
    inv[fuse] = np.arange(N, dtype=np.int32)
    cls = rw[cp[:-1]] % s
    order = np.argsort(cls, kind="stable")  # new output j' = order position
    ncp, nrw, nvl = [0], [], []
    for j in order:
        nrw += inv[rw[cp[j]:cp[j + 1]]].tolist()
        nvl += vl[cp[j]:cp[j + 1]].tolist()
        ncp.append(len(nrw))
    ncp, nrw, nvl = np.array(ncp, np.int32), np.array(nrw, np.int32), np.array(nvl, np.float32)
    values, row_idx, ptr = _ref_bcsc(tmp_path, N, ncp, nrw, nvl)
    assert np.all(np.diff(ptr) == 2)  # exactly two 16-row blocks per block column, as the reference kernel needs
    got = triples(*cd.ref_bcsc_to_csc(N, values, row_idx, ptr, fuse))
    want = triples(ncp, fuse[nrw], nvl)
    assert got == want


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref/ref_driver not built")
def test_reference_bcsc_builder_quirk_on_random_rows(tmp_path):
    """For non-fused rows the reference builder stores an entry that opens a new block in a
    later column at offset 0 of that block (utils/matrix.h:1248-1252), i.e. in the block
    column's first output.  Our reader decodes the arrays as the kernel would read them:
    nothing is lost, and every misplaced entry sits in a first output column."""
    N = 256
    cp, rw, vl = orc.gen_layer(9, orc.RANDOM, N, 32, 0)
    values, row_idx, ptr = _ref_bcsc(tmp_path, N, cp, rw, vl)
    got = triples(*cd.ref_bcsc_to_csc(N, values, row_idx, ptr, None))
    want = triples(cp, rw, vl)
    extra = set(got) - set(want)
    assert sum(got.values()) == pytest.approx(sum(want.values()))
    assert extra and all(o % 16 == 0 for (_, o) in extra)
