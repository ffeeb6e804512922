"""The drop-in command-line drivers on a reference-format dataset directory."""
import os
import subprocess

import numpy as np
import pytest

import cospdnn_amd as cd
from cospdnn_amd.dataset import write_dataset
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu
PKG = os.path.dirname(cd.__file__)


def oracle_categories(N, L, S, bias, kind, seed):
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    return np.flatnonzero(alive) + 1  # 1-based like the category files


def read_ids(path):
    return np.loadtxt(path, dtype=np.int64, ndmin=1) if os.path.getsize(path) else np.zeros(0, np.int64)


def test_single_task_cli(tmp_path):
    N, L, S, seed = 1024, 16, 3000, 41
    print(f"cli single N={N} L={L} S={S}", flush=True)
    write_dataset(str(tmp_path), N, L, S, cd.GEN_BANDED, seed)
    want = oracle_categories(N, L, S, -0.3, orc.BANDED, seed)
    for prec in ("fp16", "fp32"):
        cats = tmp_path / f"cats_{prec}.txt"
        r = subprocess.run([os.path.join(PKG, "cospdnn.out"), str(N), str(S), str(L), "1", "0", str(tmp_path),
                            f"--precision={prec}", f"--categories={cats}"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        assert "Kernel Exec Time = " in r.stdout and "[END]..." in r.stdout
        assert np.array_equal(read_ids(cats), want)


def test_single_task_cli_missing_dataset(tmp_path):
    r = subprocess.run([os.path.join(PKG, "cospdnn.out"), "1024", "10", "2", "1", "0", str(tmp_path / "nope")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "does not exists." in r.stdout  # exit(-1) as the reference


def test_concurrent_task_cli(tmp_path):
    seed = 43
    print("cli double", flush=True)
    write_dataset(str(tmp_path), 1024, 12, 2000, cd.GEN_RANDOM, seed)
    write_dataset(str(tmp_path), 4096, 12, 1500, cd.GEN_BANDED, seed)
    cats = tmp_path / "cats.txt"
    r = subprocess.run([os.path.join(PKG, "cospdnn_double.out"), "1024", "2000", "12", "4096", "1500", "12", "0",
                        str(tmp_path), f"--categories={cats}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Total Kernel Exec Time = " in r.stdout
    got = np.loadtxt(cats, dtype=np.int64, ndmin=2)
    want1 = oracle_categories(1024, 12, 2000, -0.3, orc.RANDOM, seed)
    want2 = oracle_categories(4096, 12, 1500, -0.35, orc.BANDED, seed)
    assert np.array_equal(got[got[:, 0] == 1, 1], want1)
    assert np.array_equal(got[got[:, 0] == 2, 1], want2)
    # unequal layer counts are refused like the reference unless --allow-uneven
    write_dataset(str(tmp_path), 1024, 8, 500, cd.GEN_RANDOM, seed)
    r = subprocess.run([os.path.join(PKG, "cospdnn_double.out"), "1024", "500", "8", "4096", "500", "12", "0",
                        str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert "must have the same number of layers" in r.stderr


@pytest.mark.parametrize("kind", [cd.GEN_BANDED, cd.GEN_RANDOM], ids=["banded_fused", "random"])
def test_cpp_mirror_single(tmp_path, kind):
    """graph_challenge_tc.out uses the engine only through the C++ mirror, as the reference's
    main() does (src/cospdnn.cpp:80-188): reads the dataset, builds the 16x16 BCSC arrays and the
    fuse lists, calls test_graph_challenge_tc; categories equal to the FP64 oracle's."""
    N, L, S, seed = 1024, 14, 2500, 47
    print(f"mirror single kind={kind}", flush=True)
    write_dataset(str(tmp_path), N, L, S, kind, seed)
    want = oracle_categories(N, L, S, -0.3, orc.BANDED if kind else orc.RANDOM, seed)
    cats = tmp_path / "cats.txt"
    r = subprocess.run([os.path.join(PKG, "graph_challenge_tc.out"), str(N), str(S), str(L), "1", "0", str(tmp_path),
                        f"--categories={cats}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Kernel Exec Time = " in r.stdout and "[END]..." in r.stdout
    assert np.array_equal(read_ids(cats), want)


def test_cpp_mirror_concurrent(tmp_path):
    """graph_challenge_tc_co.out: two networks through test_graph_challenge_tc_co (one engine, one
    concurrent run), as the reference's concurrent main() (src/cospdnn_double.cpp:81-197)."""
    seed = 48
    print("mirror double", flush=True)
    write_dataset(str(tmp_path), 1024, 12, 2000, cd.GEN_RANDOM, seed)
    write_dataset(str(tmp_path), 4096, 12, 1500, cd.GEN_BANDED, seed)
    cats = tmp_path / "cats.txt"
    r = subprocess.run([os.path.join(PKG, "graph_challenge_tc_co.out"), "1024", "2000", "12", "4096", "1500", "12",
                        "0", str(tmp_path), f"--categories={cats}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Total Kernel Exec Time = " in r.stdout
    got = np.loadtxt(cats, dtype=np.int64, ndmin=2)
    assert np.array_equal(got[got[:, 0] == 1, 1], oracle_categories(1024, 12, 2000, -0.3, orc.RANDOM, seed))
    assert np.array_equal(got[got[:, 0] == 2, 1], oracle_categories(4096, 12, 1500, -0.35, orc.BANDED, seed))


def test_transformed_dataset_runs_through_both_drivers(tmp_path):
    """A banded network written with natural neuron labels, run through the similarity transform
    (cospdnn_amd.transform: relabelled neurons, fuse lists), then through cospdnn.out (fuse lists
    read as the reference does) and through graph_challenge_tc.out (the reference's 16x16 BCSC
    built from the fused files): categories equal the FP64 oracle's on the raw network."""
    from cospdnn_amd import transform as tf

    N, L, S, seed = 1024, 10, 2000, 49
    print("transformed dataset", flush=True)
    src, dst = str(tmp_path / "raw"), str(tmp_path / "fused")
    write_dataset(src, N, L, S, cd.GEN_BANDED, seed, fused=False)
    tf.transform_dataset(N, L, src, dst)
    want = oracle_categories(N, L, S, -0.3, orc.BANDED, seed)
    for exe in ("cospdnn.out", "graph_challenge_tc.out"):
        cats = tmp_path / f"cats_{exe}.txt"
        r = subprocess.run([os.path.join(PKG, exe), str(N), str(S), str(L), "1", "0", dst, f"--categories={cats}"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        assert np.array_equal(read_ids(cats), want), exe
