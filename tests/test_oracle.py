"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks: the committed golden fixtures (dense FP64 numpy), brute-force dense
FP64 on random tiny nets, the reference's own CPU forward (CpuSpmmFuse built
from the reference sources into oracle/_ref) for integer biases, closed forms
of the first layer and of banded layers, exact fixed points, permutation
invariance of the category set, and agreement of the two generators.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_ctypes as orc

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_driver")


def load_fixture(path):
    z = np.load(path)
    L = int(z["L"])
    layers = [(z[f"colptr_{l}"], z[f"rows_{l}"], z[f"vals_{l}"]) for l in range(L)]
    return int(z["N"]), float(z["bias"]), layers, z["y0"], z["alive"], z["live"], z["final"]


def dense_forward(N, bias, layers, y0):
    b = np.float64(np.float32(bias))
    Y = y0.astype(np.float64)
    live = [Y.shape[0]]
    for cp, rw, vl in layers:
        W = np.zeros((N, N))
        for j in range(N):
            for q in range(cp[j], cp[j + 1]):
                W[rw[q], j] += np.float64(vl[q])
        Y = np.clip(Y @ W + b, 0.0, 32.0)
        live.append(int((Y > 0).any(axis=1).sum()))
    return (Y > 0).any(axis=1).astype(np.uint8), np.array(live), Y


def rand_layers(rng, N, L, K, values=(0.0625,)):
    layers = []
    for _ in range(L):
        cp = np.arange(N + 1, dtype=np.int32) * K
        rows = np.concatenate([np.sort(rng.choice(N, K, replace=False)) for _ in range(N)]).astype(np.int32)
        vals = rng.choice(np.array(values, dtype=np.float32), size=N * K).astype(np.float32)
        layers.append((cp, rows, vals))
    return layers


FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_matches_golden(path):
    N, bias, layers, y0, alive, live, final = load_fixture(path)
    assert np.all((y0 == 0) | (y0 == 1))
    net = orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8))
    S = y0.shape[0]
    a64, l64, f64 = orc.forward_fp64(net, 0, S, final_ids=np.arange(S))
    assert np.array_equal(a64, alive)
    assert np.array_equal(l64, live)
    np.testing.assert_allclose(f64, final, rtol=0, atol=1e-9)
    a32, l32 = orc.forward_fp32(net, 0, S)
    assert np.array_equal(a32, alive)
    assert np.array_equal(l32, live)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_oracle_matches_dense_bruteforce(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.choice([32, 64, 96]))
    L = int(rng.integers(2, 9))
    K = int(rng.choice([4, 8, 16]))
    bias = float(rng.choice([-0.3, -0.2, -0.5]))
    layers = rand_layers(rng, N, L, K, values=(0.0625, 0.125, 0.25, -0.0625))
    S = 400
    y0 = (rng.random((S, N)) < rng.uniform(0.15, 0.45, size=(S, 1))).astype(np.uint8)
    want_alive, want_live, want_final = dense_forward(N, bias, layers, y0)
    net = orc.OracleNet.explicit(N, bias, layers, y0)
    a, live, fin = orc.forward_fp64(net, 0, S, final_ids=np.arange(S))
    assert np.array_equal(a, want_alive)
    assert np.array_equal(live, want_live)
    np.testing.assert_allclose(fin, want_final, atol=1e-9, rtol=0)


def test_oracle_generated_equals_explicit():
    """The oracle's generator path and explicit arrays of the same net agree."""
    N, L, S, seed = 256, 10, 300, 77
    for kind in (orc.RANDOM, orc.BANDED):
        gen = orc.OracleNet.generated(N, L, -0.3, kind, seed)
        layers = [orc.gen_layer(seed, kind, N, 32, l) for l in range(L)]
        exp = orc.OracleNet.explicit(N, -0.3, layers, orc.gen_y0_rows(seed, N, 0, S))
        a1, l1 = orc.forward_fp64(gen, 0, S)
        a2, l2 = orc.forward_fp64(exp, 0, S)
        assert np.array_equal(a1, a2) and np.array_equal(l1, l2)
        a3, l3 = orc.forward_fp32(gen, 0, S)
        assert np.array_equal(a1, a3)


def test_first_layer_closed_form():
    """With binary input, y_j = clamp(0.0625 * |row(s) & in(j)| + b) exactly."""
    N, S, seed, bias = 512, 64, 5, -0.3
    net = orc.OracleNet.generated(N, 1, bias, orc.RANDOM, seed)
    _, _, fin = orc.forward_fp64(net, 0, S, final_ids=np.arange(S))
    cp, rw, _ = orc.gen_layer(seed, orc.RANDOM, N, 32, 0)
    y0 = orc.gen_y0_rows(seed, N, 0, S)
    cnt = np.stack([y0[:, rw[cp[j]:cp[j + 1]]].sum(axis=1) for j in range(N)], axis=1)
    want = np.clip(0.0625 * cnt + np.float64(np.float32(bias)), 0, 32)
    assert np.array_equal(fin, want)


def test_banded_outputs_equal_within_class():
    """Outputs of a banded layer that read the same residue class are equal."""
    N, S, seed = 1024, 32, 9
    net = orc.OracleNet.generated(N, 1, -0.4, orc.BANDED, seed)
    _, _, fin = orc.forward_fp64(net, 0, S, final_ids=np.arange(S))
    cp, rw, _ = orc.gen_layer(seed, orc.BANDED, N, 32, 0)
    cls = rw[cp[:-1]] % (N // 32)  # smallest row of each column identifies its class
    for c in np.unique(cls):
        cols = np.flatnonzero(cls == c)
        assert len(cols) == 32
        assert np.all(fin[:, cols] == fin[:, cols[:1]])


def test_fixed_points():
    """All-zero rows stay zero (b < 0); all-32 rows stay 32 when 32*colsum + b >= 32."""
    N, L = 64, 6
    rng = np.random.default_rng(0)
    layers = rand_layers(rng, N, L, 32)
    y0 = np.zeros((4, N), dtype=np.uint8)
    y0[2:] = 1
    net = orc.OracleNet.explicit(N, -0.3, layers, y0)
    # a row of ones: layer 1 gives 2 - 0.3 = 1.7 everywhere, then saturates at 32
    a, live, fin = orc.forward_fp64(net, 0, 4, final_ids=np.arange(4))
    assert list(a) == [0, 0, 1, 1]
    assert np.all(fin[:2] == 0) and np.all(fin[2:] == 32)
    assert list(live) == [4] + [2] * L


def test_category_set_invariant_under_neuron_relabelling():
    rng = np.random.default_rng(11)
    N, L, S = 96, 6, 300
    layers = rand_layers(rng, N, L, 16, values=(0.0625, 0.125))
    y0 = (rng.random((S, N)) < 0.3).astype(np.uint8)
    perms = [rng.permutation(N) for _ in range(L + 1)]  # perms[l]: relabelling of layer l's input neurons
    new_layers = []
    for l, (cp, rw, vl) in enumerate(layers):
        cols = [[] for _ in range(N)]
        for j in range(N):
            for q in range(cp[j], cp[j + 1]):
                cols[perms[l + 1][j]].append((perms[l][rw[q]], vl[q]))
        ncp, nrw, nvl = [0], [], []
        for j in range(N):
            for r, v in sorted(cols[j]):
                nrw.append(r)
                nvl.append(v)
            ncp.append(len(nrw))
        new_layers.append((np.array(ncp, np.int32), np.array(nrw, np.int32), np.array(nvl, np.float32)))
    y0p = np.zeros_like(y0)
    y0p[:, perms[0]] = y0
    a1, l1 = orc.forward_fp64(orc.OracleNet.explicit(N, -0.3, layers, y0), 0, S)
    a2, l2 = orc.forward_fp64(orc.OracleNet.explicit(N, -0.3, new_layers, y0p), 0, S)
    assert np.array_equal(a1, a2) and np.array_equal(l1, l2)


def test_live_counts_monotone_for_nonpositive_bias():
    net = orc.OracleNet.generated(1024, 30, -0.3, orc.RANDOM, 20261203)
    _, live = orc.forward_fp64(net, 0, 3000)
    assert live[0] == 3000
    assert np.all(np.diff(live) <= 0)


def test_sample_window_consistency():
    """Rows are independent: a window [s0, s0+S) equals the slice of the full run."""
    net = orc.OracleNet.generated(512, 20, -0.3, orc.RANDOM, 3)
    a_full, _ = orc.forward_fp64(net, 0, 1000)
    a_win, _ = orc.forward_fp64(net, 300, 200)
    assert np.array_equal(a_full[300:500], a_win)


def test_generator_structure():
    N, K = 2048, 32
    cp, rw, vl = orc.gen_layer(123, orc.RANDOM, N, K, 4)
    assert np.array_equal(cp, np.arange(N + 1) * K)
    r = rw.reshape(N, K)
    assert np.all(np.diff(r, axis=1) > 0)  # distinct, sorted
    assert r.min() >= 0 and r.max() < N
    assert np.all(vl == np.float32(0.0625))
    # rows are close to uniform: every input neuron is used about K times
    use = np.bincount(rw, minlength=N)
    assert abs(use.mean() - K) < 1e-9 and use.std() < 2 * np.sqrt(K)
    cp, rw, _ = orc.gen_layer(123, orc.BANDED, N, K, 4)
    r = rw.reshape(N, K)
    s = N // K
    assert np.all(np.diff(r, axis=1) == s)  # one residue class per column
    cls = r[:, 0]
    assert np.all(np.bincount(cls, minlength=s) == K)  # every class read by exactly K outputs
    y = orc.gen_y0_rows(123, N, 0, 400)
    dens = y.mean(axis=1)
    thr = np.array([orc.y0_row_threshold(123, s) for s in range(400)]) / 2.0**32
    assert np.all((thr >= 0.15 - 1e-9) & (thr <= 0.45 + 1e-9))
    assert np.abs(dens - thr).max() < 0.06


def test_product_generator_matches_oracle_restatement():
    import cospdnn_amd as cd

    for kind, N, seed, l in [(0, 1024, 20261203, 0), (0, 4096, 20261204, 17), (1, 16384, 20261205, 3),
                             (1, 2048, 7, 119)]:
        a = cd.generate_layer(seed, kind, N, 32, l)
        b = orc.gen_layer(seed, kind, N, 32, l)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert np.array_equal(cd.generate_input_rows(20261205, 16384, 59990, 10),
                          orc.gen_y0_rows(20261205, 16384, 59990, 10))


def _write_coo(path, cp, rw, vl):
    with open(path, "w") as f:
        for j in range(len(cp) - 1):
            for q in range(cp[j], cp[j + 1]):
                f.write(f"{rw[q] + 1}\t{j + 1}\t{vl[q]:.9g}\n")


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref/ref_driver not built")
@pytest.mark.parametrize("bias", [0, -1])
def test_oracle_matches_reference_cpu_forward(tmp_path, bias):
    """The reference's CpuSpmmFuse (its int-bias API limits this to integer biases)."""
    rng = np.random.default_rng(100 + abs(bias))
    N, L, S = 64, 4, 50
    layers = rand_layers(rng, N, L, 12, values=(0.0625, 0.125, 0.25, 0.5))
    y0 = (rng.random((S, N)) < rng.uniform(0.15, 0.45, size=(S, 1))).astype(np.uint8)
    for l, (cp, rw, vl) in enumerate(layers):
        _write_coo(tmp_path / f"w{l}.tsv", cp, rw, vl)
    net = orc.OracleNet.explicit(N, float(bias), layers, y0)
    _, _, fin = orc.forward_fp64(net, 0, S, final_ids=np.arange(S))
    y0.astype(np.float32).tofile(tmp_path / "in.bin")
    fin.astype(np.float32).tofile(tmp_path / "out.bin")
    r = subprocess.run([REF_DRIVER, "cpufuse", str(N), str(S), str(bias), str(L), str(tmp_path / "w"),
                        str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Success" in r.stdout
    # a planted mistake (one activation off by 0.01) must be caught by the same check
    bad = fin.astype(np.float32)
    bad[S // 2, N // 3] += 0.01
    bad.tofile(tmp_path / "bad.bin")
    r = subprocess.run([REF_DRIVER, "cpufuse", str(N), str(S), str(bias), str(L), str(tmp_path / "w"),
                        str(tmp_path / "in.bin"), str(tmp_path / "bad.bin")], capture_output=True, text=True)
    assert r.returncode != 0
