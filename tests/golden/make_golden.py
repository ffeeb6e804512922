"""Writes the golden fixtures of the sparse-DNN forward (tests/golden/*.npz).

Each fixture is a small network with its expected outputs, computed here by
brute-force dense FP64 numpy: Y_{l+1} = clip(Y_l @ W_l + b, 0, 32), category =
any(Y_L > 0) (the layer rule of kernels/kernel.cuh:132-136 and
utils/cpu_spmm_fuse.h:46,74 of the reference, with a float32 bias).  The
reference ships no test vectors of its own, so these fixtures are built from
the mathematics, independently of oracle/oracle.c and of the engine.

Fixture arrays: N, bias (float32), colptr_l / rows_l / vals_l per layer,
y0 [S][N] float32, alive [S] uint8, live [L+1] int64 (samples alive entering
each layer, live[0] = S), final [S][N] float64.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def dense_forward(N, bias, layers, y0):
    b = np.float64(np.float32(bias))
    Y = y0.astype(np.float64)
    S = Y.shape[0]
    live = [S]
    for cp, rw, vl in layers:
        W = np.zeros((N, N), dtype=np.float64)
        for j in range(N):
            for q in range(cp[j], cp[j + 1]):
                W[rw[q], j] += np.float64(vl[q])
        Y = np.clip(Y @ W + b, 0.0, 32.0)
        live.append(int(np.count_nonzero((Y > 0).any(axis=1))))
    alive = (Y > 0).any(axis=1).astype(np.uint8)
    return alive, np.array(live, dtype=np.int64), Y


def random_layer(rng, N, nnz_per_col, values):
    cp = [0]
    rows, vals = [], []
    for j in range(N):
        k = nnz_per_col(j)
        r = np.sort(rng.choice(N, size=k, replace=False)) if k > 0 else np.array([], dtype=np.int64)
        rows.extend(r.tolist())
        vals.extend(values(k).tolist())
        cp.append(len(rows))
    return (np.array(cp, dtype=np.int32), np.array(rows, dtype=np.int32), np.array(vals, dtype=np.float32))


def binary_input(rng, S, N, lo=0.15, hi=0.45):
    d = lo + (hi - lo) * rng.random(S)
    return (rng.random((S, N)) < d[:, None]).astype(np.float32)


def save(name, N, bias, layers, y0):
    alive, live, final = dense_forward(N, bias, layers, y0)
    arrays = {"N": np.int64(N), "bias": np.float32(bias), "L": np.int64(len(layers)), "y0": y0,
              "alive": alive, "live": live, "final": final}
    for l, (cp, rw, vl) in enumerate(layers):
        arrays[f"colptr_{l}"] = cp
        arrays[f"rows_{l}"] = rw
        arrays[f"vals_{l}"] = vl
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
    print(f"{name}: N={N} L={len(layers)} S={y0.shape[0]} categories={int(alive.sum())} live={live.tolist()}")


def main():
    rng = np.random.default_rng(20261203)
    # 1. graph-challenge shape: 32 nonzeros of 0.0625 per column, b = -0.3
    N = 128
    layers = [random_layer(rng, N, lambda j: 32, lambda k: np.full(k, 0.0625)) for _ in range(8)]
    save("gc_n128_l8", N, -0.3, layers, binary_input(rng, 700, N))
    # 2. non-uniform and negative weights, variable column counts, N not a multiple of 32
    N = 80
    layers = [random_layer(rng, N, lambda j: int(rng.integers(0, 40)),
                           lambda k: rng.choice([0.5, 0.25, 0.125, -0.125, -0.25, 0.0625], size=k))
              for _ in range(5)]
    save("mixed_n80_l5", N, -0.15, layers, binary_input(rng, 333, N))
    # 3. banded layers: output j reads (pi(j) + k*N/8) mod N, 8 per column, value 0.25
    N, K = 64, 8
    layers = []
    for _ in range(6):
        pi = rng.permutation(N)
        cp = np.arange(N + 1, dtype=np.int32) * K
        rows = np.concatenate([np.sort((pi[j] + np.arange(K) * (N // K)) % N) for j in range(N)]).astype(np.int32)
        layers.append((cp, rows, np.full(N * K, 0.25, dtype=np.float32)))
    save("banded_n64_l6", N, -0.4, layers, binary_input(rng, 513, N))
    # 4. positive bias: all-zero input rows come alive
    N = 48
    layers = [random_layer(rng, N, lambda j: 6, lambda k: np.full(k, -0.125)) for _ in range(3)]
    y0 = binary_input(rng, 90, N)
    y0[::7] = 0.0
    save("posbias_n48_l3", N, 0.2, layers, y0)
    # 5. everything dies after the first layer
    N = 64
    layers = [random_layer(rng, N, lambda j: 4, lambda k: np.full(k, 0.0625)) for _ in range(4)]
    save("alldead_n64_l4", N, -1.0, layers, binary_input(rng, 257, N))


if __name__ == "__main__":
    main()
