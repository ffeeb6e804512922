"""The sample-slab SpMM kernel (spmm_slab, plan kind SLAB) against the FP64 oracle.

Every slab size T (64 samples at N <= 1023 down to 1 at N = 65536), the sample-blocked
activation layout between slab layers and its changes to and from the MFMA kernels' layout,
one weight value and mixed values, uneven column lengths (padded octets), compaction with
deaths (per-value stores) and kept layouts, the bit-packed and the FP16 layer-0 input, and
graph replay.  Bar: categories equal to the oracle's; FP16 activations within 0.05 of FP64.
"""
import sys

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gen_net(engine, N, L, S, bias, kind, seed, max_waste=64.0):
    net = cd.Net(engine, N, L, S, bias)
    net.generate_layers(kind, seed)
    net.generate_input(seed)
    net.upload(cd.default_opts(max_tile_waste=max_waste))
    return net


def slab_T(N):
    for t in (64, 32, 16, 8, 4, 2, 1):
        if (((N + 7) // 8 * 8 + 1) * min(t, 4) * 2 + 15) // 16 * 16 <= (131072 + 256) // (t // min(t, 4)) // 16 * 16:
            return t
    return 0


def expect_slab(out, N, bits):
    """Every layer on the sample-slab kernel; layer 0 on its bit-sliced variant when it reads the
    bit-packed input (the generated and test layer-0 weights are one value, lists below 64)."""
    ks = out.layer_kernel.tolist()
    k0 = cd.KERNEL_SLAB_BITS if bits and slab_T(N) <= 16 else cd.KERNEL_SLAB
    assert ks[0] == k0, (N, bits, ks[0])
    assert set(ks[1:]) <= {cd.KERNEL_SLAB}


# N, L, S, bias, kind, seed: one case per slab size
CASES = [
    (1024, 60, 4000, -0.3, cd.GEN_RANDOM, 101),    # T = 64
    (2048, 40, 3000, -0.3, cd.GEN_RANDOM, 102),    # T = 32
    (4096, 60, 3000, -0.35, cd.GEN_RANDOM, 103),   # T = 16
    (8192, 30, 1500, -0.35, cd.GEN_RANDOM, 104),   # T = 8
    (16384, 40, 1200, -0.4, cd.GEN_RANDOM, 105),   # T = 4
    (16384, 40, 1000, -0.4, cd.GEN_BANDED, 106),   # T = 4, banded
    (32768, 10, 300, -0.45, cd.GEN_RANDOM, 107),   # T = 2
    (65536, 6, 200, -0.45, cd.GEN_RANDOM, 108),    # T = 1
]


@pytest.mark.parametrize("case", CASES, ids=[f"N{c[0]}_T{slab_T(c[0])}_{'banded' if c[4] else 'random'}" for c in CASES])
def test_slab_generated(engine, case):
    N, L, S, bias, kind, seed = case
    say(f"slab N={N} T={slab_T(N)} L={L} S={S} kind={kind}")
    # values are compared at an early layer: deeper, FP16 rows near the unstable fixed point
    # drift from FP64 by up to ~0.15 (the guard tiers decide those rows)
    d = min(L - 1, 5)
    alive, live = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, kind, seed)
    for bits in (1, 0):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, spmm_slab=1, input_bits=bits, dump_layer=d),
                     dump_cap_rows=S)[0]
        say(f"  bits={bits} categories={len(out.categories)} oracle={len(want)} flagged={out.n_flagged} "
            f"ms={out.ms_total:.2f}")
        expect_slab(out, N, bits)
        assert np.array_equal(out.categories, want), bits
        assert out.live[0] == S
        assert np.abs(out.live - live).max() <= out.n_flagged + out.n_replayed + max(3, S // 1000)
        # the dumped rows (live after layer d in the FP16 pass) against FP64
        if out.dump_rows:
            _, _, f64 = orc.forward_fp64(orc.OracleNet.generated(N, d + 1, bias, kind, seed), 0, S,
                                         final_ids=out.dump_ids)
            err = np.abs(out.dump_vals - f64).max()
            say(f"  max |dY| at layer {d}: {err:.3g} over {out.dump_rows} rows")
            assert err < 0.05
    net.close()


@pytest.mark.parametrize("N,L,S,bias,seed", [(4096, 24, 5000, -0.35, 111), (16384, 12, 1500, -0.4, 112),
                                          (32768, 8, 600, -0.45, 113), (65536, 7, 300, -0.45, 114)],
                         ids=["T16", "T4", "T2", "T1"])
def test_slab_layout_changes_and_plans(engine, N, L, S, bias, seed):
    """Slab layers next to MFMA layers (activations change between the sample-blocked and the
    neuron-major layout at every switch, in the slab kernel itself at 16 samples per slab and
    through the layout conversion kernel at 4 and fewer), kept layouts and graph replay:
    categories equal to the oracle's, the activations after layer 5 within 0.05 of FP64 (the
    plans switch layouts before it)."""
    d = 5
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, cd.GEN_RANDOM, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed)
    ref = cd.run([net], cd.default_opts(precision=cd.FP16, dump_layer=d), dump_cap_rows=S)[0]
    assert np.array_equal(ref.categories, want)
    _, _, f64 = orc.forward_fp64(orc.OracleNet.generated(N, d + 1, bias, cd.GEN_RANDOM, seed), 0, S,
                                 final_ids=np.arange(S))
    slab = cd.LayerPlan(cd.PLAN_SLAB, 0, 1)
    mfma = cd.LayerPlan(cd.PLAN_MFMA, 6, 1)
    plans = {
        "all slab": [slab] * L,
        "alternating": [slab if l % 2 == 0 else mfma for l in range(L)],
        "alternating from mfma": [slab if l % 2 == 1 else mfma for l in range(L)],
        "runs of three": [slab if (l // 3) % 2 else mfma for l in range(L)],
        "slab first and last": [slab if l in (0, L - 1) else mfma for l in range(L)],
    }
    for name, plan in plans.items():
        net.set_plan(plan)
        for keep, graph in ((0, 0), (30000, 0), (0, 1), (0, 1)):
            out = cd.run([net], cd.default_opts(precision=cd.FP16, keep_holes_ppm=keep, use_graph=graph))[0]
            ks = out.layer_kernel.tolist()
            for l in range(L):
                slab_k = ks[l] in (cd.KERNEL_SLAB, cd.KERNEL_SLAB_BITS)
                assert slab_k == (plan[l].kernel == cd.PLAN_SLAB), (name, l, ks[l])
            assert np.array_equal(out.categories, want), (name, keep, graph)
            say(f"  {name} keep={keep} graph={graph}: categories={len(out.categories)} ms={out.ms_total:.2f}")
        out = cd.run([net], cd.default_opts(precision=cd.FP16, dump_layer=d), dump_cap_rows=S)[0]
        assert np.abs(out.dump_vals - f64[out.dump_ids]).max() < 0.05, name
    net.set_plan(None)
    net.close()


def random_csc(rng, N, kmin, kmax, values):
    counts = rng.integers(kmin, kmax + 1, size=N)
    cp = np.zeros(N + 1, np.int32)
    cp[1:] = np.cumsum(counts)
    rows = np.concatenate([np.sort(rng.choice(N, c, replace=False)) for c in counts]).astype(np.int32)
    vals = rng.choice(np.asarray(values, np.float32), size=len(rows))
    return cp, rows, vals


@pytest.mark.parametrize("N", [81, 1000, 3072, 8200], ids=lambda n: f"N{n}_T{slab_T(n)}")
def test_slab_mixed_values_and_uneven_columns(engine, N):
    """Columns of 0 to 40 nonzeros (padded octets, empty columns), several weight values (the
    slab image carries FP16 values), one value on every other layer (no value array), odd N."""
    rng = np.random.default_rng(N)
    L, S, bias = 8, 700, -0.3
    layers = []
    for l in range(L):
        if l % 2:
            layers.append(random_csc(rng, N, 0, 40, [0.0625, 0.125, -0.0625, 0.03125, 0.25]))
        else:
            cp = np.arange(N + 1, dtype=np.int32) * 24
            rows = np.concatenate([np.sort(rng.choice(N, 24, replace=False)) for _ in range(N)]).astype(np.int32)
            layers.append((cp, rows, np.full(len(rows), 0.0625, np.float32)))
    y0 = (rng.random((S, N)) < 0.3).astype(np.float32)
    y0[::7] = 0
    d = 5
    alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8)), 0, S)
    _, _, f64 = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers[: d + 1], y0.astype(np.uint8)), 0, S,
                                 final_ids=np.arange(S))
    net = cd.Net(engine, N, L, S, bias)
    for l, (cp, rw, vl) in enumerate(layers):
        net.set_layer_csc(l, cp, rw, vl)
    net.set_input_dense(y0)
    net.upload(cd.default_opts(max_tile_waste=1e9))
    for bits in (1, 0):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, spmm_slab=1, input_bits=bits, dump_layer=d),
                     dump_cap_rows=S)[0]
        expect_slab(out, N, bits)
        assert np.array_equal(out.categories, np.flatnonzero(alive)), bits
        if out.dump_rows:
            err = np.abs(out.dump_vals - f64[out.dump_ids]).max()
            say(f"  N={N} bits={bits} categories={len(out.categories)} max |dY| {err:.3g}")
            assert err < 0.05
    net.close()


def test_slab_edge_shapes(engine):
    """One sample, one layer, all rows dead after layer 0, a positive bias (every row stays
    alive), a partial last slab."""
    rng = np.random.default_rng(9)
    for N, L, S, bias, K in [(1024, 3, 1, -0.3, 20), (1024, 1, 300, -0.3, 32), (4096, 3, 77, -5.0, 8),
                             (2048, 3, 130, 0.2, 6), (16384, 2, 9, -0.4, 32)]:
        say(f"slab edge N={N} L={L} S={S} bias={bias}")
        layers = []
        for _ in range(L):
            cp = np.arange(N + 1, dtype=np.int32) * K
            rows = np.concatenate([np.sort(rng.choice(N, K, replace=False)) for _ in range(N)]).astype(np.int32)
            layers.append((cp, rows, np.full(len(rows), 0.0625, np.float32)))
        y0 = (rng.random((S, N)) < 0.3).astype(np.float32)
        alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8)), 0, S)
        net = cd.Net(engine, N, L, S, bias)
        for l, (cp, rw, vl) in enumerate(layers):
            net.set_layer_csc(l, cp, rw, vl)
        net.set_input_dense(y0)
        net.upload()
        out = cd.run([net], cd.default_opts(precision=cd.FP16, spmm_slab=1))[0]
        expect_slab(out, N, True)
        assert np.array_equal(out.categories, np.flatnonzero(alive)), (N, L, S, bias)
        net.close()


@pytest.mark.parametrize("N,S", [(4096, 3000), (8192, 1500), (16384, 1100), (32768, 500), (65536, 300)],
                         ids=lambda v: str(v))
def test_bits_layer0_equals_the_slab_kernel(engine, N, S):
    """Layer 0 on the bit-packed input by bit-sliced counts (W = 16 P samples a slab: 64 at
    N <= 16384, 32 at 32768, 16 at 65536; S not a multiple of W, so the last slab is partial):
    its activations equal the slab kernel's on the FP16 input bit for bit, with layer 1 on the
    slab kernel (sample-blocked output) and on the MFMA kernel (neuron-major output);
    categories equal to the oracle's; a layer 0 with a column of more than 63 rows stays on
    the slab kernel."""
    L, bias, seed = 4, -0.4 if N >= 16384 else -0.35, 131 + N // 1024
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, cd.GEN_RANDOM, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed)
    slab = cd.LayerPlan(cd.PLAN_SLAB, 0, 1)
    mfma = cd.LayerPlan(cd.PLAN_MFMA, 1, 4)
    for name, plan in (("slab after", [slab] * L), ("mfma after", [slab, mfma, mfma, slab])):
        net.set_plan(plan)
        outs = {}
        for bits in (1, 0):
            out = cd.run([net], cd.default_opts(precision=cd.FP16, input_bits=bits, dump_layer=0), dump_cap_rows=S)[0]
            assert out.layer_kernel[0] == (cd.KERNEL_SLAB_BITS if bits else cd.KERNEL_SLAB), (name, bits)
            assert np.array_equal(out.categories, want), (name, bits)
            outs[bits] = out
        a, b = outs[1], outs[0]
        say(f"  N={N} {name}: categories={len(a.categories)} dumped rows={a.dump_rows} ms {a.ms_total:.2f} "
            f"(FP16 input {b.ms_total:.2f})")
        assert np.array_equal(a.dump_ids, b.dump_ids)
        assert np.array_equal(a.dump_vals, b.dump_vals), name
        assert np.array_equal(a.live, b.live)
    net.close()
    # a column of 80 input rows in layer 0: beyond the 6 count planes, the slab kernel instead
    rng = np.random.default_rng(N)
    layers = []
    for l in range(2):
        counts = np.full(N, 32)
        if l == 0:
            counts[5] = 80
        cp = np.zeros(N + 1, np.int32)
        cp[1:] = np.cumsum(counts)
        rows = np.concatenate([np.sort(rng.choice(N, c, replace=False)) for c in counts]).astype(np.int32)
        layers.append((cp, rows, np.full(len(rows), 0.0625, np.float32)))
    y0 = (rng.random((200, N)) < 0.3).astype(np.float32)
    alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8)), 0, 200)
    net = cd.Net(engine, N, 2, 200, bias)
    for l, (cp, rw, vl) in enumerate(layers):
        net.set_layer_csc(l, cp, rw, vl)
    net.set_input_dense(y0)
    net.upload()
    out = cd.run([net], cd.default_opts(precision=cd.FP16, spmm_slab=1))[0]
    assert out.layer_kernel[0] == cd.KERNEL_SLAB
    assert np.array_equal(out.categories, np.flatnonzero(alive))
    net.close()
