"""Weight streaming: layers whose MFMA images do not fit the weight budget keep them in pinned
host memory and the FP16 pass copies each into one of two device slots, one layer ahead of its
SpMM (the reference streams every layer so, cospdnn.cu:170-210).  A 1 MiB budget streams
nearly every layer; the results must equal the resident network's exactly."""
import sys

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[streaming] {msg}", file=sys.stderr, flush=True)


def upload(engine, N, L, S, bias, layers, y0, gen, budget_mb):
    net = cd.Net(engine, N, L, S, bias)
    if gen is not None:
        net.generate_layers(gen[0], gen[1])
        net.generate_input(gen[1])
    else:
        for l, (cp, rw, vl) in enumerate(layers):
            net.set_layer_csc(l, cp, rw, vl)
        net.set_input_dense(y0)
    net.upload(cd.default_opts(weight_budget_mb=budget_mb))
    return net


def mixed_layers(N, L, K, rng):
    out = []
    for _ in range(L):
        cp = np.arange(N + 1, dtype=np.int32) * K
        rows = np.concatenate([np.sort(rng.choice(N, K, replace=False)) for _ in range(N)]).astype(np.int32)
        vals = rng.choice(np.array([0.0625, 0.125, -0.0625], np.float32), size=len(rows))
        out.append((cp, rows, vals))
    return out


CASES = [("random", 2048, 24, 3000, -0.35, cd.GEN_RANDOM, 61), ("banded", 4096, 20, 2000, -0.35, cd.GEN_BANDED, 62),
         ("mixed", 1024, 12, 1500, -0.3, None, 63)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_streamed_weights_match_resident(engine, case):
    name, N, L, S, bias, kind, seed = case
    say(f"{name} N={N} L={L} S={S}")
    if kind is None:
        rng = np.random.default_rng(seed)
        layers = mixed_layers(N, L, 24, rng)
        y0 = (rng.random((S, N)) < 0.3).astype(np.float32)
        alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8)), 0, S)
        gen = None
    else:
        layers = y0 = None
        alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
        gen = (kind, seed)
    want = np.flatnonzero(alive)
    res = upload(engine, N, L, S, bias, layers, y0, gen, 0)
    stm = upload(engine, N, L, S, bias, layers, y0, gen, 1)
    for order in (1, 5):
        o = cd.default_opts(precision=cd.FP16, spmm_order=order, dump_layer=L - 1)
        a = cd.run([res], o, dump_cap_rows=S)[0]
        b = cd.run([stm], o, dump_cap_rows=S)[0]
        say(f"  order {order}: streamed layers {b.streamed_layers}, mfma layers {b.mfma_layers}")
        assert a.streamed_layers == 0 and b.streamed_layers >= L - 1
        assert b.mfma_layers == L == a.mfma_layers
        assert np.array_equal(a.categories, want) and np.array_equal(b.categories, want)
        ga = np.zeros((S, N), np.float32)
        gb = np.zeros((S, N), np.float32)
        ga[a.dump_ids] = a.dump_vals
        gb[b.dump_ids] = b.dump_vals
        assert np.array_equal(ga, gb)
    for graph in (1, 1, 0):
        out = cd.run([stm], cd.default_opts(precision=cd.FP16, use_graph=graph))[0]
        assert np.array_equal(out.categories, want), graph
    # the FP32 path reads no MFMA image of a streamed layer
    assert np.array_equal(cd.run([stm], cd.default_opts(precision=cd.FP32))[0].categories, want)
    res.close()
    stm.close()
