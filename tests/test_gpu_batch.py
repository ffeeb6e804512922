"""Several networks in one cospdnn_run: concurrent main passes on side streams, eagerly
and from one hipGraph, against the sequential run and the FP64 oracle."""
import sys

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu

SPECS = [(1024, 20, 3000, -0.3, cd.GEN_RANDOM, 41), (2048, 12, 5000, -0.35, cd.GEN_RANDOM, 42),
         (4096, 16, 2000, -0.4, cd.GEN_BANDED, 43), (512, 30, 4000, -0.3, cd.GEN_RANDOM, 44)]


def say(msg):
    print(f"[gpu-batch] {msg}", file=sys.stderr, flush=True)


def make(engine, N, L, S, bias, kind, seed):
    net = cd.Net(engine, N, L, S, bias)
    net.generate_layers(kind, seed)
    net.generate_input(seed)
    net.upload(cd.default_opts(max_tile_waste=64.0))
    return net


def oracle_cats(N, L, S, bias, kind, seed):
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.BANDED if kind else orc.RANDOM, seed), 0, S)
    return np.flatnonzero(alive)


def test_concurrent_batch_matches_sequential_and_oracle(engine):
    want = [oracle_cats(*s) for s in SPECS]
    nets = [make(engine, *s) for s in SPECS]
    seq = cd.run(nets, cd.default_opts(batch_concurrent=0))
    for mode, opts in [("1 chain, eager", cd.default_opts(batch_concurrent=1)),
                       ("2 chains, eager", cd.default_opts(batch_concurrent=2)),
                       ("4 chains, graph", cd.default_opts(batch_concurrent=4, use_graph=1)),
                       ("4 chains, graph again", cd.default_opts(batch_concurrent=4, use_graph=1)),
                       ("2 chains, graph", cd.default_opts(batch_concurrent=2, use_graph=1))]:
        say(f"batch {mode}")
        outs = cd.run(nets, opts)
        for s, o, w in zip(seq, outs, want):
            assert np.array_equal(o.categories, w)
            assert np.array_equal(o.live, s.live)
        # every result reports the batch's time
        assert len({round(o.ms_total, 6) for o in outs}) == 1
    for s, w in zip(seq, want):
        assert np.array_equal(s.categories, w)
    # replacing a network invalidates the captured batch graph
    nets[1].close()
    nets[1] = make(engine, *SPECS[1][:5], 45)
    want[1] = oracle_cats(*SPECS[1][:5], 45)
    outs = cd.run(nets, cd.default_opts(batch_concurrent=2, use_graph=1))
    for o, w in zip(outs, want):
        assert np.array_equal(o.categories, w)
    for n in nets:
        n.close()


def test_a_network_twice_in_one_call_is_refused(engine):
    net = make(engine, 256, 4, 100, -0.3, cd.GEN_RANDOM, 46)
    with pytest.raises(cd.CospdnnError):
        cd.run([net, net])
    net.close()


def test_c5_six_networks_full_size(engine):
    """The six uneven networks of the batched workload (N 1024..16384, L 120 or 480, 10000 to
    60000 samples each, random rows, tile images as the bench uploads them) in one call: one
    after another, in 2 and 3 concurrent chains, eager and from one hipGraph (replayed
    twice); every network's categories equal the FP64 oracle's."""
    specs = cd.config_nets("c5")
    want = []
    for sp in specs:
        say(f"c5 oracle {sp.name}: N={sp.N} L={sp.L} S={sp.S}")
        want.append(oracle_cats(sp.N, sp.L, sp.S, sp.bias, sp.kind, sp.seed))
    nets = [make(engine, sp.N, sp.L, sp.S, sp.bias, sp.kind, sp.seed) for sp in specs]
    for mode, opts in [("one after another", cd.default_opts(batch_concurrent=0)),
                       ("2 chains, eager", cd.default_opts(batch_concurrent=2)),
                       ("3 chains, graph", cd.default_opts(batch_concurrent=3, use_graph=1)),
                       ("3 chains, graph replay", cd.default_opts(batch_concurrent=3, use_graph=1)),
                       ("2 chains, graph", cd.default_opts(batch_concurrent=2, use_graph=1))]:
        outs = cd.run(nets, opts)
        say(f"c5 {mode}: {outs[0].ms_total:.1f} ms, categories {[len(o.categories) for o in outs]}")
        for sp, o, w in zip(specs, outs, want):
            assert np.array_equal(o.categories, w), (sp.name, mode)
            assert o.mfma_layers == sp.L, (sp.name, mode)
    for n in nets:
        n.close()


def test_graph_replay_reads_the_final_ids_when_the_last_layer_kills_rows(engine):
    """A replayed batch graph enqueues nothing, so the side holding each network's final ids
    must not depend on the capture: networks whose last layer kills most rows (the two id
    lists then differ), an even and an odd layer count, captured once and replayed twice,
    categories equal to the oracle's every time."""
    rng = np.random.default_rng(77)
    nets, want = [], []
    for N, L, S in ((1024, 6, 2500), (2048, 5, 1800)):
        layers = []
        for l in range(L):
            k = 8 if l == L - 1 else 32  # 8 inputs per output in the last layer: most rows die there
            cp = np.arange(N + 1, dtype=np.int32) * k
            rows = np.concatenate([np.sort(rng.choice(N, k, replace=False)) for _ in range(N)]).astype(np.int32)
            layers.append((cp, rows, np.full(len(rows), 0.0625, np.float32)))
        y0 = (rng.random((S, N)) < 0.3).astype(np.float32)
        alive, live = orc.forward_fp64(orc.OracleNet.explicit(N, -0.3, layers, y0.astype(np.uint8)), 0, S)
        assert 0 < live[-1] < live[-2] // 2  # the last layer kills most rows, not all
        want.append(np.flatnonzero(alive))
        net = cd.Net(engine, N, L, S, -0.3)
        for l, (cp, rw, vl) in enumerate(layers):
            net.set_layer_csc(l, cp, rw, vl)
        net.set_input_dense(y0)
        net.upload(cd.default_opts(max_tile_waste=64.0))
        nets.append(net)
    for call in ("capture", "replay", "replay again"):
        say(f"batch graph, last layer kills rows: {call}")
        outs = cd.run(nets, cd.default_opts(batch_concurrent=2, use_graph=1))
        for o, w in zip(outs, want):
            assert np.array_equal(o.categories, w), call
    for n in nets:
        n.close()
