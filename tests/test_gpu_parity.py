"""Parity of the HIP engine (through the C ABI) against the FP64 oracle and the golden fixtures.

Bar: the category set (and the per-layer live counts of the FP32 path) equal the
oracle's exactly; activations are compared within stated tolerances.
"""
import glob
import os
import sys

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fixture_net(engine, path):
    z = np.load(path)
    N, L, bias = int(z["N"]), int(z["L"]), float(z["bias"])
    y0 = z["y0"]
    net = cd.Net(engine, N, L, y0.shape[0], bias)
    for l in range(L):
        net.set_layer_csc(l, z[f"colptr_{l}"], z[f"rows_{l}"], z[f"vals_{l}"])
    net.set_input_dense(y0)
    net.upload(cd.default_opts(max_tile_waste=1e9))
    return net, z


@pytest.mark.parametrize("prec", [cd.FP32, cd.FP16], ids=["fp32", "fp16"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_golden_fixture(engine, path, prec):
    say(f"golden {os.path.basename(path)} prec={prec}")
    net, z = fixture_net(engine, path)
    L = int(z["L"])
    out = cd.run([net], cd.default_opts(precision=prec, dump_layer=L - 1), dump_cap_rows=net.S)[0]
    assert np.array_equal(out.categories, np.flatnonzero(z["alive"]))
    final = z["final"]
    got = np.zeros_like(final)
    got[out.dump_ids] = out.dump_vals
    if prec == cd.FP32:
        assert np.array_equal(out.live, z["live"])
        np.testing.assert_allclose(got, final, atol=1e-4, rtol=0)
    else:
        # FP16 activations: rounding 2^-11 relative per layer, amplified by the fan-in
        assert np.abs(got - final).max() < 0.05
    net.close()


def assert_live_close(got, want, out):
    """Live counts come from the first pass; rows it decides differently from FP64 (all of them
    recomputed or replayed, so the categories stay exact) can shift a layer's count by a few."""
    assert got[0] == want[0] and got[-1] == want[-1]
    assert np.abs(got - want).max() <= out.n_flagged + out.n_replayed + max(3, int(got[0]) // 1000)


def gen_net(engine, N, L, S, bias, kind, seed, max_waste=4.0):
    net = cd.Net(engine, N, L, S, bias)
    net.generate_layers(kind, seed)
    net.generate_input(seed)
    net.upload(cd.default_opts(max_tile_waste=max_waste))
    return net


CASES = [
    # N, L, S, bias, kind, seed, max_waste (MFMA image on random layers too when large)
    (1024, 120, 4096, -0.3, cd.GEN_RANDOM, 20261203, 64.0),
    (1024, 40, 3000, -0.3, cd.GEN_RANDOM, 11, 4.0),
    (4096, 60, 2048, -0.35, cd.GEN_RANDOM, 20261204, 64.0),
    (2048, 50, 5000, -0.4, cd.GEN_BANDED, 5, 4.0),
    (16384, 120, 1500, -0.4, cd.GEN_BANDED, 20261205, 4.0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"N{c[0]}_L{c[1]}_S{c[2]}_{'banded' if c[4] else 'random'}" for c in CASES])
def test_generated_parity(engine, case):
    N, L, S, bias, kind, seed, mw = case
    say(f"generated N={N} L={L} S={S} kind={kind}")
    alive, live = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, kind, seed, mw)
    o32 = cd.run([net], cd.default_opts(precision=cd.FP32))[0]
    assert np.array_equal(o32.categories, want)
    assert_live_close(o32.live, live, o32)
    o16 = cd.run([net], cd.default_opts(precision=cd.FP16))[0]
    assert np.array_equal(o16.categories, want)
    assert o16.live[0] == S
    assert_live_close(o16.live, live, o16)
    if kind == cd.GEN_BANDED or mw > 10:
        assert o16.mfma_layers == L
    net.close()


FULL = ["c1", "c2", "c3"]


def check_plans(net, want, live, L, name):
    """Every tuner candidate applied to every layer, then the exhaustive per-layer plan under
    its tuned compaction threshold (eager and graph-replayed): the plans the bench times."""
    for cand in cd.candidates():
        net.set_plan([cand] * L)
        out = cd.run([net], cd.default_opts(precision=cd.FP16))[0]
        say(f"  {name} candidate {cand}: categories={len(out.categories)} ms={out.ms_total:.2f} "
            f"mfma_layers={out.mfma_layers}")
        assert np.array_equal(out.categories, want), cand
        assert_live_close(out.live, live, out)
        if cand[0] in (1, 3):  # MFMA kernels: every layer has a tile image at waste <= 64
            assert out.mfma_layers == L, cand
    plan, _, _ = cd.tune_exhaustive(net, reps=1)
    net.set_plan(plan)
    ppm, _ = cd.tune_keep(net, candidates=(0, 30000, 1000000), reps=1)
    kinds = sorted({p[:2] for p in plan})
    say(f"  {name} tuned plan (kernel, order) {kinds}, keep_holes_ppm {ppm}")
    for k_ppm in sorted({ppm, 1000000}):
        for graph in (0, 1):
            out = cd.run([net], cd.default_opts(precision=cd.FP16, keep_holes_ppm=k_ppm, use_graph=graph))[0]
            assert np.array_equal(out.categories, want), (k_ppm, graph)
            assert_live_close(out.live, live, out)
    net.set_plan(None)


@pytest.mark.parametrize("name", FULL)
def test_full_config_categories(engine, name):
    """The full-size networks (60000 samples) against the FP64 oracle, with the tile images the
    bench uses (random layers too, waste up to 64): the default launch, every tuner candidate
    and the tuned plan on the FP16 path, and the FP32 path on the SIMT and the MFMA kernels."""
    (spec,) = cd.config_nets(name)
    say(f"full config {name}: N={spec.N} L={spec.L} S={spec.S}")
    alive, live = orc.forward_fp64(orc.OracleNet.generated(spec.N, spec.L, spec.bias, spec.kind, spec.seed), 0, spec.S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, spec.N, spec.L, spec.S, spec.bias, spec.kind, spec.seed, 64.0)
    for prec, m32 in ((cd.FP16, 0), (cd.FP32, 0), (cd.FP32, 1)):
        out = cd.run([net], cd.default_opts(precision=prec, fp32_mfma=m32))[0]
        say(f"  prec={prec} fp32_mfma={m32} categories={len(out.categories)} oracle={len(want)} "
            f"flagged={out.n_flagged} replayed={out.n_replayed} changed={out.n_changed} ms={out.ms_total:.2f} "
            f"mfma_layers={out.mfma_layers}")
        assert np.array_equal(out.categories, want)
        assert_live_close(out.live, live, out)
        if prec == cd.FP16:
            assert out.mfma_layers == spec.L
    check_plans(net, want, live, spec.L, name)
    net.close()


def test_edge_shapes(engine):
    rng = np.random.default_rng(5)
    # N = 81: an odd row count (the LDS-resident kernel's staging areas start past the tile)
    for N, L, S, bias, K in [(80, 3, 1, -0.3, 20), (80, 3, 257, -0.3, 20), (96, 1, 300, -0.3, 32),
                             (64, 4, 600, -5.0, 8), (48, 3, 90, 0.2, 6), (128, 0, 50, -0.3, 4), (81, 3, 130, -0.3, 12)]:
        say(f"edge N={N} L={L} S={S} bias={bias}")
        layers = []
        for _ in range(L):
            cp = np.arange(N + 1, dtype=np.int32) * K
            cp[N // 2 + 1:] -= K  # one empty column
            rows = np.concatenate([np.sort(rng.choice(N, K, replace=False)) for _ in range(N - 1)]).astype(np.int32)
            vals = rng.choice(np.array([0.0625, 0.125, -0.0625], np.float32), size=len(rows))
            layers.append((cp, rows, vals))
        y0 = (rng.random((S, N)) < 0.3).astype(np.float32)
        y0[::5] = 0
        alive, live = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers, y0.astype(np.uint8)), 0, S)
        net = cd.Net(engine, N, L, S, bias)
        for l, (cp, rw, vl) in enumerate(layers):
            net.set_layer_csc(l, cp, rw, vl)
        net.set_input_dense(y0)
        net.upload(cd.default_opts(max_tile_waste=1e9))
        # every MFMA item order: the XCD-partitioned ones on grids smaller than 8 XCDs too
        for prec, order in [(cd.FP32, 1), (cd.FP16, 0), (cd.FP16, 1), (cd.FP16, 2), (cd.FP16, 3), (cd.FP16, 4),
                            (cd.FP16, 5), (cd.FP16, 6), (cd.FP64, 1)]:
            out = cd.run([net], cd.default_opts(precision=prec, spmm_order=order))[0]
            assert np.array_equal(out.categories, np.flatnonzero(alive)), (N, L, S, bias, prec, order)
            if prec == cd.FP16:
                assert out.mfma_layers == L, (N, L, S, bias, order)
            if prec == cd.FP32 and L > 0:
                assert_live_close(out.live, live, out)
        net.close()


@pytest.mark.parametrize("kind", [cd.GEN_BANDED, cd.GEN_RANDOM], ids=["banded", "random"])
def test_item_orders_and_grids_agree(engine, kind):
    """Every MFMA work order and grid size, with compacted layers gathered by output position
    or stored compacted from input positions, with layer 0 reading the bit-packed or the
    FP16 input, and with the weights read from the mask image or the FP16 tiles, computes
    the same layers: categories equal to the oracle's, final activations
    identical (each output is one MFMA chain with the same operands)."""
    N, L, S, bias, seed = (2048, 24, 5000, -0.4, 31) if kind == cd.GEN_BANDED else (1024, 24, 5000, -0.3, 32)
    say(f"orders kind={kind}")
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, kind, seed, 64.0)
    ref = None
    for gather, bits, wmask in ((1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0), (0, 1, 0)):
        for order in (0, 1, 2, 3, 4, 5, 6):
            for groups in (1, 4):
                o = cd.default_opts(precision=cd.FP16, spmm_order=order, mfma_groups_per_cu=groups, dump_layer=L - 1,
                                    spmm_gather=gather, input_bits=bits, weight_mask=wmask)
                out = cd.run([net], o, dump_cap_rows=S)[0]
                assert out.mfma_layers == L
                assert np.array_equal(out.categories, np.flatnonzero(alive)), (gather, bits, wmask, order, groups)
                # the kernel each layer reports: the order's kernel (band and LDS-resident kernels
                # on the random N = 1024 net, which has band images and fits LDS; the banded
                # N = 2048 net falls back to the gather kernel), its mask variant when the
                # weights are read as a mask
                if order == 6 and kind == cd.GEN_RANDOM:
                    want_k = {cd.KERNEL_MFMA_BAND_MASK} if wmask else {cd.KERNEL_MFMA_BAND, cd.KERNEL_MFMA}
                elif order == 5 and kind == cd.GEN_RANDOM:
                    want_k = {cd.KERNEL_MFMA_LDS_MASK if wmask else cd.KERNEL_MFMA_LDS}
                else:
                    want_k = {cd.KERNEL_MFMA_MASK if wmask else cd.KERNEL_MFMA}
                assert set(out.layer_kernel.tolist()) <= want_k, (order, wmask, out.layer_kernel)
                got = np.zeros((S, N), np.float32)
                got[out.dump_ids] = out.dump_vals
                if ref is None:
                    ref = got
                assert np.array_equal(got, ref), (gather, bits, wmask, order, groups)
    net.close()


def test_fp16_refuses_inexact_values(engine):
    N, S = 64, 10
    net = cd.Net(engine, N, 1, S, -0.3)
    cp = np.arange(N + 1, dtype=np.int32) * 2
    rows = np.tile(np.array([0, 1], np.int32), N)
    net.set_layer_csc(0, cp, rows, np.full(2 * N, 0.1, np.float32))  # 0.1 is not exact in FP16
    y0 = np.zeros((S, N), np.float32)
    y0[:, :2] = 1
    net.set_input_dense(y0)
    net.upload()
    with pytest.raises(cd.CospdnnError) as e:
        cd.run([net], cd.default_opts(precision=cd.FP16))
    assert e.value.code == 5
    out = cd.run([net], cd.default_opts(precision=cd.FP32))[0]
    assert len(out.categories) == 0  # 2 * 0.1 - 0.3 < 0
    net.close()


def test_guard_recomputes_everything_at_layer_zero(engine):
    N, L, S, bias, seed = 1024, 30, 2000, -0.3, 77
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.RANDOM, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed, 64.0)
    out = cd.run([net], cd.default_opts(precision=cd.FP16, guard_layer=0))[0]
    assert out.n_flagged == S
    assert np.array_equal(out.categories, np.flatnonzero(alive))
    # the literal guard band alone (no checkpoint): rows with a pre-activation within 2^-8 of
    # 0 or 32 at any layer go to the FP32 tier (and from its checkpoint to the FP64 tier)
    lit = cd.run([net], cd.default_opts(precision=cd.FP16, guard_layer=-1, guard_band=2.0 ** -8, guard_band_from=0))[0]
    say(f"literal band: flagged={lit.n_flagged} replayed={lit.n_replayed} changed={lit.n_changed}")
    assert 0 < lit.n_flagged < S
    assert np.array_equal(lit.categories, np.flatnonzero(alive))
    net.close()


def test_graph_replay_matches_eager(engine):
    N, L, S, bias, seed = 2048, 40, 3000, -0.4, 21
    net = gen_net(engine, N, L, S, bias, cd.GEN_BANDED, seed)
    for prec in (cd.FP16, cd.FP32):
        eager = cd.run([net], cd.default_opts(precision=prec))[0]
        for _ in range(3):
            g = cd.run([net], cd.default_opts(precision=prec, use_graph=1))[0]
            assert np.array_equal(g.categories, eager.categories)
            assert np.array_equal(g.live, eager.live)
            e2 = cd.run([net], cd.default_opts(precision=prec))[0]
            assert np.array_equal(e2.categories, eager.categories)
    net.close()


@pytest.mark.parametrize("kind", [cd.GEN_BANDED, cd.GEN_RANDOM], ids=["banded", "random"])
def test_concurrent_guard_matches_sequential(engine, kind):
    """The guard tiers forked onto a side stream (device-side row counts, eager and replayed
    from a hipGraph, several replays with eager runs between them) decide every row as the
    tiers run one after the other do, and as the FP64 oracle does."""
    N, L, S, bias, seed = (2048, 40, 6000, -0.4, 41) if kind == cd.GEN_BANDED else (1024, 40, 6000, -0.3, 42)
    say(f"concurrent guard kind={kind}")
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, kind, seed, 64.0)
    seq = cd.run([net], cd.default_opts(precision=cd.FP16, guard_concurrent=0))[0]
    assert np.array_equal(seq.categories, want)
    assert seq.n_flagged > 0
    for graph in (0, 1, 1, 1):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, guard_concurrent=1, use_graph=graph))[0]
        assert np.array_equal(out.categories, want), graph
        assert np.array_equal(out.live, seq.live)
        assert out.n_flagged == seq.n_flagged
        e = cd.run([net], cd.default_opts(precision=cd.FP16, guard_concurrent=0))[0]
        assert np.array_equal(e.categories, want)
    net.close()


@pytest.mark.parametrize("kind", [cd.GEN_BANDED, cd.GEN_RANDOM], ids=["banded", "random"])
def test_kept_layout_matches_compaction(engine, kind):
    """Scans that keep the layout (dead rows left in place as all-zero holes) instead of
    compacting decide every row as the compacting pass does: categories equal to the FP64
    oracle's, the same live rows per layer and the same rows handed to the guard tiers, on the
    MFMA and the SIMT kernel, eager and replayed from a hipGraph.  With every allowed scan
    keeping, the layout carries holes; with a threshold below one row, it carries none."""
    N, L, S, bias, seed = (2048, 40, 6000, -0.4, 41) if kind == cd.GEN_BANDED else (1024, 40, 6000, -0.3, 42)
    say(f"kept layout kind={kind}")
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, kind, seed, 64.0)
    for use_mfma in (1, 0):
        base = cd.run([net], cd.default_opts(precision=cd.FP16, use_mfma=use_mfma))[0]
        assert np.array_equal(base.categories, want)
        assert np.array_equal(base.positions[:L], base.live[:L])
        for ppm, graph, conc, gather in ((1, 0, 1, 1), (1000, 0, 1, 1), (20000, 1, 1, 1), (20000, 0, 1, 0),
                                         (1000000, 0, 1, 1), (1000000, 1, 1, 0), (1000000, 0, 0, 1)):
            say(f"  use_mfma={use_mfma} ppm={ppm} graph={graph} concurrent={conc} gather={gather}")
            o = cd.default_opts(precision=cd.FP16, use_mfma=use_mfma, keep_holes_ppm=ppm, use_graph=graph,
                                guard_concurrent=conc, spmm_gather=gather)
            out = cd.run([net], o)[0]
            assert np.array_equal(out.categories, want), (use_mfma, ppm, graph)
            assert np.array_equal(out.live, base.live), (use_mfma, ppm, graph)
            assert out.n_flagged == base.n_flagged
            assert (out.positions[:L] >= out.live[:L]).all()
            assert (out.positions <= S).all()
            if ppm == 1:
                assert np.array_equal(out.positions[:L], out.live[:L])
            if ppm == 1000000:
                assert (out.positions[:L] > out.live[:L]).any()
    net.close()


def test_concurrent_guard_overflow_falls_back(engine):
    """Checkpoints at layer 1 of both guarded passes leave more rows for the FP64 tier than
    its device buffers hold (512): the run redoes the tiers one after the other and still
    matches the oracle."""
    N, L, S, bias, seed = 1024, 12, 3000, -0.3, 43
    say("concurrent guard overflow")
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, cd.GEN_RANDOM, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed, 64.0)
    for graph in (0, 1):
        o = cd.default_opts(precision=cd.FP16, guard_layer=1, guard_layer32=1, use_graph=graph)
        out = cd.run([net], o)[0]
        assert out.n_replayed > 512
        assert np.array_equal(out.categories, np.flatnonzero(alive))
    net.close()


def test_graph_recaptured_after_tier_buffers_grow(engine):
    """A captured graph holds the addresses of the guard tiers' buffers: a run that grows
    them (one-after-the-other tiers over nearly every row) must make the next graph replay
    recapture, not replay stale addresses."""
    N, L, S, bias, seed = 1024, 24, 20000, -0.3, 44
    say("graph after tier buffers grow")
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, cd.GEN_RANDOM, seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed, 64.0)
    g = cd.default_opts(precision=cd.FP16, use_graph=1)
    assert np.array_equal(cd.run([net], g)[0].categories, want)
    big = cd.run([net], cd.default_opts(precision=cd.FP16, guard_layer=1, guard_concurrent=0))[0]
    assert big.n_flagged > S // 4  # more rows than the concurrent tiers' buffers hold
    assert np.array_equal(big.categories, want)
    for _ in range(2):
        assert np.array_equal(cd.run([net], g)[0].categories, want)
    net.close()


def test_layer_timing_and_counters(engine):
    N, L, S, bias, seed = 4096, 20, 4000, -0.35, 3
    net = gen_net(engine, N, L, S, bias, cd.GEN_BANDED, seed)
    out = cd.run([net], cd.default_opts(precision=cd.FP16, layer_timing=1))[0]
    assert out.ms_spmm > 0 and abs(out.layer_ms.sum() - out.ms_spmm) < 1e-3 * max(1.0, out.ms_spmm)
    assert out.edges == S * N * 32 * L
    assert out.alg_bytes > 0 and out.edges_per_sec > 0
    net.close()


def test_dump_layer_max_abs_error(engine):
    N, S, bias, seed, dl = 1024, 1000, -0.3, 20261203, 3
    net = gen_net(engine, N, 12, S, bias, cd.GEN_RANDOM, seed, 64.0)
    o16 = cd.run([net], cd.default_opts(precision=cd.FP16, dump_layer=dl), dump_cap_rows=S)[0]
    o32 = cd.run([net], cd.default_opts(precision=cd.FP32, dump_layer=dl), dump_cap_rows=S)[0]
    _, _, ref = orc.forward_fp64(orc.OracleNet.generated(N, dl + 1, bias, orc.RANDOM, seed), 0, S,
                                 final_ids=np.arange(S))
    d16 = np.abs(o16.dump_vals - ref[o16.dump_ids]).max()
    d32 = np.abs(o32.dump_vals - ref[o32.dump_ids]).max()
    say(f"layer {dl}: max|dY| fp16={d16:.3e} fp32={d32:.3e} rows={len(o16.dump_ids)}")
    assert d32 < 1e-4
    assert d16 < 2e-2
    net.close()


def test_two_networks_one_run(engine):
    a = (1024, 30, 3000, -0.3, cd.GEN_RANDOM, 8)
    b = (2048, 45, 1200, -0.4, cd.GEN_BANDED, 9)
    nets = [gen_net(engine, *x[:6], 64.0) for x in (a, b)]
    outs = cd.run(nets, cd.default_opts(precision=cd.FP16))
    for x, o in zip((a, b), outs):
        N, L, S, bias, kind, seed = x
        alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
        assert np.array_equal(o.categories, np.flatnonzero(alive))
    for n in nets:
        n.close()


def ref_bcsc_of_banded(cp, rw, vl, N, fuse_inv):
    """Reference-style BCSC of a banded layer: block column bc = 16 outputs of one class, two
    16-row blocks over the class's 32 fused rows, values [out][in] (utils/matrix.h:1245)."""
    s = N // 32
    cls = rw[cp[:-1]] % s
    order = np.argsort(cls, kind="stable")
    values, row_idx, ptr = [], [], [0]
    for bc in range(N // 16):
        outs = order[bc * 16:(bc + 1) * 16]
        start = int(cls[outs[0]]) * 32
        for half in range(2):
            blk = np.zeros((16, 16), np.float32)
            for oc, j in enumerate(outs):
                for q in range(cp[j], cp[j + 1]):
                    fr = fuse_inv[rw[q]] - start - 16 * half
                    if 0 <= fr < 16:
                        blk[oc, fr] = vl[q]
            values += blk.ravel().tolist()
            row_idx.append(start + 16 * half)
        ptr.append(len(row_idx))
    return np.array(values, np.float32), np.array(row_idx, np.int32), np.array(ptr, np.int32)


def test_reference_bcsc_input(engine):
    """A network handed over in the reference's BCSC + fuse form (test_graph_challenge_tc's inputs);
    the oracle runs the CSC those arrays decode to (the decoding is pinned against the reference's
    own builder in test_format.py)."""
    N, L, S, bias, seed = 512, 10, 800, -0.4, 31
    s = N // 32
    fuse = np.array([(r % 32) * s + r // 32 for r in range(N)], np.int32)
    inv = np.empty(N, np.int32)
    inv[fuse] = np.arange(N, dtype=np.int32)
    net = cd.Net(engine, N, L, S, bias)
    dec = []
    for l in range(L):
        cp, rw, vl = orc.gen_layer(seed, orc.BANDED, N, 32, l)
        values, row_idx, ptr = ref_bcsc_of_banded(cp, rw, vl, N, inv)
        net.set_layer_ref_bcsc(l, values, row_idx, ptr, fuse)
        dec.append(cd.ref_bcsc_to_csc(N, values, row_idx, ptr, fuse))
    y0 = orc.gen_y0_rows(seed, N, 0, S)
    net.set_input_dense(y0.astype(np.float32))
    net.upload(cd.default_opts())
    alive, live = orc.forward_fp64(orc.OracleNet.explicit(N, bias, dec, y0), 0, S)
    for prec in (cd.FP32, cd.FP16):
        out = cd.run([net], cd.default_opts(precision=prec))[0]
        assert np.array_equal(out.categories, np.flatnonzero(alive))
        if prec == cd.FP16:
            assert out.mfma_layers == L  # fused blocks give zero-waste tiles
    net.close()


def test_fp64_precision_matches_oracle(engine):
    """The FP64 precision (layer-major double SpMM, the exact tier) equals the FP64 oracle, banded and random."""
    for N, L, S, bias, kind, seed in [(1024, 60, 700, -0.3, cd.GEN_RANDOM, 12), (2048, 40, 500, -0.4, cd.GEN_BANDED, 13)]:
        alive, live = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
        net = gen_net(engine, N, L, S, bias, kind, seed)
        out = cd.run([net], cd.default_opts(precision=cd.FP64))[0]
        assert out.n_flagged == 0 and out.n_replayed == 0
        assert np.array_equal(out.live, live)
        assert np.array_equal(out.categories, np.flatnonzero(alive))
        net.close()


def test_fp32_pass_replays_undecided_rows(engine):
    """guard_layer32=0 sends every row of the FP32 path through the FP64 replay."""
    N, L, S, bias, seed = 2048, 30, 900, -0.4, 14
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.BANDED, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, cd.GEN_BANDED, seed)
    out = cd.run([net], cd.default_opts(precision=cd.FP32, guard_layer32=0))[0]
    assert out.n_replayed == S
    assert np.array_equal(out.categories, np.flatnonzero(alive))
    m32 = cd.run([net], cd.default_opts(precision=cd.FP32, fp32_mfma=1))[0]
    assert m32.mfma_layers == L
    assert np.array_equal(m32.categories, np.flatnonzero(alive))
    net.close()


@pytest.mark.gpu
def test_guard_retires_saturated_rows_only_where_all32_is_fixed(engine):
    """The guard passes drop all-32 rows only while every remaining layer keeps all-32 fixed.
    With bias -20, layer 7 (16 inputs per column, column sum 1) maps all-32 to 12 and the
    rows die two layers later: retiring them before layer 7 would report them as categories."""
    N, S, bias = 256, 400, -20.0
    rng = np.random.default_rng(21)

    def layer(K, w):
        cp = np.arange(N + 1, dtype=np.int32) * K
        rows = np.concatenate([np.sort(rng.choice(N, K, replace=False)) for _ in range(N)]).astype(np.int32)
        return cp, rows, np.full(N * K, w, np.float32)

    layers = [layer(32, 4.0)] + [layer(32, 0.0625) for _ in range(6)] + [layer(16, 0.0625)]
    layers += [layer(32, 0.0625) for _ in range(4)]
    y0 = (rng.random((S, N)) < rng.uniform(0.2, 0.8, (S, 1))).astype(np.float32)
    for L in (7, 12):
        say(f"retire N={N} L={L} S={S}")
        alive, _ = orc.forward_fp64(orc.OracleNet.explicit(N, bias, layers[:L], y0.astype(np.uint8)), 0, S)
        want = np.flatnonzero(alive)
        assert (len(want) > 0) if L == 7 else (len(want) == 0)
        net = cd.Net(engine, N, L, S, bias)
        for l, (cp, rw, vl) in enumerate(layers[:L]):
            net.set_layer_csc(l, cp, rw, vl)
        net.set_input_dense(y0)
        net.upload(cd.default_opts(max_tile_waste=1e9))
        for opts in (cd.default_opts(precision=cd.FP32, guard_layer32=0),
                     cd.default_opts(precision=cd.FP16, guard_layer=0, guard_layer32=0),
                     cd.default_opts(precision=cd.FP16)):
            out = cd.run([net], opts)[0]
            assert np.array_equal(out.categories, want), (L, opts.precision, opts.guard_layer)
        net.close()


def test_short_net_late_deaths_are_rechecked(engine):
    """A row that FP64 keeps alive to the last layer but that the FP16 pass lets die in one
    of the last layers (N=2048 random, L=12, seed 45 has one: FP64 max pre-activation
    +0.0096 at layer 11, the FP16 trajectory -0.10) is recomputed: deaths in the last
    guard_tail layers are flagged like undecided rows."""
    N, L, S, bias, seed = 2048, 12, 5000, -0.35, 45
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.RANDOM, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed, max_waste=64.0)
    out = cd.run([net], cd.default_opts(precision=cd.FP16))[0]
    assert np.array_equal(out.categories, np.flatnonzero(alive))
    assert out.n_flagged > 0
    net.close()


def test_c4_shape_on_a_sample_slice(engine):
    """The largest network (N=65536, L=480: ~1e9 weights) on the first 12000 samples (188
    64-sample tiles per layer at first, 4 or fewer after layer 7), with the tile images the
    bench uses (waste ~32): the default MFMA launch, the row-gather kernel on every layer,
    the sample-slab kernel on every layer (one sample per slab: the 768-thread variant on the
    layers before the checkpoint, layer 12, the 512-thread one after), and the tuned per-layer plan
    (the bench's: MFMA on layer 0, the slab kernel after it), eager and graph-replayed;
    categories equal to the FP64 oracle's."""
    spec = cd.config_nets("c4")[0]
    S = 12000
    say(f"c4 shape N={spec.N} L={spec.L} S={S}")
    alive, live = orc.forward_fp64(orc.OracleNet.generated(spec.N, spec.L, spec.bias, orc.RANDOM, spec.seed), 0, S)
    want = np.flatnonzero(alive)
    net = gen_net(engine, spec.N, spec.L, S, spec.bias, cd.GEN_RANDOM, spec.seed, 64.0)
    out = cd.run([net], cd.default_opts(precision=cd.FP16))[0]
    say(f"  default: categories={len(out.categories)} oracle={len(want)} mfma_layers={out.mfma_layers} "
        f"ms={out.ms_total:.2f}")
    assert np.array_equal(out.categories, want)
    assert_live_close(out.live, live, out)
    assert out.mfma_layers == spec.L
    net.set_plan([(cd.PLAN_ROWS, 1, 5)] * spec.L)
    out = cd.run([net], cd.default_opts(precision=cd.FP16))[0]
    assert np.array_equal(out.categories, want)
    net.set_plan([(cd.PLAN_SLAB, 0, 1)] * spec.L)
    for graph in (0, 1):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, use_graph=graph))[0]
        say(f"  all slab, graph={graph}: categories={len(out.categories)} ms={out.ms_total:.2f}")
        ks = out.layer_kernel.tolist()  # layer 0 on the bit-packed input: the bit-sliced variant
        assert ks[0] == cd.KERNEL_SLAB_BITS and set(ks[1:]) == {cd.KERNEL_SLAB}
        assert np.array_equal(out.categories, want), graph
        assert_live_close(out.live, live, out)
    plan, _, _ = cd.tune_exhaustive(net, reps=1)
    net.set_plan(plan)
    kinds = {}
    for p in plan:
        kinds[p[:2]] = kinds.get(p[:2], 0) + 1
    say(f"  tuned plan layers per (kernel, order): {kinds}")
    for graph in (0, 1):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, use_graph=graph))[0]
        assert np.array_equal(out.categories, want), graph
        assert_live_close(out.live, live, out)
    net.close()


def test_lds_resident_kernel_loops_over_tiles(engine):
    """The LDS-resident kernel (order 5) with more 64-sample tiles than work-groups (N=1024,
    S=40000: 625 tiles on a grid of at most one work-group per CU), so work-groups refill
    their LDS tile after a barrier; every layer on it, final activations equal to the plain
    MFMA kernel's and categories to the oracle's."""
    N, L, S, bias, seed = 1024, 20, 40000, -0.3, 51
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, orc.RANDOM, seed), 0, S)
    net = gen_net(engine, N, L, S, bias, cd.GEN_RANDOM, seed, 64.0)
    ref = None
    for order in (5, 1):
        out = cd.run([net], cd.default_opts(precision=cd.FP16, spmm_order=order, dump_layer=L - 1,
                                            spmm_gather=0), dump_cap_rows=S)[0]
        assert out.mfma_layers == L
        assert np.array_equal(out.categories, np.flatnonzero(alive)), order
        got = np.zeros((S, N), np.float32)
        got[out.dump_ids] = out.dump_vals
        if ref is None:
            ref = got
        assert np.array_equal(got, ref), order
    net.close()
