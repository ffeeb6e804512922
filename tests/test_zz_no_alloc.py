"""The allocation-free entry point (cospdnn_run_noalloc, through the C ABI) against
cospdnn_run on the same networks and against the FP64 oracle: two calls in a row, the second
with a smaller input, then the first network again (the fixed array keeps the previous
call's contents); networks too large for the array and unsupported options are refused."""
import sys

import numpy as np
import pytest

import cospdnn_amd as cd
from oracle import oracle_ctypes as orc

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[noalloc] {msg}", file=sys.stderr, flush=True)


def make(engine, N, L, S, bias, kind, seed):
    net = cd.Net(engine, N, L, S, bias)
    net.generate_layers(kind, seed)
    net.generate_input(seed)
    net.upload(cd.default_opts(max_tile_waste=64.0))
    return net


def check_same(got, ref, want):
    assert np.array_equal(got.categories, want)
    assert np.array_equal(got.categories, ref.categories)
    assert np.array_equal(got.live, ref.live)
    assert np.array_equal(got.positions, ref.positions)
    assert got.n_flagged == ref.n_flagged and got.n_replayed == ref.n_replayed
    assert got.mfma_layers == ref.mfma_layers


def test_noalloc_matches_run(engine):
    # a random network on the MFMA tile images, then a smaller banded one
    specs = [(1024, 30, 12000, -0.3, cd.GEN_RANDOM, 61), (2048, 20, 3000, -0.4, cd.GEN_BANDED, 62)]
    nets, refs, wants = [], [], []
    for N, L, S, bias, kind, seed in specs:
        say(f"N={N} L={L} S={S} kind={kind}")
        alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, kind, seed), 0, S)
        net = make(engine, N, L, S, bias, kind, seed)
        # the allocation-free call runs the guard tiers one after the other, as this does
        ref = cd.run([net], cd.default_opts(precision=cd.FP16, guard_concurrent=0))[0]
        got = cd.run_noalloc(net, cd.default_opts(precision=cd.FP16))
        say(f"  categories {len(got.categories)} flagged {got.n_flagged} replayed {got.n_replayed} "
            f"ms {got.ms_total:.2f} (cospdnn_run {ref.ms_total:.2f})")
        check_same(got, ref, np.flatnonzero(alive))
        assert ref.n_flagged > 0
        nets.append(net)
        refs.append(ref)
        wants.append(np.flatnonzero(alive))
    # the first network again after the smaller one, and with other launch options
    again = cd.run_noalloc(nets[0], cd.default_opts(precision=cd.FP16))
    check_same(again, refs[0], wants[0])
    o = cd.default_opts(precision=cd.FP16, spmm_order=5, keep_holes_ppm=30000)
    check_same(cd.run_noalloc(nets[0], o), cd.run([nets[0]], cd.default_opts(
        precision=cd.FP16, spmm_order=5, keep_holes_ppm=30000, guard_concurrent=0))[0], wants[0])
    # the sample-slab kernel on every layer (64 and 32 samples per slab)
    for net, want in zip(nets, wants):
        say(f"slab N={net.N}")
        check_same(cd.run_noalloc(net, cd.default_opts(precision=cd.FP16, spmm_slab=1)),
                   cd.run([net], cd.default_opts(precision=cd.FP16, spmm_slab=1, guard_concurrent=0))[0], want)
    for n in nets:
        n.close()


def test_noalloc_slab_layers_next_to_mfma_layers(engine):
    """4 samples per slab (N = 16384), slab layers alternating with MFMA layers: cospdnn_run
    converts the activation layout beside every slab layer through a third buffer, the fixed
    array holds none, so there the slab kernel reads and writes the neuron-major layout
    itself; the same results either way."""
    N, L, S, bias, seed = 16384, 10, 200, -0.4, 65  # the fixed array also holds both tiers' buffers
    alive, _ = orc.forward_fp64(orc.OracleNet.generated(N, L, bias, cd.GEN_RANDOM, seed), 0, S)
    net = make(engine, N, L, S, bias, cd.GEN_RANDOM, seed)
    net.set_plan([cd.LayerPlan(cd.PLAN_SLAB, 0, 1) if l % 2 == 0 else cd.LayerPlan(cd.PLAN_MFMA, 1, 4)
                  for l in range(L)])
    ref = cd.run([net], cd.default_opts(precision=cd.FP16, guard_concurrent=0))[0]
    got = cd.run_noalloc(net, cd.default_opts(precision=cd.FP16))
    say(f"N={N} alternating slab/MFMA: categories {len(got.categories)} ms {got.ms_total:.2f} "
        f"(cospdnn_run {ref.ms_total:.2f})")
    # layer 0 reads the bit-packed input: the slab plan's bit-sliced variant there
    assert set(got.layer_kernel.tolist()) == {cd.KERNEL_SLAB_BITS, cd.KERNEL_SLAB, cd.KERNEL_MFMA_MASK}
    check_same(got, ref, np.flatnonzero(alive))
    net.close()


def test_noalloc_refuses_what_it_cannot_hold(engine):
    # 4 * N * S bytes of FP16 activations alone exceed the fixed array
    net = make(engine, 4096, 4, 20000, -0.35, cd.GEN_RANDOM, 63)
    with pytest.raises(cd.CospdnnError) as e:
        cd.run_noalloc(net)
    assert e.value.code == 1
    net.close()
    small = make(engine, 512, 4, 500, -0.3, cd.GEN_RANDOM, 64)
    for o in (cd.default_opts(precision=cd.FP32), cd.default_opts(layer_timing=1), cd.default_opts(dump_layer=1)):
        with pytest.raises(cd.CospdnnError) as e:
            cd.run_noalloc(small, o)
        assert e.value.code == 1
    assert len(cd.run_noalloc(small).categories) == len(cd.run([small])[0].categories)
    small.close()
