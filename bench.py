"""Benchmark of the sparse-DNN inference path on one MI355X.

A step is one full forward of the network (all L layers, compaction and the
FP32 / FP64 guard recompute) over the synthetic input already resident in HBM.
Before the timed steps the per-layer launch plan and the compaction threshold are
autotuned (exhaustive search, the compaction threshold on whole forwards, then the
cost-model MLP trained on those timings picks the plan that is timed).
Default workload: N=16384, L=120, 60000 samples, banded connectivity, bias -0.4
(the configuration the FP16-MFMA target is stated on).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--cpu-samples 1500] [--tune 1] [--graph 1]
                    [--tuning-out FILE | --tuning-in FILE]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "inference edges/sec (samples×nnz(W)×layers/s), FP16-MFMA vs plain FP32, 1 GPU"
HBM_PEAK_GBS = 8000.0
MLP_FALLBACK_REGRET = 0.02


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_value(prefix, config, key, kernel):
    """The figure `key` of `kernel` on workload `config` from a committed rocprofv3 summary
    (profiles/<prefix>_<config>_*.md), each of which ends with machine-readable lines
    `<key> <kernel> <value>`, or (None, None)."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{prefix}_{config}_*.md"))):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == key and parts[1] == kernel:
                    return float(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def profiled_traffic(config, kernel):
    """HBM bytes per launch of `kernel` on this workload from a rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (profiles/hbm_traffic_<config>_*.md), or None."""
    return profiled_value("hbm_traffic", config, "traffic_bytes_per_launch", kernel)


def profiled_mfma_util(config, kernel):
    """MfmaUtil (%) of `kernel` on this workload from a rocprofv3 counter summary
    (profiles/mfma_util_<config>_*.md), or None."""
    return profiled_value("mfma_util", config, "mfma_util_percent", kernel)


def kernel_runs(cd, layer_kernel):
    """The SpMM kernel of each layer as runs [name, first layer, last layer]."""
    runs = []
    for l, k in enumerate(layer_kernel):
        name = cd.kernel_name(int(k)) or "none"
        if runs and runs[-1][0] == name and runs[-1][2] == l - 1:
            runs[-1][2] = l
        else:
            runs.append([name, l, l])
    return runs


def kernel_mix(cd, np, spec, out, layer_us):
    """SpMM time by kernel over the layers of one forward: {name: {layers, ms, alg_bytes}},
    with the algorithmic bytes of layer l = 2 x live_l x N x 2 B (FP16 activations read and
    written) + nnz(W_l) x 2 B."""
    mix = {}
    for l in range(spec.L):
        name = cd.kernel_name(int(out.layer_kernel[l])) or "none"
        m = mix.setdefault(name, {"layers": 0, "ms": 0.0, "alg_bytes": 0.0})
        m["layers"] += 1
        m["ms"] += float(layer_us[l]) * 1e-3
        m["alg_bytes"] += 2.0 * float(out.live[l]) * spec.N * 2 + float(spec.N * spec.K) * 2
    return mix


def timed_runs(cd, net, opts, steps, warmup):
    for _ in range(warmup):
        cd.run([net], opts)
    outs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        outs.append(cd.run([net], opts)[0])
    wall = time.perf_counter() - t0
    return outs, wall


def cpu_baseline(spec, samples, gpu_cats):
    """Repo-style FP32 CPU forward (the oracle's port, all host threads) on the first `samples` rows."""
    import numpy as np

    from oracle import oracle_ctypes as orc

    net = orc.OracleNet.generated(spec.N, spec.L, spec.bias, spec.kind, spec.seed)
    orc.forward_fp32(net, 0, 64)  # warm the layer cache and the thread pool
    t0 = time.perf_counter()
    alive, _ = orc.forward_fp32(net, 0, samples)
    dt = time.perf_counter() - t0
    edges = samples * spec.N * spec.K * spec.L
    gpu_in_sample = gpu_cats[gpu_cats < samples]
    return {
        "value": edges / dt,
        "unit": "edges/s",
        "cores": orc.num_threads(),
        "kind": "port",
        "sample": f"first {samples} of {spec.S} samples, all {spec.L} layers, FP32 CPU forward ({dt:.1f} s)",
        "categories_match_gpu": bool(np.array_equal(np.flatnonzero(alive), gpu_in_sample)),
    }


def max_abs_dy(cd, np, net, spec, samples):
    """max |Y_l - Y_l(FP64)| of the FP16 pass over the first `samples` samples, after
    layers l in {3, 6, 12, L} (the activations of the FP16 pass before any guard
    recompute, against the FP64 oracle truncated to l layers).  Returns {l: value}."""
    from oracle import oracle_ctypes as orc

    out = {}
    for l in sorted({min(x, spec.L) for x in (3, 6, 12, spec.L)}):
        o = cd.default_opts(precision=cd.FP16, dump_layer=l - 1)
        r = cd.run([net], o, dump_cap_rows=samples)[0]
        ids = r.dump_ids
        keep = ids < samples
        got = np.zeros((samples, spec.N), dtype=np.float64)
        got[ids[keep]] = r.dump_vals[keep]
        onet = orc.OracleNet.generated(spec.N, l, spec.bias, spec.kind, spec.seed)
        _, _, want = orc.forward_fp64(onet, 0, samples, final_ids=np.arange(samples))
        out[l] = float(np.abs(got - want).max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--cpu-samples", type=int, default=1500)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=1, help="1: replay the FP16 layer loop from a hipGraph")
    ap.add_argument("--tune", type=int, default=1, help="1: autotune the per-layer launch plan before timing")
    ap.add_argument("--tuning-out", default=None, help="write the tuned plan and compaction threshold (JSON)")
    ap.add_argument("--tuning-in", default=None,
                    help="use the plan and compaction threshold of a --tuning-out file instead of tuning "
                         "(profiled runs time the same configuration without the tuning runs)")
    ap.add_argument("--dy-samples", type=int, default=1000, help="samples for max |dY| against the FP64 oracle (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by samples: each GPU would run this engine on its share)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cospdnn_amd as cd

    specs = cd.config_nets(args.config)
    if len(specs) != 1:
        return bench_batch(cd, np, args, specs)
    spec = specs[0]
    log(f"[bench] {cd.version()} config {args.config}: N={spec.N} L={spec.L} S={spec.S} "
        f"{'banded' if spec.kind else 'random'} bias={spec.bias}")
    eng = cd.Engine(0)
    cu, hbm = eng.info()
    t0 = time.perf_counter()
    net = cd.Net(eng, spec.N, spec.L, spec.S, spec.bias)
    net.generate_layers(spec.kind, spec.seed, K=spec.K, wval=spec.wval)
    net.generate_input(spec.seed)
    # tile images up to waste 64 (the default): random layers may then take the MFMA kernel
    # when the tuner finds it faster (c4: 128 MB of tile image per layer, 61 GB in all)
    net.upload(cd.default_opts())
    log(f"[bench] upload {time.perf_counter() - t0:.1f}s, {cu} CUs, {hbm / 2**30:.0f} GiB HBM")

    # Autotuning (untimed).  1: exhaustive per-layer search over the launch candidates with
    # compaction at every scan, and the cost-model MLP (14-64-64-1, trained by the library's
    # HIP kernels on those per-layer timings) picks a plan.  2: the compaction threshold
    # (keep_holes_ppm) on whole forwards with that plan.  3: the launch plan again under that
    # threshold (a kept layout changes which layers see shifted tiles), the MLP retrained on
    # both tables; its plan is the one the timed runs use.
    tuning = None
    keep_ppm = 0
    if args.tuning_in:
        with open(args.tuning_in) as f:
            saved = json.load(f)
        if saved["config"] != args.config or len(saved["plan"]) != spec.L:
            raise SystemExit(f"bench.py: {args.tuning_in} holds a tuning of {saved['config']}, not {args.config}")
        net.set_plan([tuple(p) for p in saved["plan"]])
        keep_ppm = int(saved["keep_holes_ppm"])
        tuning = {"from_file": args.tuning_in, "keep_holes_ppm": keep_ppm}
        log(f"[bench] plan and compaction threshold ({keep_ppm} ppm) from {args.tuning_in}")
    elif args.tune:
        t0 = time.perf_counter()
        cands = cd.candidates()
        _, table0, pos0 = cd.tune_exhaustive(net, reps=2)
        X0, y0 = cd.tune_samples(net, table0, pos0)
        mlp = cd.Mlp(eng, cd.TUNE_FEATURES, 64, seed=1)
        mlp.train(X0, y0, steps=3000, lr=2e-3)
        net.set_plan(mlp.plan(net, pos0)[0])
        keep_ppm, keep_ms = cd.tune_keep(net, cd.default_opts(precision=cd.FP16, use_graph=1 if args.graph else 0))
        ex_plan, table, pos = cd.tune_exhaustive(net, cd.default_opts(precision=cd.FP16, keep_holes_ppm=keep_ppm), reps=2)
        X1, y1 = cd.tune_samples(net, table, pos)
        mlp = cd.Mlp(eng, cd.TUNE_FEATURES, 64, seed=1)
        loss = mlp.train(np.concatenate([X0, X1]), np.concatenate([y0, y1]), steps=3000, lr=2e-3)
        mlp_plan, _ = mlp.plan(net, pos)
        pick = lambda plan: float(sum(table[cands.index(p), l] for l, p in enumerate(plan)))  # noqa: E731
        default = [cands.index((cd.PLAN_MFMA, 1, 4))] * spec.L
        tuning = {
            "candidates": [list(c) for c in cands],
            "keep_holes_ppm": keep_ppm,
            "forward_ms_per_keep_ppm": {str(c): float(m) for c, m in zip(cd.KEEP_CANDIDATES, keep_ms)},
            "training_samples": int(len(y0) + len(y1)), "training_loss_ln_time": loss,
            "spmm_ms_exhaustive": pick(ex_plan), "spmm_ms_mlp": pick(mlp_plan),
            "spmm_ms_default": float(sum(table[k, l] for l, k in enumerate(default))),
            "spmm_ms_default_compacting": float(sum(table0[k, l] for l, k in enumerate(default))),
            "tuning_s": time.perf_counter() - t0,
        }
        tuning["mlp_regret"] = tuning["spmm_ms_mlp"] / tuning["spmm_ms_exhaustive"] - 1.0
        # exhaustive tuning is the cost model's fallback: its plan is timed when the MLP's
        # plan is more than 2% slower on the measured table
        use_mlp = tuning["mlp_regret"] <= MLP_FALLBACK_REGRET
        plan = mlp_plan if use_mlp else ex_plan
        tuning["plan_timed"] = "mlp" if use_mlp else "exhaustive (MLP regret above 2%)"
        net.set_plan(plan)
        if args.tuning_out:
            with open(args.tuning_out, "w") as f:
                json.dump({"config": args.config, "plan": [list(p) for p in plan], "keep_holes_ppm": keep_ppm}, f)
        log(f"[bench] compaction threshold {keep_ppm} ppm (forward ms per candidate "
            f"{ {k: round(v, 3) for k, v in tuning['forward_ms_per_keep_ppm'].items()} })")
        log(f"[bench] tuning: exhaustive {tuning['spmm_ms_exhaustive']:.2f} ms, MLP plan {tuning['spmm_ms_mlp']:.2f} ms "
            f"(regret {tuning['mlp_regret']:.3f}), default {tuning['spmm_ms_default']:.2f} ms "
            f"(always compacting {tuning['spmm_ms_default_compacting']:.2f} ms), {tuning['tuning_s']:.1f} s; "
            f"timed plan: {tuning['plan_timed']}")

    # FP16-MFMA path with per-layer HIP events on the engine's stream: times the SpMM kernel
    # of every layer inside the timed steps (roofline of the dominant kernel).
    o16 = cd.default_opts(precision=cd.FP16, layer_timing=1, keep_holes_ppm=keep_ppm)
    outs16, wall16 = timed_runs(cd, net, o16, args.steps, args.warmup)
    ms_timed = wall16 * 1e3 / args.steps
    edges = outs16[0].edges
    dev_ms = float(np.median([o.ms_total for o in outs16]))
    spmm_ms = float(np.mean([o.ms_spmm for o in outs16]))
    alg_bytes = float(outs16[0].alg_bytes)
    launches = spec.L
    achieved = alg_bytes / (spmm_ms * 1e-3) / 1e9
    cats16 = outs16[0].categories
    log(f"[bench] fp16 (per-layer events): {ms_timed:.2f} ms/step (device {dev_ms:.2f} ms, spmm sum {spmm_ms:.2f} ms), "
        f"categories {len(cats16)}, recomputed {outs16[0].n_flagged}, replayed {outs16[0].n_replayed}, "
        f"guard tiers {outs16[0].ms_recompute:.2f} ms, mfma layers {outs16[0].mfma_layers}")

    # headline: the same forward without per-layer events, the layer loop replayed from a
    # hipGraph (use --graph 0 for eager launches)
    og = cd.default_opts(precision=cd.FP16, use_graph=1 if args.graph else 0, keep_holes_ppm=keep_ppm)
    outs_g, wall_g = timed_runs(cd, net, og, args.steps, args.warmup)
    ms_step = wall_g * 1e3 / args.steps
    value = edges / (ms_step * 1e-3)
    if not all(np.array_equal(o.categories, cats16) for o in outs_g):
        raise SystemExit("bench.py: graph-replayed categories differ from the eager run")
    log(f"[bench] fp16{' graph' if args.graph else ''}: {ms_step:.2f} ms/step, {value:.3e} edges/s")

    # FP32 baselines on the GPU: the plain SIMT gather path, and the FP32 MFMA path
    # (v_mfma_f32_16x16x4_f32 on the layers with an FP32 tile image, SIMT elsewhere); the
    # FP16 speed-up is stated against the faster of the two
    fp32 = None
    if not args.no_fp32:
        fp32 = {}
        for key, m32 in (("simt", 0), ("mfma", 1)):
            o32 = cd.default_opts(precision=cd.FP32, layer_timing=1, fp32_mfma=m32)
            outs32, wall32 = timed_runs(cd, net, o32, args.steps, args.warmup)
            ms32 = wall32 * 1e3 / args.steps
            spmm32 = float(np.mean([o.ms_spmm for o in outs32]))
            fp32[key] = {
                "value": edges / (ms32 * 1e-3),
                "ms_per_step": ms32,
                "spmm_ms": spmm32,
                "hbm_gbs": outs32[0].alg_bytes / (spmm32 * 1e-3) / 1e9,
                "mfma_layers": int(outs32[0].mfma_layers),
                "categories_equal_fp16": bool(np.array_equal(outs32[0].categories, cats16)),
            }
            log(f"[bench] fp32 {key}: {ms32:.2f} ms/step, {fp32[key]['value']:.3e} edges/s, SpMM "
                f"{fp32[key]['hbm_gbs']:.0f} GB/s, mfma layers {outs32[0].mfma_layers}, "
                f"categories equal: {fp32[key]['categories_equal_fp16']}")
        fp32["faster"] = max(("simt", "mfma"), key=lambda k: fp32[k]["value"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(spec, min(args.cpu_samples, spec.S), cats16)
        log(f"[bench] cpu: {cpu['value']:.3e} edges/s on {cpu['cores']} threads")

    dy = None
    if args.dy_samples > 0:
        t0 = time.perf_counter()
        dy = max_abs_dy(cd, np, net, spec, min(args.dy_samples, spec.S))
        log(f"[bench] max |dY| of the FP16 pass vs FP64 over {min(args.dy_samples, spec.S)} samples, by layer: "
            + ", ".join(f"{l}: {v:.3e}" for l, v in dy.items()) + f" ({time.perf_counter() - t0:.1f}s)")

    # 2 flops per live edge: live rows entering each layer x that layer's nonzeros
    alg_flops = 2.0 * float(sum(int(outs16[0].live[l]) for l in range(spec.L))) * spec.K * spec.N
    layer_us = np.median(np.stack([o.layer_ms for o in outs16]), axis=0) * 1e3
    # the roofline is stated for the kernel that takes most of the SpMM time
    mix = kernel_mix(cd, np, spec, outs16[0], layer_us)
    dom = max(mix, key=lambda k: mix[k]["ms"])
    dm = mix[dom]
    dom_gbs = dm["alg_bytes"] / (dm["ms"] * 1e-3) / 1e9
    traffic, traffic_src = profiled_traffic(args.config, dom)
    mfma_util, mfma_src = profiled_mfma_util(args.config, dom)
    log("[bench] SpMM by kernel: " + ", ".join(
        f"{k} {v['layers']} layers {v['ms']:.2f} ms ({v['alg_bytes'] / (v['ms'] * 1e-3) / 1e9:.0f} GB/s)"
        for k, v in sorted(mix.items(), key=lambda kv: -kv[1]["ms"])))
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: N={spec.N}, L={spec.L}, samples={spec.S}, "
                        f"{'banded' if spec.kind else 'random'} 32 nnz/column of {spec.wval}, bias {spec.bias}",
            "neurons": spec.N, "layers": spec.L, "samples": spec.S, "nnz_per_column": spec.K,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": dom_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": dom,
            "kernel_layers": dm["layers"],
            "avg_launch_ms": dm["ms"] / dm["layers"],
            "alg_bytes_per_launch": dm["alg_bytes"] / dm["layers"],
            "spmm_all_layers_gbs": achieved,
            "spmm_by_kernel": {k: {"layers": v["layers"], "ms": round(v["ms"], 3),
                                   "gbs": round(v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in mix.items()},
            "alg_flops_per_launch": alg_flops / launches,
            "mfma_tflops": alg_flops / (spmm_ms * 1e-3) / 1e12,
            "mfma_util_percent": mfma_util,
            "mfma_util_source": mfma_src,
        },
        "cpu_baseline": cpu,
        "fp32_gpu": fp32,
        "fp16_over_fp32": (value / fp32[fp32["faster"]]["value"]) if fp32 else None,
        "fp16_launch": "hipGraph replay" if args.graph else "eager",
        "tuning": tuning,
        "fp16_with_layer_events_ms_per_step": ms_timed,
        "device_ms_per_step": float(np.median([o.ms_total for o in outs_g])),
        "categories": int(len(cats16)),
        "recomputed_rows_fp32": int(outs16[0].n_flagged),
        "replayed_rows_fp64": int(outs16[0].n_replayed),
        "guard_tiers_ms": float(outs16[0].ms_recompute),
        "live_per_layer": [int(x) for x in outs16[0].live],
        "spmm_us_by_layer": [round(float(x), 1) for x in layer_us],
        "spmm_kernel_by_layer": kernel_runs(cd, outs16[0].layer_kernel),
        "max_abs_dy_fp16_vs_fp64": max(dy.values()) if dy else None,
        "max_abs_dy_by_layer": {str(k): v for k, v in dy.items()} if dy else None,
        "dy_samples": min(args.dy_samples, spec.S) if dy is not None else 0,
    }
    print(json.dumps(line), flush=True)


def bench_batch(cd, np, args, specs):
    """Several networks (c5): per-network plans from exhaustive tuning, then the aggregate
    edges/s of one concurrent run (main passes on side streams, one hipGraph) against the
    same networks run one after another, with identical category sets."""
    eng = cd.Engine(0)
    t0 = time.perf_counter()
    nets = []
    for sp in specs:
        net = cd.Net(eng, sp.N, sp.L, sp.S, sp.bias)
        net.generate_layers(sp.kind, sp.seed, K=sp.K, wval=sp.wval)
        net.generate_input(sp.seed)
        net.upload(cd.default_opts(max_tile_waste=64.0))
        nets.append(net)
    log(f"[bench] {args.config}: {len(nets)} networks uploaded in {time.perf_counter() - t0:.1f}s")
    tuning = None
    if args.tune:
        t0 = time.perf_counter()
        for net in nets:
            plan, _, _ = cd.tune_exhaustive(net, reps=1)
            net.set_plan(plan)
        tuning = {"method": "exhaustive per network", "tuning_s": time.perf_counter() - t0}
        log(f"[bench] tuned in {tuning['tuning_s']:.1f}s")
    edges = float(sum(sp.edges for sp in specs))
    # one after another: each network's own pass (replayed from its hipGraph, and launched
    # eagerly); the faster is the sequential time
    graphs = (1, 0) if args.graph else (0,)
    seq_best = None
    sequential = {}
    for g in graphs:
        seq_opts = cd.default_opts(precision=cd.FP16, use_graph=g, batch_concurrent=0)
        outs_q, wall_q = timed_runs_many(cd, nets, seq_opts, args.steps, args.warmup)
        sequential["graph" if g else "eager"] = wall_q * 1e3 / args.steps
        log(f"[bench] one after another, {'graph' if g else 'eager'}: {wall_q * 1e3 / args.steps:.2f} ms")
        if seq_best is None or wall_q < seq_best[1]:
            seq_best = (outs_q, wall_q, g)
    outs_s, wall_s, seq_graph = seq_best
    seq_opts = cd.default_opts(precision=cd.FP16, use_graph=seq_graph, batch_concurrent=0)
    best = None
    concurrent = {}
    settings = [(chains, share, graphs[0]) for chains in (2, 3) for share in (0, 1)]
    if args.graph:
        settings += [(2, 0, 0), (3, 0, 0), (4, 0, 0)]  # the chains launched eagerly (4: the engine's cap)
    for chains, share, g in settings:
        bat_opts = cd.default_opts(precision=cd.FP16, use_graph=g, batch_concurrent=chains, batch_cu_share=share)
        outs_c, wall_c = timed_runs_many(cd, nets, bat_opts, args.steps, args.warmup)
        key = f"{chains} chains, cu_share {share}, {'graph' if g else 'eager'}"
        concurrent[key] = wall_c * 1e3 / args.steps
        log(f"[bench] {key}: {wall_c * 1e3 / args.steps:.2f} ms")
        if best is None or wall_c < best[1]:
            best = (outs_c, wall_c, chains, share, g)
    outs_b, wall_b, chains, share, con_graph = best
    same = all(np.array_equal(a.categories, b.categories) for a, b in zip(outs_s[0], outs_b[0]))
    ms_s = wall_s * 1e3 / args.steps
    ms_b = wall_b * 1e3 / args.steps
    # each network alone (where the batch's time goes)
    per_net = []
    for sp, net in zip(specs, nets):
        _, wall_n = timed_runs_many(cd, [net], seq_opts, args.steps, args.warmup)
        per_net.append({"network": f"N={sp.N}, L={sp.L}, samples={sp.S}", "ms_per_step": wall_n * 1e3 / args.steps})
    log("[bench] networks alone: " + ", ".join(f"{x['network']}: {x['ms_per_step']:.2f} ms" for x in per_net))
    # the batch's throughput is that of the faster mode
    mode = "concurrent" if ms_b < ms_s else "sequential"
    ms_step = min(ms_b, ms_s)
    value = edges / (ms_step * 1e-3)
    log(f"[bench] sequential {ms_s:.2f} ms ({edges / (ms_s * 1e-3):.3e} edges/s), "
        f"concurrent {ms_b:.2f} ms ({edges / (ms_b * 1e-3):.3e} edges/s), categories equal: {same}; value: {mode}")
    line = {
        "metric": METRIC + " (aggregate over the batch)",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: " + ", ".join(f"(N={sp.N}, L={sp.L}, samples={sp.S})" for sp in specs)
                   + ", random 32 nnz/column of 0.0625", "networks": len(specs), "parallelism": "single GPU"},
        "batch_mode": mode,
        "sequential": {"ms_per_step": ms_s, "value": edges / (ms_s * 1e-3), "graph": bool(seq_graph)},
        "concurrent": {"ms_per_step": ms_b, "value": edges / (ms_b * 1e-3), "graph": bool(con_graph)},
        "networks_alone": per_net,
        "concurrent_chains": chains,
        "concurrent_cu_share": share,
        "sequential_ms_by_setting": sequential,
        "concurrent_ms_by_setting": concurrent,
        "concurrent_over_sequential": ms_s / ms_b,
        "categories_equal": same,
        "tuning": tuning,
        "categories": [int(len(o.categories)) for o in outs_b[0]],
    }
    print(json.dumps(line), flush=True)


def timed_runs_many(cd, nets, opts, steps, warmup):
    for _ in range(warmup):
        cd.run(nets, opts)
    outs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        outs.append(cd.run(nets, opts))
    return outs, time.perf_counter() - t0


if __name__ == "__main__":
    main()
