"""Markdown summaries of rocprofv3 result databases (kernel trace and HBM counters).

    python -m cospdnn_amd.prof_report trace <results.db> <title> <command>
    python -m cospdnn_amd.prof_report traffic <kernel-substring> <alg-bytes-per-launch>
        <fetch.db> <write.db> <copy-fetch.db> <copy-write.db> <title> <commands>
    python -m cospdnn_amd.prof_report mfma <kernel-substring> <alg-flops-per-launch>
        <counter_collection.csv> <kernel_trace.csv> <simd-count> <title> <commands>
    python -m cospdnn_amd.prof_report layers <bench.json> <copy-fetch.csv> <copy-write.csv>
        <title> <commands> <counter_collection.csv>...

The layers summary maps the FP16 main-pass SpMM dispatches of the last forward in each
counter file (a bench.py run with --graph 0) to layers 0..L-1, checks their kernel names
against the bench line's spmm_kernel_by_layer, and tabulates the counters per layer group
beside the bench line's per-layer times and algorithmic bytes.

The traffic summary scales each counter's per-dispatch reading by what the same counter
reads for hbm_copy.out (512 MiB read and 512 MiB written per dispatch) and prints one
machine-readable line, `traffic_bytes_per_launch <kernel> <bytes>`, that bench.py reads.
"""
import sqlite3
import sys

COPY_KIB = 512 * 1024  # hbm_copy.out moves 512 MiB each way per dispatch
XCDS = 8  # MI355X: GRBM_GUI_ACTIVE reads one count per XCD, summed in the CSV


def trace_table(db, title, cmd):
    c = sqlite3.connect(db)
    rows = list(c.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration), max(grid_x), "
        "max(workgroup_x), max(vgpr_count), max(accum_vgpr_count), max(lds_size) from kernels group by name "
        "order by sum(duration) desc"))
    tot = sum(r[2] for r in rows) or 1
    out = [f"# {title}", "", "Command (from the repo root, on an MI355X):", "", "```", cmd, "```", "",
           "| kernel | calls | total ms | avg us | min us | max us | % | grid | wg | vgpr | agpr | lds B |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for n, k, s, a, mn, mx, gx, wx, vg, ag, lds in rows:
        n = n.replace("|", "/")
        if len(n) > 70:
            n = n[:70] + "…"
        out.append(f"| `{n}` | {k} | {s / 1e6:.3f} | {a / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                   f"{100 * s / tot:.1f} | {gx} | {wx} | {vg} | {ag} | {lds} |")
    return "\n".join(out)


def counter_avg(db, kernel):
    """Mean per-dispatch counter value (KiB) over the dispatches whose name contains `kernel`."""
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, counter_value from pmc_events"))
    vals = [v for n, v in rows if kernel in n]
    if not vals:
        raise SystemExit(f"no dispatch of {kernel} in {db}")
    return sum(vals) / len(vals), len(vals)


def traffic_table(kernel, alg_bytes, fetch_db, write_db, cfetch_db, cwrite_db, title, cmds):
    f_kib, nf = counter_avg(fetch_db, kernel)
    w_kib, nw = counter_avg(write_db, kernel)
    cf, _ = counter_avg(cfetch_db, "hbm_copy_kernel")
    cw, _ = counter_avg(cwrite_db, "hbm_copy_kernel")
    ff, fw = COPY_KIB / cf, COPY_KIB / cw
    fetch_b = f_kib * ff * 1024
    write_b = w_kib * fw * 1024
    total = fetch_b + write_b
    out = [f"# {title}", "", "Commands (from the repo root, on an MI355X):", "", "```", cmds, "```", "",
           "| quantity | value |", "|---|---|",
           f"| FETCH_SIZE of hbm_copy_kernel (512 MiB read) | {cf:.0f} KiB -> factor {ff:.3f} |",
           f"| WRITE_SIZE of hbm_copy_kernel (512 MiB written) | {cw:.0f} KiB -> factor {fw:.3f} |",
           f"| `{kernel}` dispatches | {nf} (fetch run), {nw} (write run) |",
           f"| FETCH_SIZE per dispatch (raw / scaled) | {f_kib:.0f} KiB / {fetch_b / 1e6:.1f} MB |",
           f"| WRITE_SIZE per dispatch (raw / scaled) | {w_kib:.0f} KiB / {write_b / 1e6:.1f} MB |",
           f"| HBM traffic per dispatch | {total / 1e6:.1f} MB |",
           f"| algorithmic bytes per dispatch | {alg_bytes / 1e6:.1f} MB |",
           f"| traffic / algorithmic | {total / alg_bytes:.3f} |", "",
           f"traffic_bytes_per_launch {kernel} {total:.0f}"]
    return "\n".join(out)


def _csv_rows(path):
    """Rows of a rocprofv3 CSV file, keys lower-cased."""
    import csv
    with open(path, newline="") as f:
        return [{k.strip().lower(): v for k, v in r.items()} for r in csv.DictReader(f)]


def mfma_table(kernel, flops_per_launch, counter_csv, trace_csv, simd_num, title, cmd):
    """MFMA utilisation of the dispatches whose name contains `kernel`.

    counter_csv: rocprofv3 --output-format csv counter file of a run with
    --pmc MfmaUtil SQ_VALU_MFMA_BUSY_CYCLES SQ_INSTS_VALU_MFMA_MOPS_F16 GRBM_GUI_ACTIVE;
    trace_csv: the kernel trace of the same run (dispatch durations).
    MfmaUtil is rocprofv3's derived counter, SQ_VALU_MFMA_BUSY_CYCLES summed over SIMDs
    / (GRBM_GUI_ACTIVE * SIMD_NUM); MOPS_F16 counts MFMA F16 operations in units of 512.
    """
    per = {}
    for r in _csv_rows(counter_csv):
        if kernel not in r.get("kernel_name", ""):
            continue
        d = per.setdefault(r["dispatch_id"], {})
        d[r["counter_name"]] = d.get(r["counter_name"], 0.0) + float(r["counter_value"])
    if not per:
        raise SystemExit(f"no dispatch of {kernel} in {counter_csv}")
    dur = {}
    for r in _csv_rows(trace_csv):
        if kernel in r.get("kernel_name", ""):
            dur[r["dispatch_id"]] = (float(r["end_timestamp"]) - float(r["start_timestamp"])) * 1e-9
    n = len(per)
    mean = lambda name: sum(d.get(name, 0.0) for d in per.values()) / n  # noqa: E731
    util = mean("MfmaUtil")
    busy = mean("SQ_VALU_MFMA_BUSY_CYCLES")
    gui = mean("GRBM_GUI_ACTIVE")
    mops = mean("SQ_INSTS_VALU_MFMA_MOPS_F16")
    flops = mops * 512.0
    t = sum(dur.get(k, 0.0) for k in per) / max(1, sum(1 for k in per if k in dur))
    out = [f"# {title}", "", "Commands (from the repo root, on an MI355X):", "", "```", cmd, "```", "",
           "| quantity | value |", "|---|---|",
           f"| `{kernel}` dispatches | {n} |",
           f"| mean duration under counter collection | {t * 1e6:.1f} us |",
           f"| MfmaUtil (rocprofv3 derived: MFMA busy cycles / (GRBM_GUI_ACTIVE x {simd_num} SIMDs)) | {util:.2f} % |",
           f"| SQ_VALU_MFMA_BUSY_CYCLES per dispatch (all SIMDs) | {busy:.4g} |",
           f"| GRBM_GUI_ACTIVE per dispatch (summed over the {XCDS} XCDs) | {gui:.4g} "
           f"({gui / XCDS / max(1e-12, t) / 1e9:.2f} GHz per XCD) |",
           f"| busy / (GRBM_GUI_ACTIVE / {XCDS} x {simd_num}) | {100 * busy / max(1.0, gui / XCDS * simd_num):.2f} % |",
           f"| MFMA F16 flops per dispatch (MOPS_F16 x 512) | {flops:.4g} |",
           f"| algorithmic flops per dispatch (2 per live edge) | {flops_per_launch:.4g} |",
           f"| MFMA flops / algorithmic flops (tile padding) | {flops / max(1.0, flops_per_launch):.3f} |",
           f"| MFMA F16 rate | {flops / max(1e-12, t) / 1e12:.1f} TFLOP/s "
           f"({100 * flops / max(1e-12, t) / 2.5e15:.2f} % of 2.5 PFLOP/s dense) |",
           "", f"mfma_util_percent {kernel} {util:.3f}"]
    return "\n".join(out)


FP16_SPMM = ("spmm_mfma_kernel", "spmm_mfma_mask_kernel", "spmm_mfma_lds_kernel", "spmm_mfma_lds_mask_kernel",
             "spmm_mfma_band_kernel", "spmm_mfma_band_mask_kernel", "spmm_wave_rows_kernel")


def fp16_spmm_name(name):
    """The engine's kernel name of an FP16 main-pass SpMM dispatch, or None (guard tiers run
    the FP32 and FP64 instances, which never match).  rocprofv3 reports some names demangled
    (`cospdnn::spmm_mfma_kernel(...)`, `void cospdnn::spmm_simt_kernel<_Float16, _Float16>(...)`)
    and some mangled (`_ZN7cospdnn21spmm_simt_rows_kernelIDF16_DF16_EEv...`)."""
    n = name.replace("void ", "")
    base = n.split("(")[0].split("::")[-1]
    # the sample-slab kernels: spmm_slab[_uni|_band|_uni_wide]_kernel_<T>, and
    # spmm_bits[_wide]_kernel_<T>
    if base.startswith("spmm_bits_") or "cospdnn" in n and "spmm_bits_" in n:
        return "spmm_bits_kernel"
    if base.startswith("spmm_slab_") or "cospdnn" in n and "spmm_slab_" in n:
        return "spmm_slab_kernel"
    for k in FP16_SPMM:
        if base == k or f"{len(k)}{k}E" in n:
            return k
    for k in ("spmm_simt_kernel", "spmm_simt_rows_kernel"):
        if base == k + "<_Float16, _Float16>" or f"{len(k)}{k}IDF16_DF16_E" in n:
            return k
    return None


def per_layer_counters(csv_path, L):
    """{counter: [per-layer value]} and the kernel names of the last L FP16 SpMM dispatches."""
    per = {}
    names = {}
    for r in _csv_rows(csv_path):
        k = fp16_spmm_name(r.get("kernel_name", ""))
        if k is None:
            continue
        d = int(r["dispatch_id"])
        names[d] = k
        c = per.setdefault(d, {})
        c[r["counter_name"]] = c.get(r["counter_name"], 0.0) + float(r["counter_value"])
    ids = sorted(per)[-L:]
    if len(ids) < L:
        raise SystemExit(f"{csv_path}: {len(ids)} FP16 SpMM dispatches, expected at least {L}")
    counters = sorted({c for d in ids for c in per[d]})
    return {c: [per[d].get(c, 0.0) for d in ids] for c in counters}, [names[d] for d in ids]


def copy_factor(csv_path, counter):
    vals = [float(r["counter_value"]) for r in _csv_rows(csv_path)
            if "hbm_copy_kernel" in r.get("kernel_name", "") and r["counter_name"] == counter]
    if not vals:
        raise SystemExit(f"no {counter} of hbm_copy_kernel in {csv_path}")
    return COPY_KIB / (sum(vals) / len(vals))


def layer_groups(L):
    """Layers 0..11 one by one, then [12, 20), [20, 40), [40, 80), ... doubling."""
    groups = [(l, l + 1) for l in range(min(L, 12))]
    lo, hi = 12, 20
    while lo < L:
        groups.append((lo, min(hi, L)))
        lo, hi = hi, hi * 2
    return groups


def layers_table(bench_json, copy_fetch, copy_write, title, cmds, counter_csvs):
    import json

    with open(bench_json) as f:
        line = json.loads([x for x in f.read().splitlines() if x.startswith("{")][-1])
    L = line["config"]["layers"]
    N = line["config"]["neurons"]
    K = line["config"]["nnz_per_column"]
    live = line["live_per_layer"]
    us = line["spmm_us_by_layer"]
    kname = [None] * L
    for name, a, b in line["spmm_kernel_by_layer"]:
        for l in range(a, b + 1):
            kname[l] = name
    alg = [2.0 * live[l] * N * 2 + N * K * 2.0 for l in range(L)]
    cnt = {}
    for path in counter_csvs:
        c, names = per_layer_counters(path, L)
        bad = [l for l in range(L) if names[l] != kname[l]]
        if bad:
            raise SystemExit(f"{path}: layer {bad[0]} ran {names[bad[0]]}, the bench line says {kname[bad[0]]}")
        cnt.update(c)
    ff = copy_factor(copy_fetch, "FETCH_SIZE")
    fw = copy_factor(copy_write, "WRITE_SIZE")
    fetch = [v * ff * 1024 for v in cnt["FETCH_SIZE"]] if "FETCH_SIZE" in cnt else None
    write = [v * fw * 1024 for v in cnt["WRITE_SIZE"]] if "WRITE_SIZE" in cnt else None
    hit = cnt.get("TCC_HIT_sum")
    miss = cnt.get("TCC_MISS_sum")
    wait = cnt.get("SQ_WAIT_INST_ANY")
    wcyc = cnt.get("SQ_WAVE_CYCLES")
    out = [f"# {title}", "", "Commands (from the repo root, on an MI355X):", "", "```", cmds, "```", "",
           f"FETCH_SIZE scaled by {ff:.3f} and WRITE_SIZE by {fw:.3f} (what each reads for hbm_copy.out, "
           "512 MiB read and 512 MiB written per dispatch). Algorithmic bytes of layer l: 2 x live_l x N x 2 B "
           "(FP16 activations read and written) + 32 N x 2 B (weight values). SpMM us: the bench line's "
           "per-layer HIP events (no counters).", "",
           "| layers | live rows entering | kernel | SpMM us per layer | alg. GB/s | HBM MB per layer (fetch + write) "
           "| traffic / alg. | L2 hit % | wave wait % |",
           "|---|---|---|---|---|---|---|---|---|"]
    for lo, hi in layer_groups(L):
        r = range(lo, hi)
        n = hi - lo
        t = sum(us[l] for l in r) / n
        a = sum(alg[l] for l in r) / n
        ks = sorted({kname[l] for l in r})
        tr = (sum(fetch[l] + write[l] for l in r) / n) if fetch and write else None
        h = (100 * sum(hit[l] for l in r) / max(1.0, sum(hit[l] + miss[l] for l in r))) if hit and miss else None
        w = (100 * sum(wait[l] for l in r) / max(1.0, sum(wcyc[l] for l in r))) if wait and wcyc else None
        lv = f"{live[lo]}" if n == 1 else f"{live[lo]} .. {live[hi - 1]}"
        out.append(f"| {lo if n == 1 else f'{lo}-{hi - 1}'} | {lv} | {', '.join(f'`{k}`' for k in ks)} | {t:.1f} | "
                   f"{a / (t * 1e-6) / 1e9:.0f} | " + (f"{tr / 1e6:.1f} | {tr / a:.2f}" if tr else "- | -") + " | "
                   + (f"{h:.0f}" if h is not None else "-") + " | " + (f"{w:.0f}" if w is not None else "-") + " |")
    out.append("")
    tot_us = sum(us)
    out.append(f"SpMM over the forward: {tot_us / 1e3:.2f} ms, {sum(alg) / (tot_us * 1e-6) / 1e9:.0f} GB/s of "
               f"algorithmic bytes ({100 * sum(alg) / (tot_us * 1e-6) / 8e12:.1f} % of 8 TB/s)"
               + (f", HBM traffic / algorithmic {sum(fetch[l] + write[l] for l in range(L)) / sum(alg):.2f}"
                  if fetch and write else "") + ".")
    out.append("")
    for k in sorted(set(kname)):
        ls = [l for l in range(L) if kname[l] == k]
        t = sum(us[l] for l in ls)
        a = sum(alg[l] for l in ls)
        out.append(f"- `{k}`: {len(ls)} layers, {t / 1e3:.2f} ms, {a / (t * 1e-6) / 1e9:.0f} GB/s algorithmic"
                   + (f", traffic / algorithmic {sum(fetch[l] + write[l] for l in ls) / a:.2f}"
                      if fetch and write else ""))
    if fetch and write:
        out.append("")
        for k in sorted(set(kname)):
            ls = [l for l in range(L) if kname[l] == k]
            out.append(f"traffic_bytes_per_launch {k} {sum(fetch[l] + write[l] for l in ls) / len(ls):.0f}")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 4 and argv[0] == "trace":
        print(trace_table(argv[1], argv[2], argv[3]))
    elif len(argv) >= 9 and argv[0] == "traffic":
        print(traffic_table(argv[1], float(argv[2]), *argv[3:7], argv[7], argv[8]))
    elif len(argv) >= 7 and argv[0] == "layers":
        print(layers_table(argv[1], argv[2], argv[3], argv[4], argv[5], argv[6:]))
    elif len(argv) >= 8 and argv[0] == "mfma":
        print(mfma_table(argv[1], float(argv[2]), argv[3], argv[4], int(argv[5]), argv[6], argv[7]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
