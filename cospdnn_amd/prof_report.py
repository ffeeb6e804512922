"""Markdown summaries of rocprofv3 result databases (kernel trace and HBM counters).

    python -m cospdnn_amd.prof_report trace <results.db> <title> <command>
    python -m cospdnn_amd.prof_report traffic <kernel-substring> <alg-bytes-per-launch>
        <fetch.db> <write.db> <copy-fetch.db> <copy-write.db> <title> <commands>
    python -m cospdnn_amd.prof_report mfma <kernel-substring> <alg-flops-per-launch>
        <counter_collection.csv> <kernel_trace.csv> <simd-count> <title> <commands>

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
    out = [f"# {title}", "", "Command (from the repo root on the GPU box):", "", "```", cmd, "```", "",
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
    out = [f"# {title}", "", "Commands (from the repo root on the GPU box):", "", "```", cmds, "```", "",
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
    out = [f"# {title}", "", "Commands (from the repo root on the GPU box):", "", "```", cmd, "```", "",
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


def main(argv):
    if len(argv) >= 4 and argv[0] == "trace":
        print(trace_table(argv[1], argv[2], argv[3]))
    elif len(argv) >= 9 and argv[0] == "traffic":
        print(traffic_table(argv[1], float(argv[2]), *argv[3:7], argv[7], argv[8]))
    elif len(argv) >= 8 and argv[0] == "mfma":
        print(mfma_table(argv[1], float(argv[2]), argv[3], argv[4], int(argv[5]), argv[6], argv[7]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
