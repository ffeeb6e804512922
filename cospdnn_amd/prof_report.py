"""Markdown summaries of rocprofv3 result databases (kernel trace and HBM counters).

    python -m cospdnn_amd.prof_report trace <results.db> <title> <command>
    python -m cospdnn_amd.prof_report traffic <kernel-substring> <alg-bytes-per-launch>
        <fetch.db> <write.db> <copy-fetch.db> <copy-write.db> <title> <commands>

The traffic summary scales each counter's per-dispatch reading by what the same counter
reads for hbm_copy.out (512 MiB read and 512 MiB written per dispatch) and prints one
machine-readable line, `traffic_bytes_per_launch <kernel> <bytes>`, that bench.py reads.
"""
import sqlite3
import sys

COPY_KIB = 512 * 1024  # hbm_copy.out moves 512 MiB each way per dispatch


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


def main(argv):
    if len(argv) >= 4 and argv[0] == "trace":
        print(trace_table(argv[1], argv[2], argv[3]))
    elif len(argv) >= 9 and argv[0] == "traffic":
        print(traffic_table(argv[1], float(argv[2]), *argv[3:7], argv[7], argv[8]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
