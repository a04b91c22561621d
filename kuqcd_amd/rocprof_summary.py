"""Short text summaries of rocprofv3 CSV output (kernel statistics and HBM counters).

    python -m kuqcd_amd.rocprof_summary stats <run>_kernel_stats.csv [--top N] [--filter S] [--command "..."]
    python -m kuqcd_amd.rocprof_summary phases <run>_kernel_trace.csv [--gap-ms G] [--min-ms M] [--command "..."]
    python -m kuqcd_amd.rocprof_summary traffic --fetch <fetch_run>_counter_collection.csv \
        --write <write_run>_counter_collection.csv --kernel SUBSTRING [--copy-bytes B] [--command "..."]

`phases` splits a kernel trace at idle gaps longer than G ms and, for every phase
longer than M ms (e.g. one solve), lists its kernels by total time.
`traffic` reports FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the kernel whose
name contains SUBSTRING, each scaled by the factor that makes the same counter of
the plain copy kernel (k_copy16, which reads and writes copy-bytes per launch)
equal to its known size, and prints `traffic_bytes_per_launch: <bytes>`.
"""
from __future__ import annotations

import argparse
import csv
import statistics
import sys
from pathlib import Path


def _col(header, *names):
    low = [h.strip().lower() for h in header]
    for n in names:
        if n.lower() in low:
            return low.index(n.lower())
    raise KeyError(f"none of {names} in {header}")


def _rows(path):
    with open(path, newline="") as f:
        r = csv.reader(f)
        header = next(r)
        return header, list(r)


def _short(name):
    # drop the namespaces first: "(anonymous namespace)::" would otherwise end the name at its "("
    n = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("kq::", "")
    n = n.split("(")[0]
    return n if len(n) <= 70 else n[:67] + "..."


def stats(path, top=15, filt=""):
    header, rows = _rows(path)
    i_name = _col(header, "Name", "KernelName", "Kernel_Name")
    i_calls = _col(header, "Calls")
    i_tot = _col(header, "TotalDurationNs", "Total_Duration_Ns")
    i_avg = _col(header, "AverageNs", "Average_Ns")
    i_pct = _col(header, "Percentage")
    out = [f"{'kernel':<90} {'calls':>7} {'avg us':>10} {'total ms':>10} {'%':>6}"]
    if filt:
        rows = [r for r in rows if filt in r[i_name]]
    rows.sort(key=lambda r: -float(r[i_tot]))
    for r in rows[:top]:
        name = r[i_name]
        name = name if len(name) <= 90 else name[:87] + "..."
        out.append(f"{name:<90} {int(r[i_calls]):>7} {float(r[i_avg]) / 1e3:>10.2f} "
                   f"{float(r[i_tot]) / 1e6:>10.3f} {float(r[i_pct]):>6.2f}")
    return "\n".join(out)


def phases(path, gap_ms=0.3, min_ms=20.0, top=20):
    header, rows = _rows(path)
    i_name = _col(header, "Kernel_Name", "KernelName", "Name")
    i_s = _col(header, "Start_Timestamp")
    i_e = _col(header, "End_Timestamp")
    rows = sorted(rows, key=lambda r: int(r[i_s]))
    if not rows:
        return "(empty trace)"
    groups, cur = [], [rows[0]]
    for a, b in zip(rows, rows[1:]):
        if int(b[i_s]) - int(a[i_e]) > gap_ms * 1e6:
            groups.append(cur)
            cur = []
        cur.append(b)
    groups.append(cur)
    t0 = int(rows[0][i_s])
    out = []
    for g in groups:
        s0, e0 = int(g[0][i_s]), int(g[-1][i_e])
        if (e0 - s0) / 1e6 < min_ms:
            continue
        busy = sum(int(r[i_e]) - int(r[i_s]) for r in g)
        out.append(f"phase at {(s0 - t0) / 1e6:.1f} ms: wall {(e0 - s0) / 1e6:.2f} ms, kernels busy {busy / 1e6:.2f} ms, "
                   f"{len(g)} dispatches")
        agg = {}
        for r in g:
            k = _short(r[i_name])
            n, t = agg.get(k, (0, 0))
            agg[k] = (n + 1, t + int(r[i_e]) - int(r[i_s]))
        for k, (n, t) in sorted(agg.items(), key=lambda x: -x[1][1])[:top]:
            out.append(f"    {k:<70} {n:>6} {t / 1e6:>9.2f} ms {t / n / 1e3:>9.1f} us")
    return "\n".join(out)


def _per_kernel(path, counter, substr):
    header, rows = _rows(path)
    i_name = _col(header, "Kernel_Name", "KernelName", "Name")
    i_cn = _col(header, "Counter_Name", "CounterName")
    i_cv = _col(header, "Counter_Value", "CounterValue")
    i_d = _col(header, "Dispatch_Id", "DispatchId", "Dispatch_ID")
    per = {}
    for r in rows:
        if r[i_cn].strip() != counter or substr not in r[i_name]:
            continue
        per[r[i_d]] = per.get(r[i_d], 0.0) + float(r[i_cv])  # sums per-instance rows of one dispatch
    if not per:
        raise ValueError(f"no {counter} rows for kernels containing {substr!r} in {path}")
    vals = list(per.values())
    return statistics.median(vals), len(vals)


def traffic(fetch_csv, write_csv, kernel, copy_bytes=1 << 30, copy_kernel="k_copy16"):
    f_k, n_fk = _per_kernel(fetch_csv, "FETCH_SIZE", kernel)
    w_k, n_wk = _per_kernel(write_csv, "WRITE_SIZE", kernel)
    f_c, _ = _per_kernel(fetch_csv, "FETCH_SIZE", copy_kernel)
    w_c, _ = _per_kernel(write_csv, "WRITE_SIZE", copy_kernel)
    sf = copy_bytes / (f_c * 1024.0)
    sw = copy_bytes / (w_c * 1024.0)
    fb, wb = f_k * 1024.0 * sf, w_k * 1024.0 * sw
    lines = [
        f"kernel: {kernel} ({n_fk} FETCH_SIZE dispatches, {n_wk} WRITE_SIZE dispatches; median per dispatch)",
        f"FETCH_SIZE: {f_k:.1f} KiB/dispatch raw, copy-kernel factor {sf:.4f} -> {fb / 1e6:.2f} MB",
        f"WRITE_SIZE: {w_k:.1f} KiB/dispatch raw, copy-kernel factor {sw:.4f} -> {wb / 1e6:.2f} MB",
        f"copy kernel {copy_kernel}: FETCH_SIZE {f_c:.1f} KiB, WRITE_SIZE {w_c:.1f} KiB for {copy_bytes} B each way",
        f"traffic_bytes_per_launch: {fb + wb:.6e}",
    ]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=15)
    s.add_argument("--filter", default="")
    s.add_argument("--command", default="")
    ph = sub.add_parser("phases")
    ph.add_argument("csv")
    ph.add_argument("--gap-ms", type=float, default=0.3)
    ph.add_argument("--min-ms", type=float, default=20.0)
    ph.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    t.add_argument("--command", default="")
    a = ap.parse_args(argv)
    head = f"command: {a.command}\n" if a.command else ""
    if a.cmd == "stats":
        print(head + stats(Path(a.csv), a.top, a.filter))
    elif a.cmd == "phases":
        print(head + phases(Path(a.csv), a.gap_ms, a.min_ms))
    else:
        print(head + traffic(Path(a.fetch), Path(a.write), a.kernel, a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
