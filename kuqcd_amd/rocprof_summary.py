"""Short text summaries of rocprofv3 CSV output (kernel statistics and HBM counters).

    python -m kuqcd_amd.rocprof_summary stats <run>_kernel_stats.csv [--top N] [--command "..."]
    python -m kuqcd_amd.rocprof_summary traffic --fetch <fetch_run>_counter_collection.csv \
        --write <write_run>_counter_collection.csv --kernel SUBSTRING [--copy-bytes B] [--command "..."]

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


def stats(path, top=15):
    header, rows = _rows(path)
    i_name = _col(header, "Name", "KernelName", "Kernel_Name")
    i_calls = _col(header, "Calls")
    i_tot = _col(header, "TotalDurationNs", "Total_Duration_Ns")
    i_avg = _col(header, "AverageNs", "Average_Ns")
    i_pct = _col(header, "Percentage")
    out = [f"{'kernel':<90} {'calls':>7} {'avg us':>10} {'total ms':>10} {'%':>6}"]
    rows.sort(key=lambda r: -float(r[i_tot]))
    for r in rows[:top]:
        name = r[i_name]
        name = name if len(name) <= 90 else name[:87] + "..."
        out.append(f"{name:<90} {int(r[i_calls]):>7} {float(r[i_avg]) / 1e3:>10.2f} "
                   f"{float(r[i_tot]) / 1e6:>10.3f} {float(r[i_pct]):>6.2f}")
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
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    t.add_argument("--command", default="")
    a = ap.parse_args(argv)
    head = f"command: {a.command}\n" if a.command else ""
    if a.cmd == "stats":
        print(head + stats(Path(a.csv), a.top))
    else:
        print(head + traffic(Path(a.fetch), Path(a.write), a.kernel, a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
