"""Summaries of rocprofv3 CSV output: per-kernel durations and HBM traffic.

``traffic`` turns two counter runs of the same command (``--pmc FETCH_SIZE``,
then ``--pmc WRITE_SIZE``; both in KiB per dispatch) into bytes per launch of
one kernel.  Each counter is scaled by what it reads for a plain copy of known
size (``bench.py --copy-calibration MIB`` copies MIB MiB three times), so the
result is in real HBM bytes whatever unit the counter counts in.

    python -m ttnumpy_amd.profiling stats  --csv DIR/x_kernel_stats.csv
    python -m ttnumpy_amd.profiling traffic --kernel eng_solve_kernel \\
        --fetch F.csv --write W.csv --cal-fetch CF.csv --cal-write CW.csv --cal-mib 1024
"""

from __future__ import annotations

import argparse
import csv
import json
from collections import defaultdict


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "")


def kernel_stats(path: str):
    """Rows of a ``--stats`` kernel summary: name, calls, total/avg/min/max ns, share."""
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append({"kernel": _short(r["Name"]), "calls": int(r["Calls"]),
                     "total_ms": float(r["TotalDurationNs"]) / 1e6,
                     "avg_us": float(r["AverageNs"]) / 1e3, "min_us": float(r["MinNs"]) / 1e3,
                     "max_us": float(r["MaxNs"]) / 1e3, "percent": float(r["Percentage"])})
    return rows


def counter_per_dispatch(path: str):
    """Mean counter value per dispatch for every kernel of a ``--pmc`` CSV."""
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        acc[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def traffic(kernel: str, fetch_csv: str, write_csv: str, cal_fetch_csv: str, cal_write_csv: str,
            cal_mib: int, copy_kernel: str = "ttx::copy_nd_kernel"):
    """HBM bytes per launch of ``kernel`` from FETCH_SIZE and WRITE_SIZE runs."""
    cal_kib = cal_mib * 1024.0
    cf, _ = counter_per_dispatch(cal_fetch_csv)[copy_kernel]
    cw, _ = counter_per_dispatch(cal_write_csv)[copy_kernel]
    f_scale, w_scale = cal_kib / cf, cal_kib / cw
    match = lambda d: next(v for k, v in d.items() if k.endswith(kernel))  # noqa: E731
    fk, nf = match(counter_per_dispatch(fetch_csv))
    wk, nw = match(counter_per_dispatch(write_csv))
    read_b, write_b = fk * f_scale * 1024.0, wk * w_scale * 1024.0
    return {"kernel": kernel, "dispatches": [nf, nw],
            "fetch_kib_per_dispatch": fk, "write_kib_per_dispatch": wk,
            "copy_calibration_mib": cal_mib, "fetch_scale": f_scale, "write_scale": w_scale,
            "read_bytes": read_b, "write_bytes": write_b, "bytes_per_launch": read_b + write_b}


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--csv", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    for k in ("fetch", "write", "cal-fetch", "cal-write"):
        t.add_argument(f"--{k}", required=True)
    t.add_argument("--cal-mib", type=int, required=True)
    a = ap.parse_args()
    if a.cmd == "stats":
        print("| kernel | calls | total ms | avg us | min us | max us | % |")
        print("|---|---|---|---|---|---|---|")
        for r in kernel_stats(a.csv):
            print(f"| {r['kernel']} | {r['calls']} | {r['total_ms']:.1f} | {r['avg_us']:.1f} | "
                  f"{r['min_us']:.1f} | {r['max_us']:.1f} | {r['percent']:.2f} |")
    else:
        print(json.dumps(traffic(a.kernel, a.fetch, a.write, a.cal_fetch, a.cal_write, a.cal_mib),
                         indent=1))


if __name__ == "__main__":
    main()
