"""Summaries of rocprofv3 CSV output (-f csv): per-kernel launch statistics from a
--kernel-trace --stats run, and per-kernel mean counter values per dispatch from a --pmc run.

    python -m godunov_five_eq_amd.rocprof_summary stats DIR           # markdown table
    python -m godunov_five_eq_amd.rocprof_summary pmc DIR COUNTER     # json {kernel: mean}
"""
import csv
import glob
import json
import os
import sys


def _rows(directory, suffix):
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def kernel_stats(directory):
    """[(name, calls, total_ns, avg_ns, percent)] sorted by total time."""
    out = []
    for r in _rows(directory, "kernel_stats.csv"):
        out.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                    float(r["Percentage"])))
    return sorted(out, key=lambda x: -x[2])


def counter_means(directory, counter):
    """{kernel name: (dispatches, mean counter value per dispatch)}"""
    acc = {}
    for r in _rows(directory, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        name = r["Kernel_Name"]
        n, s = acc.get(name, (0, 0.0))
        acc[name] = (n + 1, s + float(r["Counter_Value"]))
    return {k: (n, s / n) for k, (n, s) in acc.items()}


def main(argv):
    if argv[0] == "stats":
        print("| kernel | calls | total ms | average us | % |")
        print("|---|---:|---:|---:|---:|")
        for name, calls, tot, avg, pct in kernel_stats(argv[1])[:12]:
            print(f"| `{name[:110]}` | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.1f} | {pct:.1f} |")
    else:
        print(json.dumps(counter_means(argv[1], argv[2]), indent=1))


if __name__ == "__main__":
    main(sys.argv[1:])
