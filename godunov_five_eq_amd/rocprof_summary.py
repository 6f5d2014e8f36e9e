"""Summaries of rocprofv3 CSV output (-f csv): per-kernel launch statistics from a
--kernel-trace --stats run, and per-kernel mean counter values per dispatch from a --pmc run.

    python -m godunov_five_eq_amd.rocprof_summary stats DIR           # markdown table
    python -m godunov_five_eq_amd.rocprof_summary pmc DIR COUNTER     # json {kernel: mean}
    python -m godunov_five_eq_amd.rocprof_summary traffic KERNEL CELLS NVAR FETCH WRITE CAL_FETCH \
        CAL_WRITE [REQ CAL_REQ]                                        # json: HBM bytes per launch

--last N (stats, pmc, traffic): the last N dispatches of each kernel only -- the steps that run
after the launch-geometry autotuner's trial launches of the same kernel, which use other
geometries (stats then reads the kernel trace of the same --kernel-trace run).

traffic: FETCH_SIZE and WRITE_SIZE (KiB per dispatch, one --pmc run each) of the kernel whose name
contains KERNEL, each scaled by what the same counter reads for a plain device copy of known size
(godunov_five_eq_amd.calibrate: 1 GiB read and written per launch), against the algorithmic
2 * NVAR * 8 B per cell.  With REQ / CAL_REQ (a --pmc run of TCC_EA0_RDREQ_32B, TCC_EA0_RDREQ_64B
and TCC_EA0_RDREQ_128B over the kernel and over the copy), the read bytes are cross-checked by the
memory read requests counted by size, scaled the same way by the copy.
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
    """[(name, calls, total_ns, avg_ns, percent)] sorted by total time; with --last N, from the
    kernel trace: the last N dispatches of each kernel only"""
    out = []
    if LAST:
        acc = {}
        for r in sorted(_rows(directory, "kernel_trace.csv"), key=lambda r: int(r["Dispatch_Id"])):
            acc.setdefault(r["Kernel_Name"], []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        tot = {k: sum(v[-LAST:]) for k, v in acc.items()}
        whole = sum(tot.values())
        for k, v in acc.items():
            v = v[-LAST:]
            out.append((k, len(v), float(tot[k]), tot[k] / len(v), 100.0 * tot[k] / whole))
        return sorted(out, key=lambda x: -x[2])
    for r in _rows(directory, "kernel_stats.csv"):
        out.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                    float(r["Percentage"])))
    return sorted(out, key=lambda x: -x[2])


LAST = None  # --last N: only the last N dispatches of each kernel


def counter_means(directory, counter):
    """{kernel name: (dispatches, mean counter value per dispatch)}, in dispatch order per kernel"""
    acc = {}
    for r in sorted(_rows(directory, "counter_collection.csv"), key=lambda r: int(r.get("Dispatch_Id") or 0)):
        if r.get("Counter_Name") != counter:
            continue
        acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    out = {}
    for k, v in acc.items():
        v = v[-LAST:] if LAST else v
        out[k] = (len(v), sum(v) / len(v))
    return out


def _one(directory, counter, kernel):
    hits = [(k, v) for k, v in counter_means(directory, counter).items() if kernel in k]
    if not hits:
        raise SystemExit(f"no dispatch of a kernel containing {kernel!r} for {counter} in {directory}")
    return hits[0]


def _req_bytes(directory, kernel):
    sizes = {"TCC_EA0_RDREQ_32B": 32, "TCC_EA0_RDREQ_64B": 64, "TCC_EA0_RDREQ_128B": 128}
    out = {}
    for c, b in sizes.items():
        _, (n, v) = _one(directory, c, kernel)
        out[c] = v
    return out, sum(out[c] * b for c, b in sizes.items())


def traffic(kernel, cells, nvar, fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, req_dir=None, cal_req_dir=None,
            copy_kernel="copyBuffer", copy_bytes=1 << 30):
    name, (nf, fetch) = _one(fetch_dir, "FETCH_SIZE", kernel)
    _, (nw, write) = _one(write_dir, "WRITE_SIZE", kernel)
    _, (_, cfetch) = _one(cal_fetch_dir, "FETCH_SIZE", copy_kernel)
    _, (_, cwrite) = _one(cal_write_dir, "WRITE_SIZE", copy_kernel)
    fs, ws = copy_bytes / (cfetch * 1024), copy_bytes / (cwrite * 1024)
    rd, wr = fetch * 1024 * fs, write * 1024 * ws
    alg = 2 * nvar * 8 * cells
    out = {"kernel": name, "fetch_kib_per_dispatch": fetch, "write_kib_per_dispatch": write, "dispatches": [nf, nw],
           "copy_calibration": {"bytes_read": copy_bytes, "bytes_written": copy_bytes, "fetch_kib": cfetch,
                                "write_kib": cwrite, "fetch_scale": fs, "write_scale": ws},
           "bytes_per_launch": rd + wr, "algorithmic_bytes_per_launch": alg, "traffic_over_algorithmic": (rd + wr) / alg,
           "read_bytes_per_cell": rd / cells, "write_bytes_per_cell": wr / cells}
    if req_dir and cal_req_dir:
        counts, rb = _req_bytes(req_dir, kernel)
        ccounts, crb = _req_bytes(cal_req_dir, copy_kernel)
        rscale = copy_bytes / crb
        out["read_cross_check"] = {"counters": counts, "copy_counters": ccounts, "request_bytes": rb,
                                   "copy_request_bytes": crb, "copy_scale": rscale,
                                   "read_bytes_per_launch": rb * rscale,
                                   "read_bytes_per_cell": rb * rscale / cells,
                                   "over_fetch_size_estimate": rb * rscale / rd}
    return out


def main(argv):
    global LAST
    if "--last" in argv:
        k = argv.index("--last")
        LAST = int(argv[k + 1])
        argv = argv[:k] + argv[k + 2:]
    if argv[0] == "traffic":
        kernel, cells, nvar = argv[1], int(argv[2]), int(argv[3])
        print(json.dumps(traffic(kernel, cells, nvar, *argv[4:]), indent=1))
        return
    if argv[0] == "stats":
        print("| kernel | calls | total ms | average us | % |")
        print("|---|---:|---:|---:|---:|")
        for name, calls, tot, avg, pct in kernel_stats(argv[1])[:12]:
            print(f"| `{name[:110]}` | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.1f} | {pct:.1f} |")
    else:
        print(json.dumps(counter_means(argv[1], argv[2]), indent=1))


if __name__ == "__main__":
    main(sys.argv[1:])
