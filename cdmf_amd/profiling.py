"""Summaries of rocprofv3 output for the CCD++ kernels.

  python -m cdmf_amd.profiling stats <rocprof output dir>
      per-kernel calls / total / average duration (kernel_stats.csv or the
      rocpd SQLite database)
  python -m cdmf_amd.profiling traffic <fetch dir> <write dir> [--probe-bytes N]
      per-kernel FETCH_SIZE and WRITE_SIZE (KiB per dispatch, from two
      `rocprofv3 --pmc` runs), each scaled by what the same counter reads for
      the stream-copy calibration kernel of known size (bench.py --probe),
      printed as JSON keyed by solver kernel class.

Kernel classes (the benchmark runs H first, so the first-swept layout is the
CSC copy with 32-bit row indices and the second the CSR copy with 16-bit
column indices): sweep_f / sweep_s (plain sweeps), fused_f / fused_s
(residual update fused with the first sweep of a rank).
"""
from __future__ import annotations

import csv
import glob
import json
import os
import re
import sqlite3
import sys


def _files(d, pattern):
    return sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))


def kernel_stats(d):
    """[(name, calls, total_ns, avg_ns)] from a rocprofv3 output directory."""
    rows = []
    for f in _files(d, "*kernel_stats.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"])))
    if rows:
        return rows
    for f in _files(d, "*.db"):
        db = sqlite3.connect(f)
        for name, calls, total, avg, _pct in db.execute("select * from top_kernels"):
            rows.append((name, int(calls), float(total) * 1e3, float(avg) * 1e3))  # view reports microseconds
    return rows


def kernel_class(name: str):
    """Solver kernel class of a demangled kernel name, or None.

    sweep_kernel<IdxT, TF, TA, SUB, ADD, SWEEP> and
    tiled_kernel<IdxT, SIGMA, TF, TA, SUB, ADD, SWEEP, XCUR>: the first
    template argument is the index type (unsigned short, ccd::Idx24 or
    unsigned int), the first three booleans are SUB, ADD, SWEEP."""
    m = re.search(r"(sweep_kernel|tiled_kernel)<([^>]*)>", name)
    if not m:
        if "copy_kernel" in name:
            return "copy_probe"
        if "commit_kernel" in name:
            return "copy"
        if "span_kernel" in name or "finalize_kernel" in name or "split_finish_kernel" in name:
            return "finalize"
        if "rmse" in name:
            return "rmse"
        return None
    args = [a.strip() for a in m.group(2).split(",")]
    bools = [a == "true" for a in args if a in ("true", "false")]
    if len(bools) < 3:
        return None
    sub, add, sweep = bools[:3]
    if not sweep:
        return "residual"
    # the first-swept layout (CSC) has 32-bit or 3-byte indices, the second
    # (CSR of a matrix with at most 65,536 columns) 16-bit ones
    layout = "s" if args[0] == "unsigned short" else "f"
    return ("fused_" if (sub or add) else "sweep_") + layout


def counters(d, counter):
    """{kernel class: [values per dispatch]} for one --pmc counter."""
    out = {}
    for f in _files(d, "*counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                cls = kernel_class(r["Kernel_Name"])
                if cls:
                    out.setdefault(cls, []).append(float(r["Counter_Value"]))
    return out


def traffic(fetch_dir, write_dir, probe_bytes=256 << 20):
    fetch = counters(fetch_dir, "FETCH_SIZE")
    write = counters(write_dir, "WRITE_SIZE")
    kib = probe_bytes / 1024.0
    sf = kib / (sum(fetch["copy_probe"]) / len(fetch["copy_probe"])) if fetch.get("copy_probe") else 1.0
    sw = kib / (sum(write["copy_probe"]) / len(write["copy_probe"])) if write.get("copy_probe") else 1.0
    res = {"scale_fetch": sf, "scale_write": sw}
    for cls in sorted(set(fetch) | set(write)):
        if cls == "copy_probe":
            continue
        f = sum(fetch.get(cls, [0.0])) / max(len(fetch.get(cls, [])), 1)
        w = sum(write.get(cls, [0.0])) / max(len(write.get(cls, [])), 1)
        res[cls] = {"fetch_kib": f, "write_kib": w,
                    "hbm_bytes_per_launch": 1024.0 * (f * sf + w * sw)}
    return res


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        rows = sorted(kernel_stats(argv[1]), key=lambda r: -r[2])
        tot = sum(r[2] for r in rows) or 1.0
        print("%-100s %8s %12s %10s %6s" % ("kernel", "calls", "total_us", "avg_us", "pct"))
        for name, calls, total, avg in rows:
            print("%-100s %8d %12.1f %10.2f %6.2f" % (name[:100], calls, total / 1e3, avg / 1e3, 100 * total / tot))
        return 0
    if len(argv) >= 3 and argv[0] == "traffic":
        pb = 256 << 20
        if "--probe-bytes" in argv:
            pb = int(argv[argv.index("--probe-bytes") + 1])
        print(json.dumps(traffic(argv[1], argv[2], pb), indent=1))
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
