"""Summaries of rocprofv3 CSV output for the profile notes.

  python -m exploringsycl_amd.rocprof_summary stats DIR
      per-kernel launches, average and total device time from the
      --kernel-trace --stats run (DIR/*_kernel_stats.csv, summed over the
      processes of the run).

  python -m exploringsycl_amd.rocprof_summary counters DIR [DIR ...] [--kernel REGEX]
      mean of every counter per dispatch, per kernel and grid size, from
      --pmc runs (DIR/*counter_collection.csv).

  python -m exploringsycl_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR \
      --kernel REGEX --copy-kernel REGEX --copy-bytes B --algorithmic-bytes A
      HBM bytes per launch of the kernel from a FETCH_SIZE run and a
      WRITE_SIZE run (KiB per dispatch), each scaled by what the same counter
      reads for a plain copy kernel that reads and writes B bytes; prints JSON.
"""

import argparse
import csv
import glob
import json
import os
import re
import sys


def _all(pattern):
    files = sorted(glob.glob(pattern))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files


def kernel_stats(d):
    """(name, calls, average us, total ms, % of all kernel time), summed over
    every process's stats file in d (one per process with -o name_%pid%)."""
    calls, total = {}, {}
    for path in _all(os.path.join(d, "*kernel_stats.csv")):
        with open(path) as f:
            for r in csv.DictReader(f):
                calls[r["Name"]] = calls.get(r["Name"], 0) + int(r["Calls"])
                total[r["Name"]] = total.get(r["Name"], 0.0) + float(r["TotalDurationNs"])
    grand = sum(total.values()) or 1.0
    rows = [(k, calls[k], total[k] / calls[k] / 1e3, total[k] / 1e6, 100.0 * total[k] / grand)
            for k in calls]
    rows.sort(key=lambda t: -t[3])
    return rows


def short(name):
    name = name.replace("tea::(anonymous namespace)::", "").replace("tea::", "")
    name = name.replace("void ", "")
    depth, out = 0, []
    for ch in name:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out)


def counter_means(d, counter):
    """{kernel name: mean counter value per dispatch}."""
    acc = {}
    for path in _all(os.path.join(d, "*counter_collection.csv")):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                key = (path, r["Kernel_Name"], r["Dispatch_Id"])
                acc[key] = acc.get(key, 0.0) + float(r["Counter_Value"])
    per = {}
    for (_, name, _), v in acc.items():
        per.setdefault(name, []).append(v)
    return {k: sum(v) / len(v) for k, v in per.items()}


def all_counter_means(d):
    """{(kernel, grid size): {counter: mean value per dispatch}}."""
    acc = {}
    for path in _all(os.path.join(d, "*counter_collection.csv")):
        with open(path) as f:
            for r in csv.DictReader(f):
                key = (short(r["Kernel_Name"]), r["Grid_Size"], r["Counter_Name"])
                acc.setdefault(key, {})
                disp = (path, r["Dispatch_Id"])
                acc[key][disp] = acc[key].get(disp, 0.0) + float(r["Counter_Value"])
    out = {}
    for (name, grid, counter), per in acc.items():
        out.setdefault((name, grid), {})[counter] = sum(per.values()) / len(per)
    return out


def _match(table, regex):
    hits = [k for k in table if re.search(regex, k)]
    if len(hits) != 1:
        raise SystemExit(f"{regex!r} matches {len(hits)} kernels: {hits[:4]}")
    return table[hits[0]]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=12)
    cc = sub.add_parser("counters")
    cc.add_argument("dirs", nargs="+")
    cc.add_argument("--kernel", default=".")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True,
                   help="bytes the copy kernel reads (and writes) per launch")
    t.add_argument("--algorithmic-bytes", type=float, required=True)
    a = ap.parse_args(argv)

    if a.cmd == "stats":
        print(f"{'kernel':64s} {'calls':>7s} {'avg us':>10s} {'total ms':>10s} {'%':>6s}")
        for name, calls, avg, tot, pct in kernel_stats(a.dir)[:a.top]:
            print(f"{short(name)[:64]:64s} {calls:7d} {avg:10.1f} {tot:10.2f} {pct:6.2f}")
        return 0

    if a.cmd == "counters":
        for d in a.dirs:
            print(f"== {d}")
            for (name, grid), cs in sorted(all_counter_means(d).items()):
                if re.search(a.kernel, name):
                    vals = " ".join(f"{k}={v:.4g}" for k, v in sorted(cs.items()))
                    print(f"  {name[:48]:48s} grid={grid:>9s} {vals}")
        return 0

    fetch = counter_means(a.fetch_dir, "FETCH_SIZE")
    write = counter_means(a.write_dir, "WRITE_SIZE")
    copy_f = _match(fetch, a.copy_kernel) * 1024.0
    copy_w = _match(write, a.copy_kernel) * 1024.0
    sf, sw = a.copy_bytes / copy_f, a.copy_bytes / copy_w
    kf = _match(fetch, a.kernel) * 1024.0 * sf
    kw = _match(write, a.kernel) * 1024.0 * sw
    out = {
        "kernel": a.kernel,
        "fetch_bytes_per_launch": kf,
        "write_bytes_per_launch": kw,
        "hbm_bytes_per_launch": kf + kw,
        "algorithmic_bytes_per_launch": a.algorithmic_bytes,
        "traffic_over_algorithmic": (kf + kw) / a.algorithmic_bytes,
        "scale_fetch": sf,
        "scale_write": sw,
        "copy_kernel_raw_kib": {"FETCH_SIZE": copy_f / 1024.0, "WRITE_SIZE": copy_w / 1024.0},
    }
    json.dump(out, sys.stdout, indent=1)
    print()
    return 0


if __name__ == "__main__":
    sys.exit(main())
