"""Summaries of rocprofv3 CSV output for the library's kernels.

    python -m bkma_amd.rocprof_summary DIR --kernel SUBSTRING [--last N]

DIR is a rocprofv3 output directory (``-d DIR --output-format csv``).  For a
kernel trace (``*kernel_trace.csv``) it prints calls and average / min / max
duration of every kernel whose name contains SUBSTRING; for a counter
collection (``*counter_collection.csv``) the average of every counter over
those kernels' dispatches.  ``--last N`` keeps only the last N dispatches of
each kernel (the timed passes of a run, after warm-up and tuning dispatches
of the same kernel).
"""
import argparse
import collections
import csv
import glob
import os
import sys


def _rows(pattern):
    for path in sorted(glob.glob(pattern, recursive=True)):
        with open(path, newline="") as f:
            yield from csv.DictReader(f)


def _kname(row):
    return row.get("Kernel_Name") or row.get("Kernel-Name") or ""


def kernel_trace(d, sub, last):
    by = collections.defaultdict(list)  # name -> [(start, duration ns)]
    for r in _rows(os.path.join(d, "**", "*kernel_trace.csv")):
        k = _kname(r)
        if sub in k:
            s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            by[k].append((s, e - s))
    out = []
    for k, v in by.items():
        v.sort()
        if last:
            v = v[-last:]
        ms = [x[1] / 1e6 for x in v]
        out.append((k, len(ms), sum(ms) / len(ms), min(ms), max(ms)))
    return out


def counters(d, sub, last):
    # one row per (dispatch, counter); the dispatch id orders the dispatches
    disp = collections.defaultdict(dict)  # (kernel, dispatch) -> counter -> value
    res = {}  # kernel -> its resources as the dispatch reports them
    for r in _rows(os.path.join(d, "**", "*counter_collection.csv")):
        k = _kname(r)
        if sub not in k:
            continue
        res.setdefault(k, {f: r.get(f, "?") for f in ("VGPR_Count", "Accum_VGPR_Count", "SGPR_Count",
                                                       "LDS_Block_Size", "Scratch_Size", "Workgroup_Size")})
        did = int(r.get("Dispatch_Id", r.get("Dispatch-Id", 0)))
        name = r["Counter_Name"]
        disp[(k, did)][name] = disp[(k, did)].get(name, 0.0) + float(r["Counter_Value"])
    by = collections.defaultdict(list)
    for (k, did), cs in disp.items():
        by[k].append((did, cs))
    out = []
    for k, v in by.items():
        v.sort(key=lambda x: x[0])
        if last:
            v = v[-last:]
        names = sorted({n for _, cs in v for n in cs})
        out.append((k, len(v), {n: sum(cs.get(n, 0.0) for _, cs in v) / len(v) for n in names}, res[k]))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("dir")
    ap.add_argument("--kernel", default="", help="substring of the kernel names to keep")
    ap.add_argument("--last", type=int, default=0, help="keep the last N dispatches of each kernel")
    a = ap.parse_args(argv)
    found = False
    for k, n, avg, lo, hi in kernel_trace(a.dir, a.kernel, a.last):
        found = True
        print(f"| `{k}` | {n} | {avg:.4f} | {lo:.4f} | {hi:.4f} |")
    for k, n, cs, rs in counters(a.dir, a.kernel, a.last):
        found = True
        print(f"`{k}` ({n} dispatches; " + ", ".join(f"{f} {v}" for f, v in rs.items()) + ")")
        print("| counter | average per dispatch |")
        print("|---|---|")
        for c, v in cs.items():
            print(f"| {c} | {v:.6g} |")
    if not found:
        print(f"no kernel containing {a.kernel!r} under {a.dir}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
