"""Markdown tables of a sweep run (the JSON lines `python -m bkma_amd.sweep` prints).

    python -m bkma_amd.sweep_report FILE [--command "..."]

Prints the advection table (one row per shape, implementation and storage
type), the Conv1d table and the performance-portability summary, in the
order the sweep ran them.
"""
import argparse
import json
import sys


def _rows(path):
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("{"):
                yield json.loads(line)


def _fmt(v):
    if v is None:
        return ""
    if isinstance(v, float) and v != 0 and abs(v) < 1e-3:
        return f"{v:.2e}"
    return str(v)


def report(path, command=None):
    recs = list(_rows(path))
    out = []
    if command:
        out += ["Command (from `parallel-advection_amd/`):", "", "    " + command, ""]
    adv = [r for r in recs if r.get("kind") == "advection"]
    if adv:
        out += ["| shape (n0, n1, n2) | impl | w | dtype | ms/step | Gcell/s | GB/s | arch | app | L1 | note |",
                "|---|---|---|---|---|---|---|---|---|---|---|"]
        for r in adv:
            shape = f"{r['size']} ({r['n0']}, {r['n1']}, {r['n2']})"
            if r.get("status") != "ok":
                note = r.get("status", "")
                out.append(f"| {shape} | {r['impl']} | {_fmt(r.get('pref_wg_size'))} | {r['dtype']} "
                           f"| | | | | | | {note} |")
                continue
            out.append(f"| {shape} | {r['impl']} | {_fmt(r.get('pref_wg_size'))} | {r['dtype']} | "
                       f"{r['ms_per_step']} | {r['gcell_s']} | {r['GBps']} | {r['arch_eff']} | {r.get('app_eff', '')} | "
                       f"{_fmt(r.get('l1_err'))} | {r.get('kernel', '')} |")
        out.append("")
    conv = [r for r in recs if r.get("kind") == "conv1d"]
    if conv:
        out += ["| config (batch, length, K, channels) | ms | Gcell/s | GB/s | result | inv |",
                "|---|---|---|---|---|---|"]
        for r in conv:
            cfg = f"({r['batch_size']}, {r['input_length']}, {r['kernel_size']}, {r['channels']})"
            if r.get("status") != "ok":
                out.append(f"| {cfg} | | | | | {r.get('status', '')} |")
                continue
            out.append(f"| {cfg} | {r['ms_per_step']} | {r['gcell_s']} | {r['GBps']} | {r['result']:.2f} | "
                       f"{r['batch_invariant']} |")
        out.append("")
    summ = [r for r in recs if r.get("kind") == "summary"]
    if summ:
        s = summ[-1]
        out += ["| impl | arch | app |", "|---|---|---|"]
        for kind in ("advection_pp", "conv1d_pp"):
            for k, v in s.get(kind, {}).items():
                out.append(f"| {kind.split('_')[0]} {k} | {v['arch']} | {v['app']} |")
        out.append("")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("file")
    ap.add_argument("--command", default=None, help="the sweep command, printed above the tables")
    a = ap.parse_args(argv)
    print(report(a.file, a.command))
    return 0


if __name__ == "__main__":
    sys.exit(main())
