"""Kernel-time summary of a rocprofv3 --kernel-trace database (the
*_results.db it writes), grouped by kernel family, as the markdown tables
kept under profiles/.

    python -m jet_amd.profsum <run_results.db> --partitions N [--title T]

Families group the template instances of one kernel; the profiler families of
bench.py ("lp_select", "afterburner", "update", "weak_select") are the
kernels k_lp_select, k_afterburner_*, k_update_* and k_weak_select.
"""
import argparse
import re
import sqlite3
import sys
from collections import defaultdict

FAMILY_OF = [
    (re.compile(r"\bk_lp_select\b"), "lp_select"),
    (re.compile(r"\bk_afterburner_"), "afterburner"),
    (re.compile(r"\bk_update_(small|big|hubs)\b"), "update"),
    (re.compile(r"\bk_weak_select\b"), "weak_select"),
]


def family(name):
    for rx, fam in FAMILY_OF:
        if rx.search(name):
            return fam
    return None


def base_name(name):
    name = re.sub(r"^void ", "", name)
    name = re.sub(r"[<(].*$", "", name)
    return re.sub(r"^jet::", "", name)


def load(db_path):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, start, end from kernels order by start").fetchall()
    return [(n, s, e) for n, s, e in rows]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--partitions", type=int, required=True, help="partitions the traced command ran")
    ap.add_argument("--title", default="Kernel trace")
    ap.add_argument("--top", type=int, default=25)
    args = ap.parse_args()
    rows = load(args.db)
    if not rows:
        sys.exit("no kernel dispatches in the database")
    P = args.partitions
    tot = sum(e - s for _, s, e in rows)
    by_base = defaultdict(lambda: [0, 0])
    by_fam = defaultdict(lambda: [0, 0])
    for n, s, e in rows:
        d = e - s
        b = by_base[base_name(n)]
        b[0] += d
        b[1] += 1
        f = family(n)
        if f:
            by_fam[f][0] += d
            by_fam[f][1] += 1
    out = [f"# {args.title}", ""]
    out.append(f"Kernel time per partition: {tot / 1e6 / P:.3f} ms over {len(rows) / P:.0f} dispatches "
               f"({P} partitions traced).")
    out.append("")
    out.append("By kernel (template instances summed):")
    out.append("")
    out.append("| kernel | ms / partition | dispatches / partition | average us | share |")
    out.append("|---|---:|---:|---:|---:|")
    for name, (d, c) in sorted(by_base.items(), key=lambda kv: -kv[1][0])[:args.top]:
        out.append(f"| `{name}` | {d / 1e6 / P:.3f} | {c / P:.1f} | {d / 1e3 / c:.2f} | {100.0 * d / tot:.1f}% |")
    out.append("")
    out.append("Profiler families (bench.py `kernel_families`):")
    out.append("")
    out.append("| family | ms / partition | dispatches / partition | average us per dispatch |")
    out.append("|---|---:|---:|---:|")
    for f in ("lp_select", "afterburner", "update", "weak_select"):
        d, c = by_fam.get(f, (0, 0))
        if c:
            out.append(f"| {f} | {d / 1e6 / P:.3f} | {c / P:.1f} | {d / 1e3 / c:.2f} |")
    print("\n".join(out))


if __name__ == "__main__":
    main()
