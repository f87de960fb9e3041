"""Kernel-time summary of a rocprofv3 --kernel-trace database (the
*_results.db it writes), grouped by kernel family, as the markdown tables
kept under profiles/.

    python -m jet_amd.profsum <run_results.db> --partitions N [--title T]
    python -m jet_amd.profsum --traffic --family F --fetch <db> --write <db> \
        --copy-fetch <db> --copy-write <db>

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


def counter_rows(db_path, counter):
    db = sqlite3.connect(db_path)
    return db.execute("select kernel_name, value from counters_collection where counter_name = ? order by dispatch_id",
                      (counter,)).fetchall()


def copy_scale(db_path, counter, nbytes):
    """nbytes / the counter's bytes for the plain copy kernel (the elementwise
    add of bench.py --calibrate-copy, 1 GiB read and 1 GiB written)."""
    vals = [v for n, v in counter_rows(db_path, counter) if "add" in n]
    if not vals:
        sys.exit(f"no copy dispatches in {db_path}")
    return nbytes / (1024.0 * float(sorted(vals)[len(vals) // 2]))


def traffic(args):
    """HBM bytes per dispatch of a family, from one FETCH_SIZE and one
    WRITE_SIZE pass, each scaled by the same counter over the plain copy."""
    fs = copy_scale(args.copy_fetch, "FETCH_SIZE", 1 << 30)
    ws = copy_scale(args.copy_write, "WRITE_SIZE", 1 << 30)
    out = []
    for counter, db_path, scale in (("FETCH_SIZE", args.fetch, fs), ("WRITE_SIZE", args.write, ws)):
        vals = [v for n, v in counter_rows(db_path, counter) if family(n) == args.family]
        if not vals:
            sys.exit(f"no {args.family} dispatches in {db_path}")
        kib = sum(vals) / len(vals)
        out.append((counter, len(vals), kib, scale, kib * 1024.0 * scale))
    print(f"# HBM traffic of the {args.family} family")
    print("")
    print("| counter | dispatches | KiB per dispatch (raw) | scale from the 1 GiB copy | bytes per dispatch |")
    print("|---|---:|---:|---:|---:|")
    for counter, nd, kib, scale, b in out:
        print(f"| {counter} | {nd} | {kib:.1f} | {scale:.3f} | {b:.0f} |")
    print("")
    print(f"HBM bytes per launch: {round(out[0][4] + out[1][4])}")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic":
        ap = argparse.ArgumentParser()
        ap.add_argument("--traffic", action="store_true")
        ap.add_argument("--family", required=True)
        ap.add_argument("--fetch", required=True, help="results db of a --pmc FETCH_SIZE pass")
        ap.add_argument("--write", required=True, help="results db of a --pmc WRITE_SIZE pass")
        ap.add_argument("--copy-fetch", required=True, help="FETCH_SIZE pass over bench.py --calibrate-copy")
        ap.add_argument("--copy-write", required=True, help="WRITE_SIZE pass over bench.py --calibrate-copy")
        return traffic(ap.parse_args())
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
