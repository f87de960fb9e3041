"""Autotuner for the result-neutral tuning parameters (launch shapes and path
thresholds, include/jet_capi.h): coordinate descent over each parameter's
candidate values, timing full partitions of representative device-resident
workloads, and checking that every setting gives the partitions of the
built-in defaults bit for bit. Writes a tuning table ("name value" lines)
that the library applies at context creation when it is installed as
tuning/<arch>.txt beside libjet_amd.so.

    python -m jet_amd.autotune --out tuning/gfx950.txt

Needs a GPU; the workloads are generated on the host and copied to HBM once.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

import jet_amd

# candidate values per parameter, searched in this order
SPACE = {
    "small_level_n": [8192, 32768, 131072],
    "ab_lanes_small": [1, 2, 4],
    "upd_lanes": [1, 2, 4],
    "upd_big_slices": [1, 4, 16, 64],
    "lp_one_pass_n": [131072, 524288, 2097152],
    "ct_wave_max": [32, 64, 128],
    "ct_block_max": [1024, 4096],
    "wg_max_n": [0, 1024, 4096, 16384],
    "wg_big_deg": [0, 8, 32],
    "redraw_lds_max": [0, 8192],
    "sel_max_blocks": [512, 1024, 2048, 4096, 8192],
    "sel_copy_wgs": [32, 64, 128, 256],
}

# (name, builder, k, imbalance, coarsening algorithm, weight of its time)
WORKLOADS = [
    ("stencil3d-128 k=64 HEC", lambda: jet_amd.stencil3d(128, 128, 128, seed=0, wmax=10), 64, 1.03, 1, 1.0),
    ("rmat-18 k=64 HEC", lambda: jet_amd.rmat(18, 16, seed=1), 64, 1.10, 1, 1.0),
    ("grid2d-512 k=16 2-hop", lambda: jet_amd.grid2d(512, 512), 16, 1.03, 0, 1.0),
    ("stencil3d-24 k=8 HEC", lambda: jet_amd.stencil3d(24, 24, 24, seed=2, wmax=4), 8, 1.03, 1, 4.0),
]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class Workload:
    def __init__(self, name, g, k, imb, alg, weight):
        import torch

        self.name, self.weight = name, weight
        self.n = len(g["rowptr"]) - 1
        dev = torch.device("cuda:0")
        self.rp = torch.from_numpy(g["rowptr"]).to(dev)
        self.col = torch.from_numpy(g["col"]).to(dev)
        self.ew = torch.from_numpy(g["ew"]).to(dev)
        self.vw = torch.from_numpy(g["vw"]).to(dev)
        self.part = torch.empty(self.n, dtype=torch.int32, device=dev)
        self.uniform = bool(g.get("uniform", False))
        self.cfg = jet_amd.Config(num_parts=k, max_imb_ratio=imb, coarsening_alg=alg)

    def run(self):
        return jet_amd.partition_device(self.n, self.rp.data_ptr(), self.col.data_ptr(), self.ew.data_ptr(),
                                        self.vw.data_ptr(), self.part.data_ptr(), self.cfg, seed=0,
                                        uniform=self.uniform)[0]

    def digest(self):
        return hashlib.sha256(self.part.cpu().numpy().tobytes()).hexdigest()

    def time(self, reps):
        import torch

        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.run()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))


def check(wls, expect):
    """Runs every workload once and checks its partition against `expect`
    (name -> (cut, digest))."""
    for w in wls:
        got = (w.run(), w.digest())
        if got != expect[w.name]:
            raise RuntimeError(f"{w.name}: a tuning setting changed the partition ({got[0]} vs "
                               f"{expect[w.name][0]}): {jet_amd.tuning_all()}")


def compare(wls, name, cands, rounds, expect):
    """Weighted partition time of each candidate value of `name`: the
    candidates are timed in interleaved rounds (one partition of every
    workload per candidate per round), so that drift of the device's clocks
    over the search falls on all of them alike; each workload's median over
    the rounds is weighted and summed."""
    import torch

    for v in cands:
        jet_amd.tuning_set(name, v)
        check(wls, expect)
    samples = {v: {w.name: [] for w in wls} for v in cands}
    for _ in range(rounds):
        for v in cands:
            jet_amd.tuning_set(name, v)
            for w in wls:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                w.run()
                torch.cuda.synchronize()
                samples[v][w.name].append(time.perf_counter() - t0)
    return {v: sum(w.weight * float(np.median(samples[v][w.name])) for w in wls) for v in cands}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="tuning table to write")
    ap.add_argument("--reps", type=int, default=7, help="interleaved timing rounds per comparison")
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--min-gain", type=float, default=0.02, help="relative improvement needed to change a value")
    ap.add_argument("--report", help="JSON file for the search record")
    args = ap.parse_args()
    # the search starts from the built-in defaults: no installed table is
    # applied when the first context is created
    os.environ["JET_TUNING_FILE"] = os.devnull
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    # search from the built-in defaults, not from an installed table
    defaults = jet_amd.tuning_all()
    wls = [Workload(name, mk(), k, imb, alg, wt) for name, mk, k, imb, alg, wt in WORKLOADS]
    expect = {}
    for w in wls:
        cut = w.run()
        expect[w.name] = (cut, w.digest())
    best = dict(defaults)
    for n, v in best.items():
        jet_amd.tuning_set(n, v)
    base = compare(wls, "wg_max_n", [best["wg_max_n"]], args.reps, expect)[best["wg_max_n"]]
    say(f"# {arch}: defaults {best}: {1e3 * base:.2f} ms")
    record = {"arch": arch, "defaults": defaults, "default_ms": round(1e3 * base, 3), "steps": []}
    for sweep in range(args.sweeps):
        changed = False
        for name, cands in SPACE.items():
            cands = sorted(set(cands) | {best[name]})
            results = compare(wls, name, cands, args.reps, expect)
            for v in cands:
                record["steps"].append({"sweep": sweep, "name": name, "value": v, "ms": round(1e3 * results[v], 3)})
                say(f"#   {name}={v}: {1e3 * results[v]:.2f} ms")
            v_best = min(results, key=results.get)
            if v_best != best[name] and results[v_best] < (1 - args.min_gain) * results[best[name]]:
                best[name] = v_best
                changed = True
            jet_amd.tuning_set(name, best[name])
            say(f"# {name} -> {best[name]}")
        if not changed:
            break
    final = compare(wls, "wg_max_n", [best["wg_max_n"]], args.reps, expect)[best["wg_max_n"]]
    record.update({"tuned": best, "tuned_ms": round(1e3 * final, 3)})
    with open(args.out, "w") as f:
        f.write(f"# tuning table for {arch}: written by jet_amd.autotune\n")
        f.write(f"# weighted partition time {1e3 * base:.2f} ms with the defaults, {1e3 * final:.2f} ms tuned, over "
                + "; ".join(w.name for w in wls) + "\n")
        for n in jet_amd.TUNING_NAMES:
            f.write(f"{n} {best[n]}\n")
    if args.report:
        with open(args.report, "w") as f:
            json.dump(record, f, indent=1)
    print(json.dumps({"arch": arch, "default_ms": record["default_ms"], "tuned_ms": record["tuned_ms"],
                      "tuned": best}), flush=True)


if __name__ == "__main__":
    main()
