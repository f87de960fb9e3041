#!/usr/bin/env python3
"""Throughput of the MI355X Jet partitioner on one GPU.

One step = one full multilevel k-way partition (coarsen + initial partition +
uncoarsen/refine, i.e. the reference's Measurement::Total,
partitioner.hpp:70-111) of a synthetic graph already resident in HBM.
Default workload: 3-D 7-point stencil 128^3 (2,097,152 vertices, 6,242,304
undirected edges), edge weights uniform in [1, 10], k = 64, imbalance 1.03,
HEC coarsening. Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "jet-partitioner-official-fork_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import jet_amd  # noqa: E402

WORKLOADS = {
    "stencil3d-128": dict(desc="3-D 7-pt stencil 128^3, w in [1,10], k=64, imb 1.03, HEC", k=64, imb=1.03, alg=1),
    "grid2d-256": dict(desc="2-D grid 256x256, unit weights, k=8, imb 1.03, 2-hop", k=8, imb=1.03, alg=0),
    "rmat-22": dict(desc="R-MAT scale 22 ef 16 (.57,.19,.19), unit weights, k=256, imb 1.10, HEC", k=256, imb=1.10,
                    alg=1),
}


def make_graph(name):
    if name == "stencil3d-128":
        return jet_amd.stencil3d(128, 128, 128, seed=0, wmax=10)
    if name == "grid2d-256":
        return jet_amd.grid2d(256, 256)
    if name == "rmat-22":
        return jet_amd.rmat(22, 16, seed=0)
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="stencil3d-128", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prof", default="lp_select", help="kernel family timed with HIP events")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (each box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    print("# importing torch", file=sys.stderr, flush=True)
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    wl = WORKLOADS[args.workload]
    t_gen = time.time()
    g = make_graph(args.workload)
    n = len(g["rowptr"]) - 1
    m = len(g["col"])
    print(f"# workload {args.workload}: n={n} m={m} (generated in {time.time() - t_gen:.1f}s)", file=sys.stderr,
          flush=True)
    dev = torch.device("cuda:0")
    d_rp = torch.from_numpy(g["rowptr"]).to(dev)
    d_col = torch.from_numpy(g["col"]).to(dev)
    d_ew = torch.from_numpy(g["ew"]).to(dev)
    d_vw = torch.from_numpy(g["vw"]).to(dev)
    d_part = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    cfg = jet_amd.Config(num_parts=wl["k"], max_imb_ratio=wl["imb"], coarsening_alg=wl["alg"])
    uniform = bool(g.get("uniform", False))

    def step():
        return jet_amd.partition_device(n, d_rp.data_ptr(), d_col.data_ptr(), d_ew.data_ptr(), d_vw.data_ptr(),
                                        d_part.data_ptr(), cfg, seed=args.seed, uniform=uniform)

    for w in range(args.warmup):
        step()
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    prof_ok = jet_amd.lib().jet_prof_enable(args.prof.encode()) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cut = None
    st = None
    for s in range(args.steps):
        cut, st = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {dt:.3f}s, cut {cut}", file=sys.stderr, flush=True)
    for L in range(st.num_levels):
        lv = st.levels[L]
        print(f"#   level n={lv.num_vertices:9d} m={lv.num_edges:10d} iters={lv.total_iterations:3d} "
              f"lp={lv.lp_iterations:3d} refine_ms={1e3 * lv.total_ref_time:8.3f}", file=sys.stderr, flush=True)
    ms = 1e3 * dt / args.steps
    edges = m // 2
    value = edges * args.steps / dt

    roofline = None
    if prof_ok:
        launches, tot_ms, nbytes = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)
        if jet_amd.lib().jet_prof_read(ctypes.byref(launches), ctypes.byref(tot_ms), ctypes.byref(nbytes)) == 0 \
                and launches.value > 0 and tot_ms.value > 0:
            avg_s = tot_ms.value / 1e3 / launches.value
            per_launch = nbytes.value / launches.value
            achieved = per_launch / avg_s / 1e9
            roofline = {"kernel": args.prof, "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(achieved / 8000.0, 4), "traffic": None,
                        "launches": launches.value, "avg_us": round(avg_s * 1e6, 2),
                        "bytes_per_launch": round(per_launch)}

    part = d_part.cpu().numpy()
    sizes = np.bincount(part, weights=g["vw"], minlength=wl["k"])
    opt = (int(g["vw"].sum()) + wl["k"] - 1) // wl["k"]
    balanced = bool(sizes.max() <= int(opt * wl["imb"]))

    cpu = None
    cut_ratio = None
    if not args.no_cpu_baseline:
        import oracle

        t1 = time.perf_counter()
        opart, ocut, _ = oracle.partition(g, wl["k"], wl["imb"], alg=wl["alg"], seed=args.seed)
        cdt = time.perf_counter() - t1
        cpu = {"value": round(edges / cdt, 1), "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": f"one full partition of the same {args.workload} graph, same seed, serial oracle "
                         f"({cdt:.2f}s)"}
        cut_ratio = round(cut / ocut, 6) if ocut else None
        print(f"# cpu oracle {cdt:.2f}s cut {ocut}; gpu/oracle partitions identical: "
              f"{bool(np.array_equal(part, opart))}", file=sys.stderr, flush=True)

    out = {
        "metric": "edges/s partitioned (1 GPU) at cut <=1.01x ref",
        "value": round(value, 1),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": wl["desc"], "vertices": n, "undirected_edges": edges,
                   "k": wl["k"], "imbalance": wl["imb"], "coarsening_alg": wl["alg"], "seed": args.seed},
        "edge_cut": cut,
        "cut_vs_oracle": cut_ratio,
        "balanced": balanced,
        "phases_s": {"coarsen": round(st.coarsen_s, 4), "init": round(st.init_partition_s, 4),
                     "refine": round(st.refine_s, 4), "levels": st.num_levels} if st else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
