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
import hashlib
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
                    alg=1, cpu_sample="rmat-20", golden="rmat22_k256_golden.json"),
}

# The serial oracle needs minutes for R-MAT scale 22; its CPU baseline is timed
# on the same generator at scale 20 (same edge factor, k, imbalance, seed).
SAMPLES = {"rmat-20": "R-MAT scale 20 ef 16 (.57,.19,.19), k=256, imb 1.10, HEC, seed 0"}


def counted_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE,
    each scaled by the same counter over a plain copy of known size), or None
    when no summary exists for that pair."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("HBM bytes per launch:"):
                return int(line.split(":")[1])
    return None


def cpu_threads():
    """Threads for the CPU baseline: the CPUs this process may run on, at most
    16 (one GPU's share of the host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def host_cpu():
    """Model and core counts of the host, from lscpu."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[key] = val
    except Exception:  # lscpu missing: report what Python knows
        info["CPU(s)"] = str(os.cpu_count())
    info["OMP_PROC_BIND"] = os.environ.get("OMP_PROC_BIND")
    info["OMP_PLACES"] = os.environ.get("OMP_PLACES")
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def copy_bandwidth(torch):
    """This device's plain-copy bandwidth (read + write bytes / time) over a
    1 GiB int32 copy, best of three, timed with events on torch's stream."""
    a = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    best = None
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1) / 1e3
        best = t if best is None else min(best, t)
    del a, b
    return round(2 * (4 << 28) / best / 1e9, 1)


def block_partition(name, k):
    """The geometric block partition of a grid or stencil workload (k split
    over the axes as evenly as powers of two allow), or None for other
    workloads: the quality anchor printed beside the edge cut."""
    dims = {"stencil3d-128": (128, 128, 128), "grid2d-256": (256, 256)}.get(name)
    if dims is None or k & (k - 1):
        return None
    splits = [1] * len(dims)
    b = 0
    while (1 << b) < k:
        splits[b % len(dims)] *= 2
        b += 1
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.int64)
    part = np.zeros(n, dtype=np.int64)
    stride = 1
    mult = 1
    for d, s in zip(dims, splits):
        coord = (idx // stride) % d
        part += (coord * s // d) * mult
        stride *= d
        mult *= s
    return part.astype(np.int32)


def edge_cut(g, part):
    rows = np.repeat(np.arange(len(g["rowptr"]) - 1), np.diff(g["rowptr"]))
    return int(g["ew"][part[rows] != part[g["col"]]].astype(np.int64).sum() // 2)


def kernel_rooflines(step, args, ms_per_step, torch):
    """Per-family rooflines of the partition: for each kernel family, the
    same --steps partitions again with every launch of that family made with
    its own start and end HIP events (hipExtLaunchKernelGGL on the refinement
    stream, bounded below by a marker event recorded before the launch: the
    interval is the kernel's execution, as in a profiler's dispatch record), then one partition in counting mode for every family's
    algorithmic bytes (the computation is deterministic, so its launches are
    the timed ones). Returns (roofline of the family with the most kernel
    time, all families, end-to-end figure: modeled bytes of all families /
    partition time)."""
    fams = {}
    for fam in jet_amd.PROF_FAMILIES:
        jet_amd.prof_enable(fam)
        torch.cuda.synchronize()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        units, tot_ms, _ = jet_amd.prof_read(fam)
        fams[fam] = {"units": units, "ms": tot_ms}
    jet_amd.prof_enable("all:count")
    step()
    torch.cuda.synchronize()
    total_bytes = 0.0
    for fam in jet_amd.PROF_FAMILIES:
        cu, _, nbytes = jet_amd.prof_read(fam)
        f = fams[fam]
        f["units_per_partition"] = cu
        f["bytes_per_partition"] = round(nbytes)
        total_bytes += nbytes
        if cu > 0 and f["units"] == cu * args.steps and f["ms"] > 0:
            avg_s = f["ms"] / 1e3 / f["units"]
            f["avg_us"] = round(avg_s * 1e6, 3)
            f["bytes_per_unit"] = round(nbytes / cu)
            f["achieved_GBps"] = round(nbytes / cu / avg_s / 1e9, 1)
            f["frac"] = round(nbytes / cu / avg_s / 8e12, 4)
            f["ms_per_partition"] = round(f["ms"] / args.steps, 3)
    jet_amd.prof_enable("none")
    timed = {k: v for k, v in fams.items() if "frac" in v}
    roofline = None
    if timed:
        top = max(timed, key=lambda k: timed[k]["ms"])
        f = timed[top]
        roofline = {"kernel": top, "bound": "hbm", "achieved": f["achieved_GBps"], "peak": 8000.0,
                    "copy_GBps": copy_bandwidth(torch), "unit": "GB/s", "frac": f["frac"],
                    "traffic": counted_traffic(top, args.workload),
                    "launches": f["units"], "avg_us": f["avg_us"], "bytes_per_launch": f["bytes_per_unit"]}
    e2e = {"modeled_bytes_per_partition": round(total_bytes), "families": list(jet_amd.PROF_FAMILIES),
           "achieved_GBps": round(total_bytes / (ms_per_step / 1e3) / 1e9, 1),
           "frac": round(total_bytes / (ms_per_step / 1e3) / 8e12, 5)}
    return roofline, fams, e2e


def make_graph(name):
    if name == "stencil3d-128":
        return jet_amd.stencil3d(128, 128, 128, seed=0, wmax=10)
    if name == "grid2d-256":
        return jet_amd.grid2d(256, 256)
    if name == "rmat-22":
        return jet_amd.rmat(22, 16, seed=0)
    if name == "rmat-20":
        return jet_amd.rmat(20, 16, seed=0)
    raise SystemExit(f"unknown workload {name}")


def batch_graphs(count, seed):
    """The batched mix: half 2-D/3-D grids, half R-MAT (a=.57, b=.19, c=.19,
    edge factor 8); vertex counts log-uniform in [1K, 200K] (R-MAT: the power
    of two nearest the draw); edge weights uniform in [1, 4]; per graph k from
    {2..64}, imbalance from {1.01, 1.03, 1.10}, 2-hop or HEC coarsening."""
    rs = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = float(np.exp(rs.uniform(np.log(1e3), np.log(2e5))))
        if i % 2 == 0:
            if rs.random() < 0.5:
                side = max(2, int(round(n ** 0.5)))
                g = jet_amd.grid2d(side, side)
            else:
                side = max(2, int(round(n ** (1 / 3))))
                g = jet_amd.stencil3d(side, side, side, seed=i, wmax=0)
        else:
            g = jet_amd.rmat(max(10, int(round(np.log2(n)))), 8, seed=i)
        g["ew"] = rs.integers(1, 5, len(g["col"])).astype(np.int32)
        rp, col = g["rowptr"], g["col"]
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        # symmetric weights: w(u, v) from the unordered pair
        lo, hi = np.minimum(rows, col).astype(np.int64), np.maximum(rows, col).astype(np.int64)
        g["ew"] = (1 + ((lo * 0x9E3779B1 + hi * 0x85EBCA77 + i) % 4)).astype(np.int32)
        g["uniform"] = False
        cfg = jet_amd.Config(num_parts=int(rs.choice([2, 4, 8, 16, 32, 64])),
                             max_imb_ratio=float(rs.choice([1.01, 1.03, 1.10])),
                             coarsening_alg=int(rs.integers(0, 2)))
        out.append((g, cfg, 1000 + i))
    return out


def run_batch(args):
    """Batched mode: the same graphs partitioned one call at a time (plain
    loop) and through jet_partition_batch; graphs/s for both."""
    import torch

    count = args.graphs
    t_gen = time.time()
    jobs = batch_graphs(count, args.seed)
    print(f"# {count} graphs generated in {time.time() - t_gen:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    ns = [len(g["rowptr"]) - 1 for g, _, _ in jobs]
    ms = [len(g["col"]) for g, _, _ in jobs]
    voff = np.zeros(count + 1, dtype=np.int64)
    voff[1:] = np.cumsum(ns)
    eoff = np.zeros(count + 1, dtype=np.int64)
    eoff[1:] = np.cumsum(ms)
    d_rp = torch.from_numpy(np.concatenate([g["rowptr"].astype(np.int64) for g, _, _ in jobs])).to(dev)
    d_col = torch.from_numpy(np.concatenate([g["col"] for g, _, _ in jobs])).to(dev)
    d_ew = torch.from_numpy(np.concatenate([g["ew"] for g, _, _ in jobs])).to(dev)
    d_vw = torch.ones(int(voff[-1]), dtype=torch.int32, device=dev)
    d_part = torch.empty(int(voff[-1]), dtype=torch.int32, device=dev)
    d_part2 = torch.empty(int(voff[-1]), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    cfgs = (jet_amd.Config * count)(*[c for _, c, _ in jobs])
    seeds = np.array([s for _, _, s in jobs], dtype=np.uint64)
    uni = np.zeros(count, dtype=np.int32)
    ctx = jet_amd.Context(workers=args.workers)
    cuts_b = np.empty(count, dtype=np.int64)
    L = jet_amd.lib()
    i4 = torch.tensor([], dtype=torch.int32).element_size()

    def plain():
        cuts = []
        for i in range(count):
            c, _ = jet_amd.partition_device(ns[i], d_rp.data_ptr() + 8 * int(voff[i] + i),
                                            d_col.data_ptr() + i4 * int(eoff[i]), d_ew.data_ptr() + i4 * int(eoff[i]),
                                            d_vw.data_ptr() + i4 * int(voff[i]), d_part.data_ptr() + i4 * int(voff[i]),
                                            cfgs[i], seed=int(seeds[i]))
            cuts.append(c)
        return cuts

    def batched():
        jet_amd._check(L.jet_partition_batch(ctx._h, count, cfgs, seeds.ctypes.data, voff.ctypes.data,
                                             eoff.ctypes.data, d_rp.data_ptr(), d_col.data_ptr(), d_ew.data_ptr(),
                                             d_vw.data_ptr(), uni.ctypes.data, jet_amd.PTR_DEVICE,
                                             d_part2.data_ptr(), cuts_b.ctypes.data, None))

    for w in range(args.warmup):
        plain()
        batched()
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cuts_p = plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.steps
    print(f"# plain loop {t_plain:.3f}s per pass", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    print(f"# batched {t_batch:.3f}s per pass", file=sys.stderr, flush=True)
    same = bool(np.array_equal(np.array(cuts_p), cuts_b) and torch.equal(d_part, d_part2))
    ctx.close()
    out = {
        "metric": "batched graphs/s (1 GPU)",
        "value": round(count / t_batch, 2),
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_batch, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"batch-{count}", "graphs": count, "vertices": int(voff[-1]),
                   "undirected_edges": int(eoff[-1]) // 2, "workers": args.workers, "seed": args.seed},
        "plain_loop_graphs_per_s": round(count / t_plain, 2),
        "batched_vs_plain": round(t_plain / t_batch, 3),
        "identical_to_plain": same,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def serve_pool(size, seed):
    """Graphs requests draw from: grids and R-MAT (edge factor 8), vertex
    counts log-uniform in [10K, 4M]."""
    rs = np.random.default_rng(seed)
    pool = []
    for i in range(size):
        n = float(np.exp(rs.uniform(np.log(1e4), np.log(4e6))))
        if i % 2 == 0:
            side = int(round(n ** 0.5))
            g = jet_amd.grid2d(side, side)
        else:
            g = jet_amd.rmat(int(round(np.log2(n))), 8, seed=i)
        g["uniform"] = True
        pool.append(g)
    return pool


def run_serve(args):
    """Serving: Poisson arrivals of partition jobs (each a pooled graph with a
    per-request seed and k), admitted under an HBM budget; latency from
    submission to completion."""
    import torch

    rs = np.random.default_rng(args.seed)
    t_gen = time.time()
    pool = serve_pool(args.pool, args.seed)
    print(f"# pool of {len(pool)} graphs generated in {time.time() - t_gen:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    dpool = []
    for g in pool:
        dpool.append((len(g["rowptr"]) - 1, torch.from_numpy(g["rowptr"]).to(dev), torch.from_numpy(g["col"]).to(dev)))
    torch.cuda.synchronize()
    nreq = args.requests
    gaps = rs.exponential(1.0 / args.rate, nreq)
    which = rs.integers(0, len(pool), nreq)
    ks = rs.choice([8, 16, 32, 64, 128, 256], nreq)
    cfgs = [jet_amd.Config(num_parts=int(k), max_imb_ratio=1.03, coarsening_alg=1) for k in ks]
    parts = [torch.empty(dpool[w][0], dtype=torch.int32, device=dev) for w in which]
    ctx = jet_amd.Context(workers=args.workers, budget=int(args.budget_gb * (1 << 30)))
    tickets = [None] * nreq
    done = {}
    t0 = time.perf_counter()
    next_t = 0.0
    collected = 0
    for r in range(nreq):
        next_t += gaps[r]
        while True:  # collect finished jobs while waiting for the next arrival
            now = time.perf_counter() - t0
            if now >= next_t:
                break
            for q in range(collected, r):
                if q in done or tickets[q] is None:
                    continue
                res = ctx.poll(tickets[q], wait_ms=0)
                if res is not None:
                    done[q] = res
            time.sleep(min(0.001, max(0.0, next_t - now)))
        n, d_rp, d_col = dpool[which[r]]
        tickets[r] = ctx.submit_device(n, d_rp.data_ptr(), d_col.data_ptr(), None, None, parts[r].data_ptr(),
                                       cfgs[r], seed=r, uniform=True)
        if (r + 1) % 100 == 0:
            print(f"# {r + 1}/{nreq} submitted, {len(done)} done", file=sys.stderr, flush=True)
    for q in range(nreq):
        if q not in done:
            done[q] = ctx.poll(tickets[q], wait_ms=-1)
        if (q + 1) % 100 == 0:
            print(f"# {q + 1}/{nreq} collected", file=sys.stderr, flush=True)
    t_all = time.perf_counter() - t0
    lat = np.array([1e3 * (done[q][3].end_s - done[q][3].submit_s) for q in range(nreq)])
    first = min(done[q][3].submit_s for q in range(nreq))
    last = max(done[q][3].end_s for q in range(nreq))
    # invariance: a sample of requests re-run alone on the plain path
    same = True
    for q in range(0, nreq, max(1, nreq // 8)):
        n, d_rp, d_col = dpool[which[q]]
        alone = torch.empty(n, dtype=torch.int32, device=dev)
        c, _ = jet_amd.partition_device(n, d_rp.data_ptr(), d_col.data_ptr(), None, None, alone.data_ptr(), cfgs[q],
                                        seed=q, uniform=True)
        same = same and c == done[q][1] and bool(torch.equal(alone, parts[q]))
    ctx.close()
    out = {
        "metric": "serving p99 latency ms (1 GPU)",
        "value": round(float(np.percentile(lat, 99)), 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": nreq,
        "warmup": 0,
        "ms_per_step": round(1e3 * t_all / nreq, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "serve", "requests": nreq, "rate_per_s": args.rate, "pool": len(pool),
                   "workers": args.workers, "budget_gb": args.budget_gb, "seed": args.seed},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "graphs_per_s": round(nreq / (last - first), 3),
        "same_result_alone": same,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="stencil3d-128", choices=sorted(WORKLOADS) + ["batch", "serve"])
    ap.add_argument("--graphs", type=int, default=512, help="batch: number of graphs")
    ap.add_argument("--workers", type=int, default=None,
                    help="batch/serve: concurrent GPU workers (default 16)")
    ap.add_argument("--requests", type=int, default=400, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: Poisson arrival rate per second")
    ap.add_argument("--pool", type=int, default=16, help="serve: distinct graphs requests draw from")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget for admitted jobs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (reads 1 GiB, writes 1 GiB) three times, for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU-baseline runs after one warm-up")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="set a result-neutral tuning parameter before the run (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (each box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    print("# importing torch", file=sys.stderr, flush=True)
    import torch

    # the library is loaded after torch, so both use the HIP runtime torch loaded
    for kv in args.tune:
        name, _, val = kv.partition("=")
        jet_amd.tuning_set(name, int(val))
        print(f"# tuning {name} = {jet_amd.tuning_get(name)}", file=sys.stderr, flush=True)

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 << 28, "copy_bytes_written": 4 << 28, "launches": 3}), flush=True)
        return
    if args.workload == "batch":
        if args.workers is None:
            args.workers = 16
        return run_batch(args)
    if args.workload == "serve":
        if args.workers is None:
            args.workers = 16
        return run_serve(args)
    wl = WORKLOADS[args.workload]
    t_gen = time.time()
    g = make_graph(args.workload)
    n = len(g["rowptr"]) - 1
    m = len(g["col"])
    deg = np.diff(g["rowptr"])
    max_deg, isolated = int(deg.max()), int((deg == 0).sum())
    print(f"# workload {args.workload}: n={n} m={m} max degree {max_deg} isolated {isolated} "
          f"(generated in {time.time() - t_gen:.1f}s)", file=sys.stderr, flush=True)
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

    roofline, families, e2e = kernel_rooflines(step, args, ms, torch)

    part = d_part.cpu().numpy()

    # secondary figure: the same partition from host arrays, so the call
    # copies the graph to HBM and the partition back (PCIe included); not `value`
    t2 = time.perf_counter()
    hpart, hcut = jet_amd.partition_csr(g, cfg, seed=args.seed, uniform=uniform)
    host_dt = time.perf_counter() - t2
    pcie = {"value": round(edges / host_dt, 1), "unit": "edges/s", "ms": round(1e3 * host_dt, 3),
            "same_partition": bool(np.array_equal(hpart, part) and hcut == cut)}
    sizes = np.bincount(part, weights=g["vw"], minlength=wl["k"])
    opt = (int(g["vw"].sum()) + wl["k"] - 1) // wl["k"]
    balanced = bool(sizes.max() <= int(opt * wl["imb"]))

    cpu = None
    cut_ratio = None
    blk = block_partition(args.workload, wl["k"])
    block_cut = edge_cut(g, blk) if blk is not None else None
    if "golden" in wl:
        # the oracle's partition of this exact graph, recorded once
        with open(os.path.join(ROOT, "tests", "golden", wl["golden"])) as f:
            gold = json.load(f)
        cut_ratio = round(cut / gold["cut"], 6)
        same = hashlib.sha256(part.astype("<i4").tobytes()).hexdigest() == gold["part_sha256"]
        print(f"# oracle golden cut {gold['cut']}; gpu partition identical to it: {same}", file=sys.stderr,
              flush=True)
    if not args.no_cpu_baseline:
        import oracle

        # the OpenMP build of the CPU restatement (identical results to the
        # serial build), on this process's CPU share, bound to cores. The
        # binding is set only now: libgomp pins the thread that loads it, so
        # set before torch (whose own OpenMP runtime loads with it) it would
        # pin the main thread, and the host threads of every partition call,
        # to one core.
        threads = cpu_threads()
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        OL, threads = oracle.omp_lib(threads)
        sname = wl.get("cpu_sample")
        sg = make_graph(sname) if sname else g
        sedges = len(sg["col"]) // 2
        what = f"{SAMPLES[sname]} ({sedges} edges)" if sname else f"the same {args.workload} graph, same seed"
        print(f"# cpu baseline: OpenMP restatement on {threads} threads, "
              f"{'the ' + sname + ' sample' if sname else 'the full graph'}, 1 warm-up + median of "
              f"{args.cpu_runs}", file=sys.stderr, flush=True)
        times = []
        opart = ocut = None
        for r in range(args.cpu_runs + 1):
            t1 = time.perf_counter()
            opart, ocut, _ = oracle.partition(sg, wl["k"], wl["imb"], alg=wl["alg"], seed=args.seed, L=OL)
            if r > 0:
                times.append(time.perf_counter() - t1)
        cdt = float(np.median(times))
        cpu = {"value": round(sedges / cdt, 1), "unit": "edges/s", "cores": threads, "kind": "port",
               "build": "oracle/libjet_oracle_omp.so (OpenMP restatement of jet_host; same results as the serial "
                        "oracle)",
               "host": host_cpu(),
               "sample": f"full partitions of {what}: 1 warm-up, then the median of {args.cpu_runs} "
                         f"({cdt:.2f}s; runs {', '.join(f'{t:.2f}' for t in times)}s)"}
        if not sname:
            cut_ratio = round(cut / ocut, 6) if ocut else None
            print(f"# cpu baseline {cdt:.2f}s cut {ocut}; gpu/oracle partitions identical: "
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
                   "max_degree": max_deg, "isolated_vertices": isolated,
                   "k": wl["k"], "imbalance": wl["imb"], "coarsening_alg": wl["alg"], "seed": args.seed},
        "edge_cut": cut,
        "cut_vs_oracle": cut_ratio,
        "block_partition_cut": block_cut,
        "cut_vs_block_partition": round(cut / block_cut, 4) if block_cut else None,
        "balanced": balanced,
        "phases_s": {"coarsen": round(st.coarsen_s, 4), "init": round(st.init_partition_s, 4),
                     "refine": round(st.refine_s, 4), "levels": st.num_levels} if st else None,
        "roofline": roofline,
        "kernel_families": families,
        "end_to_end": e2e,
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
