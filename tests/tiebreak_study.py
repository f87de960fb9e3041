"""Cut effect of the tie-break between parts of equal connectivity, on the CPU
oracle: the lowest-part-id rule (the GPU path's) against the first maximum in
table-slot order (the reference's loops, jet_refiner.hpp:226-233 and :603-611,
with the Serial backend's slot layout). Prints per-seed cuts and the geometric
mean ratio for the 256^2 grid (k=8, 2-hop matching) and the 128^3 stencil
(k=64, HEC), 10 seeds each.

Run from the repository root: python tests/tiebreak_study.py [--small] [--grid-seeds N]
"""
import math
import os
import sys

sys.path[:0] = [os.path.join(os.path.dirname(os.path.abspath(__file__)), p) for p in ("", "../oracle")]

import graphs  # noqa: E402
import oracle  # noqa: E402


def study(name, g, k, imb, alg, seeds):
    ratios = []
    for s in seeds:
        oracle.set_slot_order_ties(False)
        _, c_low, _ = oracle.partition(g, k, imb, alg=alg, seed=s)
        oracle.set_slot_order_ties(True)
        _, c_slot, _ = oracle.partition(g, k, imb, alg=alg, seed=s)
        oracle.set_slot_order_ties(False)
        ratios.append(c_low / c_slot)
        print(f"{name} seed {s}: lowest-id {c_low}  slot-order {c_slot}  ratio {c_low / c_slot:.4f}", flush=True)
    gm = math.exp(sum(math.log(r) for r in ratios) / len(ratios))
    print(f"{name}: geomean cut ratio lowest-id / slot-order over {len(seeds)} seeds = {gm:.4f}", flush=True)
    return gm


if __name__ == "__main__":
    seeds = range(10)
    gs = int(sys.argv[sys.argv.index("--grid-seeds") + 1]) if "--grid-seeds" in sys.argv else 10
    study("grid2d-256 k=8 2-hop", graphs.grid2d(256, 256), 8, 1.03, 0, range(gs))
    if "--small" not in sys.argv:
        study("stencil3d-128 k=64 HEC", graphs.stencil3d(128, 128, 128, seed=0), 64, 1.03, 1, seeds)
