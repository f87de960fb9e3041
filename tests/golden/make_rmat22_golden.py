"""Writes rmat22_k256_golden.json: the CPU oracle's partition of the R-MAT
scale-22 workload (edge factor 16, a/b/c = .57/.19/.19, seed 0; k = 256,
imbalance 1.10, HEC, partition seed 0) summarised as cut, per-level records
and a SHA-256 of the int32 partition. The oracle runs with its exact
fixed-point shortcuts (shared repeated levels; an unchanged repeated level is
not refined again), which tests check against its literal level loop; the
literal loop would contract the 69.5M-edge fixed-point graph 195 times.
Run from the repository root: python tests/golden/make_rmat22_golden.py"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "jet-partitioner-official-fork_amd")]

import numpy as np  # noqa: E402

import jet_amd  # noqa: E402  (host-side generator only)
import oracle  # noqa: E402


def main():
    t = time.time()
    g = jet_amd.rmat(22, 16, seed=0)
    g["uniform"] = True
    print(f"generated in {time.time() - t:.1f}s", flush=True)
    t = time.time()
    # the OpenMP build of the oracle: the same results as the serial build
    # (tests/test_oracle_omp.py), in a fraction of the time
    part, cut, levels = oracle.partition(g, 256, 1.10, alg=1, seed=0, repeat_shortcut=True, L=oracle.omp_lib()[0])
    print(f"oracle partition in {time.time() - t:.1f}s, cut {cut}, {len(levels)} levels", flush=True)
    out = {
        "workload": "rmat-22 ef16 (.57,.19,.19) seed 0; k=256 imb 1.10 HEC seed 0",
        "n": int(len(g["rowptr"]) - 1),
        "m": int(len(g["col"])),
        "cut": int(cut),
        "part_sha256": hashlib.sha256(np.ascontiguousarray(part, dtype=np.int32).tobytes()).hexdigest(),
        "levels": [[int(x) for x in r[:6]] for r in levels],
    }
    with open(os.path.join(ROOT, "tests", "golden", "rmat22_k256_golden.json"), "w") as f:
        json.dump(out, f)
    print("written", flush=True)


if __name__ == "__main__":
    main()
