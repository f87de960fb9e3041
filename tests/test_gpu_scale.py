"""Full-size parity: the two single-graph workloads partitioned on the GPU
equal the CPU oracle's partitions. The 3-D stencil 128^3 is checked against
the oracle directly (a few seconds of CPU); R-MAT scale 22 against the
committed golden record tests/golden/rmat22_k256_golden.json (written by
tests/golden/make_rmat22_golden.py)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import jet_amd
import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_stencil128_k64_equals_oracle():
    say("stencil3d-128 k=64 HEC: generating")
    g = jet_amd.stencil3d(128, 128, 128, seed=0, wmax=10)
    cfg = jet_amd.Config(num_parts=64, max_imb_ratio=1.03, coarsening_alg=1)
    part, cut, st = jet_amd.partition_csr(g, cfg, seed=0, stats=True)
    say("gpu cut", cut, "oracle running")
    ep, ec, elv = oracle.partition(g, 64, 1.03, alg=1, seed=0)
    assert cut == ec
    assert np.array_equal(part, ep)
    assert st.num_levels == len(elv)
    for L in range(st.num_levels):
        assert 2 * st.levels[L].edge_cut == elv[L][0] and st.levels[L].total_iterations == elv[L][4]
    sizes = np.bincount(part, minlength=64)
    assert sizes.max() <= int(((len(part) + 63) // 64) * 1.03)


def test_rmat22_k256_equals_golden():
    gold = json.load(open(os.path.join(HERE, "golden", "rmat22_k256_golden.json")))
    say("rmat-22 k=256 HEC: generating")
    g = jet_amd.rmat(22, 16, seed=0)
    assert len(g["rowptr"]) - 1 == gold["n"] and len(g["col"]) == gold["m"]
    cfg = jet_amd.Config(num_parts=256, max_imb_ratio=1.10, coarsening_alg=1)
    part, cut, st = jet_amd.partition_csr(g, cfg, seed=0, uniform=True, stats=True)
    say("gpu cut", cut, "golden", gold["cut"])
    assert cut == gold["cut"]
    assert hashlib.sha256(np.ascontiguousarray(part, dtype=np.int32).tobytes()).hexdigest() == gold["part_sha256"]
    assert st.num_levels == len(gold["levels"])
    for L, r in enumerate(gold["levels"]):
        lv = st.levels[L]
        assert [2 * lv.edge_cut, lv.num_vertices, lv.num_edges, lv.total_iterations, lv.lp_iterations] == \
            [r[0], r[2], r[3], r[4], r[5]], L
    sizes = np.bincount(part, minlength=256)
    assert sizes.max() <= int(((len(part) + 255) // 256) * 1.10)


def test_rmat21_k256_same_partition_every_run():
    """Three partitions of R-MAT scale 21 (k=256) in one process are identical:
    the multilevel initial partition refines host-built levels with thousands
    of hub rows, and the later calls reuse the device memory of the first."""
    say("rmat-21 k=256 HEC: generating")
    g = jet_amd.rmat(21, 16, seed=0)
    cfg = jet_amd.Config(num_parts=256, max_imb_ratio=1.10, coarsening_alg=1)
    runs = []
    for rep in range(3):
        part, cut = jet_amd.partition_csr(g, cfg, seed=0, uniform=True)
        digest = hashlib.sha256(np.ascontiguousarray(part, dtype=np.int32).tobytes()).hexdigest()
        say("run", rep, "cut", cut, digest[:16])
        runs.append((cut, digest))
    assert runs[0] == runs[1] == runs[2]
