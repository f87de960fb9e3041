"""The multilevel initial partition of large coarsest graphs (initpart.h:
host coarsening of the coarsest graph, recursive bisection of its smallest
level, Jet refinement of each level on the device) against the oracle, which
runs the same scheme with its CPU refinement: partitions, cuts and per-level
records must be identical. The hub-and-leaf graphs reach the weighted-HEC
coarsening fixed point at once, so their coarsest graph is the input itself
(over a million adjacency entries; with leaf pairs, one level later)."""
import sys

import numpy as np
import pytest

import graphs
import jet_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    ("hubs 60 x leaves 12000, k=64", lambda: graphs.heavy_bipartite(60, 12000, seed=5, density=0.8), 64, 1.10, 2),
    ("hubs 60 x leaves 20000 with leaf pairs, k=32",
     lambda: graphs.heavy_bipartite(60, 20000, seed=6, density=0.9, leaf_pairs=True), 32, 1.03, 3),
]


@pytest.mark.parametrize("name,mk,k,imb,seed", CASES, ids=[c[0] for c in CASES])
def test_multilevel_initial_partition_equals_oracle(name, mk, k, imb, seed):
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    g = mk()
    say("multilevel init", name)
    cfg = jet_amd.Config(num_parts=k, max_imb_ratio=imb, coarsening_alg=1)
    part, cut, st = jet_amd.partition_csr(g, cfg, seed=seed, stats=True)
    epart, ecut, levels = oracle.partition(g, k, imb, alg=1, seed=seed, repeat_shortcut=True)
    assert cut == ecut
    assert np.array_equal(part, epart)
    assert st.num_levels == len(levels)
    for L in range(st.num_levels):
        assert st.levels[L].num_vertices == levels[L][2]
        assert st.levels[L].total_iterations == levels[L][4]
        assert 2 * st.levels[L].edge_cut == levels[L][0]
    assert cut == oracle.total_cut(g, part) // 2
