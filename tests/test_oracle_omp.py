"""The OpenMP build of the oracle (the multi-core CPU baseline) computes
exactly what the serial build computes: same partition, cut and per-level
records, on graphs that take every parallel loop (contraction ranges, the
row-wise update_small of >= 4096 moves, update_large rebuilds, weak and strong
rebalancing, 2-hop, matching and HEC coarsening, ultra settings)."""
import numpy as np
import pytest

import graphs
import oracle

CASES = [
    ("stencil 40x40x30 k=64 HEC", lambda: graphs.stencil3d(40, 40, 30, seed=1), 64, 1.03, 1, False),
    ("grid 256x256 k=8 2-hop", lambda: graphs.grid2d(256, 256), 8, 1.03, 0, False),
    ("rmat-14 k=32 HEC", lambda: graphs.rmat(14, 8, seed=3), 32, 1.10, 1, False),
    ("weighted rmat-13 k=16 matching imb 1.01",
     lambda: graphs.with_weights(graphs.rmat(13, 8, seed=5), seed=6, wmax=9, vwmax=4), 16, 1.01, 2, False),
    ("stencil 24x20x18 k=8 ultra", lambda: graphs.stencil3d(24, 20, 18, seed=2), 8, 1.03, 1, True),
]


@pytest.mark.parametrize("name,mk,k,imb,alg,ultra", CASES, ids=[c[0] for c in CASES])
def test_omp_build_equals_serial(name, mk, k, imb, alg, ultra):
    g = mk()
    L, threads = oracle.omp_lib(4)
    assert threads == 4
    for seed in (0, 1):
        p1, c1, r1 = oracle.partition(g, k, imb, alg=alg, seed=seed, ultra=ultra)
        p2, c2, r2 = oracle.partition(g, k, imb, alg=alg, seed=seed, ultra=ultra, L=L)
        assert c1 == c2
        assert np.array_equal(p1, p2)
        assert np.array_equal(r1, r2)


def test_omp_row_wise_update_matches_on_a_big_lp_step():
    # a random start on a 48 K-vertex stencil: the first LP steps move well
    # over 4096 vertices but fewer than n/10, so the row-wise update_small runs
    g = graphs.stencil3d(40, 40, 30, seed=4)
    n = len(g["rowptr"]) - 1
    rs = np.random.default_rng(2)
    # a blocky start with noise: LP moves are many but not n/10
    p0 = ((np.arange(n) // (n // 16)).clip(0, 15)).astype(np.int32)
    flip = rs.random(n) < 0.06
    p0[flip] = rs.integers(0, 16, int(flip.sum()))
    moves, _ = oracle.refine_step(0, g, p0, 16, 1.03)
    assert 4096 <= len(moves) <= n // 10
    L, _ = oracle.omp_lib(4)
    a = oracle.refine_level(g, p0, 16, 1.03)
    orig = oracle._lib
    try:
        oracle._lib = L
        b = oracle.refine_level(g, p0, 16, 1.03)
    finally:
        oracle._lib = orig
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
