"""Pins the oracle's coarsening and contraction."""
import numpy as np
import pytest
import scipy.sparse as sp

import graphs
import oracle
import refsim


def to_scipy(g):
    n = len(g["rowptr"]) - 1
    return sp.csr_matrix((g["ew"].astype(np.int64), g["col"], g["rowptr"]), shape=(n, n))


SMALL = [
    ("grid", lambda: graphs.grid2d(13, 11)),
    ("wgrid", lambda: graphs.with_weights(graphs.grid2d(9, 17), seed=3, wmax=5, vwmax=3)),
    ("stencil", lambda: graphs.stencil3d(6, 5, 4, seed=2)),
    ("rmat", lambda: graphs.rmat(8, 8, seed=1)),
    ("wrmat", lambda: graphs.with_weights(graphs.rmat(8, 6, seed=5), seed=1, wmax=3, vwmax=4)),
    ("star", lambda: graphs.star(40)),
    ("clique", lambda: graphs.clique(12)),
    ("path", lambda: graphs.path(50)),
    ("random", lambda: graphs.random_graph(120, 300, seed=4, wmax=6)),
]


@pytest.mark.parametrize("name,mk", SMALL)
def test_contract_equals_galerkin_product(name, mk):
    g = mk()
    n = len(g["rowptr"]) - 1
    rs = np.random.default_rng(7)
    for nc in (1, max(1, n // 3), n):
        mp = rs.integers(0, nc, n).astype(np.int32)
        mp[:nc] = np.arange(nc)  # onto
        rs.shuffle(mp)
        gc = oracle.contract(g, mp, nc)
        P = sp.csr_matrix((np.ones(n, dtype=np.int64), (np.arange(n), mp)), shape=(n, nc))
        C = (P.T @ to_scipy(g) @ P).tolil()
        C.setdiag(0)
        C = C.tocsr()
        C.eliminate_zeros()
        C.sort_indices()
        assert np.array_equal(gc["rowptr"], C.indptr)
        assert np.array_equal(gc["col"], C.indices)
        assert np.array_equal(gc["ew"], C.data)
        assert np.array_equal(gc["vw"], np.bincount(mp, weights=g["vw"], minlength=nc).astype(np.int32))


@pytest.mark.parametrize("name,mk", SMALL)
def test_hec_heavy_neighbour_rule(name, mk):
    g = mk()
    rp, col, ew, vw = g["rowptr"], g["col"], g["ew"], g["vw"]
    n = len(rp) - 1
    hn = oracle.hec_hn(g, False, 0, 0)
    max_allowed = 6 * int(vw.sum()) // n
    for i in range(n):
        row = range(rp[i], rp[i + 1])
        elig = [j for j in row if vw[col[j]] < max_allowed]
        if len(row) == 0 or vw[i] >= max_allowed or not elig:
            assert hn[i] == i
        else:
            wmax = max(ew[j] for j in elig)
            last = [j for j in elig if ew[j] == wmax][-1]
            assert hn[i] == col[last]
    hu = oracle.hec_hn(g, True, 11, 2)
    for i in range(n):
        d = int(rp[i + 1] - rp[i])
        r = oracle.rng(11, 2, 1, 0, i)
        assert hu[i] == (col[int(rp[i]) + r % d] if d else r % n)


def check_map(mp, nc, n, min_root=False):
    mp = np.asarray(mp)
    assert len(mp) == n and mp.min() == 0 and mp.max() == nc - 1
    assert len(np.unique(mp)) == nc
    if min_root:  # matching roots are the smaller vertex of each pair
        first = np.full(nc, n)
        np.minimum.at(first, mp, np.arange(n))
        assert np.all(np.diff(first) > 0)  # ids dense in increasing root order


def assert_hn_connected(mp, nc, hn):
    n = len(mp)
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for u in range(n):
        if mp[u] == mp[hn[u]]:
            parent[find(u)] = find(int(hn[u]))
    comps = {}
    for u in range(n):
        comps.setdefault(int(mp[u]), set()).add(find(u))
    assert all(len(c) == 1 for c in comps.values())


@pytest.mark.parametrize("name,mk", SMALL)
def test_map_construct_equals_literal_schedule(name, mk):
    g = mk()
    n = len(g["rowptr"]) - 1
    for uniform in (True, False):
        hn = oracle.hec_hn(g, uniform, 5, 0)
        mp, nc = oracle.map_construct(hn)
        ref, rnc = refsim.map_construct([int(x) for x in hn])
        assert nc == rnc and list(mp) == ref
        check_map(mp, nc, n)
        # every aggregate is connected through heavy-neighbour pointer edges
        assert_hn_connected(mp, nc, hn)


@pytest.mark.parametrize("name,mk", SMALL)
@pytest.mark.parametrize("alg", [0, 2])
def test_matching_equals_literal_schedule(name, mk, alg):
    g = mk()
    n = len(g["rowptr"]) - 1
    for seed in (0, 9):
        mp, nc = oracle.coarsen_map(g, alg, g["uniform"], seed, level=1)
        ref, rnc = refsim.coarsen_match(g, g["uniform"], seed, 1, alg == 0)
        assert nc == rnc and list(mp) == ref
        check_map(mp, nc, n, min_root=True)
        sizes = np.bincount(mp)
        assert sizes.max() <= 2
        if alg == 2:  # pure matching: every pair is an edge
            rp, col = g["rowptr"], g["col"]
            for c in np.where(sizes == 2)[0]:
                a, b = np.where(mp == c)[0]
                assert b in col[rp[a]:rp[a + 1]]


@pytest.mark.parametrize("name,mk", SMALL)
def test_hec_map_invariants(name, mk):
    g = mk()
    n = len(g["rowptr"]) - 1
    mp, nc = oracle.coarsen_map(g, 1, g["uniform"], 3, level=0)
    check_map(mp, nc, n)
    if not g["uniform"]:
        vw = g["vw"]
        max_allowed = 6 * int(vw.sum()) // n
        heavy = vw >= max_allowed
        # a heavy vertex is never chosen as a heavy neighbour, so it only grows
        # an aggregate that others join through non-heavy vertices
        hn = oracle.hec_hn(g, False, 3, 0)
        assert np.all(~heavy[hn] | (hn == np.arange(n)))


@pytest.mark.parametrize("pairs", [False, True])
def test_hec_fixed_point_runs_to_level_limit(pairs):
    """Weighted HEC leaves every vertex alone once leaves only see hubs heavier
    than 6x the mean vertex weight (heuristics.hpp:217-243), so the level loop
    stops only at its 200-level limit (contract.hpp:508-527): 201 levels, all
    of one size after the first merge."""
    g = graphs.heavy_bipartite(8, 900, seed=4, leaf_pairs=pairs)
    part, cut, lv = oracle.partition(g, 4, 1.10, alg=1, seed=0)
    assert len(lv) == 201
    sizes = {int(r[2]) for r in lv[:-1]}
    assert len(sizes) == 1
    n = len(g["rowptr"]) - 1
    assert int(lv[-1][2]) == n
    assert (int(lv[0][2]) == n) == (not pairs)
    assert cut == oracle.total_cut(g, part) // 2


@pytest.mark.parametrize("name,mk,k,imb", [
    ("bipartite", lambda: graphs.heavy_bipartite(8, 900, seed=4), 4, 1.10),
    ("bipartite-late", lambda: graphs.heavy_bipartite(8, 900, seed=4, leaf_pairs=True), 4, 1.10),
    ("rmat16", lambda: graphs.rmat(16, 16, seed=5), 64, 1.01),
    ("wrmat13", lambda: graphs.with_weights(graphs.rmat(13, 8, seed=1), seed=1, wmax=4), 8, 1.01),
    ("grid", lambda: graphs.grid2d(60, 50), 8, 1.03),
])
def test_repeat_shortcut_equals_literal_loop(name, mk, k, imb):
    """The oracle's fixed-point shortcuts (levels shared, an unchanged repeated
    level not refined again) give exactly the literal loop's partition, cut
    and per-level records."""
    g = mk()
    for seed in (0, 3):
        p1, c1, l1 = oracle.partition(g, k, imb, alg=1, seed=seed)
        p2, c2, l2 = oracle.partition(g, k, imb, alg=1, seed=seed, repeat_shortcut=True)
        assert c1 == c2 and np.array_equal(p1, p2)
        assert np.array_equal(l1, l2)
