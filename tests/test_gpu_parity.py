"""GPU parity: every hot-path stage through the C-ABI against the CPU oracle
on the same inputs, bit-exact (integer work), then end-to-end partitions."""
import os
import subprocess
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


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


GRAPHS = [
    ("grid", lambda: graphs.grid2d(23, 19)),
    ("wgrid", lambda: graphs.with_weights(graphs.grid2d(30, 17), seed=3, wmax=5, vwmax=3)),
    ("stencil", lambda: graphs.stencil3d(10, 9, 8, seed=2)),
    ("rmat", lambda: graphs.rmat(11, 8, seed=1)),
    ("wrmat", lambda: graphs.with_weights(graphs.rmat(10, 6, seed=5), seed=1, wmax=3, vwmax=4)),
    ("star", lambda: graphs.star(5000)),
    ("clique", lambda: graphs.clique(70)),
    ("path", lambda: graphs.path(300)),
    ("random", lambda: graphs.random_graph(2000, 9000, seed=4, wmax=6)),
]


def test_gain_bucket_gpu_equals_oracle():
    rs = np.random.default_rng(0)
    g = np.concatenate([rs.integers(-10**7, 10**7, 20000), rs.integers(-300, 300, 20000), [0, -1, 1, -(10**9)]])
    w = np.concatenate([rs.integers(1, 100, 20000), rs.integers(1, 5, 20000), [1, 1, 1, 1]])
    got = jet_amd.gain_bucket_batch(g, w)
    exp = np.array([oracle.gain_bucket(int(a), int(b)) for a, b in zip(g, w)])
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("name,mk", GRAPHS)
def test_contract_bit_exact(name, mk):
    g = mk()
    n = len(g["rowptr"]) - 1
    rs = np.random.default_rng(11)
    for nc in sorted({1, 3, max(1, n // 50), max(1, n // 3), n}):
        say("contract", name, n, nc)
        mp = rs.integers(0, nc, n).astype(np.int32)
        mp[:nc] = np.arange(nc)
        rs.shuffle(mp)
        got = jet_amd.contract(g, mp, nc)
        exp = oracle.contract(g, mp, nc)
        for key in ("rowptr", "col", "ew", "vw"):
            assert np.array_equal(got[key], exp[key]), (name, nc, key)


def test_contract_hub_spill_path():
    g = graphs.with_weights(graphs.rmat(13, 16, seed=2), seed=7, wmax=9)
    n = len(g["rowptr"]) - 1
    deg = np.diff(g["rowptr"])
    mp = np.arange(n, dtype=np.int32) // 2
    order = np.argsort(-deg)
    mp[order[:64]] = 0  # one huge aggregate of the 64 largest hubs
    _, mp = np.unique(mp, return_inverse=True)
    mp = mp.astype(np.int32)
    nc = int(mp.max()) + 1
    say("contract spill", n, nc)
    got = jet_amd.contract(g, mp, nc)
    exp = oracle.contract(g, mp, nc)
    for key in ("rowptr", "col", "ew", "vw"):
        assert np.array_equal(got[key], exp[key]), key


@pytest.mark.parametrize("name,mk", GRAPHS)
def test_hec_heavy_neighbours_bit_exact(name, mk):
    g = mk()
    say("hec_hn", name)
    for uniform in (False, True):
        assert np.array_equal(jet_amd.hec_hn(g, uniform, 5, 2), oracle.hec_hn(g, uniform, 5, 2))


@pytest.mark.parametrize("name,mk", GRAPHS)
def test_map_construct_bit_exact(name, mk):
    g = mk()
    say("map_construct", name)
    for uniform in (False, True):
        hn = oracle.hec_hn(g, uniform, 9, 0)
        got, nc = jet_amd.map_construct(hn)
        exp, enc = oracle.map_construct(hn)
        assert nc == enc and np.array_equal(got, exp)


@pytest.mark.parametrize("name,mk", GRAPHS)
@pytest.mark.parametrize("alg", [0, 1, 2])
def test_coarsen_map_bit_exact(name, mk, alg):
    g = mk()
    say("coarsen_map", name, alg)
    for seed in (0, 3):
        got, nc = jet_amd.coarsen_map(g, alg, g["uniform"], seed, level=1)
        exp, enc = oracle.coarsen_map(g, alg, g["uniform"], seed, level=1)
        assert nc == enc and np.array_equal(got, exp), (name, alg, seed)


def canonical_tables(off, sp, sv, sz):
    rows = []
    for i in range(len(sz)):
        seg = list(zip(sp[off[i]:off[i] + sz[i]].tolist(), sv[off[i]:off[i] + sz[i]].tolist()))
        rows.append(sorted((p, v) for p, v in seg if p >= 0))
    return rows


@pytest.mark.parametrize("name,mk", GRAPHS)
@pytest.mark.parametrize("k", [2, 7, 64])
def test_conn_tables_bit_exact(name, mk, k):
    g = mk()
    n = len(g["rowptr"]) - 1
    part = np.random.default_rng(k).integers(0, k, n).astype(np.int32)
    say("conn", name, k)
    off, sp, sv, sz = jet_amd.conn_tables(g, part, k)
    eoff, ep, ev, esz = oracle.conn_build(g, part, k)
    assert np.array_equal(sz, esz)
    got = canonical_tables(off, sp, sv, sz)
    for i in range(n):
        assert got[i] == list(zip(ep[eoff[i]:eoff[i + 1]].tolist(), ev[eoff[i]:eoff[i + 1]].tolist())), i


@pytest.mark.parametrize("k", [300, 5000, 8192])
def test_conn_tables_large_k(k):
    """Workgroup-built rows with table capacity up to k = 8192 slots (two
    tables of 8192 int2 in LDS: 128 KB of the 160 KB per workgroup)."""
    rs = np.random.default_rng(k)
    hubs, n = 3, 30000
    u = np.concatenate([np.full(n - hubs, h) for h in range(hubs)] + [rs.integers(hubs, n, 20000)])
    v = np.concatenate([np.arange(hubs, n) for _ in range(hubs)] + [rs.integers(hubs, n, 20000)])
    g = graphs.from_edges(n, u, v, w=rs.integers(1, 9, len(u)))
    part = rs.integers(0, k, n).astype(np.int32)
    say("conn-large-k", k)
    off, sp, sv, sz = jet_amd.conn_tables(g, part, k)
    eoff, ep, ev, esz = oracle.conn_build(g, part, k)
    assert np.array_equal(sz, esz)
    got = canonical_tables(off, sp, sv, sz)
    for i in range(n):
        assert got[i] == list(zip(ep[eoff[i]:eoff[i + 1]].tolist(), ev[eoff[i]:eoff[i + 1]].tolist())), i


@pytest.mark.parametrize("n", [4096, 4097, 16384, 16385, 30011, 32768, 32769, 70001])
def test_conn_offsets_are_capacity_prefix_sums(n):
    """Table offsets = exclusive prefix sums of min(degree, k), across the
    scan's one-tile, single-workgroup multi-tile and multi-workgroup sizes."""
    rs = np.random.default_rng(n)
    m = 3 * n
    g = graphs.from_edges(n, rs.integers(0, n, m), rs.integers(0, n, m), w=rs.integers(1, 9, m))
    k = 5
    part = rs.integers(0, k, n).astype(np.int32)
    say("conn-offsets", n)
    off, _, _, _ = jet_amd.conn_tables(g, part, k)
    exp = np.concatenate([[0], np.cumsum(np.minimum(np.diff(g["rowptr"]), k))])
    assert np.array_equal(off, exp)


def rand_part(n, k, seed, skew=0.0):
    rs = np.random.default_rng(seed)
    pr = np.exp(-skew * np.arange(k))
    return rs.choice(k, n, p=pr / pr.sum()).astype(np.int32)


@pytest.mark.parametrize("name,mk", GRAPHS)
@pytest.mark.parametrize("kind,skew", [(0, 0.0), (1, 0.7), (1, 2.5), (2, 0.7), (2, 2.5)])
def test_refine_step_bit_exact(name, mk, kind, skew):
    g = mk()
    n = len(g["rowptr"]) - 1
    for k in (4, 32):
        part = rand_part(n, k, 17 + kind, skew)
        say("step", name, kind, skew, k)
        mv, de = jet_amd.refine_step(kind, g, part, k, 1.03, 0.75)
        emv, ede = oracle.refine_step(kind, g, part, k, 1.03, 0.75)
        assert np.array_equal(mv, emv) and np.array_equal(de, ede), (name, kind, k)


# Small levels are refined by one workgroup running the whole iteration loop
# (wg_max_n); "multi" forces the multi-kernel path on every level.
PATHS = [("wg", 16384), ("multi", 0)]


@pytest.fixture(params=PATHS, ids=[p[0] for p in PATHS])
def refine_path(request):
    with jet_amd.tuned(wg_max_n=request.param[1]):
        yield request.param[0]


@pytest.mark.parametrize("name,mk", GRAPHS)
@pytest.mark.parametrize("skew", [0.0, 1.5])
def test_refine_level_bit_exact(name, mk, skew, refine_path):
    g = mk()
    n = len(g["rowptr"]) - 1
    for k in (3, 16):
        part = rand_part(n, k, 23, skew)
        say("refine_level", refine_path, name, skew, k)
        p, rec = jet_amd.refine_level(g, part, k, 1.03, uniform=g["uniform"])
        ep, erec = oracle.refine_level(g, part, k, 1.03, uniform=g["uniform"])
        assert np.array_equal(rec, erec), (rec, erec)
        assert np.array_equal(p, ep)


def test_project_cut_sizes():
    g = graphs.rmat(12, 8, seed=3)
    n = len(g["rowptr"]) - 1
    mp = (np.arange(n) * 7 % 1000).astype(np.int32)
    coarse = np.random.default_rng(1).integers(0, 9, 1000).astype(np.int32)
    assert np.array_equal(jet_amd.project(mp, coarse), coarse[mp])
    part = coarse[mp]
    assert jet_amd.edge_cut2(g, part) == oracle.total_cut(g, part)
    vw = np.random.default_rng(2).integers(1, 5, n).astype(np.int32)
    assert np.array_equal(jet_amd.part_sizes(vw, part, 9), np.bincount(part, weights=vw, minlength=9).astype(np.int64))


E2E = [
    ("C1-grid256", lambda: graphs.grid2d(256, 256), 8, 1.03, 0),
    ("grid-hec", lambda: graphs.grid2d(100, 80), 16, 1.03, 1),
    ("stencil-hec", lambda: graphs.stencil3d(24, 24, 24, seed=0), 64, 1.03, 1),
    ("stencil-match", lambda: graphs.stencil3d(20, 20, 20, seed=1), 8, 1.01, 2),
    ("rmat-hec", lambda: graphs.rmat(14, 16, seed=0), 256, 1.10, 1),
    ("rmat-2hop", lambda: graphs.rmat(13, 8, seed=3), 32, 1.03, 0),
    ("wrmat-hec", lambda: graphs.with_weights(graphs.rmat(12, 8, seed=4), seed=2, wmax=4, vwmax=4), 16, 1.03, 1),
    ("wrmat-tight", lambda: graphs.with_weights(graphs.rmat(13, 8, seed=1), seed=1, wmax=4), 8, 1.01, 1),
    ("rmat16-tight", lambda: graphs.rmat(16, 16, seed=5), 64, 1.01, 1),
    ("hec-fixed-point", lambda: graphs.heavy_bipartite(6, 600, seed=3), 4, 1.10, 1),
    ("hec-fixed-point-late", lambda: graphs.heavy_bipartite(8, 900, seed=4, leaf_pairs=True), 4, 1.10, 1),
]


@pytest.mark.parametrize("name,mk,k,imb,alg", E2E)
def test_partition_end_to_end_bit_exact(name, mk, k, imb, alg, refine_path):
    g = mk()
    n = len(g["rowptr"]) - 1
    for seed in (0, 1):
        say("e2e", refine_path, name, seed)
        cfg = jet_amd.Config(num_parts=k, max_imb_ratio=imb, coarsening_alg=alg)
        part, cut, st = jet_amd.partition_csr(g, cfg, seed=seed, stats=True)
        epart, ecut, elv = oracle.partition(g, k, imb, alg=alg, seed=seed)
        assert cut == ecut
        assert np.array_equal(part, epart)
        assert st.num_levels == len(elv)
        for L in range(st.num_levels):
            assert st.levels[L].num_vertices == elv[L][2]
            assert st.levels[L].total_iterations == elv[L][4]
            assert 2 * st.levels[L].edge_cut == elv[L][0]
        assert cut == oracle.total_cut(g, part) // 2
        sizes = np.bincount(part, weights=g["vw"], minlength=k)
        opt = (int(g["vw"].sum()) + k - 1) // k
        assert sizes.max() <= int(opt * imb) or elv[-1][1] > 0


def test_attempts_keep_best_cut():
    g = graphs.grid2d(64, 64)
    cuts = []
    for seed in range(3):
        _, c = jet_amd.partition_csr(g, jet_amd.Config(num_parts=4), seed=seed)
        cuts.append(c)
    _, best, st = jet_amd.partition_csr(g, jet_amd.Config(num_parts=4, num_iter=3), seed=0, stats=True)
    assert best == min(cuts) and st.attempts == 3


def test_invalid_inputs_rejected():
    g = graphs.grid2d(8, 8)
    bad = dict(g)
    bad["col"] = g["col"].copy()
    bad["col"][3] = 1000
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.partition_csr(bad, jet_amd.Config(num_parts=2))
    assert e.value.code == jet_amd.JET_EINVAL
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.partition_csr(g, jet_amd.Config(num_parts=0))
    assert e.value.code == jet_amd.JET_EINVAL
    neg = dict(g)
    neg["ew"] = g["ew"].copy()
    neg["ew"][0] = -1
    with pytest.raises(jet_amd.JetError):
        jet_amd.partition_csr(neg, jet_amd.Config(num_parts=2))


def _fnv(part):
    h = 1469598103934665603
    for v in part.tolist():
        h = ((h ^ (v & 0xFFFFFFFF)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_jet_h_nine_entry_points(tmp_path):
    """include/jet.h: the nine reference names (header/jet.h:46-107), host and
    device variants, all widths, give the partition the C-ABI gives."""
    from test_capi_cpu import build_jet_h_example
    exe = build_jet_h_example(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln.split() for ln in r.stdout.strip().splitlines()]
    assert [ln[0] for ln in lines] == ["partition_host", "partition_serial", "partition_host_big",
                                       "partition_serial_big", "partition_host_biggest", "partition_serial_biggest",
                                       "partition", "partition_big", "partition_biggest"]
    g = graphs.grid2d(96, 64)
    rp, col = g["rowptr"], g["col"]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    g["ew"] = (1 + (7 * np.minimum(rows, col) + np.maximum(rows, col)) % 4).astype(np.int32)
    part, cut = jet_amd.partition_csr(g, jet_amd.Config(num_parts=8, coarsening_alg=1), seed=5, uniform=False)
    for ln in lines:
        assert int(ln[1]) == cut and int(ln[2]) == _fnv(part)


def _numpy_quality(g, part, k):
    rp, col, ew = g["rowptr"], g["col"], g["ew"]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    ext = part[rows] != part[col]
    cut = int(ew[ext].sum()) // 2
    cut_pp = np.bincount(part[rows][ext], weights=ew[ext], minlength=k).astype(np.int64)
    pairs = np.unique(np.stack([rows[ext], part[col][ext]]), axis=1)
    sizes = np.bincount(part, minlength=k)
    opt = (len(part) + k - 1) // k
    return cut, pairs.shape[1], int(cut_pp.max()), sizes.max() / opt, sizes.min() / opt


@pytest.mark.parametrize("k", [2, 16, 300])
def test_partition_quality_metrics(k):
    g = graphs.with_weights(graphs.rmat(12, 8, seed=6), seed=6, wmax=7)
    part = np.random.default_rng(k).integers(0, k, len(g["rowptr"]) - 1).astype(np.int32)
    q = jet_amd.partition_quality(g, part, k)
    cut, comm, mpc, mx, mn = _numpy_quality(g, part, k)
    assert (q.edge_cut, q.comm_size, q.max_part_cut) == (cut, comm, mpc)
    assert abs(q.max_imb - mx) < 1e-12 and abs(q.min_imb - mn) < 1e-12


@pytest.mark.parametrize("k", [1, 8, 64, 100])
def test_cut_heatmap(k):
    """part_stat.hpp:116-129: heat[p, q] = weight of the edges from part p to
    part q != p, against a bincount over the edge list; row sums give the
    max part cut. k = 8 and 64 take the LDS path, k = 100 the global one."""
    g = graphs.with_weights(graphs.rmat(12, 8, seed=9), seed=4, wmax=9)
    n = len(g["rowptr"]) - 1
    part = np.random.default_rng(k + 1).integers(0, k, n).astype(np.int32)
    heat = jet_amd.cut_heatmap(g, part, k)
    src = np.repeat(np.arange(n), np.diff(g["rowptr"]))
    pu, pv = part[src], part[g["col"]]
    cut = pu != pv
    exp = np.bincount((pu[cut].astype(np.int64) * k + pv[cut]), weights=g["ew"][cut], minlength=k * k)
    assert np.array_equal(heat.ravel(), exp.astype(np.int64))
    assert np.array_equal(heat, heat.T) and not np.diag(heat).any()
    assert heat.sum(axis=1).max() == jet_amd.partition_quality(g, part, k).max_part_cut
    assert heat.sum() == 2 * jet_amd.partition_quality(g, part, k).edge_cut


def test_part_ids_out_of_range_rejected():
    g = graphs.grid2d(16, 16)
    part = np.zeros(256, dtype=np.int32)
    part[7] = 4
    for call in (lambda: jet_amd.part_sizes(np.ones(256, dtype=np.int32), part, 4),
                 lambda: jet_amd.partition_quality(g, part, 4),
                 lambda: jet_amd.cut_heatmap(g, part, 4),
                 lambda: jet_amd.conn_tables(g, part, 4),
                 lambda: jet_amd.refine_level(g, part, 4, 1.03)):
        with pytest.raises(jet_amd.JetError) as e:
            call()
        assert e.value.code == jet_amd.JET_EINVAL


def test_map_and_hn_entries_out_of_range_rejected():
    # jet_contract indexes per-coarse-vertex arrays by map[i] and
    # jet_map_construct indexes claim arrays by hn[u]: both are range-checked
    # on the device before any kernel uses them
    g = graphs.grid2d(16, 16)
    n = 256
    mp = (np.arange(n) // 2).astype(np.int32)
    for bad in (128, -1, 1 << 30):
        m2 = mp.copy()
        m2[37] = bad
        say("contract with map entry", bad)
        with pytest.raises(jet_amd.JetError) as e:
            jet_amd.contract(g, m2, 128)
        assert e.value.code == jet_amd.JET_EINVAL
    for bad in (n, -5):
        hn = np.arange(n, dtype=np.int32)[::-1].copy()
        hn[11] = bad
        say("map_construct with hn entry", bad)
        with pytest.raises(jet_amd.JetError) as e:
            jet_amd.map_construct(hn)
        assert e.value.code == jet_amd.JET_EINVAL
    # the checks pass valid inputs through unchanged
    cg = jet_amd.contract(g, mp, 128)
    assert len(cg["rowptr"]) == 129


def test_cli_jet_and_pstat(tmp_path):
    from test_io_cpu import write_metis
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jet-partitioner-official-fork_amd")
    g = graphs.with_weights(graphs.grid2d(64, 48), seed=2, wmax=4)
    write_metis(tmp_path / "g.graph", g, True)
    (tmp_path / "c.txt").write_text("1\n8\n3\n1.03\n")
    r = subprocess.run([os.path.join(pkg, "jet"), str(tmp_path / "g.graph"), str(tmp_path / "c.txt"),
                        str(tmp_path / "p.txt"), str(tmp_path / "m.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    best = min(jet_amd.partition_csr(g, jet_amd.Config(num_parts=8, coarsening_alg=1), seed=s)[1] for s in range(3))
    assert f"min edgecut found is {best}" in r.stdout
    part = jet_amd.read_partition(str(tmp_path / "p.txt"), len(g["rowptr"]) - 1)
    assert jet_amd.edge_cut2(g, part) // 2 == best
    assert len((tmp_path / "m.json").read_text().splitlines()) == 3
    r = subprocess.run([os.path.join(pkg, "pstat"), str(tmp_path / "g.graph"), str(tmp_path / "p.txt"), "8"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cut, comm, mpc, mx, mn = _numpy_quality(g, part, 8)
    assert f"Cutsize: {cut}" in r.stdout and f"Comm size: {comm}" in r.stdout and f"Max part cut: {mpc}" in r.stdout
