"""Coarse-hierarchy dump (binary_dump.hpp:61-97, uncoarsen.hpp:94-111) and
replay (app/import_coarse.cpp) on the GPU.

The dump's content is checked independently of the build's coarsening: level 0
is the input graph, every map is onto the next level, and every coarser level
equals PᵀAP of the previous one without its diagonal (scipy), with summed
vertex weights; the dumped partition is that of the first balanced level.
Replaying the files must give the partition and cut of the run that wrote
them (the replay starts refinement from the same partition, cut and part
sizes), and that run equals the oracle (test_gpu_parity)."""
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import graphs
import jet_amd
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "jet-partitioner-official-fork_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


WIDTHS = {jet_amd.DUMP_32: (np.int32, np.int32, np.int32), jet_amd.DUMP_BIG: (np.int32, np.int64, np.int64),
          jet_amd.DUMP_BIGGEST: (np.int64, np.int64, np.int64)}


def read_hierarchy(path, kind):
    ot, et, vt = WIDTHS[kind]
    data = open(path, "rb").read()
    pos = 0

    def take(dt, count):
        nonlocal pos
        a = np.frombuffer(data, dtype=dt, count=count, offset=pos)
        pos += a.nbytes
        return a.astype(np.int64)

    levels = []
    count = int(take(np.int32, 1)[0])
    for i in range(count):
        n = int(take(ot, 1)[0])
        m = int(take(et, 1)[0])
        L = {"rowptr": take(et, n + 1), "col": take(ot, m), "ew": take(vt, m), "vw": take(vt, n)}
        if i > 0:
            L["map"] = take(ot, levels[-1]["vw"].size)
        levels.append(L)
    assert pos == len(data)
    return levels


def adjacency(L):
    n = len(L["rowptr"]) - 1
    return sp.csr_matrix((L["ew"], L["col"], L["rowptr"]), shape=(n, n))


def check_dump(tmp, g, k, imb, kind):
    levels = read_hierarchy(os.path.join(tmp, "coarse_graphs.out"), kind)
    part = np.fromfile(os.path.join(tmp, "coarse_part.out"), dtype=np.int32)
    fine = levels[0]
    assert np.array_equal(fine["rowptr"], g["rowptr"])
    assert np.array_equal(fine["col"], g["col"])
    assert np.array_equal(fine["ew"], g["ew"])
    assert np.array_equal(fine["vw"], g.get("vw", np.ones(len(g["rowptr"]) - 1, dtype=np.int64)))
    for a, b in zip(levels[:-1], levels[1:]):
        na, nb = len(a["rowptr"]) - 1, len(b["rowptr"]) - 1
        mp = b["map"]
        assert mp.min() >= 0 and mp.max() < nb and np.unique(mp).size == nb
        P = sp.csr_matrix((np.ones(na), (np.arange(na), mp)), shape=(na, nb))
        C = (P.T @ adjacency(a) @ P).tocsr()
        C.setdiag(0)
        C.eliminate_zeros()
        got = adjacency(b)
        assert (abs(C - got)).max() == 0 if got.nnz else C.nnz == 0
        assert np.array_equal(np.bincount(mp, weights=a["vw"], minlength=nb).astype(np.int64), b["vw"])
    last = levels[-1]
    assert part.size == len(last["rowptr"]) - 1 and part.min() >= 0 and part.max() < k
    sizes = np.bincount(part, weights=last["vw"], minlength=k)
    opt = (int(last["vw"].sum()) + k - 1) // k
    assert sizes.max() / opt <= imb
    return levels, part


CASES = [
    ("stencil-hec", lambda: graphs.stencil3d(16, 16, 16, seed=0), 16, 1.03, 1),
    ("wrmat-2hop", lambda: graphs.with_weights(graphs.rmat(12, 8, seed=2), seed=3, wmax=4, vwmax=3), 8, 1.03, 0),
    ("wgrid-match", lambda: graphs.with_weights(graphs.grid2d(60, 50), seed=1, wmax=5), 8, 1.01, 2),
    ("hec-fixed-point-late", lambda: graphs.heavy_bipartite(8, 900, seed=4, leaf_pairs=True), 4, 1.10, 1),
]
ENTRIES = [("csr32", jet_amd.DUMP_32), ("csr_big", jet_amd.DUMP_BIG), ("csr_biggest", jet_amd.DUMP_BIGGEST),
           ("csr", jet_amd.DUMP_BIG)]


def run_entry(entry, g, cfg, seed):
    if entry == "csr":
        part, cut = jet_amd.partition_csr(g, cfg, seed=seed, uniform=False)
        return part, cut
    big = {"csr32": False, "csr_big": True, "csr_biggest": "biggest"}[entry]
    part, cut, _ = jet_amd._partition_named(g, cfg, False, seed, big)
    return part, cut


@pytest.mark.parametrize("name,mk,k,imb,alg", CASES)
@pytest.mark.parametrize("entry,kind", ENTRIES)
def test_dump_then_replay_gives_the_same_partition(tmp_path, monkeypatch, name, mk, k, imb, alg, entry, kind):
    g = mk()
    say("dump", name, entry)
    monkeypatch.chdir(tmp_path)
    cfg = jet_amd.Config(num_parts=k, max_imb_ratio=imb, coarsening_alg=alg)
    part0, cut0 = run_entry(entry, g, cfg, 0)
    cfg.dump_coarse = 1
    part1, cut1 = run_entry(entry, g, cfg, 0)
    assert cut1 == cut0 and np.array_equal(part1, part0)  # dumping changes nothing
    if entry == "csr":
        epart, ecut, _ = oracle.partition(g, k, imb, alg=alg, seed=0)
        assert cut1 == ecut and np.array_equal(part1, epart)
    levels, cpart = check_dump(str(tmp_path), g, k, imb, kind)
    h = jet_amd.Hierarchy(str(tmp_path / "coarse_graphs.out"), kind)
    assert h.levels == len(levels)
    cp = jet_amd.load_coarse_part(str(tmp_path / "coarse_part.out"), h.vertices(h.levels - 1))
    assert np.array_equal(cp, cpart)
    cfg.dump_coarse = 0
    part2, cut2, st = h.uncoarsen(cfg, cp, stats=True)
    assert cut2 == cut1 and np.array_equal(part2, part1)
    assert st.num_levels == h.levels
    h.close()


def test_no_dump_when_no_level_is_balanced(tmp_path, monkeypatch):
    """uncoarsen.hpp:94-111 dumps before the first balanced level only; this
    graph's partition stays at imbalance 1.21 > 1.10 on every level."""
    g = graphs.heavy_bipartite(6, 600, seed=3)
    monkeypatch.chdir(tmp_path)
    cfg = jet_amd.Config(num_parts=4, max_imb_ratio=1.10, coarsening_alg=1)
    cfg.dump_coarse = 1
    part, cut = jet_amd.partition_csr(g, cfg, seed=0)
    epart, ecut, _ = oracle.partition(g, 4, 1.10, alg=1, seed=0)
    assert cut == ecut and np.array_equal(part, epart)
    assert not (tmp_path / "coarse_graphs.out").exists() and not (tmp_path / "coarse_part.out").exists()


def test_replay_of_a_single_level(tmp_path, monkeypatch):
    """A hierarchy of one level: refinement of the given partition only,
    equal to jet_refine_level on the same graph and partition."""
    g = graphs.with_weights(graphs.grid2d(40, 40), seed=2, wmax=3)
    n = len(g["rowptr"]) - 1
    rs = np.random.default_rng(0)
    start = rs.integers(0, 4, n).astype(np.int32)
    with open(tmp_path / "h.out", "wb") as f:
        f.write(np.int32(1).tobytes())
        f.write(np.int32(n).tobytes())
        f.write(np.int32(len(g["col"])).tobytes())
        f.write(np.asarray(g["rowptr"], dtype=np.int32).tobytes())
        f.write(np.asarray(g["col"], dtype=np.int32).tobytes())
        f.write(np.asarray(g["ew"], dtype=np.int32).tobytes())
        f.write(np.ones(n, dtype=np.int32).tobytes())
    h = jet_amd.Hierarchy(str(tmp_path / "h.out"))
    part, cut = h.uncoarsen(jet_amd.Config(num_parts=4, max_imb_ratio=1.03), start)
    epart, erec = jet_amd.refine_level(g, start, 4, 1.03)
    assert np.array_equal(part, epart) and cut == erec[0] // 2
    with pytest.raises(jet_amd.JetError):
        h.uncoarsen(jet_amd.Config(num_parts=4), np.full(n, 4, dtype=np.int32))


def test_batch_rejects_dump():
    g = graphs.grid2d(20, 20)
    cfg = jet_amd.Config(num_parts=4)
    cfg.dump_coarse = 1
    with jet_amd.Context(workers=1) as ctx:
        with pytest.raises(jet_amd.JetError) as e:
            ctx.submit(g, cfg)
        assert e.value.code == jet_amd.JET_EINVAL


def test_export_then_import_cli(tmp_path):
    from test_io_cpu import write_metis
    g = graphs.with_weights(graphs.grid2d(48, 40), seed=5, wmax=4)
    write_metis(tmp_path / "g.graph", g, True)
    (tmp_path / "c.txt").write_text("1 8 1 1.03\n")
    say("cli export")
    r = subprocess.run([os.path.join(PKG, "jet_export"), str(tmp_path / "g.graph"), str(tmp_path / "c.txt"),
                        str(tmp_path / "a.part")], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    exp_cut = int(r.stdout.split("min edgecut found is")[1].split()[0])
    assert (tmp_path / "coarse_graphs.out").exists() and (tmp_path / "coarse_part.out").exists()
    say("cli import")
    r = subprocess.run([os.path.join(PKG, "jet_import"), str(tmp_path / "c.txt"), str(tmp_path / "b.part"),
                        str(tmp_path / "m.json")], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Importing" in r.stdout
    assert int(r.stdout.split("min edgecut found is")[1].split()[0]) == exp_cut
    assert (tmp_path / "a.part").read_text() == (tmp_path / "b.part").read_text()
