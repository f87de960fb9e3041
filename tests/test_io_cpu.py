"""File formats of the reference's drivers (app/io.hpp:50-244) and the
command-line tools' argument and error handling, on the CPU."""
import os
import subprocess

import numpy as np
import pytest

import graphs
import jet_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "jet-partitioner-official-fork_amd")


def write_metis(path, g, weights, trailing="\n", comments=True):
    rp, col, ew = g["rowptr"], g["col"], g["ew"]
    n = len(rp) - 1
    lines = ["% a METIS graph"] if comments else []
    lines.append(f"{n} {len(col) // 2}" + (" 1" if weights else ""))
    for i in range(n):
        row = []
        for j in range(rp[i], rp[i + 1]):
            row.append(str(col[j] + 1))
            if weights:
                row.append(str(ew[j]))
        lines.append(" ".join(row))
    with open(path, "w") as f:
        f.write("\n".join(lines) + trailing)


@pytest.mark.parametrize("weights", [False, True])
@pytest.mark.parametrize("trailing", ["\n", "", "\n\n\n"])
def test_metis_round_trip(tmp_path, weights, trailing):
    g = graphs.with_weights(graphs.grid2d(9, 7), seed=3, wmax=5)
    p = tmp_path / "g.graph"
    write_metis(p, g, weights, trailing)
    h = jet_amd.load_metis(str(p))
    assert np.array_equal(h["rowptr"], g["rowptr"]) and np.array_equal(h["col"], g["col"])
    assert np.array_equal(h["ew"], g["ew"] if weights else np.ones(len(g["col"]), dtype=np.int32))
    assert h["uniform"] == (not weights)


def test_metis_isolated_vertices_and_comment_lines(tmp_path):
    p = tmp_path / "g.graph"
    p.write_text("%c\n5 2\n2\n1\n\n% inner comment\n5\n4\n")
    h = jet_amd.load_metis(str(p))
    assert list(h["rowptr"]) == [0, 1, 2, 2, 3, 4]
    assert list(h["col"]) == [1, 0, 4, 3]


@pytest.mark.parametrize("text,code", [
    ("3 2 10\n1 2\n1\n1\n", jet_amd.JET_EUNSUPPORTED),  # vertex weights
    ("3 2 0 2\n2\n1\n\n", jet_amd.JET_EUNSUPPORTED),    # ncon
    ("3 2\n2\n1\n", jet_amd.JET_EINVAL),                # too few lines / nonzeros
    ("3 1\n4\n\n\n", jet_amd.JET_EINVAL),               # neighbour out of range
    ("3 4611686018427387904\n2\n1\n\n", jet_amd.JET_EINVAL),  # m that no reservation can hold
    ("3 99999999999999999999999\n2\n1\n\n", jet_amd.JET_EINVAL),  # header value overflows 64 bits
    ("2 1 1\n2 99999999999\n1 99999999999\n", jet_amd.JET_EINVAL),  # edge weight beyond int32
    ("2 1\n99999999999999999999999999\n1\n", jet_amd.JET_EINVAL),  # neighbour id overflows
])
def test_metis_errors(tmp_path, text, code):
    p = tmp_path / "g.graph"
    p.write_text(text)
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.load_metis(str(p))
    assert e.value.code == code


def test_config_files(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("0\n64\n21\n1.03\n")
    c = jet_amd.load_config(str(p))
    assert (c.coarsening_alg, c.num_parts, c.num_iter, c.max_imb_ratio, c.ultra_settings) == (0, 64, 21, 1.03, 0)
    assert c.refine_tolerance == 0.999
    p.write_text("1 8 1 1.10 1")
    c = jet_amd.load_config(str(p))
    assert (c.coarsening_alg, c.num_parts, c.num_iter, c.ultra_settings) == (1, 8, 1, 1)
    p.write_text("1\n8\n3.14\n")  # fewer than four values
    with pytest.raises(jet_amd.JetError):
        jet_amd.load_config(str(p))


def test_partition_file_round_trip(tmp_path):
    part = np.random.default_rng(2).integers(0, 17, 1000).astype(np.int32)
    p = tmp_path / "part.txt"
    jet_amd.write_partition(str(p), part)
    assert p.read_text().splitlines()[:3] == [str(x) for x in part[:3]]
    assert np.array_equal(jet_amd.read_partition(str(p), 1000), part)
    with pytest.raises(jet_amd.JetError):
        jet_amd.read_partition(str(p), 1001)


def test_cli_usage_and_no_device(tmp_path):
    r = subprocess.run([os.path.join(PKG, "jet")], capture_output=True, text=True)
    assert r.returncode != 0 and "Usage:" in r.stderr
    r = subprocess.run([os.path.join(PKG, "pstat")], capture_output=True, text=True)
    assert r.returncode != 0 and "Usage:" in r.stderr
    g = graphs.grid2d(8, 8)
    write_metis(tmp_path / "g.graph", g, False)
    (tmp_path / "c.txt").write_text("0\n4\n1\n1.03\n")
    r = subprocess.run([os.path.join(PKG, "jet"), str(tmp_path / "g.graph"), str(tmp_path / "missing.txt")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "config" in r.stderr
    if jet_amd.device_count() == 0:
        r = subprocess.run([os.path.join(PKG, "jet"), str(tmp_path / "g.graph"), str(tmp_path / "c.txt")],
                           capture_output=True, text=True)
        assert r.returncode != 0 and "no HIP device" in r.stderr
        assert "vertices: 64; edges: 112" in r.stdout


# ---------------------------------------------------------------- coarse-hierarchy files
WIDTHS = {jet_amd.DUMP_32: (np.int32, np.int32, np.int32), jet_amd.DUMP_BIG: (np.int32, np.int64, np.int64),
          jet_amd.DUMP_BIGGEST: (np.int64, np.int64, np.int64)}


def write_hierarchy(path, levels, kind):
    """binary_dump.hpp:61-97 layout: int32 count, then per level n, m, row
    offsets, columns, edge weights, vertex weights and (after the first) the
    map from the previous level, at the widths of `kind`."""
    ot, et, vt = WIDTHS[kind]
    with open(path, "wb") as f:
        f.write(np.int32(len(levels)).tobytes())
        for i, L in enumerate(levels):
            n = len(L["rowptr"]) - 1
            f.write(np.array([n], dtype=ot).tobytes())
            f.write(np.array([len(L["col"])], dtype=et).tobytes())
            f.write(np.asarray(L["rowptr"], dtype=et).tobytes())
            f.write(np.asarray(L["col"], dtype=ot).tobytes())
            f.write(np.asarray(L["ew"], dtype=vt).tobytes())
            f.write(np.asarray(L["vw"], dtype=vt).tobytes())
            if i > 0:
                f.write(np.asarray(L["map"], dtype=ot).tobytes())


def two_levels():
    """An 8x8 grid and its contraction by 2x2 blocks into a 4x4 grid."""
    fine = graphs.grid2d(8, 8)
    fine["vw"] = np.ones(64, dtype=np.int32)
    mp = np.array([(y // 2) * 4 + x // 2 for y in range(8) for x in range(8)], dtype=np.int32)
    coarse = graphs.grid2d(4, 4)
    coarse["ew"] = np.full(len(coarse["col"]), 2, dtype=np.int32)
    coarse["vw"] = np.full(16, 4, dtype=np.int32)
    coarse["map"] = mp
    return [fine, coarse]


@pytest.mark.parametrize("kind", sorted(WIDTHS))
def test_hierarchy_file_loads_at_each_width(tmp_path, kind):
    lv = two_levels()
    write_hierarchy(tmp_path / "h.out", lv, kind)
    h = jet_amd.Hierarchy(str(tmp_path / "h.out"), kind)
    assert h.levels == 2
    assert (h.vertices(0), h.edges(0)) == (64, len(lv[0]["col"]))
    assert (h.vertices(1), h.edges(1)) == (16, len(lv[1]["col"]))
    assert h.vertices(2) == -1
    h.close()


@pytest.mark.parametrize("case", ["truncated", "map_range", "col_range", "weight", "offsets", "count"])
def test_hierarchy_file_errors(tmp_path, case):
    lv = two_levels()
    kind = jet_amd.DUMP_32
    if case == "map_range":
        lv[1]["map"] = lv[1]["map"].copy()
        lv[1]["map"][5] = 16
    elif case == "col_range":
        lv[0]["col"] = lv[0]["col"].copy()
        lv[0]["col"][3] = 64
    elif case == "weight":
        lv[1]["vw"] = lv[1]["vw"].copy()
        lv[1]["vw"][0] = 0
    elif case == "offsets":
        lv[0]["rowptr"] = lv[0]["rowptr"].copy()
        lv[0]["rowptr"][10], lv[0]["rowptr"][11] = lv[0]["rowptr"][11], lv[0]["rowptr"][10]
    write_hierarchy(tmp_path / "h.out", lv, kind)
    if case == "truncated":
        data = (tmp_path / "h.out").read_bytes()
        (tmp_path / "h.out").write_bytes(data[:-7])
    if case == "count":
        (tmp_path / "h.out").write_bytes(np.int32(0).tobytes())
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.Hierarchy(str(tmp_path / "h.out"), kind)
    assert e.value.code == jet_amd.JET_EINVAL
    with pytest.raises(jet_amd.JetError):
        jet_amd.Hierarchy(str(tmp_path / "missing.out"), kind)
    with pytest.raises(jet_amd.JetError):
        jet_amd.Hierarchy(str(tmp_path / "h.out"), 7)


def test_coarse_part_file(tmp_path):
    p = np.arange(10, dtype=np.int32) % 3
    (tmp_path / "coarse_part.out").write_bytes(p.tobytes())
    assert np.array_equal(jet_amd.load_coarse_part(str(tmp_path / "coarse_part.out"), 10), p)
    with pytest.raises(jet_amd.JetError):
        jet_amd.load_coarse_part(str(tmp_path / "coarse_part.out"), 11)


def test_replay_without_device_is_a_clean_error(tmp_path):
    if jet_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    write_hierarchy(tmp_path / "h.out", two_levels(), jet_amd.DUMP_32)
    h = jet_amd.Hierarchy(str(tmp_path / "h.out"))
    with pytest.raises(jet_amd.JetError) as e:
        h.uncoarsen(jet_amd.Config(num_parts=2), np.zeros(16, dtype=np.int32))
    assert e.value.code == jet_amd.JET_EDEVICE
    cfg = jet_amd.Config(num_parts=2)
    cfg.dump_coarse = 1
    with pytest.raises(jet_amd.JetError) as e:
        h.uncoarsen(cfg, np.zeros(16, dtype=np.int32))
    assert e.value.code == jet_amd.JET_EINVAL


def test_import_cli_usage_and_missing_files(tmp_path):
    r = subprocess.run([os.path.join(PKG, "jet_import")], capture_output=True, text=True)
    assert r.returncode != 0 and "Usage:" in r.stderr
    (tmp_path / "c.txt").write_text("1 4 1 1.03\n")
    r = subprocess.run([os.path.join(PKG, "jet_import"), str(tmp_path / "c.txt")], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode != 0 and "coarse_graphs.out" in r.stderr
