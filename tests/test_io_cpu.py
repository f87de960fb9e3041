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
