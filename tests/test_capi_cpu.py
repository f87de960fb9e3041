"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
its host generators produce the measured workloads, and it reports a clean
error (no crash) when no GPU is present."""
import os
import re
import subprocess

import numpy as np
import pytest

import graphs
import jet_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "jet_capi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(jet_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = jet_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s


def test_config_defaults_match_reference():
    c = jet_amd.Config()
    jet_amd.lib().jet_config_default(c)
    assert (c.coarsening_alg, c.num_iter, c.max_imb_ratio, c.num_parts, c.refine_tolerance, c.dump_coarse,
            c.verbose, c.ultra_settings) == (0, 1, 1.03, 2, 0.999, 0, 0, 0)


def test_grid_generator_matches_numpy_builder():
    g = jet_amd.grid2d(37, 23)
    h = graphs.grid2d(37, 23)
    assert np.array_equal(g["rowptr"], h["rowptr"]) and np.array_equal(g["col"], h["col"])


def test_stencil_generator_matches_numpy_builder():
    g = jet_amd.stencil3d(9, 7, 5, seed=3, wmax=10)
    h = graphs.stencil3d(9, 7, 5, seed=3)
    for key in ("rowptr", "col", "ew"):
        assert np.array_equal(g[key], h[key])
    assert g["ew"].min() >= 1 and g["ew"].max() <= 10


def test_stencil_c2_shape():
    g = jet_amd.stencil3d(128, 128, 128, seed=0)
    assert len(g["rowptr"]) - 1 == 2097152
    assert len(g["col"]) == 12484608


def test_rmat_generator_is_simple_symmetric():
    g = jet_amd.rmat(12, 8, seed=1)
    rp, col = g["rowptr"], g["col"]
    n = len(rp) - 1
    src = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(src != col)
    a = src.astype(np.int64) * n + col
    b = col.astype(np.int64) * n + src
    assert np.array_equal(np.sort(a), np.sort(b))
    assert len(np.unique(a)) == len(a)
    for i in range(0, n, 97):
        assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)
    deg = np.diff(rp)
    assert deg.max() > 20 * deg.mean()  # power-law hubs
    assert (deg == 0).mean() > 0.2       # many isolated vertices at edge factor 8


def test_missing_gpu_is_a_clean_error():
    if jet_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    g = graphs.grid2d(8, 8)
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.partition_csr(g, jet_amd.Config(num_parts=2))
    assert e.value.code == jet_amd.JET_EDEVICE


def build_jet_h_example(tmp_path):
    """Compiles tests/cpp/jet_h_partitions.cpp against include/jet.h."""
    src = os.path.join(ROOT, "tests", "cpp", "jet_h_partitions.cpp")
    exe = str(tmp_path / "jet_h_partitions")
    pkg = os.path.join(ROOT, "jet-partitioner-official-fork_amd")
    cmd = ["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), src, "-o", exe, "-L" + pkg, "-ljet_amd", "-Wl,-rpath," + pkg,
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_jet_h_mirror_compiles_and_reports_no_device(tmp_path):
    if jet_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    exe = build_jet_h_example(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert r.stdout.startswith("error -4 ")


def test_batch_context_without_gpu_is_a_clean_error():
    if jet_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(jet_amd.JetError) as e:
        jet_amd.Context(workers=2)
    assert e.value.code == jet_amd.JET_EDEVICE


def test_loop_free_gain_bucket_equals_the_loop(tmp_path):
    """The kernels' loop-free gain_bucket (threshold tables) against the
    literal loop of jet_refiner.hpp:326-356, both built for the host from the
    device header: dense integers, every threshold's neighbouring floats and
    random 64-bit pairs. A one-ulp change to any threshold fails it."""
    src = os.path.join(ROOT, "tests", "cpp", "gain_bucket_check.cpp")
    exe = str(tmp_path / "gain_bucket_check")
    csrc = os.path.join(ROOT, "jet-partitioner-official-fork_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + csrc, src,
                    "-o", exe], check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert int(r.stdout.split()[-1]) > 7_000_000


def test_tuning_table_parsing(tmp_path):
    import jet_amd

    old = jet_amd.tuning_all()
    try:
        good = tmp_path / "t.txt"
        good.write_text("# a table\nupd_lanes 1   # lanes per move\n\nct_wave_max 64\n")
        jet_amd.tuning_load(str(good))
        assert jet_amd.tuning_get("upd_lanes") == 1 and jet_amd.tuning_get("ct_wave_max") == 64
        for text in ("no_such_knob 3\n", "upd_lanes 9\n", "upd_lanes\n", "ct_block_max x\n"):
            bad = tmp_path / "bad.txt"
            bad.write_text("upd_lanes 4\n" + text)
            with pytest.raises(jet_amd.JetError):
                jet_amd.tuning_load(str(bad))
            assert jet_amd.tuning_get("upd_lanes") == 1  # nothing applied
        with pytest.raises(jet_amd.JetError):
            jet_amd.tuning_load(str(tmp_path / "missing.txt"))
    finally:
        for k, v in old.items():
            jet_amd.tuning_set(k, v)


def test_installed_tuning_tables_are_valid():
    import glob

    import jet_amd

    old = jet_amd.tuning_all()
    try:
        tables = glob.glob(os.path.join(os.path.dirname(jet_amd.__file__), "..", "tuning", "*.txt"))
        for t in tables:
            jet_amd.tuning_load(t)
    finally:
        for k, v in old.items():
            jet_amd.tuning_set(k, v)
