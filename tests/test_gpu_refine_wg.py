"""The one-workgroup refinement of small levels against the oracle and the
independent restatement: the refsim battery (weak and strong steps, gain
buckets up to 50, strong re-draws, tolerances), multi-chunk ranking (more than
1024 ranked members), update_large rebuilds mid-level, ultra settings and the
refine_tolerance variants of jet4 / jet2, each on both refinement paths."""
import sys

import numpy as np
import pytest

import graphs
import jet_amd
import oracle
from test_refsim_refine import BATTERY, SKEWS, rand_part

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")


PATHS = [("wg", 16384), ("multi", 0)]


@pytest.fixture(params=PATHS, ids=[p[0] for p in PATHS])
def refine_path(request):
    with jet_amd.tuned(wg_max_n=request.param[1]):
        yield request.param[0]


@pytest.mark.parametrize("name,mk,k,imb,tol", BATTERY)
def test_battery_refine_level(name, mk, k, imb, tol, refine_path):
    g = mk()
    n = len(g["rowptr"]) - 1
    for skew in SKEWS:
        p0 = rand_part(n, k, 6, skew)
        say("battery", refine_path, name, skew)
        p, rec = jet_amd.refine_level(g, p0, k, imb, tol=tol, uniform=g["uniform"])
        ep, erec, _ = oracle.refine_level_trace(g, p0, k, imb, tol=tol, uniform=g["uniform"])
        assert list(rec) == list(erec), (rec, erec)
        assert np.array_equal(p, ep)


@pytest.mark.parametrize("case", ["strong-many-ranked", "weak-many-ranked"])
def test_many_ranked_members(case, refine_path):
    # one part holds most of the weight, so thousands of its vertices are
    # ranked in one rebalancing step (several chunks of 1024)
    g = graphs.with_weights(graphs.grid2d(110, 100), seed=4, wmax=3)
    n = len(g["rowptr"]) - 1
    k = 4 if case.startswith("strong") else 6
    p0 = rand_part(n, k, 9, 3.0)
    say("many ranked", refine_path, case)
    p, rec = jet_amd.refine_level(g, p0, k, 1.01, uniform=g["uniform"])
    ep, erec, tr = oracle.refine_level_trace(g, p0, k, 1.01, uniform=g["uniform"])
    assert any(t[0] != 0 for t in tr)
    assert list(rec) == list(erec)
    assert np.array_equal(p, ep)


@pytest.mark.parametrize("ultra,tol", [(1, 0.999), (0, 0.9999), (0, 0.99)])
def test_ultra_and_tolerance_levels(ultra, tol, refine_path):
    g = graphs.with_weights(graphs.rmat(11, 8, seed=7), seed=2, wmax=5)
    n = len(g["rowptr"]) - 1
    for k in (4, 16):
        p0 = rand_part(n, k, 3, 0.5)
        say("ultra/tol level", refine_path, ultra, tol, k)
        p, rec = jet_amd.refine_level(g, p0, k, 1.03, tol=tol, ultra=bool(ultra), uniform=False)
        ep, erec = oracle.refine_level(g, p0, k, 1.03, tol=tol, ultra=bool(ultra), uniform=False)
        assert list(rec) == list(erec)
        assert np.array_equal(p, ep)


@pytest.mark.parametrize("ultra,tol,alg", [(1, 0.999, 1), (0, 0.9999, 0), (0, 0.99, 2)])
def test_ultra_and_tolerance_end_to_end(ultra, tol, alg, refine_path):
    g = graphs.stencil3d(20, 18, 16, seed=3)
    for k in (8, 32):
        say("ultra/tol e2e", refine_path, ultra, tol, alg, k)
        cfg = jet_amd.Config(num_parts=k, max_imb_ratio=1.03, coarsening_alg=alg, refine_tolerance=tol,
                             ultra_settings=ultra)
        part, cut = jet_amd.partition_csr(g, cfg, seed=5)
        epart, ecut, _ = oracle.partition(g, k, 1.03, alg=alg, seed=5, tol=tol, ultra=bool(ultra))
        assert cut == ecut
        assert np.array_equal(part, epart)


def test_tuning_parameter_api():
    old = jet_amd.tuning_get("wg_max_n")
    with jet_amd.tuned(wg_max_n=123):
        assert jet_amd.tuning_get("wg_max_n") == 123
    assert jet_amd.tuning_get("wg_max_n") == old
    for bad in (("wg_max_n", -1), ("no_such_knob", 1)):
        with pytest.raises(jet_amd.JetError):
            jet_amd.tuning_set(*bad)


def test_weak_bucket_50_shared_with_next_part(refine_path):
    # edge weights up to 60000 give weak-rebalancing gains past the last gain
    # bucket: the uncapped bucket 50 of part q shares its index with the first
    # gain bucket of part q + 1, and positions inside it count the members of
    # both parts; two sections per part (n >= 4096 k)
    g = graphs.with_weights(graphs.random_graph(9000, 18000, seed=13, wmax=1), seed=8, wmax=60000, vwmax=3)
    n = len(g["rowptr"]) - 1
    for k, skew in ((2, 1.0), (2, 0.0), (8, 1.5)):
        p0 = rand_part(n, k, 5, skew)
        say("bucket 50", refine_path, k, skew)
        p, rec = jet_amd.refine_level(g, p0, k, 1.03, uniform=False)
        ep, erec, tr = oracle.refine_level_trace(g, p0, k, 1.03, uniform=False)
        assert list(rec) == list(erec), (rec, erec)
        assert np.array_equal(p, ep)


@pytest.mark.parametrize("name", ["vheavy", "grid-300-vw25-k64"])
def test_long_redraw_list_path(name, refine_path):
    # redraw_lds_max=0 sends every strong re-draw list through the path for
    # lists longer than the LDS holds (marks over all vertex ids, compacted in
    # id order in global memory); results must not change
    if name == "vheavy":
        (_, mk, k, imb, tol), = [c for c in BATTERY if c[0] == "vheavy"]
        g = mk()
        starts = [rand_part(len(g["rowptr"]) - 1, k, 6, s) for s in SKEWS]
    else:
        g = graphs.with_weights(graphs.grid2d(300, 300), seed=3, wmax=5, vwmax=25)
        k, imb, tol = 64, 1.01, 0.999
        starts = [rand_part(len(g["rowptr"]) - 1, k, 7, 2.0)]
    for p0 in starts:
        ep, erec, tr = oracle.refine_level_trace(g, p0, k, imb, tol=tol, uniform=False)
        assert any(t[0] == 2 for t in tr)
        for lds_max in (0, 8192):
            say("long re-draw list", refine_path, name, lds_max)
            with jet_amd.tuned(redraw_lds_max=lds_max):
                p, rec = jet_amd.refine_level(g, p0, k, imb, tol=tol, uniform=False)
            assert list(rec) == list(erec)
            assert np.array_equal(p, ep)
