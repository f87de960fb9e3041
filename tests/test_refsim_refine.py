"""Pins the oracle's refinement (jet_refine and everything it calls) against
an independent Python restatement of src/jet_refiner.hpp (tests/refsim_refine.py),
iteration by iteration, and shows that the comparison has teeth: each of the
oracle's planted-mistake builds (oracle/Makefile, ORC_MUTANT = 1..12) disagrees
with the restatement on at least one case of the same battery."""
import numpy as np
import pytest

import graphs
import oracle
import refsim_refine as R


def rand_part(n, k, seed, skew=0.0):
    rs = np.random.default_rng(seed)
    p = np.arange(k)
    probs = np.exp(-skew * p)
    probs /= probs.sum()
    return rs.choice(k, n, p=probs).astype(np.int32)


# (name, graph, k, imbalance, refine tolerance). "heavy" has edge weights up to
# 60000, so rebalancing gains reach the last gain buckets (49 and the
# uncapped 50); "vheavy" has vertex weights up to 25 and imbalance 1.01, so the
# strong rebalancer re-draws destinations.
BATTERY = [
    ("grid", lambda: graphs.grid2d(20, 17), 4, 1.03, 0.999),
    ("wgrid", lambda: graphs.with_weights(graphs.grid2d(16, 16), seed=2, wmax=5, vwmax=3), 8, 1.03, 0.999),
    ("stencil", lambda: graphs.stencil3d(8, 7, 6, seed=1), 16, 1.03, 0.999),
    ("rmat", lambda: graphs.rmat(9, 8, seed=2), 8, 1.10, 0.99),
    ("wrmat", lambda: graphs.with_weights(graphs.rmat(9, 6, seed=3), seed=4, wmax=3, vwmax=4), 32, 1.03, 0.9999),
    ("random", lambda: graphs.random_graph(300, 1500, seed=5, wmax=7), 5, 1.03, 0.999),
    ("heavy", lambda: graphs.with_weights(graphs.random_graph(400, 2400, seed=9, wmax=1), seed=3, wmax=60000,
                                          vwmax=3), 8, 1.03, 0.999),
    ("path", lambda: graphs.path(200), 4, 1.01, 0.999),
    ("vheavy", lambda: graphs.with_weights(graphs.random_graph(500, 3000, seed=11, wmax=9), seed=5, wmax=9,
                                           vwmax=25), 16, 1.01, 0.99),
    ("tol", lambda: graphs.with_weights(graphs.grid2d(24, 20), seed=7, wmax=4), 6, 1.05, 0.95),
]
SKEWS = (0.0, 1.5, 3.0)
MUTANTS = range(1, 14)

_cache = {}


def run_case(name, mk, k, imb, tol, skew):
    key = (name, skew)
    if key not in _cache:
        g = mk()
        n = len(g["rowptr"]) - 1
        p0 = rand_part(n, k, 6, skew)
        part, rec, trace = R.refine_level(g, p0, k, imb, tol=tol, uniform=g["uniform"])
        _cache[key] = (g, p0, part, rec, trace)
    return _cache[key]


@pytest.mark.parametrize("skew", SKEWS)
@pytest.mark.parametrize("name,mk,k,imb,tol", BATTERY)
def test_refine_level_equals_restatement(name, mk, k, imb, tol, skew):
    g, p0, part, rec, trace = run_case(name, mk, k, imb, tol, skew)
    opart, orec, otrace = oracle.refine_level_trace(g, p0, k, imb, tol=tol, uniform=g["uniform"])
    assert otrace == trace  # every iteration: kind, moves, cut, imbalance, best-state decision, counter
    assert list(orec) == list(rec)
    assert np.array_equal(opart, part)
    assert rec[0] == oracle.total_cut(g, part)


def test_battery_exercises_every_branch():
    kinds, takes = set(), set()
    for name, mk, k, imb, tol in BATTERY:
        for skew in SKEWS:
            trace = run_case(name, mk, k, imb, tol, skew)[4]
            kinds |= {t[0] for t in trace}
            takes |= {t[4] for t in trace}
    assert kinds == {0, 1, 2} and takes == {0, 1, 2}


@pytest.mark.parametrize("mut", MUTANTS)
def test_planted_mistake_is_detected(mut):
    """ORC_MUTANT: 1 strong bucket without its +1 shift, 2 strong gain without
    the own-part connectivity, 3 evict_adjust with > for >=, 4 cookie cutter
    without its fallback, 5 strong re-draws in decreasing id, 6 LP filter with
    > for >=, 7 weak eligibility with <= for <, 8 cut change without its
    second half, 9 update_large threshold n/5, 10 best-state balance test with
    < for <=, 11 best-state cut test with <= for <, 12 tolerance ignored, 13
    weak positions counted only among members whose own part straddles their
    bucket (leaves out the next part's members sharing bucket 50's index)."""
    L = oracle.mutant(mut)
    caught = []
    for name, mk, k, imb, tol in BATTERY:
        for skew in SKEWS:
            g, p0, part, rec, trace = run_case(name, mk, k, imb, tol, skew)
            mp, mrec, mtrace = oracle.refine_level_trace(g, p0, k, imb, tol=tol, uniform=g["uniform"], L=L)
            if mtrace != trace or list(mrec) != list(rec) or not np.array_equal(mp, part):
                caught.append((name, skew))
    assert caught, f"mutant {mut} agrees with the restatement on every case"


def test_gain_bucket_restatement():
    rs = np.random.default_rng(1)
    gains = np.concatenate([np.arange(-3000, 3000), rs.integers(-2**40, 2**40, 4000), [-(1 << 31), 25251, -25251]])
    vws = np.concatenate([np.ones(6000, dtype=np.int64), rs.integers(1, 1000, 4003)])
    for gx, vw in zip(gains.tolist(), vws.tolist()):
        assert R.gain_bucket(gx, vw) == oracle.gain_bucket(gx, vw), (gx, vw)


def test_slot_order_tie_break_differs_on_hashed_rows():
    """Vertex 0 (part 0) sees parts 5 and 2 with equal connectivity in a table
    of 4 slots: part 5 sits in slot 1 and part 2 in slot 2, so the reference's
    first-maximum rule (jet_refiner.hpp:226-233) picks 5, the lowest-id rule 2."""
    g = {"rowptr": np.array([0, 4, 5, 6, 7, 8], dtype=np.int64),
         "col": np.array([1, 2, 3, 4, 0, 0, 0, 0], dtype=np.int32),
         "ew": np.ones(8, dtype=np.int32), "vw": np.ones(5, dtype=np.int32), "uniform": True}
    part = np.array([0, 5, 2, 5, 2], dtype=np.int32)
    try:
        mv, de = oracle.refine_step(0, g, part, 8, 1.03)
        assert dict(zip(mv.tolist(), de.tolist()))[0] == 2
        oracle.set_slot_order_ties(True)
        mv, de = oracle.refine_step(0, g, part, 8, 1.03)
        assert dict(zip(mv.tolist(), de.tolist()))[0] == 5
    finally:
        oracle.set_slot_order_ties(False)
