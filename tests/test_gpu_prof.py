"""Kernel-family timing and counting (jet_prof_enable / jet_prof_read), as
bench.py uses them for its roofline fields: a timed or counted partition
must equal the plain partition, every family's launch count must be the same
in timing and counting runs, and the times and bytes must be positive."""
import numpy as np
import pytest

import graphs
import jet_amd

pytestmark = pytest.mark.gpu


def test_family_timing_and_counting_leave_results_unchanged():
    g = graphs.with_weights(graphs.stencil3d(20, 20, 20, seed=4), seed=4, wmax=9)
    cfg = jet_amd.Config(num_parts=16, max_imb_ratio=1.03, coarsening_alg=1)
    base, cut = jet_amd.partition_csr(g, cfg, seed=1)
    timed = {}
    try:
        for fam in jet_amd.PROF_FAMILIES:
            jet_amd.prof_enable(fam)
            p, c = jet_amd.partition_csr(g, cfg, seed=1)
            assert c == cut and np.array_equal(p, base), fam
            units, ms, _ = jet_amd.prof_read(fam)
            timed[fam] = (units, ms)
        jet_amd.prof_enable("all:count")
        p, c = jet_amd.partition_csr(g, cfg, seed=1)
        assert c == cut and np.array_equal(p, base)
        for fam in jet_amd.PROF_FAMILIES:
            units, _, nbytes = jet_amd.prof_read(fam)
            assert units == timed[fam][0], fam
            if units:
                assert timed[fam][1] > 0 and nbytes > 0, fam
        assert timed["update"][0] > 0 and timed["lp_select"][0] > 0
    finally:
        jet_amd.prof_enable("none")
    p, c = jet_amd.partition_csr(g, cfg, seed=1)
    assert c == cut and np.array_equal(p, base)


def test_unknown_family_is_rejected():
    with pytest.raises(jet_amd.JetError):
        jet_amd.prof_enable("no_such_family")
