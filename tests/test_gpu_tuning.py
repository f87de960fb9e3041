"""Every value the autotuner may choose for a tuning parameter gives the same
partitions as the built-in defaults, bit for bit (the parameters change
launch shapes and paths, never results)."""
import sys

import numpy as np
import pytest

import graphs
import jet_amd
from jet_amd import autotune

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    ("stencil 40x40x30 k=64 HEC", lambda: graphs.stencil3d(40, 40, 30, seed=1), 64, 1.03, 1),
    ("rmat-15 k=32 HEC (hub rows)", lambda: graphs.rmat(15, 12, seed=3), 32, 1.10, 1),
    ("grid 200x180 k=8 2-hop", lambda: graphs.grid2d(200, 180), 8, 1.03, 0),
    ("weighted rmat-13 k=16 matching", lambda: graphs.with_weights(graphs.rmat(13, 8, seed=5), seed=6, wmax=9,
                                                                   vwmax=4), 16, 1.01, 2),
]


@pytest.fixture(scope="module")
def reference():
    if jet_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    out = []
    for name, mk, k, imb, alg in CASES:
        g = mk()
        cfg = jet_amd.Config(num_parts=k, max_imb_ratio=imb, coarsening_alg=alg)
        out.append((name, g, cfg, jet_amd.partition_csr(g, cfg, seed=3)))
    return out


@pytest.mark.parametrize("knob", sorted(autotune.SPACE))
def test_every_candidate_value_is_result_neutral(knob, reference):
    for v in autotune.SPACE[knob]:
        extra = {"wg_max_n": 16384} if knob == "wg_big_deg" else {}
        with jet_amd.tuned(**{knob: v}, **extra):
            for name, g, cfg, (p0, c0) in reference:
                say(knob, v, name)
                p, c = jet_amd.partition_csr(g, cfg, seed=3)
                assert c == c0, (knob, v, name)
                assert np.array_equal(p, p0), (knob, v, name)
