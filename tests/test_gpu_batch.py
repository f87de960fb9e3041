"""Batched and served partitioning (jet_partition_batch, jet_submit /
jet_poll): every job's result must equal the same graph partitioned alone
through jet_partition_csr, whatever else runs beside it."""
import sys
import threading

import numpy as np
import pytest

import graphs
import jet_amd
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_jobs(count, seed=0, max_n=20000):
    rs = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        kind = i % 3
        if kind == 0:
            side = int(rs.integers(16, int(max_n ** 0.5)))
            g = graphs.grid2d(side, int(rs.integers(16, side + 1)))
        elif kind == 1:
            g = graphs.with_weights(graphs.rmat(int(rs.integers(9, 14)), 8, seed=i), seed=i, wmax=4)
        else:
            s3 = int(rs.integers(6, 20))
            g = graphs.stencil3d(s3, s3, s3, seed=i)
        cfg = jet_amd.Config(num_parts=int(rs.choice([2, 4, 8, 16, 32, 64])),
                             max_imb_ratio=float(rs.choice([1.01, 1.03, 1.10])), coarsening_alg=int(rs.integers(0, 3)))
        jobs.append((g, cfg, 1000 + i))
    return jobs


def test_batch_equals_single_calls():
    jobs = mixed_jobs(24, seed=1)
    alone = [jet_amd.partition_csr(g, cfg, seed=s) for g, cfg, s in jobs]
    say("batch: 24 graphs, 4 workers")
    with jet_amd.Context(workers=4) as ctx:
        parts, cuts = ctx.batch([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs])
    for i, (p, c) in enumerate(alone):
        assert cuts[i] == c, i
        assert np.array_equal(parts[i], p), i


@pytest.mark.parametrize("wg_max", [2048, 16384])
def test_graph_tagged_levels_equal_single_calls(wg_max):
    """16 workers: the levels of up to wg_max vertices of all running jobs are
    refined together by the graph-tagged one-workgroup kernel; every result
    must still equal the plain call's (which uses the multi-kernel path)."""
    jobs = mixed_jobs(40, seed=3)
    alone = [jet_amd.partition_csr(g, cfg, seed=s) for g, cfg, s in jobs]
    say(f"batch: 40 graphs, 16 workers, batch_wg_max_n {wg_max}")
    with jet_amd.tuned(batch_wg_max_n=wg_max):
        with jet_amd.Context(workers=16) as ctx:
            parts, cuts = ctx.batch([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs])
    for i, (p, c) in enumerate(alone):
        assert cuts[i] == c, i
        assert np.array_equal(parts[i], p), i


def test_batch_matches_oracle():
    jobs = mixed_jobs(6, seed=2, max_n=4000)
    with jet_amd.Context(workers=3) as ctx:
        parts, cuts = ctx.batch([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs])
    for i, (g, cfg, s) in enumerate(jobs):
        say("batch vs oracle", i)
        ep, ec, _ = oracle.partition(g, cfg.num_parts, cfg.max_imb_ratio, alg=cfg.coarsening_alg, seed=s,
                                     uniform=bool(g.get("uniform", False)))
        assert cuts[i] == ec and np.array_equal(parts[i], ep), i


def test_submit_poll_from_threads_equals_alone():
    jobs = mixed_jobs(18, seed=3)
    alone = [jet_amd.partition_csr(g, cfg, seed=s) for g, cfg, s in jobs]
    results = {}
    say("serve: 18 jobs from 3 threads")
    with jet_amd.Context(workers=4) as ctx:
        def client(idx):
            tickets = [(i, ctx.submit(jobs[i][0], jobs[i][1], seed=jobs[i][2])) for i in idx]
            for i, t in tickets:
                part, cut, st, tm = ctx.poll(t)
                assert tm.end_s >= tm.start_s >= tm.submit_s
                results[i] = (part, cut)
        th = [threading.Thread(target=client, args=(list(range(r, 18, 3)),)) for r in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert sorted(results) == list(range(18))
    for i, (p, c) in enumerate(alone):
        assert results[i][1] == c and np.array_equal(results[i][0], p), i


def test_budget_admission_and_rejection():
    g = graphs.grid2d(120, 100)
    cfg = jet_amd.Config(num_parts=8)
    ref_part, ref_cut = jet_amd.partition_csr(g, cfg, seed=7)
    # a budget below one job's footprint: a lone job is rejected
    with jet_amd.Context(workers=2, budget=1 << 20) as ctx:
        with pytest.raises(jet_amd.JetError) as e:
            ctx.submit(g, cfg, seed=7)
        assert e.value.code == jet_amd.JET_ENOMEM
    # a budget of about one job: jobs are admitted one at a time, all finish
    n, m = len(g["rowptr"]) - 1, len(g["col"])
    one = 128 * n + 48 * m + (1 << 20)
    with jet_amd.Context(workers=3, budget=int(1.5 * one)) as ctx:
        tickets = [ctx.submit(g, cfg, seed=7) for _ in range(5)]
        spans = []
        for t in tickets:
            part, cut, st, tm = ctx.poll(t)
            assert cut == ref_cut and np.array_equal(part, ref_part)
            spans.append((tm.start_s, tm.end_s))
        spans.sort()
        for a, b in zip(spans, spans[1:]):
            assert b[0] >= a[1] - 1e-6  # never two at once under this budget


def test_poll_errors_and_nonblocking():
    g = graphs.grid2d(64, 64)
    with jet_amd.Context(workers=1) as ctx:
        with pytest.raises(jet_amd.JetError) as e:
            ctx.poll(12345)
        assert e.value.code == jet_amd.JET_EINVAL
        t = ctx.submit(g, jet_amd.Config(num_parts=4), seed=1)
        r = ctx.poll(t, wait_ms=0)
        while r is None:
            r = ctx.poll(t, wait_ms=50)
        bad = dict(g)
        bad["col"] = g["col"].copy()
        bad["col"][0] = 10 ** 6
        t = ctx.submit(bad, jet_amd.Config(num_parts=4))
        with pytest.raises(jet_amd.JetError) as e:
            ctx.poll(t)
        assert e.value.code == jet_amd.JET_EINVAL
