"""Serving loop: every request served in a continuously batched mix returns
the same records as the same queries searched alone on the plain path."""
import random

import pytest

import casoff
import helpers as H
import workloads

pytestmark = pytest.mark.gpu


def test_served_results_match_alone(tmp_path):
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=21, n_contigs=5, total_len=2_000_000, n_frac=0.01), gp, 0)
    ctx = casoff.Context(0)
    ctx.load(gp)
    reqs = workloads.serving_stream(16, 1000.0, gp, seed=3, max_guides=24)
    queries = [[ctx.query(t) for _, t in groups] for _, groups, _ in reqs]
    alone = []
    for qs in queries:
        r = ctx.search(qs, casoff.MODE_PLAIN)
        alone.append([r.format(i) for i in range(len(qs))])
    srv = casoff.Server(ctx, mem_budget_bytes=4 << 20, max_batch_queries=8)
    tickets = [srv.submit(qs) for qs in queries]
    lat, n_records = [], 0
    for t, exp in zip(tickets, alone):
        res, ms = srv.poll(t, wait=True)
        lat.append(ms)
        n_records += len(res.records())
        got = [res.format(i) for i in range(len(exp))]
        assert [H.record_lines(x) for x in got] == [H.record_lines(x) for x in exp]
    st = srv.stats()
    srv.stop()
    assert all(x > 0 for x in lat)
    # every request and query was served once, and the merged batches' hits
    # are the records of the split results
    assert st["requests"] == len(queries)
    assert st["queries"] == sum(len(qs) for qs in queries)
    assert st["hits"] == n_records > 0
    assert st["lane_hits"] == 1 << 20  # the smallest lane: 4 MiB is below its share
    assert 1 <= st["batches"] <= len(queries) and st["hits_per_key"] >= 1.0
    ctx.close()


def test_server_poll_pending_and_stop(tmp_path):
    gp = str(tmp_path / "g2.fa")
    casoff.synth_write(casoff.synth_spec(seed=22, n_contigs=2, total_len=300_000), gp, 0)
    ctx = casoff.Context(0)
    ctx.load(gp)
    q = ctx.query(workloads.plain_10mbp(gp))
    srv = casoff.Server(ctx)
    t = srv.submit([q])
    out = srv.poll(t, wait=True)
    assert out is not None and out[1] > 0
    exp = H.record_lines(ctx.search([q], casoff.MODE_BATCHED).format(0))
    assert H.record_lines(out[0].format(0)) == exp
    srv.stop()
    ctx.close()
