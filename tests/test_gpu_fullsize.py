"""Parity at full size.  The 3.1 Gbp synthetic genome (seed 1, 24 contigs)
stays resident; each workload runs over all of it once per mode (plain: the
reference's launch model, cas-offinder.cpp:254-393; batched), and the records
that print a location inside a genome region must equal the oracle's records
on that region's text.  Regions: the start of chr1; the plain path's
2^31-start slab boundary (inside chr15); the chr12/chr13 junction (windows
must not cross it, cas-offinder.cl:64-67); the end of the last contig
(positions above 2^31 up to the genome end).  Workloads: 100 NGG guides
(genome_3g_ngg), 20 NRG guides with DNA bulge 2 / RNA bulge 1 (bulge_3g_nrg),
2,000 mixed-PAM guides in 16 (PAM, D, R) groups (mixed_pam_2000, one
reference input file per group: first-(id, threshold)-wins is scoped to one
file, cas-offinder.cpp:14-18, main.cpp:43-70), and the seed-3 serving stream
of 1-256 guides per request through the continuous-batching server."""
import pytest

import casoff
import helpers as H
import workloads

pytestmark = pytest.mark.gpu

MODES = (casoff.MODE_PLAIN, casoff.MODE_BATCHED)
SLAB = 1 << 31  # window starts per plain finder launch (cof_ctx::slab_positions)


@pytest.fixture(scope="module")
def ctx():
    c = casoff.Context(0)
    c.synth(workloads.genome_3g())
    yield c
    c.close()


def regions(ctx):
    cs = ctx.contigs()
    k = next(i for i, (_, gpos, clen, _) in enumerate(cs) if gpos <= SLAB < gpos + clen)
    return {
        "prefix": [H.Piece(ctx, 0, 0, 2_000_000)],
        "slab": [H.Piece(ctx, k, SLAB - cs[k][1] - 1_000_000, 2_000_000)],
        "junction": [H.Piece(ctx, 11, cs[11][2] - 1_000_000, 1_000_000), H.Piece(ctx, 12, 0, 1_000_000)],
        "last": [H.Piece(ctx, len(cs) - 1, cs[-1][2] - 1_500_000, 1_500_000)],
    }


def texts_of(workload):
    if workload == "mixed_pam_2000":
        return [t for _, t in workloads.mixed_pam_2000("synthetic")]
    return [getattr(workloads, workload)("synthetic")]


def margin_of(text):
    """Padded window length plus the largest printed offset (D + bulge size)."""
    f = text.split("\n")[1].split()
    D, R = (int(f[1]), int(f[2])) if len(f) == 3 else (0, 0)
    return len(f[0]) + 2 * D + 2 * R + 8


@pytest.fixture(scope="module")
def searched(ctx):
    """Per workload: input texts, queries and the hit records of each mode
    over the whole genome (searched once, compared region by region)."""
    cache = {}

    def get(workload):
        if workload not in cache:
            texts = texts_of(workload)
            qs = [ctx.query(t) for t in texts]
            res = {mode: ctx.search(qs, mode).records() for mode in MODES}
            assert len(res[casoff.MODE_PLAIN]) == len(res[casoff.MODE_BATCHED])
            cache[workload] = (texts, qs, res)
        return cache[workload]

    yield get
    cache.clear()


def oracle_lines(tmp_path, gp, text, pieces, margin, name):
    lines = H.record_lines(H.oracle_records(tmp_path, gp + "\n" + text.split("\n", 1)[1], name=name))
    return H.keep_in_pieces(lines, pieces, margin, rebase=True)


def product_lines(ctx, q, hits, pieces, margin):
    sub = H.hits_near_pieces(hits, pieces, margin)
    return H.keep_in_pieces(H.record_lines(ctx.format_hits(q, sub)), pieces, margin)


@pytest.mark.parametrize("region", ["prefix", "slab", "junction", "last"])
@pytest.mark.parametrize("workload", ["genome_3g_ngg", "bulge_3g_nrg", "mixed_pam_2000"])
def test_region_matches_oracle(ctx, searched, tmp_path, workload, region):
    texts, qs, res = searched(workload)
    pieces = regions(ctx)[region]
    gp = str(tmp_path / "region.fa")
    H.write_pieces(ctx, gp, pieces)
    n = 0
    for qi, text in enumerate(texts):
        margin = margin_of(text)
        exp = oracle_lines(tmp_path, gp, text, pieces, margin, "in%d.txt" % qi)
        for mode in MODES:
            hits = res[mode][res[mode]["query"] == qi]
            got = product_lines(ctx, qs[qi], hits, pieces, margin)
            assert got == exp, "%s %s query %d mode %d: %d vs %d records" % (
                workload, region, qi, mode, len(got), len(exp))
        n += len(exp)
    assert n > 0, "no records in the region"


def test_serving_stream_matches_oracle(ctx, tmp_path):
    """The seed-3 stream (1-256 guides per request, thresholds 2-6, DNA/RNA
    bulges up to 2, one query per (PAM, D, R) group) submitted at 200 QPS so
    that requests share batched passes: every served response equals the
    oracle on a 1.5 Mbp prefix of chr1, and equals the same queries searched
    alone, record for record over the whole genome."""
    n_req = 32
    reqs = workloads.serving_stream(n_req, 200.0, "synthetic", seed=3)
    queries = [[ctx.query(t) for _, t in groups] for _, groups, _ in reqs]
    srv = casoff.Server(ctx, mem_budget_bytes=8 << 30, max_batch_queries=64)
    try:
        import time
        t0 = time.perf_counter()
        tickets = []
        for (arrival, _, _), qs in zip(reqs, queries):
            wait = arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            tickets.append(srv.submit(qs))
        served = [srv.poll(t, wait=True)[0] for t in tickets]
    finally:
        srv.stop()
    pieces = [H.Piece(ctx, 0, 0, 1_500_000)]
    gp = str(tmp_path / "prefix.fa")
    H.write_pieces(ctx, gp, pieces)
    n = 0
    for ri, ((_, groups, _), qs, res) in enumerate(zip(reqs, queries, served)):
        rec = res.records()
        alone = ctx.search(qs, casoff.MODE_BATCHED).records()
        assert len(rec) == len(alone) and (rec == alone).all(), "request %d differs from its search alone" % ri
        for qi, (_, text) in enumerate(groups):
            margin = margin_of(text)
            exp = oracle_lines(tmp_path, gp, text, pieces, margin, "r%dq%d.txt" % (ri, qi))
            got = product_lines(ctx, qs[qi], rec[rec["query"] == qi], pieces, margin)
            assert got == exp, "request %d query %d: %d vs %d records" % (ri, qi, len(got), len(exp))
            n += len(exp)
    assert n > 0
