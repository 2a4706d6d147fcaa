"""HIP graph replay of the batched launch sequence: every replay, with other
GPU work and synchronizations between replays, gives the records of eager
launches (a context with COF_GRAPH=0)."""
import os

import pytest

import casoff
import workloads

pytestmark = pytest.mark.gpu


def _context(graph):
    old = os.environ.pop("COF_GRAPH", None)
    if not graph:
        os.environ["COF_GRAPH"] = "0"
    try:
        return casoff.Context(0)
    finally:
        os.environ.pop("COF_GRAPH", None)
        if old is not None:
            os.environ["COF_GRAPH"] = old


def test_graph_replays_match_eager(tmp_path):
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=21, n_contigs=5, total_len=3_000_000, n_frac=0.01), gp, 0)
    texts = [t for _, t in workloads.group_inputs(workloads.mixed_guides(4, 60, 3, 6, 0.5, 1), gp)]
    eager, graph = _context(False), _context(True)
    try:
        for c in (eager, graph):
            c.load(gp)
        qe = [eager.query(t) for t in texts]
        qg = [graph.query(t) for t in texts]
        want_all = eager.search(qe, casoff.MODE_BATCHED).records()
        assert eager.stats()["graph"] == 0
        want_one = eager.search(qe[:1], casoff.MODE_BATCHED).records()
        assert len(want_all) > 0
        modes = []
        for rep in range(4):
            got = graph.search(qg, casoff.MODE_BATCHED).records()
            modes.append(graph.stats()["graph"])
            assert len(got) == len(want_all) and (got == want_all).all(), "replay %d" % rep
            # other GPU work and synchronizations between replays
            assert (eager.search(qe[:1], casoff.MODE_BATCHED).records() == want_one).all()
            graph.text(0, 100)
        assert modes[0] == 1 and all(m == 2 for m in modes[1:]), modes
        # a different search recaptures; returning to the first recaptures again
        got = graph.search(qg[:1], casoff.MODE_BATCHED).records()
        assert graph.stats()["graph"] == 1 and (got == want_one).all()
        got = graph.search(qg, casoff.MODE_BATCHED).records()
        assert graph.stats()["graph"] == 1 and (got == want_all).all()
        got = graph.search(qg, casoff.MODE_BATCHED).records()
        assert graph.stats()["graph"] == 2 and (got == want_all).all()
    finally:
        eager.close()
        graph.close()


def test_plan_reused_across_searches_and_genomes(tmp_path):
    """The lane keeps the tables built for the last query set and reuses them
    when the same query objects are searched again, also after another
    genome is loaded; a different query set in between rebuilds them.  Every
    search equals the same queries searched on a fresh context."""
    ga, gb = str(tmp_path / "a.fa"), str(tmp_path / "b.fa")
    casoff.synth_write(casoff.synth_spec(seed=31, n_contigs=3, total_len=1_500_000, n_frac=0.01), ga, 0)
    casoff.synth_write(casoff.synth_spec(seed=32, n_contigs=4, total_len=2_000_000, n_frac=0.01), gb, 0)
    texts = [t for _, t in workloads.group_inputs(workloads.mixed_guides(5, 50, 3, 6, 0.5, 1), ga)]

    def fresh(genome, sel):
        c = casoff.Context(0)
        try:
            c.load(genome)
            return c.search([c.query(texts[i]) for i in sel], casoff.MODE_BATCHED).records()
        finally:
            c.close()

    ctx = casoff.Context(0)
    try:
        qs = [ctx.query(t) for t in texts]
        sel_all, sel_one = list(range(len(texts))), [len(texts) - 1]
        for genome in (ga, gb, ga):
            ctx.load(genome)
            for sel in (sel_all, sel_one, sel_all, sel_all):
                got = ctx.search([qs[i] for i in sel], casoff.MODE_BATCHED).records()
                want = fresh(genome, sel)
                assert len(got) == len(want) and (got == want).all(), (genome, sel)
    finally:
        ctx.close()


def test_plain_search_between_batched_searches(tmp_path):
    """A plain search writes its per-key tables into the lane's key buffer;
    the batched search of the same query objects that follows must upload its
    own tables again instead of replaying the captured graph over the plain
    path's.  Bulge groups are in the set: their kernels take variant-block
    offsets from the tables.  Every search equals the oracle."""
    import helpers as H
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=23, n_contigs=4, total_len=1_200_000, n_frac=0.01), gp, 0)
    texts = [t for _, t in workloads.group_inputs(workloads.mixed_guides(6, 40, 3, 6, 0.5, 1), gp)]
    assert any(len(t.split("\n")[1].split()) == 3 for t in texts)  # some group has DNA/RNA bulges
    want = [H.record_lines(H.oracle_records(tmp_path, t, name="in%d.txt" % i)) for i, t in enumerate(texts)]
    ctx = _context(True)
    try:
        ctx.load(gp)
        qs = [ctx.query(t) for t in texts]
        modes = []
        for mode in (casoff.MODE_BATCHED, casoff.MODE_PLAIN, casoff.MODE_BATCHED, casoff.MODE_PLAIN,
                     casoff.MODE_BATCHED, casoff.MODE_BATCHED):
            r = ctx.search(qs, mode)
            modes.append((mode, ctx.stats()["graph"]))
            for i in range(len(qs)):
                assert H.record_lines(r.format(i)) == want[i], (mode, i)
        # the batched search right after a plain one recaptured; the last one replayed
        assert modes[2] == (casoff.MODE_BATCHED, 1) and modes[4] == (casoff.MODE_BATCHED, 1), modes
        assert modes[5] == (casoff.MODE_BATCHED, 2), modes
    finally:
        ctx.close()


def test_hit_copy_forms_agree(tmp_path):
    """Hits cross to the host packed (5 bytes and a per-key first-hit table)
    below 2^32 genome positions, and as 10-byte records otherwise
    (COF_HIT_PACK=0 forces that form here): both decode to the same records,
    over several copy chunks (a workload of more than 4 M hits)."""
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=24, n_contigs=3, total_len=12_000_000, n_frac=0.01), gp, 0)
    # all-N pattern, 8 comparable positions at threshold 3: ~6% of the windows
    # hit each guide, on both strands (~6 M hits)
    guides = ["ACGTACGT", "TTGCACGT", "CCGTACGA", "GATTACAG", "AACCGGTT", "TGCATGCA", "CATGACTA", "GGACTTCA"]
    text = gp + "\nNNNNNNNNNNNNNNNN\n" + "".join("%sNNNNNNNN 3\n" % g for g in guides)
    got = {}
    for pack in ("1", "0"):
        old = os.environ.get("COF_HIT_PACK")
        os.environ["COF_HIT_PACK"] = pack
        try:
            c = casoff.Context(0)
        finally:
            if old is None:
                os.environ.pop("COF_HIT_PACK")
            else:
                os.environ["COF_HIT_PACK"] = old
        try:
            c.load(gp)
            q = c.query(text)
            got[pack] = [c.search([q], m).records() for m in (casoff.MODE_BATCHED, casoff.MODE_PLAIN)]
        finally:
            c.close()
    assert len(got["1"][0]) > (1 << 22) + 1000
    for a, b in zip(got["1"], got["0"]):
        assert len(a) == len(b) and (a == b).all()
    assert (got["1"][0] == got["1"][1]).all()
