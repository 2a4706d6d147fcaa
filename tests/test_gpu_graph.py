"""HIP graph replay of the batched launch sequence: every replay, with other
GPU work and synchronizations between replays, gives the records of eager
launches (a context with COF_GRAPH=0)."""
import os
import random

import pytest

import casoff
import helpers as H
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


def _mutate_guides(text):
    """The same input with the first A/C/G/T base of every guide changed:
    tables of the same shapes (launch arguments alike) but other contents."""
    lines = text.split("\n")
    out = lines[:2]
    for l in lines[2:]:
        if not l.strip():
            out.append(l)
            continue
        g, rest = l.split(" ", 1)
        j = next(i for i, c in enumerate(g) if c in "ACGT")
        out.append(g[:j] + "ACGT"["ACGT".index(g[j]) - 1] + g[j + 1:] + " " + rest)
    return "\n".join(out)


def test_replay_with_new_query_objects(tmp_path):
    """Graphs replay for the same query objects (the plan's tables cannot
    have changed) and for new objects whose tables equal the captured ones
    word for word; new objects with other tables (guides of one changed
    base: the same table shapes and launch arguments) recapture, and so does
    the first set searched again after them.  Every search equals an eager
    context's."""
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=25, n_contigs=4, total_len=1_500_000, n_frac=0.01), gp, 0)
    texts_a = [t for _, t in workloads.group_inputs(workloads.mixed_guides(8, 48, 3, 6, 0.5, 1), gp)]
    texts_b = [_mutate_guides(t) for t in texts_a]
    eager, graph = _context(False), _context(True)
    try:
        for c in (eager, graph):
            c.load(gp)
        want_a = eager.search([eager.query(t) for t in texts_a], casoff.MODE_BATCHED).records()
        want_b = eager.search([eager.query(t) for t in texts_b], casoff.MODE_BATCHED).records()
        assert len(want_a) > 0 and len(want_b) > 0
        qa = [graph.query(t) for t in texts_a]
        steps = [(qa, want_a, 1), (qa, want_a, 2), ([graph.query(t) for t in texts_a], want_a, 2),
                 ([graph.query(t) for t in texts_b], want_b, 1), ([graph.query(t) for t in texts_b], want_b, 2),
                 (qa, want_a, 1), (qa, want_a, 2)]
        for i, (qs, want, mode) in enumerate(steps):
            got = graph.search(qs, casoff.MODE_BATCHED).records()
            assert len(got) == len(want) and (got == want).all(), "step %d" % i
            assert graph.stats()["graph"] == mode, ("step %d" % i, graph.stats()["graph"])
    finally:
        eager.close()
        graph.close()


def _ctx_with_env(env):
    """A context opened with the variables of `env` set (lanes read them at open)."""
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return casoff.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k)
            else:
                os.environ[k] = v


def test_hit_copy_forms_agree(tmp_path):
    """Batched hits are collected launch by launch (sorted and copied on a
    second stream while later launches run) or after the last launch
    (COF_PIPELINE=0); there they cross to the host packed (5 bytes and a
    per-key first-hit table) below 2^32 genome positions, and as 10-byte
    records otherwise (COF_HIT_PACK=0 forces that form here).  Every form
    decodes to the same records, over several copy chunks and a hit buffer
    that overflows on the first pass (a workload of more than 4 M hits)."""
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=24, n_contigs=3, total_len=12_000_000, n_frac=0.01), gp, 0)
    # all-N pattern, 8 comparable positions at threshold 3: ~6% of the windows
    # hit each guide, on both strands (~6 M hits)
    guides = ["ACGTACGT", "TTGCACGT", "CCGTACGA", "GATTACAG", "AACCGGTT", "TGCATGCA", "CATGACTA", "GGACTTCA"]
    text = gp + "\nNNNNNNNNNNNNNNNN\n" + "".join("%sNNNNNNNN 3\n" % g for g in guides)
    got = {}
    for pack, pipe in (("1", "1"), ("1", "0"), ("0", "0")):
        c = _ctx_with_env({"COF_HIT_PACK": pack, "COF_PIPELINE": pipe})
        try:
            c.load(gp)
            q = c.query(text)
            got[pack + pipe] = [c.search([q], m).records() for m in (casoff.MODE_BATCHED, casoff.MODE_PLAIN)]
        finally:
            c.close()
    ref = got["11"]
    assert len(ref[0]) > (1 << 22) + 1000
    for form in ("10", "00"):
        for a, b in zip(ref, got[form]):
            assert len(a) == len(b) and (a == b).all(), form
    assert (ref[0] == ref[1]).all()


def test_launch_order_differs_from_query_order(tmp_path):
    """Key groups launch before bulge groups and queries of one pattern share
    a launch, so the launch-by-launch collection stores query runs out of
    query order: records, per-query formatting and the summary must still
    come out per query as each query searched alone on the plain path."""
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=25, n_contigs=4, total_len=3_000_000, n_frac=0.01), gp, 0)
    r = random.Random(25)
    rand = lambda n: "".join(r.choice("ACGT") for _ in range(n))
    texts = [gp + "\nNNNNNNNNNNNNNNNNNNNNNGG 1 1\n" + "".join("%sNNN 4\n" % rand(20) for _ in range(3)),
             gp + "\nNNNNNNNNNNNNNNNNNNNNNGG\n" + "".join("%sNNN 5\n" % rand(20) for _ in range(5)),
             gp + "\nNNNNNNNNNNNNNNNNNNNNNRG\n" + "".join("%sNNN 4\n" % rand(20) for _ in range(4)),
             gp + "\nNNNNNNNNNNNNNNNNNNNNNGG\n" + "".join("%sNNN 4\n" % rand(20) for _ in range(6))]
    ctx = casoff.Context(0)
    try:
        ctx.load(gp)
        qs = [ctx.query(t) for t in texts]
        res = ctx.search(qs, casoff.MODE_BATCHED)
        assert ctx.stats()["n_launches"] == 3  # NGG keys (queries 1 and 3), NRG keys, the NGG bulge group
        alone = [ctx.search([q], casoff.MODE_PLAIN) for q in qs]
        recs = res.records()
        assert len(recs) == sum(len(a.records()) for a in alone) > 0
        assert (recs["query"][1:] >= recs["query"][:-1]).all()  # query order
        at = 0
        for i, a in enumerate(alone):
            ar = a.records()
            got = recs[at:at + len(ar)]
            assert (got["query"] == i).all()
            for f in ("pos", "key", "mm", "strand"):
                assert (got[f] == ar[f]).all(), (i, f)
            at += len(ar)
            assert H.record_lines(res.format(i)) == H.record_lines(a.format(0)), i
    finally:
        ctx.close()
