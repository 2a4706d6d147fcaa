"""Context contract on the device: a result formats only against the genome
it was searched on, a failed genome load leaves the loaded genome in place,
and an attached server keeps the genome from being replaced or the context
from being closed."""
import pytest

import casoff
import helpers as H
import workloads

pytestmark = pytest.mark.gpu


def _genomes(tmp_path):
    ga, gb = str(tmp_path / "a.fa"), str(tmp_path / "b.fa")
    casoff.synth_write(casoff.synth_spec(seed=41, n_contigs=3, total_len=800_000, n_frac=0.01), ga, 0)
    casoff.synth_write(casoff.synth_spec(seed=42, n_contigs=1, total_len=200_000, n_frac=0.01), gb, 0)
    return ga, gb


def test_result_formats_only_on_its_genome(tmp_path):
    ga, gb = _genomes(tmp_path)
    ctx = casoff.Context(0)
    try:
        ctx.load(ga)
        q = ctx.query(workloads.plain_10mbp(ga))
        r = ctx.search([q])
        want = H.record_lines(r.format(0))
        assert len(want) > 0
        ctx.load(gb)  # a smaller genome: the old hits would index past its planes
        with pytest.raises(casoff.CofError, match="another genome"):
            r.format(0)
        ctx.load(ga)  # the same file again is a new genome: the old result stays refused
        with pytest.raises(casoff.CofError, match="another genome"):
            r.format(0)
        assert H.record_lines(ctx.search([q]).format(0)) == want
    finally:
        ctx.close()


def test_failed_load_keeps_the_genome(tmp_path):
    ga, _ = _genomes(tmp_path)
    ctx = casoff.Context(0)
    try:
        ctx.load(ga)
        info = ctx.info()
        q = ctx.query(workloads.plain_10mbp(ga))
        r = ctx.search([q])
        want = H.record_lines(r.format(0))
        with pytest.raises(casoff.CofError):
            ctx.load(str(tmp_path / "missing"))
        assert ctx.info() == info
        assert H.record_lines(r.format(0)) == want
        for mode in (casoff.MODE_PLAIN, casoff.MODE_BATCHED):
            assert H.record_lines(ctx.search([q], mode).format(0)) == want
    finally:
        ctx.close()


def test_attached_server_blocks_reload_and_close(tmp_path):
    ga, gb = _genomes(tmp_path)
    ctx = casoff.Context(0)
    ctx.load(ga)
    q = ctx.query(workloads.plain_10mbp(ga))
    srv = casoff.Server(ctx)
    try:
        with pytest.raises(casoff.CofError, match="server is attached"):
            ctx.load(gb)
        with pytest.raises(casoff.CofError, match="server is attached"):
            ctx.synth(casoff.synth_spec(seed=1, total_len=10_000))
        with pytest.raises(casoff.CofError, match="server is attached"):
            ctx.close()
        res, _ = srv.poll(srv.submit([q]), wait=True)
        assert H.record_lines(res.format(0)) == H.record_lines(ctx.search([q]).format(0))
    finally:
        srv.stop()
    ctx.load(gb)  # detached: reloads work again
    ctx.close()
