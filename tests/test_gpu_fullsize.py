"""Parity at the benchmark's real size: the 3.1 Gbp synthetic genome stays
resident, the workloads of bench.py run over all of it, and the records that
fall inside a 4 Mbp prefix of the first contig must equal the oracle's records
on that prefix (same cut applied to both: locations that leave room for the
whole window plus bulge slack)."""
import pytest

import casoff
import helpers as H
import workloads

pytestmark = pytest.mark.gpu

PREFIX = 4_000_000


@pytest.fixture(scope="module")
def ctx():
    c = casoff.Context(0)
    c.synth(workloads.genome_3g())
    yield c
    c.close()


def _inside(lines, name, limit):
    out = []
    for l in lines:
        f = l.split("\t")
        if f[4] == name and int(f[5]) <= limit:
            out.append(l)
    return out


@pytest.mark.parametrize("workload", ["genome_3g_ngg", "bulge_3g_nrg"])
def test_prefix_of_full_genome_matches_oracle(ctx, tmp_path, workload):
    text = getattr(workloads, workload)("synthetic")
    q = ctx.query(text)
    name, pos, clen, _ = ctx.contigs()[0]
    gp = str(tmp_path / "prefix.fa")
    seq = ctx.text(pos, PREFIX)
    with open(gp, "w") as f:
        f.write(">" + name + "\n")
        for i in range(0, len(seq), 60):
            f.write(seq[i:i + 60] + "\n")
    P = len(text.split("\n")[1].split()[0]) + 2  # pattern + DNA-bulge slack
    limit = PREFIX - P - 8
    exp = _inside(H.record_lines(H.oracle_records(tmp_path, gp + "\n" + text.split("\n", 1)[1])), name, limit)
    assert len(exp) > 0
    for mode in (casoff.MODE_BATCHED, casoff.MODE_PLAIN):
        got = _inside(H.record_lines(ctx.search([q], mode).format(0)), name, limit)
        assert got == exp, "%s mode %d: %d vs %d records" % (workload, mode, len(got), len(exp))
