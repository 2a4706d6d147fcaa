"""The region cut that the full-size GPU tests use to compare a search over a
whole genome with oracle runs over slices of it: on a small genome, the
oracle's records over the whole genome, cut to a region, equal the oracle's
records over the region's own FASTA (locations moved back to contig
coordinates), for slices inside a contig, at a contig junction and at the
genome end, with DNA and RNA bulges (records printed up to D + bulge size
past the window start)."""
import pytest

import casoff
import helpers as H


@pytest.fixture(scope="module")
def genome(tmp_path_factory):
    d = tmp_path_factory.mktemp("regions")
    gp = str(d / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=51, n_contigs=3, total_len=600_000, n_frac=0.01, n_run_min=50,
                                         n_run_max=400), gp, 0)
    ctx = casoff.Context(-1)
    ctx.load(gp)
    yield ctx, gp
    ctx.close()


CASES = [
    ("N" * 21 + "RG", 0, 0, 5),
    ("N" * 21 + "RG", 2, 1, 4),
    ("N" * 20 + "NNAGAAW", 1, 2, 5),
    ("TTTV" + "N" * 23, 1, 1, 5),  # reversed PAM
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_region_cut_equals_whole_genome_cut(genome, tmp_path, case):
    ctx, gp = genome
    pat, D, R, t = CASES[case]
    r = H.rng(100 + case)
    cs = ctx.contigs()
    regions = [
        [H.Piece(ctx, 0, 60_000, 50_000)],
        [H.Piece(ctx, 0, cs[0][2] - 40_000, 40_000), H.Piece(ctx, 1, 0, 40_000)],
        [H.Piece(ctx, 2, cs[2][2] - 50_000, 50_000)],
    ]
    # guides planted from windows inside the regions (and near their edges),
    # two bases changed, the pattern's fixed positions N
    guides = []
    for pieces in regions:
        for p in pieces:
            # windows inside, and windows crossing each edge of the piece
            # (found over the whole genome, cut away by the margin)
            for at in (p.off + 5, p.off + p.n // 2, p.off + p.n - len(pat) - 3, p.off - 2,
                       p.off + p.n - len(pat) + 3):
                # the nearest start at or after `at` whose window passes the PAM
                for at in range(at, at + 200):
                    if at < 0 or at + len(pat) > p.clen:
                        continue
                    w = list(ctx.text(p.gpos + at, len(pat)))
                    if "N" not in w and H.window_passes(pat, w):
                        break
                else:
                    continue
                for _ in range(2):
                    w[r.randrange(len(w))] = r.choice("ACGT")
                guides.append("".join("N" if q != "N" else c for q, c in zip(pat, w)))
    body = "%s %d %d\n" % (pat, D, R) + "".join("%s %d\n" % (g, t) for g in guides)
    whole = H.record_lines(H.oracle_records(tmp_path, gp + "\n" + body, name="whole.txt"))
    margin = len(pat) + 2 * D + 2 * R + 8
    n = 0
    for k, pieces in enumerate(regions):
        rp = str(tmp_path / ("r%d.fa" % k))
        H.write_pieces(ctx, rp, pieces)
        part = H.record_lines(H.oracle_records(tmp_path, rp + "\n" + body, name="r%d.txt" % k))
        got = H.keep_in_pieces(part, pieces, margin, rebase=True)
        want = H.keep_in_pieces(whole, pieces, margin)
        assert got == want, (case, k, len(got), len(want))
        n += len(want)
    assert n > 0
