"""Record formatting (cas-offinder.cpp:394-455).

Oracle side: every forward-PAM '+' record (X, DNA, RNA) and every D = 0
record must re-align to the genome at its printed location (uppercased DNA
string == genome text, gaps where the bulge type says, lowercase exactly at
IUPAC mismatches, lowercase count == Mismatches).  Product side: the
product's formatter on brute-force hit sets must print the oracle's records
byte for byte, quirk branches included."""
import random

import numpy as np
import pytest

import casoff
import helpers as H


def genome_and_input(tmp_path, seed, pattern, D, R, n_guides=3, thr=4, rev=False):
    r = random.Random(seed)
    recs = [("chrA", H.random_seq(r, 1500, n_runs=2)), ("chrB x", H.random_seq(r, 900, "ACGTN")),
            ("chrC", H.random_seq(r, 40))]
    P0 = len(pattern)
    guides = []
    for _ in range(n_guides):
        c = recs[r.randrange(2)][1]
        rc = H.revcomp(pattern)
        xs = [x for x in range(len(c) - P0) if H.window_passes(pattern, c[x:x + P0]) or H.window_passes(rc, c[x:x + P0])]
        x = r.choice(xs)
        g = list(c[x:x + P0])
        if not H.window_passes(pattern, c[x:x + P0]):
            g = list(H.revcomp("".join(g)))
        g[r.randrange(P0)] = r.choice("ACGT")
        g = ["N" if pc != "N" else (gc if gc != "N" else "A") for pc, gc in zip(pattern, g)]
        guides.append("".join(g))
    gp = str(tmp_path / ("g%d.fa" % seed))
    H.write_fasta(gp, recs)
    text = "%s\n%s %d %d\n" % (gp, pattern, D, R) + "".join("%s %d q%d\n" % (g, thr, i) for i, g in enumerate(guides))
    return recs, gp, text


def realign_checks(recs, lines, pattern, D, rev):
    seqs = {n: s for n, s in recs}
    glen = len(pattern)
    checked = 0
    for l in lines:
        gid, typ, rna, dna, chrom, loc, dr, mm, bs = l.split("\t")
        loc, mm, bs = int(loc), int(mm), int(bs)
        seq = seqs[chrom]
        plain = dna.replace("-", "")
        # aligned branches: '+' of a forward PAM (all types), '+' X / RNA of a
        # reversed PAM, and '-' records whose key has no leading slack
        # (X with D = 0, DNA with bulge size D); other '-' records keep the
        # slack in their strings (cas-offinder.cpp:407-409)
        clean_minus = dr == "-" and not rev and ((typ == "X" and D == 0) or (typ == "DNA" and bs == D))
        if dr == "+" and (not rev or typ in ("X", "RNA")) or clean_minus:
            if dr == "+":
                ref = seq[loc:loc + len(plain)]
            else:
                ref = H.revcomp(seq[loc:loc + len(plain)])
            assert plain.upper() == ref, l
            lower = [i for i, c in enumerate(dna) if c.islower()]
            assert len(lower) == mm, l
            if typ == "X":
                assert len(rna) == len(dna) and "-" not in rna
                assert lower == [i for i, (k, c) in enumerate(zip(rna, dna.upper())) if H.set_mismatch(k, c)], l
            if typ == "DNA" and not rev:
                assert rna.count("-") == bs and len(rna) == len(dna) == glen + bs
                assert rna.replace("-", "") == rna.replace("-", "")[:glen]
            if typ == "RNA":
                assert dna.count("-") == bs and len(rna) == len(dna) == glen
            checked += 1
    return checked


@pytest.mark.parametrize("D,R", [(0, 0), (1, 0), (0, 1), (2, 1), (1, 2)])
def test_oracle_records_realign_forward_pam(tmp_path, D, R):
    pattern = "N" * 20 + "NRG"
    recs, gp, text = genome_and_input(tmp_path, 10 + 3 * D + R, pattern, D, R)
    lines = H.record_lines(H.oracle_records(tmp_path, text))
    assert len(lines) > 0
    assert realign_checks(recs, lines, pattern, D, False) > 0


@pytest.mark.parametrize("D,R", [(0, 0), (1, 1), (2, 0)])
def test_oracle_records_realign_reversed_pam(tmp_path, D, R):
    pattern = "TTTV" + "N" * 20
    recs, gp, text = genome_and_input(tmp_path, 20 + D + R, pattern, D, R, thr=5, rev=True)
    lines = H.record_lines(H.oracle_records(tmp_path, text))
    assert len(lines) > 0
    realign_checks(recs, lines, pattern, D, True)


def product_format_brute(ctx, recs, text):
    q = ctx.query(text)
    P, D, R, pat, keys = H.parse_dump(q.dump())
    contigs = {nm: pos for nm, pos, _, _ in ctx.contigs()}
    bh = H.brute_hits(recs, pat, [(k, t) for k, _, t, _ in keys])
    hits = np.zeros(len(bh), dtype=casoff.HIT_DTYPE)
    for i, (ki, ci, x, s, mm) in enumerate(bh):
        hits[i] = (contigs[recs[ci][0]] + x, ki, mm, s, 0)
    return H.record_lines(ctx.format_hits(q, hits))


CASES = [("N" * 20 + "NRG", 0, 0, False), ("N" * 20 + "NRG", 2, 1, False), ("N" * 20 + "NGG", 1, 2, False),
         ("TTTV" + "N" * 20, 1, 1, True), ("TTTV" + "N" * 20, 2, 2, True), ("N" * 10, 1, 1, False)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_product_formatter_matches_oracle(tmp_path, case):
    pattern, D, R, rev = CASES[case]
    recs, gp, text = genome_and_input(tmp_path, 40 + case, pattern, D, R, thr=5, rev=rev)
    exp = H.record_lines(H.oracle_records(tmp_path, text))
    ctx = casoff.Context(-1)
    ctx.load(gp)
    got = product_format_brute(ctx, recs, text)
    assert got == exp


def test_summary_table_matches_oracle(tmp_path):
    recs, gp, text = genome_and_input(tmp_path, 77, "N" * 20 + "NRG", 1, 1, thr=5)
    exp_text, exp_sum = H.oracle_records(tmp_path, text, summary=True)
    # recount the summary from the records: "id,type,bulge size,mm" -> count
    counts = {}
    for l in H.record_lines(exp_text):
        f = l.split("\t")
        k = "%s,%s,%s,%s" % (f[0], f[1], f[8], f[7])
        counts[k] = counts.get(k, 0) + 1
    want = "".join("##Id=%s;Bulge Type=%s;Bulge Size=%s;Mismatches=%s;Number of Found Targets=%d\n"
                   % (tuple(k.split(",")) + (v,)) for k, v in sorted(counts.items()))
    assert exp_sum == want
