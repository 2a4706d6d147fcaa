"""Pins the oracle's IUPAC rule, finder and comparer (cas-offinder.cl:157-286)
against independent formulations: the IUPAC base sets, a brute-force window
search over each contig, strand symmetry and planted sites."""
import os

import pytest

import helpers as H
import oracle

CODES = list("ACGTRYSWKMBDHVN") + ["X", "U", "-", "n", "a"]
GENOME_CHARS = list("ACGTN") + ["X", "R", "n", "-", "Y"]


def test_iupac_table_matches_base_sets():
    for p in CODES:
        for c in GENOME_CHARS:
            assert oracle.iupac_mismatch(p, c) == H.set_mismatch(p, c), (p, c)


def test_iupac_special_cases():
    # a genome N mismatches every concrete base and matches every degenerate code
    for p in "ACGT":
        assert oracle.iupac_mismatch(p, "N")
    for p in "RYSWKMBDHV":
        assert not oracle.iupac_mismatch(p, "N")
    # NRG accepts ...NG (N at the degenerate R position)
    assert not oracle.iupac_mismatch("R", "N") and not oracle.iupac_mismatch("G", "G")


def test_revcomp_table():
    assert oracle.revcomp("ACGTRYMKHDBVNSW") == H.revcomp("ACGTRYMKHDBVNSW")
    assert oracle.revcomp("AACG") == "CGTT"


def _records_to_tuples(text, recs, id_of_key):
    """(key index, contig index, location, strand, mm, dna) for X records."""
    names = [n for n, _ in recs]
    out = []
    for l in H.record_lines(text):
        f = l.split("\t")
        assert f[1] == "X"
        out.append((id_of_key[f[0]], names.index(f[4]), int(f[5]), 0 if f[6] == "+" else 1, int(f[7]), f[3], f[2]))
    return sorted(out)


CASES = [
    # pattern, guide length (pattern length), n guides
    ("NNNNNNNNNNNNNNNNNNNNNRG", 23),
    ("NNNNNNNNNNNNNNNNNNNNNGG", 23),
    ("NNNNNNNNNNNNNNNNNNNNNNNGRRT", 27),
    ("TTTVNNNNNNNNNNNNNNNNNNNNNNN", 27),   # reversed (5') PAM, D = 0
    ("NNNNNNNNNNNN", 12),                    # all-N: every in-contig window
    ("NNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNGMTT", 44),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_finder_comparer_brute_force(tmp_path, case):
    pattern, P = CASES[case]
    r = H.rng(100 + case)
    recs = [("c1 first", H.random_seq(r, 700, n_runs=3)), ("c2", H.random_seq(r, 400, "ACGTACGTN")),
            ("c3", H.random_seq(r, P - 1)), ("c4", H.random_seq(r, P)), ("c5", H.random_seq(r, 300, "ACGTRY"))]
    gpath = str(tmp_path / "g.fa")
    H.write_fasta(gpath, recs)
    guides = []
    for gi in range(5):
        # guides are windows of the genome with edits, so hits exist at every threshold
        ci = r.randrange(len(recs))
        seq = recs[ci][1]
        if len(seq) < P:
            seq = H.random_seq(r, P)
        x = r.randrange(0, len(seq) - P + 1)
        g = list(seq[x:x + P])
        for _ in range(r.randint(0, 3)):
            g[r.randrange(P)] = r.choice("ACGTRYN")
        g = ["N" if pc != "N" else gc for pc, gc in zip(pattern, g)]  # N at PAM positions
        if gi == 4:
            g = [c if c != "N" else "N" for c in g]
        guides.append("".join(c if c in "ACGTRYN" else "A" for c in g))
    thr = [r.randint(0, 6) for _ in guides]
    text = "%s\n%s\n" % (gpath, pattern) + "".join("%s %d id%d\n" % (g, t, i) for i, (g, t) in enumerate(zip(guides, thr)))
    out = H.oracle_records(tmp_path, text)
    # key -> first id / threshold, as add_compare keeps them
    keys, id_of_key, first = [], {}, {}
    for i, (g, t) in enumerate(zip(guides, thr)):
        if g not in first:
            first[g] = (len(keys), "id%d" % i, t)
            keys.append((g, t))
            id_of_key["id%d" % i] = first[g][0]
    exp = []
    for ki, ci, x, s, mm in H.brute_hits(recs, pattern, keys):
        seq = recs[ci][1][x:x + P]
        k = keys[ki][0]
        if s:
            seq = H.revcomp(seq)
        dna = "".join(chr(ord(c) + 32) if H.set_mismatch(kc, c) else c for kc, c in zip(k, seq))
        exp.append((ki, ci, x, s, mm, dna, k))
    got = _records_to_tuples(out, recs, id_of_key)
    assert got == sorted(exp)


def test_strand_symmetry(tmp_path):
    r = H.rng(7)
    recs = [("a", H.random_seq(r, 900, n_runs=2)), ("b", H.random_seq(r, 500))]
    rc_recs = [(n, H.revcomp(s)) for n, s in recs]
    pattern = "N" * 20 + "NRG"
    guides = [recs[0][1][100:120] + "NNN", H.revcomp(recs[1][1][50:73])[:20] + "NNN"]
    lines = []
    for rr, nm in ((recs, "f.fa"), (rc_recs, "r.fa")):
        gp = str(tmp_path / nm)
        H.write_fasta(gp, rr)
        text = "%s\n%s\n" % (gp, pattern) + "".join("%s 6\n" % g for g in guides)
        lines.append(H.record_lines(H.oracle_records(tmp_path, text, name=nm + ".in")))
    fwd, rev = lines
    lens = {n: len(s) for n, s in recs}
    mapped = []
    for l in fwd:
        f = l.split("\t")
        loc = lens[f[4]] - 23 - int(f[5])
        f[5] = str(loc)
        f[6] = "-" if f[6] == "+" else "+"
        mapped.append("\t".join(f))
    assert sorted(mapped) == rev
    assert len(fwd) > 0


def test_planted_sites(tmp_path):
    r = H.rng(11)
    guide = "GACGTTACCGATTGCAAGCT"
    seq = list(H.random_seq(r, 3000))
    plants = {}
    for k, x in enumerate([200, 900, 1700, 2500]):
        site = list(guide)
        for p in r.sample(range(20), k):
            site[p] = {"A": "C", "C": "G", "G": "T", "T": "A"}[site[p]]
        seq[x:x + 23] = site + list("AGG")
        plants[x] = k
    gp = str(tmp_path / "p.fa")
    H.write_fasta(gp, [("chrP", "".join(seq))])
    out = H.oracle_records(tmp_path, "%s\n%s\n%s 3\n" % (gp, "N" * 21 + "GG", guide + "NNN"))
    found = {(int(l.split("\t")[5]), l.split("\t")[6]): int(l.split("\t")[7]) for l in H.record_lines(out)}
    for x, k in plants.items():
        if k <= 3:
            assert found.get((x, "+")) == k
        else:
            assert (x, "+") not in found


def test_window_never_crosses_contigs(tmp_path):
    site = "GACGTTACCGATTGCAAGCTAGG"
    recs = [("a", "ACGT" * 10 + site[:12]), ("b", site[12:] + "ACGT" * 10), ("c", "TTTT" + site + "TTTT")]
    gp = str(tmp_path / "x.fa")
    H.write_fasta(gp, recs)
    out = H.oracle_records(tmp_path, "%s\n%s\n%s 0\n" % (gp, "N" * 21 + "GG", site[:20] + "NNN"))
    got = [(l.split("\t")[4], l.split("\t")[5]) for l in H.record_lines(out)]
    assert got == [("c", "4")]


def test_all_n_pattern_counts_every_window(tmp_path):
    recs = [("a", "ACGTNACGTTGCA"), ("b", "ACG"), ("c", "TTTTT")]
    gp = str(tmp_path / "n.fa")
    H.write_fasta(gp, recs)
    out = H.oracle_records(tmp_path, "%s\nNNNN\nNNNN 0\n" % gp)
    lines = H.record_lines(out)
    exp = 2 * ((13 - 3) + 0 + (5 - 3))
    assert len(lines) == exp
