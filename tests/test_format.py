"""Record formatting (cas-offinder.cpp:394-455).

Oracle side: every forward-PAM '+' record (X, DNA, RNA) and the reversed-PAM
'+' X / RNA records must re-align to the genome at their printed location
(uppercased DNA string == genome text, gaps where the bulge type says,
lowercase exactly at IUPAC mismatches, lowercase count == Mismatches).  The
records of the reversed-orientation branch (reversed-PAM records of both
strands, forward-PAM '-' records) are checked against the genome window at
their location and a parsed compare key of their guide (properties W1-W4
below), and planted mistakes show that each property bites.  Product side:
the product's formatter on brute-force hit sets must print the oracle's
records byte for byte, quirk branches included."""
import random

import numpy as np
import pytest

import casoff
import helpers as H
import oracle


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
    recs, gp, text = genome_and_input(tmp_path, 20 + D + R, pattern, D, R, n_guides=6, thr=5, rev=True)
    lines = H.record_lines(H.oracle_records(tmp_path, text))
    assert len(lines) > 0
    assert realign_checks(recs, lines, pattern, D, True) > 0


# ---- records of the reversed-orientation branch (cas-offinder.cpp:403-449:
# reversed-PAM records of both strands, forward-PAM '-' records).  Each record
# is checked against the genome and against a compare key of its guide as the
# (separately pinned) parser lists it, with properties that fix every
# character except the reference's choice of columns:
#   W1  the printed location is the window start: the P-window at it (reverse
#       complemented on '-') is the hit window, its mismatches against the
#       key summing to Mismatches (W3);
#   W2  the DNA string without gaps, uppercased, is that window cut at the
#       trim of the record's type (X: D, DNA: none, RNA: D + bulge size);
#   W3  lowercase sits exactly on the IUPAC mismatches of the aligned columns
#       (columns where both strings hold a letter), and the lowercase count
#       plus the key's mismatches in the trimmed window tail is Mismatches;
#   W4  the crRNA without gap characters is the key minus the info's columns:
#       X: the D-trim, DNA: the bulge columns [j, j + size) (where the gaps
#       sit), RNA: the bulge bases at [j, j + size) and the D + size trim.

def _info_kind(first):
    return ("X" if int(first) == 0 else "DNA", int(first)) if first.isdigit() else ("RNA", len(first))


def rv_record_ok(seqs, P, D, key, info, f):
    """None when the record fields f satisfy W1-W4 for (key, info), else the failed property."""
    gid, typ, rna, dna, chrom, loc, dr, mm, bs = f
    first, sec = info
    kind, size = _info_kind(first)
    if (kind, size) != (typ, bs):
        return "type"
    j = abs(sec)
    seq = seqs[chrom]
    if loc < 0 or loc + P > len(seq):
        return "W1"
    W = seq[loc:loc + P] if dr == "+" else H.revcomp(seq[loc:loc + P])
    plain = dna.replace("-", "")
    n = {"X": P - D, "DNA": P, "RNA": P - D - bs}[typ]
    if len(plain) != n or plain.upper() != W[:n]:
        return "W2"
    if len(rna) != len(dna):
        return "W3"
    for r_c, d_c in zip(rna, dna):
        if r_c != "-" and d_c != "-" and d_c.islower() != H.set_mismatch(r_c, d_c.upper()):
            return "W3"
    tail = sum(H.set_mismatch(key[k], W[k]) for k in range(n, P))
    if sum(c.islower() for c in dna) + tail != mm:
        return "W3"
    if typ == "X":
        ok = rna == key[:P - D]
    elif typ == "DNA":
        ok = rna[j:j + bs] == "-" * bs and rna.replace("-", "") == key[:j] + key[j + bs:]
    else:
        ok = "-" not in rna and rna[j:j + bs] == first and rna[:j] + rna[j + bs:] == key[:P - D - bs]
    return None if ok else "W4"


def rv_checks(recs, text, lines, rev):
    """Checks every reversed-orientation record; returns their count."""
    seqs = {n: s for n, s in recs}
    ok, d = oracle.parse_dump(text)
    assert ok, d
    P, D, R, _, keys = H.parse_dump(d)
    by_id = {}
    for k, gid, t, infos in keys:
        by_id.setdefault(gid, []).extend((k, inf) for inf in infos)
    checked = 0
    for l in lines:
        f = l.split("\t")
        f[5], f[7], f[8] = int(f[5]), int(f[7]), int(f[8])
        if not (rev or f[6] == "-"):
            continue
        fails = [rv_record_ok(seqs, P, D, k, inf, f) for k, inf in by_id[f[0]]]
        assert None in fails, (l, sorted(set(fails), key=str))
        checked += 1
    return checked


RV_CASES = [("N" * 20 + "NRG", 0, 0, False), ("N" * 20 + "NRG", 1, 0, False), ("N" * 20 + "NRG", 2, 1, False),
            ("N" * 20 + "NGG", 1, 2, False), ("TTTV" + "N" * 20, 0, 0, True), ("TTTV" + "N" * 20, 1, 1, True),
            ("TTTV" + "N" * 20, 2, 2, True), ("TTTV" + "N" * 20, 0, 2, True)]


@pytest.mark.parametrize("case", range(len(RV_CASES)))
def test_oracle_reversed_orientation_records(tmp_path, case):
    pattern, D, R, rev = RV_CASES[case]
    recs, gp, text = genome_and_input(tmp_path, 60 + case, pattern, D, R, n_guides=4, thr=5, rev=rev)
    lines = H.record_lines(H.oracle_records(tmp_path, text))
    n = rv_checks(recs, text, lines, rev)
    assert n > 0
    # every (type, bulge size) the input asks for is among the checked records
    kinds = {(l.split("\t")[1], int(l.split("\t")[8])) for l in lines if rev or l.split("\t")[6] == "-"}
    assert ("X", 0) in kinds
    assert any(t == "DNA" for t, _ in kinds) == (D > 0) and any(t == "RNA" for t, _ in kinds) == (R > 0)


def _flip_case(dna):
    k = next(i for i, c in enumerate(dna) if c.isalpha() and c.upper() != "N")
    return dna[:k] + dna[k].swapcase() + dna[k + 1:]


def _move_gap(rna):
    """The gap block one column to the right (the base after it moves before it)."""
    g, b = rna.index("-"), rna.count("-")
    return rna[:g] + rna[g + b] + "-" * b + rna[g + b + 1:]


PLANTS = {
    "loc+1": lambda f: f[:5] + [str(int(f[5]) + 1)] + f[6:],
    "case flip": lambda f: f[:3] + [_flip_case(f[3])] + f[4:],
    "trim kept": lambda f: f[:2] + [f[2] + "N", f[3] + "A"] + f[4:],
    "gap moved": lambda f: f[:2] + [_move_gap(f[2])] + f[3:],
    "mm+1": lambda f: f[:7] + [str(int(f[7]) + 1)] + f[8:],
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_reversed_orientation_checks_catch_planted_mistakes(tmp_path, plant):
    """Each property set rejects oracle output with one plausible mistake
    planted in every record (a shifted location, a lowercase bit, a missing
    trim, a gap one column off, a miscount)."""
    recs, gp, text = genome_and_input(tmp_path, 62, "N" * 20 + "NRG", 2, 1, n_guides=4, thr=5)
    lines = [l for l in H.record_lines(H.oracle_records(tmp_path, text)) if l.split("\t")[6] == "-"]
    if plant == "gap moved":  # DNA records whose gap has a different base on each side
        def movable(r):
            g, b = r.index("-"), r.count("-")
            return 0 < g and g + b < len(r) and r[g - 1] != r[g + b]
        lines = [l for l in lines if l.split("\t")[1] == "DNA" and movable(l.split("\t")[2])]
    assert lines
    seqs = {n: s for n, s in recs}
    ok, d = oracle.parse_dump(text)
    P, D, R, _, keys = H.parse_dump(d)
    by_id = {}
    for k, gid, t, infos in keys:
        by_id.setdefault(gid, []).extend((k, inf) for inf in infos)
    for l in lines:
        f = PLANTS[plant](l.split("\t"))
        f[5], f[7], f[8] = int(f[5]), int(f[7]), int(f[8])
        assert None not in [rv_record_ok(seqs, P, D, k, inf, f) for k, inf in by_id[f[0]]], (plant, l)


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
