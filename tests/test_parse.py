"""Input parser and bulge expansion (cas-offinder.cpp:590-694).

The oracle is pinned by closed forms (key and info counts) and by the
relation of every key to its guide (removing the DNA-bulge gap / re-inserting
the RNA-bulge bases gives the guide back); the product parser must then give
the same dump as the oracle on every input, including the error cases."""
import random

import pytest

import casoff
import helpers as H
import oracle


def both(text):
    ok, od = oracle.parse_dump(text)
    try:
        pd = casoff.Query(None, text).dump()
        pok = True
    except casoff.CofError as e:
        pd, pok = str(e), False
    return ok, od, pok, pd


def test_key_counts_closed_form():
    r = random.Random(1)
    for _ in range(30):
        g = "".join(r.choice("ACGT") for _ in range(20))
        text = "x\n%s 2 1\n%s 3\n" % ("N" * 20 + "NRG", g + "NNN")
        ok, d = oracle.parse_dump(text)
        assert ok
        P, D, R, pat, keys = H.parse_dump(d)
        assert (P, D, R) == (25, 2, 1) and pat == "NN" + "N" * 20 + "NRG"
        infos = [i for k in keys for i in k[3]]
        # D=2: 2 x 21 DNA positions (j in [0, 21)), RNA i=1: j in [0, 19)
        assert len(infos) == 2 * 21 + 19
        x_keys = [k for k in keys if any(f == "0" for f, _ in k[3])]
        assert len(x_keys) == 1 and x_keys[0][0] == "NN" + g + "NNN"
        assert [f for f, _ in x_keys[0][3]] == ["0", "0"]  # the X record prints D times
        # distinct RNA keys = distinct single-base deletions of g[0:20] at j < 19
        dels = {g[:j] + g[j + 1:] for j in range(19)}
        rna_keys = [k for k in keys if any(not f.isdigit() for f, _ in k[3])]
        assert len(rna_keys) == len(dels)


def _undo(key, info, D, rev, glen):
    """The guide a key came from: DNA bulge -> drop the gap and slack, RNA -> re-insert bases."""
    first, second = info
    j = -second if second < 0 or rev else second
    if first.isdigit():
        i = int(first)
        if i == 0:
            return key[:glen] if rev else key[D:]
        body = key[:j] + key[j + i:] if rev else key[D - i:][:j] + key[D - i:][j + i:]
        return body[:glen] if rev else body
    i = len(first)
    body = key[:glen - i] if rev else key[D + i:]
    return body[:j] + first + body[j:]


@pytest.mark.parametrize("rev", [False, True])
def test_keys_map_back_to_guides(rev):
    r = random.Random(5 + rev)
    for _ in range(20):
        D, R = r.randint(0, 2), r.randint(0, 2)
        n = 20
        guides = []
        for _ in range(3):
            core = "".join(r.choice("ACGTRY") for _ in range(n))
            guides.append(("NNNN" + core) if rev else (core + "NNN"))
        pat = ("TTTV" + "N" * n) if rev else ("N" * n + "NGG")
        text = "x\n%s %d %d\n" % (pat, D, R) + "".join("%s 3 g%d\n" % (g, i) for i, g in enumerate(guides))
        ok, d = oracle.parse_dump(text)
        assert ok, d
        P, D2, R2, padded, keys = H.parse_dump(d)
        assert P == len(pat) + D
        for key, gid, t, infos in keys:
            assert len(key) == P
            srcs = {_undo(key, inf, D, rev, len(guides[0])) for inf in infos}
            assert srcs <= set(guides), (key, infos)
            # first guide (by input order) that produces the key names it
            owner = min(guides.index(s) for s in srcs)
            assert gid == "g%d" % owner


SPECIAL = [
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN 5\nCGCCAGCGTCAGCGACAGGTNNN 5\n",
    "x\r\nNNNNNNNNNNNNNNNNNNNNNRG 2 1\r\nGGCCGACCTGTCGCTGACGCNNN 5 a\r\nCGCCAGCGTCAGCGACAGGTNNN 5 b\r\n",
    "x\nnnnnnnnnnnnnnnnnnnnnnrg 1 1\nggccgacctgtcgctgacgcnnn 3\n",          # lowercase n: reversed PAM
    "x\nTTTVNNNNNNNNNNNNNNNNNNNNNNN 1 2\nNNNNACGTACGTACGTACGTACGTACG 4\n",
    "x\nNNNNNNNNNNNNNNNNNNNNNRG 1 1\nGGCCGACCTGTCGCTGACGCNNN 5\nGGCCGACCTGTCGCTGACGCNNN 2 dup\n",
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN -1\n",              # threshold wraps to 65535
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN 5\n\nAAAA 3\n",    # blank line ends the guides
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN 5",               # no trailing newline
    "x\nNNNNNN 1 0\nNNNNNN 0\n",                                              # all-N guide, D>0: no keys
    "x\nNNNNNN 0 0\nNNNNNN 0\n",
    "x\nNNNNNNRG 0 2\nANNNNNNN 1\n",                                           # RNA loop runs past the end
    "x\nNNNNNNRG 0 1\nNNNNNNNN 1\n",                                           # all-N guide with RNA bulge
    "x\nNNNNNNNNNNNNNNNNNNNNNRG 1\nGGCCGACCTGTCGCTGACGCNNN 5\n",                # 2 tokens on the pattern line
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNN 5\n",                  # length mismatch
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN 5 a\nCGCCAGCGTCAGCGACAGGTNNN 5\n",  # inconsistent
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN\n",                   # missing threshold
    "x\n",
    "",
    "x\nNNNNNNNNNNNNNNNNNNNNNRG\n   \n",
    "x\nNNNRG 1 1\nAGNNN 0\n",
    "x\nNNNNNNNNNNNNNNNNNNNNNRG 2 2\nGGCCGAAAAGTCGCTGACGCNNN 5\nGGCCGACCTGTCGCTGACCCNNN 4\n",
]


@pytest.mark.parametrize("i", range(len(SPECIAL)))
def test_product_parser_matches_oracle(i):
    ok, od, pok, pd = both(SPECIAL[i])
    assert ok == pok, (od, pd)
    if ok:
        assert pd == od


def test_product_parser_random():
    r = random.Random(42)
    alphabet = "ACGTNRYSWKMBDHV"
    for _ in range(300):
        n = r.randint(1, 12)
        rev = r.random() < 0.3
        pat = (r.choice("ACGT") + "".join(r.choice(alphabet) for _ in range(n - 1))) if rev else \
              ("N" + "".join(r.choice(alphabet) for _ in range(n - 1)))
        D, R = r.randint(0, 3), r.randint(0, 3)
        lines = ["g", "%s %d %d" % (pat, D, R) if r.random() < 0.8 else pat]
        if len(lines[1].split()) == 1:
            D = R = 0
        for k in range(r.randint(0, 5)):
            g = "".join(r.choice(alphabet) for _ in range(n))
            lines.append("%s %d%s" % (g, r.randint(-2, 8), (" id%d" % k) if r.random() < 0.5 else ""))
        text = "\n".join(lines) + ("\n" if r.random() < 0.8 else "")
        ok, od, pok, pd = both(text)
        assert ok == pok, text
        if ok:
            assert pd == od, text
