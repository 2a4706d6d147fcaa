"""Shared test helpers: IUPAC base sets, small genome writers and an
independent brute-force search written from the IUPAC definitions (sets),
not from the oracle's if-chains."""
import os
import random

# IUPAC nucleotide codes (README.md:139-151): the bases each code stands for.
IUPAC = {
    "A": "A", "C": "C", "G": "G", "T": "T",
    "R": "AG", "Y": "CT", "S": "GC", "W": "AT", "K": "GT", "M": "AC",
    "B": "CGT", "D": "AGT", "H": "ACT", "V": "ACG",
}
COMP = {"A": "T", "T": "A", "G": "C", "C": "G", "R": "Y", "Y": "R", "M": "K", "K": "M",
        "H": "D", "D": "H", "B": "V", "V": "B"}


def set_mismatch(p, c):
    """Mismatch of pattern/key char p against genome char c from the base sets:
    a listed code mismatches a concrete base outside its set; a non-ACGT
    genome char mismatches only the four concrete codes; N and any unlisted
    code never mismatch."""
    if p not in IUPAC:
        return False
    if c in "ACGT":
        return c not in IUPAC[p]
    return p in "ACGT"


def revcomp(s):
    return "".join(COMP.get(c, c) for c in reversed(s))


def write_fasta(path, records, width=60, crlf=False, lower=False):
    nl = "\r\n" if crlf else "\n"
    with open(path, "w", newline="") as f:
        for name, seq in records:
            f.write(">" + name + nl)
            if lower:
                seq = seq.lower()
            for i in range(0, len(seq), width):
                f.write(seq[i:i + width] + nl)


def random_seq(rng, n, alphabet="ACGT", n_runs=0, run_len=(3, 12)):
    s = [rng.choice(alphabet) for _ in range(n)]
    for _ in range(n_runs):
        L = rng.randint(*run_len)
        st = rng.randint(0, max(0, n - L))
        for i in range(st, min(n, st + L)):
            s[i] = "N"
    return "".join(s)


def window_passes(pat, w):
    return not any(set_mismatch(p, c) for p, c in zip(pat, w))


def brute_hits(records, pattern, keys):
    """Every (key index, contig index, offset, strand, mm) for windows inside one
    contig.  `keys` = [(key string, threshold)]; forward windows test the
    pattern and the key, reverse windows test revcomp(pattern) / revcomp(key)."""
    P = len(pattern)
    rpat = revcomp(pattern)
    out = []
    for ci, (_, seq) in enumerate(records):
        for x in range(0, len(seq) - P + 1):
            w = seq[x:x + P]
            fok = window_passes(pattern, w)
            rok = window_passes(rpat, w)
            if not (fok or rok):
                continue
            for ki, (k, t) in enumerate(keys):
                if fok:
                    mm = sum(set_mismatch(p, c) for p, c in zip(k, w))
                    if mm <= t:
                        out.append((ki, ci, x, 0, mm))
                if rok:
                    rk = revcomp(k)
                    mm = sum(set_mismatch(p, c) for p, c in zip(rk, w))
                    if mm <= t:
                        out.append((ki, ci, x, 1, mm))
    return out


def parse_dump(d):
    lines = d.rstrip("\n").split("\n")
    P, D, R, pat = lines[0].split(" ", 3)
    keys = []
    for l in lines[1:]:
        k, i, t, infos = l.split("\t")
        keys.append((k, i, int(t), [(x.rsplit(":", 1)[0], int(x.rsplit(":", 1)[1])) for x in infos.split(",")]))
    return int(P), int(D), int(R), pat, keys


def record_lines(text):
    return sorted(l for l in text.split("\n") if l and not l.startswith("#"))


def oracle_records(tmp_path, input_text, name="in.txt", summary=False):
    import oracle
    ip = os.path.join(str(tmp_path), name)
    op = ip + ".out"
    sp = ip + ".sum" if summary else ""
    with open(ip, "w") as f:
        f.write(input_text)
    rc, err = oracle.run(ip, op, sp)
    assert rc == 0, err
    with open(op) as f:
        text = f.read()
    if summary:
        with open(sp) as f:
            return text, f.read()
    return text


def rng(seed):
    return random.Random(seed)


# ---- genome regions: oracle runs on slices of a large resident genome

class Piece:
    """Positions [off, off + n) of contig `name` (index ci, global start
    gpos, length clen) of a context's genome."""

    def __init__(self, ctx, ci, off, n):
        self.name, self.gpos, self.clen, _ = ctx.contigs()[ci]
        self.ci, self.off = ci, off
        self.n = min(n, self.clen - off)
        assert self.n > 0 and off >= 0

    def bounds(self, margin):
        """Printed locations kept from this piece: records whose whole window
        lies inside the piece on both sides.  A printed location is the
        window start plus at most D + bulge size, so `margin` >= padded
        pattern length + D + R keeps only windows that start and end inside;
        a piece edge at a contig edge needs no margin (no window crosses it)."""
        lo = self.off + (margin if self.off > 0 else 0)
        hi = self.off + self.n - (margin if self.off + self.n < self.clen else 0)
        return lo, hi


def write_pieces(ctx, path, pieces, width=60):
    """One FASTA record per piece, named as its contig."""
    with open(path, "w") as f:
        for p in pieces:
            f.write(">" + p.name + "\n")
            seq = ctx.text(p.gpos + p.off, p.n)
            for i in range(0, len(seq), width):
                f.write(seq[i:i + width] + "\n")


def keep_in_pieces(lines, pieces, margin, rebase=False):
    """Record lines whose printed location falls inside a piece's bounds; with
    rebase the lines come from a run over write_pieces' FASTA, and their
    locations are moved back to contig coordinates first."""
    by_name = {p.name: p for p in pieces}
    out = []
    for l in lines:
        f = l.split("\t")
        p = by_name.get(f[4])
        if p is None:
            continue
        loc = int(f[5]) + (p.off if rebase else 0)
        lo, hi = p.bounds(margin)
        if lo <= loc <= hi:
            if rebase:
                f[5] = str(loc)
                l = "\t".join(f)
            out.append(l)
    return sorted(out)


def hits_near_pieces(hits, pieces, margin):
    """Hits whose window start lies within `margin` of a piece (global
    coordinates): the only ones that can print a location inside it."""
    import numpy as np
    sel = np.zeros(len(hits), dtype=bool)
    for p in pieces:
        a = p.gpos + p.off
        sel |= (hits["pos"] + margin >= a) & (hits["pos"] < a + p.n + margin)
    return hits[sel]
