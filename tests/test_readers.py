"""Genome readers: the product loader (host-only context), the oracle's
readers and the reference's own read_fasta.cpp / read_twobit.cpp (oracle/_ref,
when built) must hold the same text, record names and record offsets."""
import os
import struct

import pytest

import casoff
import helpers as H
import oracle


def product_dump(path):
    c = casoff.Context(-1)
    try:
        n = c.load(path)
    except casoff.CofError:
        return None
    if n == 0:
        return 1, [], ""
    total, _, _ = c.info()
    return 0, [(p, nm) for nm, p, _, _ in c.contigs()], c.text(0, total)


def check_all(path, expect_status=0):
    o = oracle.genome_dump(path)
    p = product_dump(path)
    assert o[0] == expect_status
    assert p == o, path
    r = oracle.ref_readers(path)
    if r is not None:
        assert r == o, path
    return o


FASTA_CASES = {
    "multi": ">chr1 desc here\nACGTNNNNacgt\nGGCCA\n>chr2\nTTTT\n",
    "crlf": ">a\r\nACGT\r\nacgt\r\n>b\r\n\r\nNNAC\r\n",
    "empty_lines": ">a\n\n\nACGT\n\n>b\n\n",
    "empty_records": ">a\n>b\n>c\nAC\n>d\n",
    "odd_bytes": ">x\nAC-GT*XNRYmk;acgt\n>y y\n  ACG T\n",
    "no_trailing_newline": ">a\nACGTACGT",
    "gt_inside": ">a\nAC>GT\n",
    "header_only": ">only\n",
}


@pytest.mark.parametrize("name", sorted(FASTA_CASES))
def test_fasta_variants(tmp_path, name):
    p = str(tmp_path / (name + ".fa"))
    with open(p, "w", newline="") as f:
        f.write(FASTA_CASES[name])
    check_all(p)


def test_not_a_genome(tmp_path):
    p = str(tmp_path / "x.txt")
    with open(p, "w") as f:
        f.write("hello\n")
    assert oracle.genome_dump(p)[0] == 1
    assert product_dump(p) == (1, [], "")
    r = oracle.ref_readers(p)
    if r is not None:
        assert r[0] == 1


@pytest.mark.parametrize("seed", range(4))
def test_synthetic_fasta_and_2bit(tmp_path, seed):
    spec = casoff.synth_spec(seed=seed, n_contigs=1 + seed, total_len=5000 + 777 * seed, n_frac=0.03,
                             n_run_min=1, n_run_max=200, lower_frac=0.4, line_width=50 + seed)
    fa, tb = str(tmp_path / "g.fa"), str(tmp_path / "g.2bit")
    casoff.synth_write(spec, fa, 0)
    casoff.synth_write(spec, tb, 1)
    a = check_all(fa)
    b = check_all(tb)
    assert a == b
    # the packed generator holds the same genome as its FASTA
    c = casoff.Context(-1)
    c.synth(spec)
    assert c.text(0, c.info()[0]) == a[2]


def _write_2bit(path, recs, nblocks=None, masks=None):
    """Minimal UCSC .2bit v0 writer for handcrafted records (T/C/A/G = 0..3)."""
    code = {"T": 0, "C": 1, "A": 2, "G": 3}
    nblocks = nblocks or [[] for _ in recs]
    masks = masks or [[] for _ in recs]
    hdr = struct.pack("<IIII", 0x1A412743, 0, len(recs), 0)
    index, body = b"", b""
    off = 16 + sum(1 + len(n) + 4 for n, _ in recs)
    for i, (n, s) in enumerate(recs):
        index += struct.pack("<B", len(n)) + n.encode() + struct.pack("<I", off + len(body))
        rec = struct.pack("<II", len(s), len(nblocks[i]))
        rec += b"".join(struct.pack("<I", a) for a, _ in nblocks[i]) + b"".join(struct.pack("<I", b) for _, b in nblocks[i])
        rec += struct.pack("<I", len(masks[i])) + b"".join(struct.pack("<I", a) for a, _ in masks[i])
        rec += b"".join(struct.pack("<I", b) for _, b in masks[i]) + struct.pack("<I", 0)
        raw = bytearray((len(s) + 3) // 4)
        for j, ch in enumerate(s):
            raw[j >> 2] |= code.get(ch, 0) << ((3 - (j & 3)) * 2)
        body += rec + bytes(raw)
    with open(path, "wb") as f:
        f.write(hdr + index + body)


def test_handcrafted_2bit(tmp_path):
    p = str(tmp_path / "h.2bit")
    _write_2bit(p, [("c1", "ACGTACGTTGCAT"), ("c2", "G"), ("c3", ""), ("c4", "TTTTCCCCAAAAGGGGT")],
                nblocks=[[(2, 3)], [], [], [(0, 4), (16, 1)]], masks=[[(0, 5)], [], [], [(3, 3)]])
    o = check_all(p)
    assert o[2] == "ACNNNCGTTGCAT;G;;NNNNCCCCAAAAGGGGN"
    assert [x[1] for x in o[1]] == ["c1", "c2", "c3", "c4"]


def test_directory_load_concatenates_files(tmp_path):
    d = tmp_path / "gdir"
    d.mkdir()
    H.write_fasta(str(d / "a.fa"), [("a1", "ACGTACGT"), ("a2", "GG")])
    H.write_fasta(str(d / "b.fa"), [("b1", "TTTT")])
    with open(str(d / "readme"), "w") as f:
        f.write("no genome\n")
    c = casoff.Context(-1)
    assert c.load(str(d)) == 2
    total, nfiles, ncontigs = c.info()
    assert nfiles == 2 and ncontigs == 3
    names = sorted(nm for nm, _, _, _ in c.contigs())
    assert names == ["a1", "a2", "b1"]
    for nm, pos, ln, f in c.contigs():
        exp = {"a1": "ACGTACGT", "a2": "GG", "b1": "TTTT"}[nm]
        assert c.text(pos, ln) == exp


def test_missing_path_is_an_error(tmp_path):
    c = casoff.Context(-1)
    with pytest.raises(casoff.CofError):
        c.load(str(tmp_path / "nope"))


def _messy_fasta(path, seed, n_records, approx_len):
    """A FASTA file of several MB (the loaders split it into per-thread
    chunks): mixed LF/CRLF line ends, empty lines, lowercase, non-ACGT bytes,
    empty records, N runs and ';' bytes inside sequences."""
    r = H.rng(seed)
    parts = []
    for i in range(n_records):
        eol = "\r\n" if r.random() < 0.3 else "\n"
        parts.append(">rec%d %s%s" % (i, "x" * r.randint(0, 5), eol))
        n = 0 if r.random() < 0.05 else r.randint(1, 2 * approx_len // n_records)
        seq = list(r.choices("ACGTacgtACGTacgtNnRY-;", k=n))
        for _ in range(n // 5000):
            st, run = r.randrange(n), r.randint(30, 300)
            seq[st:st + run] = "N" * len(seq[st:st + run])
        seq = "".join(seq)
        pos = 0
        while pos < len(seq):
            w = r.choice([1, 31, 32, 33, 60, 61, 80, 997])
            parts.append(seq[pos:pos + w] + eol)
            if r.random() < 0.02:
                parts.append(eol)
            pos += w
    with open(path, "w", newline="") as f:
        f.write("".join(parts))


@pytest.mark.parametrize("seed", range(2))
def test_large_fasta_threaded_chunks(tmp_path, seed):
    p = str(tmp_path / "big.fa")
    _messy_fasta(p, seed, n_records=40 + 200 * seed, approx_len=6_000_000)
    assert os.path.getsize(p) > 5_000_000
    check_all(p)


def test_large_2bit_and_fasta_agree(tmp_path):
    spec = casoff.synth_spec(seed=7, n_contigs=5, total_len=9_000_001, n_frac=0.04, n_run_min=1,
                             n_run_max=3000, lower_frac=0.5, line_width=61)
    fa, tb = str(tmp_path / "g.fa"), str(tmp_path / "g.2bit")
    casoff.synth_write(spec, fa, 0)
    casoff.synth_write(spec, tb, 1)
    a = check_all(fa)
    b = check_all(tb)
    assert a == b


def test_symlink_opens_raw_target_against_cwd(tmp_path, monkeypatch):
    """A symbolic link is opened through its raw target text
    (read_fasta.cpp:17-22, read_twobit.cpp:58-63): a relative target resolves
    against the working directory, not the link's directory.  The file the
    link's own directory would give holds another genome, so reading it
    would fail the comparison."""
    cwd, gdir = tmp_path / "cwd", tmp_path / "gdir"
    (cwd / "data").mkdir(parents=True)
    (gdir / "data").mkdir(parents=True)
    H.write_fasta(str(cwd / "data" / "real.fa"), [("from_cwd", "ACGTNNACGT")])
    H.write_fasta(str(gdir / "data" / "real.fa"), [("from_link_dir", "TTTTTTTT")])
    _write_2bit(str(tmp_path / "abs.2bit"), [("abs2bit", "GGGACCCT")])
    os.symlink("data/real.fa", str(gdir / "rel.fa"))
    os.symlink(str(tmp_path / "abs.2bit"), str(gdir / "abs.2bit"))
    monkeypatch.chdir(str(cwd))
    o = check_all(str(gdir / "rel.fa"))
    assert o[1] == [(0, "from_cwd")] and o[2] == "ACGTNNACGT"
    o = check_all(str(gdir / "abs.2bit"))
    assert o[1] == [(0, "abs2bit")] and o[2] == "GGGACCCT"
    # a directory load reads the links (DT_LNK entries, main.cpp:141-143) the same way
    c = casoff.Context(-1)
    assert c.load(str(gdir)) == 2
    assert sorted(nm for nm, _, _, _ in c.contigs()) == ["abs2bit", "from_cwd"]
