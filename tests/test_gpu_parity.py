"""GPU parity: the HIP path (plain and batched) against the CPU oracle, record
multisets compared bytewise after a canonical sort."""
import os
import random
import subprocess

import pytest

import casoff
import helpers as H
import workloads

pytestmark = pytest.mark.gpu

MODES = (casoff.MODE_PLAIN, casoff.MODE_BATCHED)


@pytest.fixture(scope="module", params=["tuned", "0", "1", "2", "3", "4"])
def ctx(request):
    """A context whose batched launches use the autotuner's variants, or every
    launch forced to variant 0 / 1 / 2 / 3 / 4 (fused: 32- / 8-key sign blocks /
    seed index / bit-set seed index keeping keys with 1 / 2 exact ranges, the
    indexes where the table qualifies; bulge: queued / in-place exact counts /
    bulge seed index where the table qualifies, in-place counts otherwise),
    so each kernel variant meets the oracle."""
    old = os.environ.pop("COF_FORCE_VARIANT", None)
    if request.param != "tuned":
        os.environ["COF_FORCE_VARIANT"] = request.param
    try:
        c = casoff.Context(0)
    finally:
        os.environ.pop("COF_FORCE_VARIANT", None)
        if old is not None:
            os.environ["COF_FORCE_VARIANT"] = old
    yield c
    c.close()


def product_lines(ctx, genome_path, text, mode):
    ctx.load(genome_path)
    q = ctx.query(text)
    r = ctx.search([q], mode)
    return H.record_lines(r.format(0))


def check(ctx, tmp_path, genome_path, text, label=""):
    exp = H.record_lines(H.oracle_records(tmp_path, text))
    for mode in MODES:
        got = product_lines(ctx, genome_path, text, mode)
        assert got == exp, "%s mode=%d: %d vs %d records" % (label, mode, len(got), len(exp))
    return exp


def test_plain_10mbp_config(ctx, tmp_path):
    gp = str(tmp_path / "g10.fa")
    casoff.synth_write(workloads.genome_10mbp(), gp, 0)
    exp = check(ctx, tmp_path, gp, workloads.plain_10mbp(gp), "10mbp")
    assert len(exp) > 0


PATTERNS = [
    ("N" * 21 + "RG", False),
    ("N" * 20 + "NGG", False),
    ("N" * 21 + "NNGRRT", False),
    ("N" * 20 + "NNAGAAW", False),
    ("N" * 24 + "NNNNGMTT", False),
    ("TTTV" + "N" * 23, True),
    ("N" * 12, False),
    ("N" * 40 + "NGMTT", False),          # care span > 32 -> two-word kernels
    ("TTN" + "N" * 48, True),
]


def random_case(seed, tmp_path):
    r = random.Random(seed)
    pattern, rev = PATTERNS[seed % len(PATTERNS)]
    D = r.choice([0, 0, 1, 2])
    R = r.choice([0, 0, 1, 2])
    P = len(pattern)
    recs = []
    for ci in range(r.randint(1, 4)):
        n = r.choice([0, 5, P - 1, P, r.randint(50, 3000)])
        alpha = r.choice(["ACGT", "ACGT", "ACGTN", "ACGTRYX"])
        recs.append(("chr%d %s" % (ci, r.choice(["", "desc"])), H.random_seq(r, n, alpha, n_runs=r.randint(0, 2))))
    guides = []
    allseq = "".join(s for _, s in recs)
    for gi in range(r.randint(1, 6)):
        if len(allseq) >= P and r.random() < 0.8:
            x = r.randrange(0, len(allseq) - P + 1)
            g = list(allseq[x:x + P])
            if r.random() < 0.5:
                g = list(H.revcomp("".join(g)))
        else:
            g = [r.choice("ACGT") for _ in range(P)]
        for _ in range(r.randint(0, 3)):
            g[r.randrange(P)] = r.choice("ACGTRN")
        g = ["N" if pc != "N" else (gc if gc in "ACGTRYSWKMBDHV" else "A") for pc, gc in zip(pattern, g)]
        if r.random() < 0.2:
            g = [c.lower() for c in g]
        guides.append("".join(g))
    if r.random() < 0.3 and guides:
        guides.append(guides[0])  # duplicate guide: first id / threshold wins
    lines = [pattern if (D == 0 and R == 0 and r.random() < 0.5) else "%s %d %d" % (pattern, D, R)]
    for i, g in enumerate(guides):
        lines.append("%s %d s%d" % (g, r.randint(0, 5), i))
    fmt = r.choice(["fa", "fa", "2bit", "dir"])
    if fmt == "fa":
        gp = str(tmp_path / ("c%d.fa" % seed))
        H.write_fasta(gp, recs, width=r.choice([7, 60, 1000]), crlf=r.random() < 0.2, lower=r.random() < 0.2)
    elif fmt == "2bit":
        gp = str(tmp_path / ("c%d.2bit" % seed))
        spec = casoff.synth_spec(seed=seed, n_contigs=r.randint(1, 3), total_len=r.randint(P, 5000), n_frac=0.05,
                                 n_run_min=3, n_run_max=40, lower_frac=0.3)
        casoff.synth_write(spec, gp, 1)
    else:
        gp = str(tmp_path / ("d%d" % seed))
        os.makedirs(gp)
        H.write_fasta(os.path.join(gp, "a.fa"), recs[: max(1, len(recs) // 2)])
        H.write_fasta(os.path.join(gp, "b.fa"), recs[max(1, len(recs) // 2):] or [("e", "ACGT")])
        with open(os.path.join(gp, "notes.txt"), "w") as f:
            f.write("not a genome\n")
        casoff.synth_write(casoff.synth_spec(seed=seed, n_contigs=2, total_len=2000), os.path.join(gp, "c.2bit"), 1)
    text = gp + "\n" + "\n".join(lines) + "\n"
    return gp, text


@pytest.mark.parametrize("seed", range(48))
def test_random_small_cases(ctx, tmp_path, seed):
    gp, text = random_case(seed, tmp_path)
    check(ctx, tmp_path, gp, text, "seed %d" % seed)


def test_semicolon_and_odd_bytes_in_fasta(ctx, tmp_path):
    gp = str(tmp_path / "odd.fa")
    with open(gp, "w") as f:
        f.write(">x y z\nACGTACGTTTGACCAGGTAC;GGTACCGATTAGGCCAAGG\nAC-GTAC*GTNNNNRYACGTAGGCCAAGGAAGGTT\r\n\n>e\n>w\nacgtAGGtt\n")
    text = "%s\nNNNNNNNRG 1 1\nACGTACGNN 4 first\nCCAAGGANN 3 second\n" % gp
    check(ctx, tmp_path, gp, text, "odd")


@pytest.mark.parametrize("devarg", ["G", "G0:4:2", "G3:1,0"])
def test_cli_end_to_end(ctx, tmp_path, devarg):
    gp = str(tmp_path / "g.fa")
    casoff.synth_write(casoff.synth_spec(seed=5, n_contigs=3, total_len=200000, n_frac=0.01), gp, 0)
    rr = random.Random(5)
    guides = ["".join(rr.choice("ACGT") for _ in range(20)) + "NNN" for _ in range(3)]
    text = "%s\n%s 1 1\n" % (gp, "N" * 21 + "RG") + "".join("%s 4\n" % g for g in guides)
    ip = str(tmp_path / "in.txt")
    with open(ip, "w") as f:
        f.write(text)
    op = str(tmp_path / "out.txt")
    sp = str(tmp_path / "sum.txt")
    p = subprocess.run([casoff.CLI_PATH, ip, devarg, op, "--summary", sp], capture_output=True, timeout=300)
    assert p.returncode == 0, p.stderr
    exp_text, exp_sum = H.oracle_records(tmp_path, text, name="cli_in.txt", summary=True)
    with open(op) as f:
        got = f.read()
    assert got.split("\n")[:2] == exp_text.split("\n")[:2]
    assert H.record_lines(got) == H.record_lines(exp_text)
    with open(sp) as f:
        assert f.read() == exp_sum


def test_batched_multi_query_matches_single(ctx, tmp_path):
    gp = str(tmp_path / "m.fa")
    casoff.synth_write(casoff.synth_spec(seed=9, n_contigs=4, total_len=300000, n_frac=0.01), gp, 0)
    ctx.load(gp)
    texts = [t for _, t in workloads.group_inputs(workloads.mixed_guides(2, 40, 3, 6, 0.5, 1), gp)]
    qs = [ctx.query(t) for t in texts]
    rb = ctx.search(qs, casoff.MODE_BATCHED)
    for i, t in enumerate(texts):
        exp = H.record_lines(H.oracle_records(tmp_path, t, name="q%d.txt" % i))
        assert H.record_lines(rb.format(i)) == exp


def _planted_guides(r, recs, pattern, core, n, iupac=False, edge=False):
    """Guides copied from windows that pass the PAM (either strand), with a
    planted DNA bulge (a guide base deleted), RNA bulge (a base inserted) and
    up to two substitutions, so that the bulge families have sites to find.
    edge: every other window within five positions of a sequence end."""
    P = len(pattern)
    rev = pattern[0] != "N"
    rpat = H.revcomp(pattern)
    out = []
    while len(out) < n:
        _, s = r.choice(recs)
        x = r.randrange(0, len(s) - P)
        if edge and len(out) % 2:
            x = r.choice([r.randint(0, 5), max(0, len(s) - P - r.randint(0, 5))])
        w = s[x:x + P]
        if H.window_passes(rpat, w):
            w = H.revcomp(w)
        elif not H.window_passes(pattern, w):
            continue
        c = list(w[P - core:] if rev else w[:core])
        kind = r.choice(["dna", "rna", "none"])
        if kind == "dna":  # the genome holds one base more than the guide
            del c[r.randrange(1, core - 1)]
            c = c + [r.choice("ACGT")] if rev else [r.choice("ACGT")] + c
        elif kind == "rna":  # the guide holds one base more than the genome
            c.insert(r.randrange(1, core - 1), r.choice("ACGT"))
            c = c[:core] if rev else c[1:]
        for _ in range(r.randint(0, 2)):
            c[r.randrange(core)] = r.choice("ACGT")
        if iupac and r.random() < 0.5:
            c[r.randrange(core)] = r.choice("RYKMSWBDHVN")
        out.append("".join(ch if ch in "ACGTRYKMSWBDHVN" else "A" for ch in c))
    return out


BULGE_CASES = [
    # (pattern, guide core length, D, R, IUPAC codes in guides)
    ("N" * 21 + "RG", 20, 1, 0, False),
    ("N" * 21 + "RG", 20, 0, 1, False),
    ("N" * 21 + "RG", 20, 1, 1, False),
    ("N" * 21 + "RG", 20, 2, 1, True),
    ("N" * 21 + "RG", 20, 2, 2, False),
    ("N" * 21 + "RG", 20, 3, 1, False),
    ("N" * 21 + "RG", 20, 4, 1, False),     # more shifts than the bulge kernel takes: key tables
    ("TTTV" + "N" * 23, 23, 1, 1, False),
    ("TTTV" + "N" * 23, 23, 2, 0, True),
    ("TTTV" + "N" * 23, 23, 0, 2, False),
    ("N" * 40 + "NGG", 40, 1, 1, False),    # guide span > 32: two-word masks
    ("N" * 20 + "NNAGAAW", 20, 1, 1, True),
]


@pytest.mark.parametrize("case", range(len(BULGE_CASES)))
def test_bulge_families(ctx, tmp_path, case):
    """Batched (device-side bulge expansion) and plain (one launch per key)
    against the oracle on planted DNA/RNA-bulge sites."""
    pattern, core, D, R, iupac = BULGE_CASES[case]
    r = random.Random(100 + case)
    recs = [("c%d" % i, H.random_seq(r, r.randint(20000, 60000), "ACGT", n_runs=r.randint(0, 3))) for i in range(4)]
    gp = str(tmp_path / "b.fa")
    H.write_fasta(gp, recs)
    rev = pattern[0] != "N"
    tail = "N" * (len(pattern) - core)
    cores = _planted_guides(r, recs, pattern, core, 6, iupac)
    guides = [tail + c if rev else c + tail for c in cores]
    guides.append(guides[0])  # duplicate line: its keys print under the first id
    lines = ["%s %d %d" % (pattern, D, R)] + ["%s %d g%d" % (g, r.randint(1, 5), i) for i, g in enumerate(guides)]
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "bulge case %d" % case)
    assert len(exp) > 0


@pytest.mark.parametrize("seed", range(5))
def test_bulge_index_holes_and_many_guides(ctx, tmp_path, seed):
    """Bulge groups of one care class with N at shared inner guide positions
    (ranges whose hole matches every code) and 33-120 guides (index rows of
    two and four words, full and one-guide words), thresholds 0-4, DNA and
    RNA bulges."""
    r = random.Random(700 + seed)
    recs = [("c%d" % i, H.random_seq(r, r.randint(30000, 70000), "ACGT", n_runs=r.randint(0, 2))) for i in range(3)]
    gp = str(tmp_path / "bh.fa")
    H.write_fasta(gp, recs)
    pattern = "N" * 21 + "RG"
    D, R = [(1, 1), (2, 1), (1, 2), (0, 1), (2, 0)][seed]
    n = [40, 75, 120, 33, 64][seed]
    holes = r.sample(range(2, 18), seed % 3)  # none, one or two shared inner N
    cores = []
    while len(cores) < n:
        for c in _planted_guides(r, recs, pattern, 20, min(8, n - len(cores))):
            g = list(c)
            for p in holes:
                g[p] = "N"
            cores.append("".join(g))
    lines = ["%s %d %d" % (pattern, D, R)] + ["%sNNN %d g%d" % (g, r.randint(0, 4), i) for i, g in enumerate(cores)]
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "bulge index seed %d" % seed)
    assert len(exp) > 0


@pytest.mark.parametrize("seed", range(4))
def test_seed_index_shared_inner_n(ctx, tmp_path, seed):
    """Guides with N at the same inner positions form one care class whose seed
    index lists a key under every code of a range holding its N; thresholds
    0-5 give 1-6 ranges."""
    r = random.Random(500 + seed)
    recs = [("c%d" % i, H.random_seq(r, r.randint(30000, 80000), "ACGTN" if i == 0 else "ACGT", n_runs=2))
            for i in range(3)]
    gp = str(tmp_path / "s.fa")
    H.write_fasta(gp, recs)
    pattern = "N" * 20 + "NGG"
    holes = r.sample(range(1, 19), r.randint(1, 3))
    guides = []
    for core in _planted_guides(r, recs, pattern, 20, 8):
        g = list(core)
        for p in holes:
            g[p] = "N"
        guides.append("".join(g) + "NNN")
    tmax = r.randint(0, 5)
    lines = [pattern] + ["%s %d" % (g, r.randint(0, tmax)) for g in guides]
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "seed %d" % seed)
    assert len(exp) > 0


@pytest.mark.parametrize("n_guides,tmax", [(30, 2), (60, 4), (100, 4), (100, 6), (250, 3), (420, 5), (700, 6)])
def test_seed_index_many_keys(ctx, tmp_path, n_guides, tmax):
    """Single-class NGG tables of 30-700 keys (bit-set rows of 1, 2, 4, 8, 16
    and 24 words; the wider ones read in 4-word chunks) with planted
    near-sites and thresholds up to tmax: every index variant must report
    exactly the oracle's records."""
    r = random.Random(700 + n_guides + tmax)
    recs = [("c%d" % i, H.random_seq(r, r.randint(40000, 90000), "ACGTN" if i == 1 else "ACGT", n_runs=2))
            for i in range(3)]
    gp = str(tmp_path / "k.fa")
    H.write_fasta(gp, recs)
    pattern = "N" * 20 + "NGG"
    guides = []
    while len(guides) < n_guides // 2:  # copies of NGG sites with up to tmax substitutions
        _, sq = r.choice(recs)
        x = r.randrange(0, len(sq) - 23)
        if sq[x + 21:x + 23] != "GG" or "N" in sq[x:x + 20]:
            continue
        c = list(sq[x:x + 20])
        for _ in range(r.randint(0, tmax)):
            c[r.randrange(20)] = r.choice("ACGT")
        guides.append(("".join(c) + "NNN", tmax))
    guides += [("".join(r.choice("ACGT") for _ in range(20)) + "NNN", r.randint(0, tmax))
               for _ in range(n_guides - len(guides))]
    r.shuffle(guides)
    lines = [pattern] + ["%s %d" % g for g in guides]
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "many keys %d/%d" % (n_guides, tmax))
    assert len(exp) > 0


@pytest.mark.parametrize("n_guides,holes,seed", [(150, 2, 0), (300, 1, 1), (420, 0, 2)])
def test_seed_index_reversed_pam_holes_mixed_thresholds(ctx, tmp_path, n_guides, holes, seed):
    """A reversed (5' PAM) pattern, guides sharing inner N positions and
    thresholds 0-5 mixed in one table: the bit-set index orders records by
    threshold, expands the holes under every code and, for the wide tables,
    runs several tile groups per workgroup over a genome of a few tiles."""
    r = random.Random(900 + seed)
    recs = [("c%d" % i, H.random_seq(r, r.randint(30000, 70000), "ACGTN" if i == 0 else "ACGT", n_runs=2))
            for i in range(3)]
    gp = str(tmp_path / "r.fa")
    H.write_fasta(gp, recs)
    pattern = "TTTV" + "N" * 23
    hole_pos = r.sample(range(3, 20), holes)
    guides = []
    while len(guides) < n_guides // 3:  # copies of TTTV sites (forward strand) with substitutions
        _, sq = r.choice(recs)
        x = r.randrange(0, len(sq) - 27)
        w = sq[x:x + 27]
        if not H.window_passes(pattern, w) or "N" in w[4:]:
            continue
        c = list(w[4:])
        for _ in range(r.randint(0, 3)):
            c[r.randrange(23)] = r.choice("ACGT")
        guides.append(c)
    while len(guides) < n_guides:
        guides.append([r.choice("ACGT") for _ in range(23)])
    lines = [pattern]
    for g in guides:
        for p in hole_pos:
            g[p] = "N"
        lines.append("NNNN%s %d" % ("".join(g), r.randint(0, 5)))
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "reversed holes %d/%d" % (n_guides, holes))
    assert len(exp) > 0


@pytest.mark.parametrize("parts", ["10", "12"])
@pytest.mark.parametrize("n_guides,tmax,reverse,holes", [(400, 6, False, 0), (700, 3, False, 0), (500, 5, True, 2)])
def test_mask_index_ten_and_twelve_ranges(tmp_path, parts, n_guides, tmax, reverse, holes):
    """Wide tables (12-24 words of records) with thresholds 0..tmax mixed keep
    per-word counts; COF_MASK_PARTS forces 10 or 12 ranges of 1-2 positions,
    so a word needs up to 12 exact ranges (4-bit counts of 10 and 12 rows):
    plain and batched records (variant 4 forced) equal the oracle's, with
    planted near-sites at every threshold, on both PAM sides and with holes
    (N inside ranges, listed under every code)."""
    r = random.Random(1100 + n_guides + tmax + int(parts))
    recs = [("c%d" % i, H.random_seq(r, r.randint(40000, 80000), "ACGTN" if i == 1 else "ACGT", n_runs=2))
            for i in range(3)]
    gp = str(tmp_path / "m.fa")
    H.write_fasta(gp, recs)
    pattern = "TTTV" + "N" * 23 if reverse else "N" * 20 + "NGG"
    glen = 23 if reverse else 20
    hole_pos = r.sample(range(3, glen - 2), holes)
    guides = []
    while len(guides) < n_guides // 2:  # copies of PAM sites with up to tmax substitutions
        _, sq = r.choice(recs)
        x = r.randrange(0, len(sq) - len(pattern))
        w = sq[x:x + len(pattern)]
        if not H.window_passes(pattern, w):
            continue
        core = list(w[4:] if reverse else w[:20])
        if "N" in core:
            continue
        t = r.randint(0, tmax)
        for _ in range(r.randint(0, t)):
            core[r.randrange(glen)] = r.choice("ACGT")
        guides.append((core, t))
    while len(guides) < n_guides:
        guides.append(([r.choice("ACGT") for _ in range(glen)], r.randint(0, tmax)))
    r.shuffle(guides)
    lines = [pattern]
    for core, t in guides:
        for p in hole_pos:
            core[p] = "N"
        lines.append(("NNNN%s %d" if reverse else "%sNNN %d") % ("".join(core), t))
    text = gp + "\n" + "\n".join(lines) + "\n"
    saved = {k: os.environ.pop(k, None) for k in ("COF_FORCE_VARIANT", "COF_MASK_PARTS")}
    os.environ["COF_FORCE_VARIANT"] = "4"
    os.environ["COF_MASK_PARTS"] = parts
    try:
        c = casoff.Context(0)
        try:
            exp = check(c, tmp_path, gp, text, "parts %s, %d keys, tmax %d" % (parts, n_guides, tmax))
            q = c.query(text)
            c.search([q], casoff.MODE_BATCHED)
            kernels = [rec["kernel"] for rec in c.launch_info()]
        finally:
            c.close()
    finally:
        for k, v in saved.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    assert len(exp) > 0
    assert any(", 3, %s, " % parts in k for k in kernels), kernels  # the 10- or 12-row kernel ran


MERGED_CASES = [
    # (pattern, guide core length, (D, R) of the queries sharing the pattern, IUPAC codes in guides)
    ("N" * 21 + "RG", 20, [(0, 1), (1, 0), (1, 1), (2, 0), (0, 2), (2, 2), (0, 0)], False),
    ("N" * 20 + "NGG", 20, [(1, 1), (0, 1), (1, 0)], False),
    ("TTTV" + "N" * 23, 23, [(1, 0), (0, 1), (2, 1), (1, 2)], False),
    ("N" * 20 + "NGG", 20, [(3, 1), (0, 2), (1, 0)], False),  # union of shifts past five: one launch per (D, R)
    ("N" * 21 + "RG", 20, [(1, 1), (0, 1), (2, 0)], True),  # general (IUPAC) table
    ("N" * 34 + "NGG", 34, [(1, 0), (0, 1), (2, 1)], False),  # care positions past 32: two-word masks
]


@pytest.mark.parametrize("case", range(len(MERGED_CASES)))
def test_merged_bulge_sizes(ctx, tmp_path, case):
    """Queries of one pattern with different DNA/RNA bulge sizes searched in
    one call share a bulge launch (the windows of the least padding; each
    guide moves its window start by its own extra padding and needs those
    positions inside the genome and free of ';').  On short contigs, with
    ';' inside sequence lines and N runs, so that many windows touch a
    contig edge or a separator, every query's records equal the oracle's
    for that query alone."""
    pattern, core, drs, iupac = MERGED_CASES[case]
    r = random.Random(1300 + case)
    recs = []
    for i in range(40):  # about 100 kbp: several tiles, separators near their edges
        seq = list(H.random_seq(r, r.randint(150, 4800), "ACGT", n_runs=r.randint(0, 2)))
        for _ in range(r.randint(0, 2)):
            seq[r.randrange(len(seq))] = ";"
        recs.append(("s%d" % i, "".join(seq)))
    gp = str(tmp_path / "m.fa")
    H.write_fasta(gp, recs)
    clean = [(n, s.replace(";", "A")) for n, s in recs]
    rev = pattern[0] != "N"
    tail = "N" * (len(pattern) - core)
    texts = []
    for qi, (D, R) in enumerate(drs):
        cores = _planted_guides(r, clean, pattern, core, 6, iupac=iupac, edge=True)
        guides = [tail + c if rev else c + tail for c in cores]
        head = "%s %d %d" % (pattern, D, R) if D + R else pattern
        lines = [head] + ["%s %d q%dg%d" % (g, r.randint(1, 5), qi, i) for i, g in enumerate(guides)]
        texts.append(gp + "\n" + "\n".join(lines) + "\n")
    ctx.load(gp)
    qs = [ctx.query(t) for t in texts]
    res = ctx.search(qs, casoff.MODE_BATCHED)
    bulge_launches = [rec for rec in ctx.launch_info() if rec["sig"].startswith("bulge")]
    n = 0
    for qi, t in enumerate(texts):
        exp = H.record_lines(H.oracle_records(tmp_path, t, name="q%d.txt" % qi))
        got = H.record_lines(res.format(qi))
        assert got == exp, "case %d query %d (D, R) = %s: %d vs %d records" % (case, qi, drs[qi], len(got), len(exp))
        n += len(exp)
    assert n > 0
    nb = sum(1 for D, R in drs if D + R)
    merged = max(D for D, _ in drs) + max(R for _, R in drs) + 1 <= 5
    assert len(bulge_launches) == (1 if merged else nb), bulge_launches


HALVES_CASES = [
    # (pattern, guide core length, D, R): care runs of 21 and 23 positions
    # (slice kernel widths of their own) and 22 (width 24, the positions
    # past the run masked)
    ("TTTV" + "N" * 23, 23, 1, 1),
    ("TTTV" + "N" * 23, 23, 0, 1),
    ("N" * 21 + "NGG", 21, 1, 1),
    ("TTTV" + "N" * 22, 22, 1, 1),
    # bulges of two bases (RNA: a family of gap 2)
    ("TTTV" + "N" * 23, 23, 0, 2),
    ("N" * 20 + "NGG", 20, 0, 2),
    ("N" * 21 + "NGG", 21, 2, 0),
]


@pytest.mark.parametrize("case", range(len(HALVES_CASES)))
def test_bulge_halves_at_threshold(ctx, tmp_path, case, request):
    """Bulge sites (one or two bases) planted with exactly `threshold`
    substitutions (5 or 6) and the bulge at every split of the guide: the bit-sliced filter bounds
    such words on the two halves of each family's split range, and a site
    whose count equals its threshold is found only if neither half
    overcounts (positions past the care run count nothing).  Records of
    plain and batched searches equal the oracle's."""
    pattern, core, D, R = HALVES_CASES[case]
    r = random.Random(1500 + case)
    recs = [("c%d" % i, H.random_seq(r, r.randint(40000, 70000), "ACGT")) for i in range(4)]
    gp = str(tmp_path / "h.fa")
    H.write_fasta(gp, recs)
    P = len(pattern)
    rev = pattern[0] != "N"
    rpat = H.revcomp(pattern)
    tail = "N" * (P - core)
    lines = ["%s %d %d" % (pattern, D, R)]
    while len(lines) < 31:
        _, s = r.choice(recs)
        x = r.randrange(0, len(s) - P)
        w = s[x:x + P]
        if H.window_passes(rpat, w):
            w = H.revcomp(w)
        elif not H.window_passes(pattern, w):
            continue
        c = list(w[P - core:] if rev else w[:core])
        split = 1 + (len(lines) * 7) % (core - 2)  # bulges spread over the guide
        if len(lines) % 2 and R:  # the guide holds R bases more than the genome
            for _ in range(R):
                c.insert(split, r.choice("ACGT"))
            c = c[:core] if rev else c[R:]
        elif D:  # the genome holds D bases more than the guide
            del c[split:split + D]
            pad = [r.choice("ACGT") for _ in range(D)]
            c = c + pad if rev else pad + c
        t = 5 + len(lines) % 2
        for p in r.sample(range(core), t):
            c[p] = r.choice([b for b in "ACGT" if b != c[p]])
        g = "".join(c)
        lines.append("%s %d g%d" % (tail + g if rev else g + tail, t, len(lines)))
    text = gp + "\n" + "\n".join(lines) + "\n"
    exp = check(ctx, tmp_path, gp, text, "halves case %d" % case)
    assert len(exp) >= 15
    if request.node.callspec.params["ctx"] == "3":  # the slice kernel ran, at the run's width
        width = {20: 20, 21: 21, 22: 24, 23: 23}[core]
        kernels = [rec["kernel"] for rec in ctx.launch_info()]
        assert any(k.startswith("cof::bulge_slice_kernel<%d, %d, %d," % (D, R, width)) for k in kernels) or any(
            k.startswith("cof::bulge_slice_kernel<%d, %d, %d," % (R, D, width)) for k in kernels), kernels
