"""Seeded synthetic workloads: genome specs and reference-format input files.

Five workloads, by what they are:
  plain_10mbp     10 Mbp / 4 contigs (seed 0), NRG pattern, 4 guides, mm<=5
  genome_3g_ngg   3.1 Gbp / 24 contigs (seed 1), NGG, 100 guides, mm<=4
  bulge_3g_nrg    same genome, NRG with DNA bulge 2 / RNA bulge 1, 20 guides, mm<=4
  mixed_pam_2000  same genome, 2,000 guides of four Cas9 PAMs (seed 2), mixed
                  thresholds and bulges, one input per (pattern, D, R) group
  serving_stream  Poisson request stream (seed 3), 1-256 guides per request

Guides are drawn with numpy's PCG64 (np.random.default_rng) from the stated
seeds; genomes come from the library's seeded generator (cof_synth_spec).
"""
import math

import numpy as np

import casoff

BASES = np.array(list("ACGT"))

# (name, pattern without bulge slack, guide length, trailing N count)
PAMS = {
    "SpCas9": ("N" * 20 + "NGG", 20, 3),
    "SaCas9": ("N" * 21 + "NNGRRT", 21, 6),
    "StCas9": ("N" * 20 + "NNAGAAW", 20, 7),
    "NmCas9": ("N" * 24 + "NNNNGMTT", 24, 8),
}


def genome_10mbp():
    return casoff.synth_spec(seed=0, n_contigs=4, length_mode=0, total_len=10_000_000, gc=0.41, n_frac=0.005,
                             n_run_min=100, n_run_max=5000, lower_frac=0.0, line_width=60)


def genome_3g(total_len=3_100_000_000, n_contigs=24):
    """3.1 Gbp, 24 contigs log-uniform in [57, 249] Mbp rescaled, GC 41%, ~5% N, 50% soft-masked."""
    return casoff.synth_spec(seed=1, n_contigs=n_contigs, length_mode=1, total_len=total_len,
                             min_len=57_000_000, max_len=249_000_000, gc=0.41, n_frac=0.05,
                             n_run_min=100, n_run_max=5000, lower_frac=0.5, line_width=60)


def random_guides(rng, n, glen, tail):
    return ["".join(rng.choice(BASES, glen)) + "N" * tail for _ in range(n)]


def input_text(chrdir, pattern, guides, thresholds, ids=None, dna=0, rna=0):
    lines = [chrdir, pattern if (dna == 0 and rna == 0) else "%s %d %d" % (pattern, dna, rna)]
    for i, g in enumerate(guides):
        t = thresholds[i] if hasattr(thresholds, "__len__") else thresholds
        lines.append("%s %d" % (g, t) + ((" " + ids[i]) if ids else ""))
    return "\n".join(lines) + "\n"


def plain_10mbp(chrdir="genome"):
    rng = np.random.default_rng(0)
    guides = random_guides(rng, 4, 20, 3)
    return input_text(chrdir, "N" * 21 + "RG", guides, 5)


def genome_3g_ngg(chrdir="genome", n_guides=100):
    rng = np.random.default_rng(1)
    guides = random_guides(rng, n_guides, 20, 3)
    return input_text(chrdir, "N" * 20 + "NGG", guides, 4)


def bulge_3g_nrg(chrdir="genome", n_guides=20):
    rng = np.random.default_rng(11)
    guides = random_guides(rng, n_guides, 20, 3)
    return input_text(chrdir, "N" * 20 + "NRG", guides, 4, dna=2, rna=1)


def mixed_guides(seed, n, thr_lo, thr_hi, bulge_frac, bulge_max):
    """Guides of the mixed-PAM draw: per guide (pam, guide, threshold, D, R, id)."""
    rng = np.random.default_rng(seed)
    counts = {"SpCas9": 0.4, "SaCas9": 0.3, "StCas9": 0.2, "NmCas9": 0.1}
    kinds = []
    for k, f in counts.items():
        kinds += [k] * int(round(f * n))
    kinds = kinds[:n] + ["SpCas9"] * max(0, n - len(kinds))
    kinds = [kinds[i] for i in rng.permutation(n)]
    out = []
    for i, k in enumerate(kinds):
        pat, glen, tail = PAMS[k]
        g = "".join(rng.choice(BASES, glen)) + "N" * tail
        t = int(rng.integers(thr_lo, thr_hi + 1))
        if rng.random() < bulge_frac:
            d = int(rng.integers(0, bulge_max + 1))
            r = int(rng.integers(0, bulge_max + 1))
        else:
            d = r = 0
        out.append((k, g, t, d, r, "g%d" % i))
    return out


def group_inputs(guides, chrdir="genome"):
    """One reference input per (pattern, D, R) group, guides in draw order."""
    groups = {}
    for k, g, t, d, r, gid in guides:
        groups.setdefault((k, d, r), []).append((g, t, gid))
    texts = []
    for (k, d, r), items in sorted(groups.items()):
        pat = PAMS[k][0]
        texts.append(((k, d, r), input_text(chrdir, pat, [x[0] for x in items], [x[1] for x in items],
                                            [x[2] for x in items], dna=d, rna=r)))
    return texts


def mixed_pam_2000(chrdir="genome", n=2000):
    return group_inputs(mixed_guides(2, n, 3, 6, 0.25, 1), chrdir)


def serving_stream(n_requests, qps, chrdir="genome", seed=3, max_guides=256):
    """[(arrival_s, [(group_key, input_text), ...], n_guides), ...] Poisson arrivals."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / qps)
        n = int(math.floor(math.exp(rng.uniform(0, math.log(max_guides + 1)))))
        n = max(1, min(max_guides, n))
        g = mixed_guides(int(rng.integers(0, 2**31)), n, 2, 6, 0.25, 2)
        reqs.append((t, group_inputs(g, chrdir), n))
    return reqs
