/*
 * casoff_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C, CPU-only restatement of the Cas-OFFinder 3.0.0 search path
 * (genome readers -> input parser with bulge expansion -> finder -> comparer
 * -> record formatter).  Only the test-suite, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.  The product
 * library (libcasoff.so) and the cas-offinder CLI never link or call it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in the header of casoff_oracle.c.
 */
#ifndef CASOFF_ORACLE_H
#define CASOFF_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One genome file as the reference holds it: text of all records joined by
 * ';', record names and record start offsets (read_fasta.cpp:11-57). */
typedef struct {
    char *text;
    size_t len, cap;
    char **names;
    uint64_t *chrpos;
    size_t nchr, capchr;
} orc_genome;

typedef struct {
    char *first;   /* "0", "<bulge size>" or the RNA-bulge bases */
    int second;    /* bulge index, negated for reversed PAMs */
} orc_info;

typedef struct {
    char *key;     /* compare key, length P */
    char *id;
    uint16_t threshold;
    orc_info *infos;
    size_t ninfo, capinfo;
} orc_compare;

typedef struct {
    char *chrdir;
    char *pattern;       /* padded pattern (cas-offinder.cpp:619-625) */
    unsigned P;          /* m_patternlen */
    unsigned dna_bulge;  /* m_dnabulgesize */
    unsigned rna_bulge;  /* m_rnabulgesize */
    orc_compare *cmp;    /* in std::map (bytewise key) order after orc_parse_input */
    size_t ncmp, capcmp;
} orc_query;

typedef struct {
    uint64_t locus;  /* text offset */
    uint16_t mm;
    char dir;        /* '+' or '-' */
} orc_hit;

/* Readers: 0 ok, 1 not this format (reference semantics). */
int orc_read_fasta(const char *path, orc_genome *g);
int orc_read_twobit(const char *path, orc_genome *g);
void orc_genome_free(orc_genome *g);

/* Parser: 0 ok, -1 error (message copied to err). */
int orc_parse_input(const char *text, size_t len, orc_query *q, char *err, size_t errlen);
void orc_query_free(orc_query *q);

/* Parsed query as text (see casoff_oracle.c); 0 ok, -1 error message. */
int orc_parse_dump(const char *text, size_t len, char **out);
void orc_free(void *p);
char *orc_genome_dump(const char *path, size_t *len);
int orc_iupac_mismatch(char p, char c);
void orc_revcomp(char *seq, size_t n);

/* Finder over the whole text; returns the candidate count; loci and flags are malloc-ed. */
size_t orc_finder(const orc_genome *g, const orc_query *q, uint64_t **loci, char **flags, int nthreads);
/* Comparer of one compare key over the candidates; returns hit count. */
size_t orc_comparer(const orc_genome *g, const orc_query *q, const char *key, uint16_t threshold,
                    const uint64_t *loci, const char *flags, size_t nloci, orc_hit **hits, int nthreads);

/* Whole reference run (main.cpp:72-157): input file, genome path from its first
 * line, output path ("-" = stdout), optional summary path ("" = none).
 * Returns 0 ok, 1 error. */
int orc_run(const char *input_path, const char *output_path, const char *summary_path, int nthreads);

/* Times finder + all comparer passes (no formatting) on the genome named in the
 * input file; returns seconds, fills candidate / pair / hit counts. */
double orc_time_search(const char *input_path, int nthreads, uint64_t *genome_len,
                       uint64_t *ncand, uint64_t *npairs, uint64_t *nhits);

#ifdef __cplusplus
}
#endif
#endif
