/*
 * casoff_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the Cas-OFFinder 3.0.0 search path.  It exists to check
 * the HIP product path (tests, smoke) and to be timed as the CPU baseline
 * (bench.py).  Nothing in the product links or calls it.
 *
 * It follows the reference text literally: the genome is one byte per base,
 * records joined by ';'; the finder and comparer are one position / one
 * candidate per iteration with the 14-way IUPAC if-chain and the early exit;
 * the formatter does the same substring surgery as compareAll.
 *
 * Parity status per function (what pins it besides this literal reading):
 *   orc_read_fasta / orc_read_twobit  pinned against the reference's own
 *       read_fasta.cpp / read_twobit.cpp compiled into oracle/_ref (when
 *       the reference tree is present; tests/test_readers.py), against
 *       handcrafted files and against the UCSC 2bit layout.
 *   iupac_mismatch      pinned by the closed-form IUPAC table test
 *       (README.md:139-151 base sets; tests/test_oracle_core.py).
 *   orc_finder / orc_comparer  pinned by brute force over the IUPAC sets,
 *       strand symmetry (reverse-complemented genome), planted k-mismatch
 *       sites and contig-separator invariants (tests/test_oracle_core.py).
 *   orc_parse_input     pinned by key-count closed forms and by the
 *       gap-removal relation of every DNA/RNA key to its guide
 *       (tests/test_parse.py).
 *   format_hit          forward-PAM '+' records and reversed-PAM '+' X / RNA
 *       records: pinned by re-aligning the crRNA/DNA strings to the genome
 *       text at the printed location.  Reversed-orientation records
 *       (reversed-PAM records of both strands, forward-PAM '-' records,
 *       cas-offinder.cpp:403-449): pinned by the window at the printed
 *       location, the trim of each type, lowercase on the IUPAC mismatches
 *       of aligned columns (with the trimmed tail's mismatches making up
 *       Mismatches) and the crRNA as the parsed key minus the info's
 *       columns (tests/test_format.py).  Unpinned beyond the literal
 *       reading: which columns the reference chose to keep there (the D or
 *       D + size trim taken from the end, the gap at the key's bulge index
 *       rather than at its inserted columns), so the slack N content of
 *       those crRNA strings is fixed by no property.
 */
#define _GNU_SOURCE
#include "casoff_oracle.h"

#include <ctype.h>
#include <dirent.h>
#include <errno.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>
#include <limits.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_VERSION "3.0.0"

/* ------------------------------------------------------------------ utils */

static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) { fprintf(stderr, "oracle: out of memory\n"); exit(2); }
    return q;
}
static char *xstrndup(const char *s, size_t n) {
    char *r = (char *)xrealloc(NULL, n + 1);
    memcpy(r, s, n); r[n] = 0; return r;
}
static char *xstrdup(const char *s) { return xstrndup(s, strlen(s)); }

typedef struct { char *s; size_t n, cap; } sbuf;
static void sb_putn(sbuf *b, const char *s, size_t n) {
    if (b->n + n + 1 > b->cap) {
        size_t c = b->cap ? b->cap : 64;
        while (b->n + n + 1 > c) c *= 2;
        b->s = (char *)xrealloc(b->s, c); b->cap = c;
    }
    memcpy(b->s + b->n, s, n); b->n += n; b->s[b->n] = 0;
}
static void sb_puts(sbuf *b, const char *s) { sb_putn(b, s, strlen(s)); }
static void sb_rep(sbuf *b, char c, size_t n) { for (size_t i = 0; i < n; i++) sb_putn(b, &c, 1); }
static void sb_clear(sbuf *b) { b->n = 0; if (b->s) b->s[0] = 0; }
/* std::string::substr(pos, n) with the reference's clamping; pos > len is an error. */
static int sb_substr(sbuf *b, const char *s, size_t slen, size_t pos, size_t n) {
    if (pos > slen) return -1;
    if (n > slen - pos) n = slen - pos;
    sb_putn(b, s + pos, n);
    return 0;
}

static double now_s(void) {
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* ------------------------------------------------------------- IUPAC rule */

/* cas-offinder.cl:171-200 (identical in :30-59, :97-142, :228-273 and
 * cas-offinder.cpp:302-315): 1 when pattern/key char p mismatches genome char c. */
static inline int iupac_mismatch(char p, char c) {
    return (p == 'R' && (c == 'C' || c == 'T')) ||
           (p == 'Y' && (c == 'A' || c == 'G')) ||
           (p == 'K' && (c == 'A' || c == 'C')) ||
           (p == 'M' && (c == 'G' || c == 'T')) ||
           (p == 'W' && (c == 'C' || c == 'G')) ||
           (p == 'S' && (c == 'A' || c == 'T')) ||
           (p == 'H' && (c == 'G')) ||
           (p == 'B' && (c == 'A')) ||
           (p == 'V' && (c == 'T')) ||
           (p == 'D' && (c == 'C')) ||
           (p == 'A' && (c != 'A')) ||
           (p == 'G' && (c != 'G')) ||
           (p == 'C' && (c != 'C')) ||
           (p == 'T' && (c != 'T'));
}

int orc_iupac_mismatch(char p, char c) { return iupac_mismatch(p, c); }

/* cas-offinder.cpp:35-58 set_complementary_sequence */
static void revcomp_inplace(char *seq, size_t n) {
    for (size_t i = 0; i < n; i++) {
        char c = seq[i];
        if (c == 'A') c = 'T'; else if (c == 'T') c = 'A';
        else if (c == 'G') c = 'C'; else if (c == 'C') c = 'G';
        else if (c == 'R') c = 'Y'; else if (c == 'Y') c = 'R';
        else if (c == 'M') c = 'K'; else if (c == 'K') c = 'M';
        else if (c == 'H') c = 'D'; else if (c == 'D') c = 'H';
        else if (c == 'B') c = 'V'; else if (c == 'V') c = 'B';
        seq[i] = c;
    }
    for (size_t i = 0; i < n / 2; i++) {
        char t = seq[i]; seq[i] = seq[n - i - 1]; seq[n - i - 1] = t;
    }
}

void orc_revcomp(char *seq, size_t n) { revcomp_inplace(seq, n); }

/* cas-offinder.cpp:60-70 set_seq_flags: indices of non-N chars, -1 after the
 * last one when any N is present. */
static void seq_flags(int *flags, const char *seq, size_t n) {
    size_t i, k = 0;
    for (i = 0; i < n; i++)
        if (seq[i] != 'N') flags[k++] = (int)i;
    if (i != k) flags[k] = -1;
}

/* ----------------------------------------------------------------- genome */

static void g_reset(orc_genome *g) {
    for (size_t i = 0; i < g->nchr; i++) free(g->names[i]);
    g->nchr = 0; g->len = 0;
    if (g->text) g->text[0] = 0;
}
static void g_append(orc_genome *g, const char *s, size_t n) {
    if (g->len + n + 1 > g->cap) {
        size_t c = g->cap ? g->cap : 1 << 16;
        while (g->len + n + 1 > c) c *= 2;
        g->text = (char *)xrealloc(g->text, c); g->cap = c;
    }
    memcpy(g->text + g->len, s, n); g->len += n; g->text[g->len] = 0;
}
static void g_addchr(orc_genome *g, char *name, uint64_t pos) {
    if (g->nchr + 1 > g->capchr) {
        g->capchr = g->capchr ? 2 * g->capchr : 16;
        g->names = (char **)xrealloc(g->names, g->capchr * sizeof(char *));
        g->chrpos = (uint64_t *)xrealloc(g->chrpos, g->capchr * sizeof(uint64_t));
    }
    g->names[g->nchr] = name; g->chrpos[g->nchr] = pos; g->nchr++;
}
void orc_genome_free(orc_genome *g) {
    g_reset(g); free(g->text); free(g->names); free(g->chrpos);
    memset(g, 0, sizeof(*g));
}

/* read_fasta.cpp:17-22, read_twobit.cpp:58-63: a symbolic link is opened
 * through its raw target text (relative targets resolve against the working
 * directory); any other path as given. */
static FILE *open_genome(const char *path) {
    char buf[PATH_MAX + 1];
    memset(buf, 0, sizeof(buf));
    ssize_t n = readlink(path, buf, PATH_MAX);
    return fopen(n >= 0 ? buf : path, "rb");
}

/* read_fasta.cpp:11-57.  One string per file, records joined by ';', each
 * sequence line uppercased, one trailing '\r' dropped, empty lines skipped,
 * name = header line after '>'.  Returns 1 when the first byte is not '>'. */
int orc_read_fasta(const char *path, orc_genome *g) {
    g_reset(g);
    FILE *f = open_genome(path);
    if (!f) return 1;
    int c0 = fgetc(f);
    if (c0 != '>') { fclose(f); return 1; }
    rewind(f);
    char *line = NULL; size_t lcap = 0; ssize_t n;
    int have_chr = 0;
    while ((n = getline(&line, &lcap, f)) >= 0) {
        if (n > 0 && line[n - 1] == '\n') n--;          /* getline strips the delimiter */
        if (n == 0) continue;                            /* :33 if (!line.empty()) */
        if (line[n - 1] == '\r') n--;                    /* :34-35 */
        if (n > 0 && line[0] == '>') {                   /* :36 */
            if (have_chr) g_append(g, ";", 1);           /* :41 */
            g_addchr(g, xstrndup(line + 1, (size_t)n - 1), g->len);
            have_chr = 1;
        } else {
            for (ssize_t i = 0; i < n; i++) line[i] = (char)toupper((unsigned char)line[i]);
            g_append(g, line, (size_t)n);                /* :45-47 */
        }
    }
    free(line);
    fclose(f);
    return 0;
}

static int rd_u32(FILE *f, uint32_t *v) { return fread(v, 4, 1, f) == 1 ? 0 : -1; }

/* read_twobit.cpp:44-131.  UCSC .2bit v0 read sequentially; T/C/A/G = 0/1/2/3,
 * MSB-first, N blocks set to 'N', mask blocks ignored, ';' between records. */
int orc_read_twobit(const char *path, orc_genome *g) {
    g_reset(g);
    FILE *f = open_genome(path);
    if (!f) return 1;
    uint32_t magic = 0, ver = 1, cnt = 0, tmp;
    if (rd_u32(f, &magic) || magic != 440477507u) { fclose(f); return 1; }   /* :66 */
    if (rd_u32(f, &ver) || ver != 0) { fclose(f); return 1; }                /* :70 */
    rd_u32(f, &cnt);
    fseek(f, 4, SEEK_CUR);
    char **names = (char **)xrealloc(NULL, (cnt ? cnt : 1) * sizeof(char *));
    for (uint32_t i = 0; i < cnt; i++) {                                      /* :76-82 */
        unsigned char nl = 0; char buf[256];
        if (fread(&nl, 1, 1, f) != 1) nl = 0;
        size_t got = fread(buf, 1, nl, f); buf[got] = 0;
        names[i] = xstrdup(buf);
        fseek(f, 4, SEEK_CUR);
    }
    static const char b2s[4] = {'T', 'C', 'A', 'G'};                          /* :30-42 */
    uint64_t pos0 = 0;
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t chrlen = 0, nbc = 0, mbc = 0;
        rd_u32(f, &chrlen); rd_u32(f, &nbc);
        uint32_t *ns = (uint32_t *)xrealloc(NULL, (nbc ? nbc : 1) * 4);
        uint32_t *nz = (uint32_t *)xrealloc(NULL, (nbc ? nbc : 1) * 4);
        for (uint32_t j = 0; j < nbc; j++) rd_u32(f, &ns[j]);
        for (uint32_t j = 0; j < nbc; j++) rd_u32(f, &nz[j]);
        rd_u32(f, &mbc);
        fseek(f, (long)mbc * 8 + 4, SEEK_CUR);                                  /* :93 */
        uint32_t raw = chrlen / 4 + ((chrlen & 3) ? 1 : 0);
        unsigned char *rb = (unsigned char *)xrealloc(NULL, raw + 1);
        size_t got = fread(rb, 1, raw, f);
        if (got < raw) memset(rb + got, 0, raw - got);
        char *cb = (char *)xrealloc(NULL, (size_t)chrlen + 1);
        for (uint32_t p = 0; p < chrlen; p++)
            cb[p] = b2s[(rb[p >> 2] >> ((3 - (p & 3)) * 2)) & 3];
        for (uint32_t j = 0; j < nbc; j++) {
            uint64_t s = ns[j], e = (uint64_t)ns[j] + nz[j];
            if (e > chrlen) e = chrlen;
            for (uint64_t p = s; p < e; p++) cb[p] = 'N';
        }
        g_addchr(g, names[i], pos0);
        g_append(g, cb, chrlen);
        if (i + 1 < cnt) { g_append(g, ";", 1); pos0 = g->len; }              /* :123-127 */
        free(cb); free(rb); free(ns); free(nz);
    }
    if (cnt == 0) g_addchr(g, NULL, 0), g->nchr = 0;
    (void)tmp;
    free(names);
    fclose(f);
    return 0;
}

/* ----------------------------------------------------------------- parser */

typedef struct { const char *s; size_t n, pos; int eof, fail; } istr;

/* std::getline semantics over a byte buffer. */
static int is_getline(istr *in, sbuf *line) {
    if (in->fail || in->eof) { if (!in->fail && in->eof) { sb_clear(line); in->fail = 1; } return 0; }
    sb_clear(line);
    if (in->pos >= in->n) { in->eof = 1; in->fail = 1; return 0; }
    size_t e = in->pos;
    while (e < in->n && in->s[e] != '\n') e++;
    sb_putn(line, in->s + in->pos, e - in->pos);
    if (e < in->n) in->pos = e + 1; else { in->pos = e; in->eof = 1; }
    return 1;
}

/* cas-offinder.cpp:570-582 parseLine.  Returns 1 line read, 0 none, -1 error. */
static int parse_line(istr *in, sbuf *line, int expect_line) {
    if (expect_line && in->eof) return -1;               /* "Unexpected end of file." */
    if (!is_getline(in, line)) return 0;
    if (line->n > 0 && line->s[line->n - 1] == '\r') { line->n--; line->s[line->n] = 0; }
    return 1;
}

/* cas-offinder.cpp:20-24 split on whitespace (istream_iterator<string>). */
static size_t split_ws(const char *s, char ***out) {
    size_t n = 0, cap = 4; char **v = (char **)xrealloc(NULL, cap * sizeof(char *));
    const char *p = s;
    while (*p) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        const char *b = p;
        while (*p && !isspace((unsigned char)*p)) p++;
        if (n == cap) { cap *= 2; v = (char **)xrealloc(v, cap * sizeof(char *)); }
        v[n++] = xstrndup(b, (size_t)(p - b));
    }
    *out = v; return n;
}
static void free_tokens(char **v, size_t n) { for (size_t i = 0; i < n; i++) free(v[i]); free(v); }

/* open-addressing index over compare keys (std::map find/insert of add_compare) */
typedef struct { size_t *slot; size_t cap; } kindex;
static uint64_t hash_str(const char *s) {
    uint64_t h = 1469598103934665603ull;
    while (*s) { h ^= (unsigned char)*s++; h *= 1099511628211ull; }
    return h;
}
static void ki_rehash(kindex *ki, orc_query *q) {
    size_t nc = ki->cap ? ki->cap * 2 : 64;
    size_t *ns = (size_t *)xrealloc(NULL, nc * sizeof(size_t));
    for (size_t i = 0; i < nc; i++) ns[i] = (size_t)-1;
    for (size_t i = 0; i < q->ncmp; i++) {
        size_t h = hash_str(q->cmp[i].key) & (nc - 1);
        while (ns[h] != (size_t)-1) h = (h + 1) & (nc - 1);
        ns[h] = i;
    }
    free(ki->slot); ki->slot = ns; ki->cap = nc;
}

/* cas-offinder.cpp:14-18 add_compare: insert key -> ((id,t),[info]) or append
 * info to an existing key, keeping the first (id,t). */
static void add_compare(orc_query *q, kindex *ki, const char *key, const char *id, uint16_t t,
                        const char *first, int second) {
    if ((q->ncmp + 1) * 2 > ki->cap) ki_rehash(ki, q);
    size_t h = hash_str(key) & (ki->cap - 1);
    while (ki->slot[h] != (size_t)-1) {
        orc_compare *c = &q->cmp[ki->slot[h]];
        if (strcmp(c->key, key) == 0) {
            if (c->ninfo == c->capinfo) {
                c->capinfo *= 2;
                c->infos = (orc_info *)xrealloc(c->infos, c->capinfo * sizeof(orc_info));
            }
            c->infos[c->ninfo].first = xstrdup(first);
            c->infos[c->ninfo].second = second;
            c->ninfo++;
            return;
        }
        h = (h + 1) & (ki->cap - 1);
    }
    if (q->ncmp == q->capcmp) {
        q->capcmp = q->capcmp ? 2 * q->capcmp : 16;
        q->cmp = (orc_compare *)xrealloc(q->cmp, q->capcmp * sizeof(orc_compare));
    }
    orc_compare *c = &q->cmp[q->ncmp];
    c->key = xstrdup(key); c->id = xstrdup(id); c->threshold = t;
    c->capinfo = 2; c->ninfo = 1;
    c->infos = (orc_info *)xrealloc(NULL, 2 * sizeof(orc_info));
    c->infos[0].first = xstrdup(first); c->infos[0].second = second;
    ki->slot[h] = q->ncmp;
    q->ncmp++;
}

static int cmp_compare(const void *a, const void *b) {
    return strcmp(((const orc_compare *)a)->key, ((const orc_compare *)b)->key);
}

void orc_query_free(orc_query *q) {
    for (size_t i = 0; i < q->ncmp; i++) {
        orc_compare *c = &q->cmp[i];
        for (size_t j = 0; j < c->ninfo; j++) free(c->infos[j].first);
        free(c->infos); free(c->key); free(c->id);
    }
    free(q->cmp); free(q->chrdir); free(q->pattern);
    memset(q, 0, sizeof(*q));
}

static size_t first_not_N(const char *s, size_t n) {
    for (size_t i = 0; i < n; i++) if (s[i] != 'N') return i;
    return (size_t)-1;
}
static size_t last_not_N(const char *s, size_t n) {
    for (size_t i = n; i > 0; i--) if (s[i - 1] != 'N') return i - 1;
    return (size_t)-1;
}

#define PERR(msg) do { snprintf(err, errlen, "%s", msg); goto fail; } while (0)

/* cas-offinder.cpp:590-694 parseInput (+ add_compare :14-18, split :20-33,
 * parseLine :570-582). */
int orc_parse_input(const char *text, size_t len, orc_query *q, char *err, size_t errlen) {
    memset(q, 0, sizeof(*q));
    istr in = {text, len, 0, 0, 0};
    sbuf line = {0}, chrdir = {0}, key = {0}, tmp = {0};
    char **tok = NULL; size_t ntok = 0;
    kindex ki = {0};
    char idbuf[32], numbuf[32];
    int rev = 0;
    if (parse_line(&in, &chrdir, 1) < 0) PERR("Unexpected end of file.");
    q->chrdir = xstrdup(chrdir.s ? chrdir.s : "");
    if (parse_line(&in, &line, 1) < 0) PERR("Unexpected end of file.");
    ntok = split_ws(line.s ? line.s : "", &tok);
    if (ntok != 1 && ntok != 3) PERR("Malformed input file.");
    if (ntok == 3) {
        int d = atoi(tok[1]), r = atoi(tok[2]);
        if (d < 0 || r < 0) PERR("negative bulge size");
        q->dna_bulge = (unsigned)d; q->rna_bulge = (unsigned)r;
    }
    {
        const char *pat = tok[0]; size_t pl = strlen(pat);
        sbuf p = {0};
        rev = (pat[0] != 'N');                                     /* :619 */
        if (rev) { sb_puts(&p, pat); sb_rep(&p, 'N', q->dna_bulge); }
        else { sb_rep(&p, 'N', q->dna_bulge); sb_puts(&p, pat); }
        for (size_t i = 0; i < p.n; i++) p.s[i] = (char)toupper((unsigned char)p.s[i]);
        q->pattern = p.s; q->P = (unsigned)p.n; (void)pl;
    }
    free_tokens(tok, ntok); tok = NULL; ntok = 0;
    {
        size_t linecnt = 0, entrycnt = 0;
        const unsigned D = q->dna_bulge, R = q->rna_bulge;
        for (;;) {
            int r = parse_line(&in, &line, 0);
            if (r <= 0) break;
            if (line.n == 0) break;                                /* :630 */
            ntok = split_ws(line.s, &tok);
            if (ntok != 2 && ntok != 3) PERR("Malformed input file.");
            char *g = tok[0]; size_t gl = strlen(g);
            if (gl + D != q->P)
                PERR("The length of target sequences should match with the length of pattern sequence.");
            for (size_t i = 0; i < gl; i++) g[i] = (char)toupper((unsigned char)g[i]);
            int thr = atoi(tok[1]);
            const char *id;
            if (ntok == 3) id = tok[2];
            else { snprintf(idbuf, sizeof idbuf, "%zu", linecnt); id = idbuf; }
            uint16_t t = (uint16_t)thr;
            if (D == 0) {
                add_compare(q, &ki, g, id, t, "0", 0);             /* :647-649 */
            } else {
                sb_clear(&tmp);
                if (rev) { sb_puts(&tmp, g); sb_rep(&tmp, 'N', D); }
                else { sb_rep(&tmp, 'N', D); sb_puts(&tmp, g); }
                size_t fl = first_not_N(g, gl), ll = last_not_N(g, gl);
                for (unsigned i = 1; i <= D; i++) {                /* :655-669 */
                    unsigned preN = D - i;
                    uint64_t lim = (uint64_t)ll + 2;
                    for (unsigned j = (unsigned)fl; (uint64_t)j < lim; j++) {
                        sb_clear(&key);
                        if (rev) {
                            sb_putn(&key, g, j); sb_rep(&key, 'N', i);
                            if (sb_substr(&key, g, gl, j, (size_t)-1)) PERR("basic_string::substr");
                            sb_rep(&key, 'N', preN);
                        } else {
                            sb_rep(&key, 'N', preN);
                            if (j > gl) PERR("basic_string::substr");
                            sb_putn(&key, g, j); sb_rep(&key, 'N', i);
                            sb_substr(&key, g, gl, j, (size_t)-1);
                        }
                        int sec = rev ? -(int)j : (int)j;
                        if (strcmp(key.s, tmp.s) == 0) add_compare(q, &ki, key.s, id, t, "0", sec);
                        else { snprintf(numbuf, sizeof numbuf, "%u", i); add_compare(q, &ki, key.s, id, t, numbuf, sec); }
                    }
                }
            }
            {
                size_t fl = first_not_N(g, gl), ll = last_not_N(g, gl);
                for (unsigned i = 1; i <= R; i++) {                /* :671-682 */
                    unsigned preN = D + i;
                    uint64_t lim = (uint64_t)ll + 1 - i;
                    for (unsigned j = (unsigned)fl; (uint64_t)j < lim; j++) {
                        sbuf bs = {0};
                        if (sb_substr(&bs, g, gl, j, i)) { free(bs.s); PERR("basic_string::substr"); }
                        sb_clear(&key);
                        if (rev) {
                            sb_putn(&key, g, j);
                            if (sb_substr(&key, g, gl, (size_t)j + i, (size_t)-1)) { free(bs.s); PERR("basic_string::substr"); }
                            sb_rep(&key, 'N', preN);
                        } else {
                            sb_rep(&key, 'N', preN);
                            sb_putn(&key, g, j);
                            if (sb_substr(&key, g, gl, (size_t)j + i, (size_t)-1)) { free(bs.s); PERR("basic_string::substr"); }
                        }
                        add_compare(q, &ki, key.s, id, t, bs.s ? bs.s : "", rev ? -(int)j : (int)j);
                        free(bs.s);
                    }
                }
            }
            if (entrycnt == 0) entrycnt = ntok;                    /* :683-687 */
            else if (entrycnt != ntok) PERR("The number of entries below 2nd line should be consistent.");
            linecnt++;
            free_tokens(tok, ntok); tok = NULL; ntok = 0;
        }
    }
    qsort(q->cmp, q->ncmp, sizeof(orc_compare), cmp_compare);     /* std::map order */
    free(line.s); free(chrdir.s); free(key.s); free(tmp.s); free(ki.slot);
    return 0;
fail:
    if (tok) free_tokens(tok, ntok);
    free(line.s); free(chrdir.s); free(key.s); free(tmp.s); free(ki.slot);
    orc_query_free(q);
    return -1;
}

/* Parsed query as text, for parser parity tests: "P D R pattern" then one
 * "key<TAB>id<TAB>threshold<TAB>first:second,..." line per key in map order.
 * Returns 0 and a malloc-ed string, or -1 and the error message. */
int orc_parse_dump(const char *text, size_t len, char **out) {
    orc_query q; char err[512];
    sbuf b = {0};
    char tmp[64];
    if (orc_parse_input(text, len, &q, err, sizeof err)) {
        sb_puts(&b, err); *out = b.s; return -1;
    }
    snprintf(tmp, sizeof tmp, "%u %u %u ", q.P, q.dna_bulge, q.rna_bulge);
    sb_puts(&b, tmp); sb_puts(&b, q.pattern); sb_puts(&b, "\n");
    for (size_t i = 0; i < q.ncmp; i++) {
        orc_compare *c = &q.cmp[i];
        sb_puts(&b, c->key); sb_puts(&b, "\t"); sb_puts(&b, c->id);
        snprintf(tmp, sizeof tmp, "\t%u\t", (unsigned)c->threshold); sb_puts(&b, tmp);
        for (size_t j = 0; j < c->ninfo; j++) {
            if (j) sb_puts(&b, ",");
            sb_puts(&b, c->infos[j].first);
            snprintf(tmp, sizeof tmp, ":%d", c->infos[j].second); sb_puts(&b, tmp);
        }
        sb_puts(&b, "\n");
    }
    orc_query_free(&q);
    *out = b.s ? b.s : xstrdup("");
    return 0;
}

void orc_free(void *p) { free(p); }

/* One genome file as the oracle's readers hold it, in oracle/_ref's dump
 * format: "status S", then for S == 0 "nchr N", "chr <pos> <name>" lines,
 * "len <n>" and the text.  Returns the malloc-ed dump; *len its size. */
char *orc_genome_dump(const char *path, size_t *len) {
    orc_genome g = {0};
    sbuf b = {0};
    char tmp[96];
    int st = 0;
    if (orc_read_fasta(path, &g) && orc_read_twobit(path, &g)) st = 1;
    snprintf(tmp, sizeof tmp, "status %d\n", st); sb_puts(&b, tmp);
    if (!st) {
        snprintf(tmp, sizeof tmp, "nchr %zu\n", g.nchr); sb_puts(&b, tmp);
        for (size_t i = 0; i < g.nchr; i++) {
            snprintf(tmp, sizeof tmp, "chr %llu ", (unsigned long long)g.chrpos[i]); sb_puts(&b, tmp);
            sb_puts(&b, g.names[i] ? g.names[i] : ""); sb_puts(&b, "\n");
        }
        snprintf(tmp, sizeof tmp, "len %zu\n", g.len); sb_puts(&b, tmp);
        sb_putn(&b, g.text ? g.text : "", g.len); sb_puts(&b, "\n");
    }
    orc_genome_free(&g);
    *len = b.n;
    return b.s;
}

/* ------------------------------------------------------- finder / comparer */

typedef struct { uint64_t *v; char *f; size_t n, cap; } lvec;
static void lv_push(lvec *l, uint64_t x, char f) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 1024;
        l->v = (uint64_t *)xrealloc(l->v, l->cap * 8); l->f = (char *)xrealloc(l->f, l->cap);
    }
    l->v[l->n] = x; l->f[l->n] = f; l->n++;
}

static int nthreads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}

/* cas-offinder.cl:157-213 finder_cpu over every start position of the text
 * (chunking in cas-offinder.cpp:210-252 covers exactly [0, len-P]). */
size_t orc_finder(const orc_genome *g, const orc_query *q, uint64_t **loci, char **flags, int nthreads) {
    const unsigned P = q->P;
    *loci = NULL; *flags = NULL;
    if (g->len < P || P == 0) return 0;
    char *pat = (char *)xrealloc(NULL, 2 * P);
    int *idx = (int *)xrealloc(NULL, 2 * P * sizeof(int));
    memcpy(pat, q->pattern, P); memcpy(pat + P, q->pattern, P);
    revcomp_inplace(pat + P, P);                                   /* cpp:710-716 */
    seq_flags(idx, pat, P); seq_flags(idx + P, pat + P, P);
    const uint64_t npos = g->len - P + 1;
    const char *chr = g->text;
    int nt = nthreads_or_default(nthreads);
    lvec *parts = (lvec *)calloc((size_t)nt, sizeof(lvec));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, tn = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num(); tn = omp_get_num_threads();
#endif
        uint64_t b = npos * (uint64_t)tid / (uint64_t)tn, e = npos * (uint64_t)(tid + 1) / (uint64_t)tn;
        for (uint64_t i = b; i < e; i++) {
            char lf = 0;
            for (unsigned j = 0; j < P; j++) {
                int k = idx[j];
                if (k == -1) break;
                if (iupac_mismatch(pat[k], chr[i + k])) lf |= 2;
                k = idx[P + j];
                if (iupac_mismatch(pat[k + P], chr[i + k])) lf |= 1;
                if (lf == 3) break;
            }
            if (lf != 3) {
                int ok = 1;
                for (unsigned j = 0; j < P; j++) if (chr[i + j] == ';') { ok = 0; break; }
                if (ok) lv_push(&parts[tid], i, lf);
            }
        }
    }
    size_t tot = 0;
    for (int t = 0; t < nt; t++) tot += parts[t].n;
    *loci = (uint64_t *)xrealloc(NULL, (tot ? tot : 1) * 8);
    *flags = (char *)xrealloc(NULL, tot ? tot : 1);
    size_t o = 0;
    for (int t = 0; t < nt; t++) {
        memcpy(*loci + o, parts[t].v, parts[t].n * 8); memcpy(*flags + o, parts[t].f, parts[t].n);
        o += parts[t].n; free(parts[t].v); free(parts[t].f);
    }
    free(parts); free(pat); free(idx);
    return tot;
}

typedef struct { orc_hit *v; size_t n, cap; } hvec;
static void hv_push(hvec *h, uint64_t l, uint16_t mm, char d) {
    if (h->n == h->cap) { h->cap = h->cap ? 2 * h->cap : 256; h->v = (orc_hit *)xrealloc(h->v, h->cap * sizeof(orc_hit)); }
    h->v[h->n].locus = l; h->v[h->n].mm = mm; h->v[h->n].dir = d; h->n++;
}

/* cas-offinder.cl:214-286 comparer_cpu for one compare key; the key upload
 * (cas-offinder.cpp:360-364) builds [key | revcomp(key)] and index lists. */
size_t orc_comparer(const orc_genome *g, const orc_query *q, const char *key, uint16_t threshold,
                    const uint64_t *loci, const char *flags, size_t nloci, orc_hit **hits, int nthreads) {
    const unsigned P = q->P;
    char *comp = (char *)xrealloc(NULL, 2 * P);
    int *idx = (int *)xrealloc(NULL, 2 * P * sizeof(int));
    memcpy(comp, key, P); memcpy(comp + P, key, P);
    revcomp_inplace(comp + P, P);
    seq_flags(idx, comp, P); seq_flags(idx + P, comp + P, P);
    const char *chr = g->text;
    int nt = nthreads_or_default(nthreads);
    hvec *parts = (hvec *)calloc((size_t)nt, sizeof(hvec));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0, tn = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num(); tn = omp_get_num_threads();
#endif
        size_t b = nloci * (size_t)tid / (size_t)tn, e = nloci * (size_t)(tid + 1) / (size_t)tn;
        for (size_t i = b; i < e; i++) {
            const char *w = chr + loci[i];
            if (flags[i] == 0 || flags[i] == 1) {
                unsigned cnt = 0;
                for (unsigned j = 0; j < P; j++) {
                    int k = idx[j];
                    if (k == -1) break;
                    if (iupac_mismatch(comp[k], w[k])) { cnt++; if (cnt > threshold) break; }
                }
                if (cnt <= threshold) hv_push(&parts[tid], loci[i], (uint16_t)cnt, '+');
            }
            if (flags[i] == 0 || flags[i] == 2) {
                unsigned cnt = 0;
                for (unsigned j = 0; j < P; j++) {
                    int k = idx[P + j];
                    if (k == -1) break;
                    if (iupac_mismatch(comp[k + P], w[k])) { cnt++; if (cnt > threshold) break; }
                }
                if (cnt <= threshold) hv_push(&parts[tid], loci[i], (uint16_t)cnt, '-');
            }
        }
    }
    size_t tot = 0;
    for (int t = 0; t < nt; t++) tot += parts[t].n;
    *hits = (orc_hit *)xrealloc(NULL, (tot ? tot : 1) * sizeof(orc_hit));
    size_t o = 0;
    for (int t = 0; t < nt; t++) {
        if (parts[t].n) memcpy(*hits + o, parts[t].v, parts[t].n * sizeof(orc_hit));
        o += parts[t].n; free(parts[t].v);
    }
    free(parts); free(comp); free(idx);
    return tot;
}

/* -------------------------------------------------------------- formatter */

typedef struct { char *k; unsigned n; } sentry;
typedef struct { sentry *v; size_t n, cap; } stable;
static void st_add(stable *s, const char *key) {
    for (size_t i = 0; i < s->n; i++) if (strcmp(s->v[i].k, key) == 0) { s->v[i].n++; return; }
    if (s->n == s->cap) { s->cap = s->cap ? 2 * s->cap : 64; s->v = (sentry *)xrealloc(s->v, s->cap * sizeof(sentry)); }
    s->v[s->n].k = xstrdup(key); s->v[s->n].n = 1; s->n++;
}
static int cmp_sentry(const void *a, const void *b) {
    return strcmp(((const sentry *)a)->k, ((const sentry *)b)->k);
}

static int is_numeric(const char *s) {                          /* cas-offinder.h:20-22 */
    for (; *s; s++) if (*s < '0' || *s > '9') return 0;
    return 1;
}

/* cas-offinder.cpp:394-455 for one hit of compare c: window copy, revcomp for
 * '-', lowercase mismatches (indicate_mismatches :299-317), contig lookup
 * (:400), then one record per bulge info with the substring surgery. */
static void format_hit(FILE *fo, stable *summary, const orc_genome *g, const orc_query *q,
                       const orc_compare *c, const orc_hit *h, char *strbuf) {
    const unsigned P = q->P, D = q->dna_bulge;
    memcpy(strbuf, g->text + h->locus, P); strbuf[P] = 0;
    if (h->dir == '-') revcomp_inplace(strbuf, P);
    for (unsigned k = 0; k < P; k++)
        if (iupac_mismatch(c->key[k], strbuf[k])) strbuf[k] = (char)(strbuf[k] + 32);
    size_t idx = 0;
    for (size_t j = 0; j < g->nchr && h->locus >= g->chrpos[j]; j++) idx = j;
    sbuf rna = {0}, dna = {0}, t1 = {0};
    const char *key = c->key; size_t kl = P;
    for (size_t ii = 0; ii < c->ninfo; ii++) {
        const orc_info *bi = &c->infos[ii];
        int rv; unsigned bidx, bs; int off; const char *btype;
        sb_clear(&dna); sb_puts(&dna, strbuf);
        if (bi->second < 0) { rv = 1; bidx = (unsigned)(-bi->second); }
        else { rv = (h->dir == '-'); bidx = (unsigned)bi->second; }
        sb_clear(&rna);
        if (is_numeric(bi->first)) {
            bs = (unsigned)atoi(bi->first);
            if (rv) {
                off = 0;
                sb_substr(&rna, key, kl, 0, bidx); sb_rep(&rna, '-', bs);
                sb_substr(&rna, key, kl, bidx + bs, (size_t)-1);
            } else {
                off = (int)(D - bs);
                sb_substr(&rna, key, kl, (size_t)off, bidx); sb_rep(&rna, '-', bs);
                sb_substr(&rna, key, kl, (size_t)off + bidx + bs, (size_t)-1);
                sb_clear(&t1); sb_substr(&t1, dna.s, dna.n, (size_t)off, (size_t)-1);
                sb_clear(&dna); sb_putn(&dna, t1.s ? t1.s : "", t1.n);
            }
            if (bs == 0) {
                btype = "X";
                if (rv) {   /* substr(0, size - D): a wrapped length keeps the whole string */
                    if (rna.n >= D) { rna.n -= D; rna.s[rna.n] = 0; }
                    if (dna.n >= D) { dna.n -= D; dna.s[dna.n] = 0; }
                }
            } else btype = "DNA";
        } else {
            bs = (unsigned)strlen(bi->first);
            if (rv) {
                off = 0;
                sb_substr(&rna, key, kl, 0, bidx); sb_puts(&rna, bi->first);
                sb_substr(&rna, key, kl, bidx, (size_t)-1);
                sb_clear(&t1); sb_substr(&t1, dna.s, dna.n, 0, bidx); sb_rep(&t1, '-', bs);
                sb_substr(&t1, dna.s, dna.n, bidx, (size_t)-1);
                sb_clear(&dna); sb_putn(&dna, t1.s, t1.n);
                if (rna.n >= D + bs) { rna.n -= D + bs; rna.s[rna.n] = 0; }
                if (dna.n >= D + bs) { dna.n -= D + bs; dna.s[dna.n] = 0; }
            } else {
                off = (int)(D + bs);
                sb_substr(&rna, key, kl, (size_t)off, bidx); sb_puts(&rna, bi->first);
                sb_substr(&rna, key, kl, (size_t)off + bidx, (size_t)-1);
                sb_clear(&t1); sb_substr(&t1, dna.s, dna.n, (size_t)off, bidx); sb_rep(&t1, '-', bs);
                sb_substr(&t1, dna.s, dna.n, (size_t)off + bidx, (size_t)-1);
                sb_clear(&dna); sb_putn(&dna, t1.s, t1.n);
            }
            btype = "RNA";
        }
        fprintf(fo, "%s\t%s\t%s\t%s\t%s\t%llu\t%c\t%u\t%u\n", c->id, btype, rna.s ? rna.s : "",
                dna.s ? dna.s : "", g->names[idx] ? g->names[idx] : "",
                (unsigned long long)(h->locus - g->chrpos[idx] + (uint64_t)(int64_t)off), h->dir,
                (unsigned)h->mm, bs);
        if (summary) {
            char kbuf[10000];
            snprintf(kbuf, sizeof kbuf, "%s,%s,%d,%d", c->id, btype, (int)bs, (int)h->mm);
            st_add(summary, kbuf);
        }
    }
    free(rna.s); free(dna.s); free(t1.s);
}

/* cas-offinder.cpp:471-500 writeSummaryTable */
static void write_summary(const char *path, stable *s) {
    static const char *hdr[] = {"Id", "Bulge Type", "Bulge Size", "Mismatches", "?", "?", "?", "?"};
    FILE *fo = strcmp(path, "-") == 0 ? stdout : fopen(path, "w");
    if (!fo) return;
    qsort(s->v, s->n, sizeof(sentry), cmp_sentry);
    for (size_t i = 0; i < s->n; i++) {
        fputs("##", fo);
        const char *p = s->v[i].k; int cnt = 0;
        const char *cm;
        while ((cm = strchr(p, ',')) != NULL) {
            fprintf(fo, "%s=%.*s;", hdr[cnt < 7 ? cnt : 7], (int)(cm - p), p); cnt++; p = cm + 1;
        }
        fprintf(fo, "%s=%s;", hdr[cnt < 7 ? cnt : 7], p);
        fprintf(fo, "Number of Found Targets=%u\n", s->v[i].n);
    }
    if (fo != stdout) fclose(fo); else fflush(fo);
}

/* main.cpp:43-70 run_cas_offinder for one genome file. */
static void run_file(const char *path, const orc_query *q, FILE *fo, stable *summary,
                     const char *summary_path, int nthreads) {
    orc_genome g = {0};
    fprintf(stderr, "Reading %s...\n", path);
    if (orc_read_fasta(path, &g) && orc_read_twobit(path, &g)) {
        fprintf(stderr, "Skipping non-acceptable file %s...\n", path);
        orc_genome_free(&g);
        return;
    }
    uint64_t *loci; char *flags;
    size_t nl = orc_finder(&g, q, &loci, &flags, nthreads);
    char *strbuf = (char *)xrealloc(NULL, q->P + 1);
    for (size_t ci = 0; ci < q->ncmp; ci++) {
        const orc_compare *c = &q->cmp[ci];
        orc_hit *hits; size_t nh = orc_comparer(&g, q, c->key, c->threshold, loci, flags, nl, &hits, nthreads);
        for (size_t i = 0; i < nh; i++)
            if (hits[i].mm <= c->threshold) format_hit(fo, summary, &g, q, c, &hits[i], strbuf);
        free(hits);
    }
    fflush(fo);
    free(strbuf); free(loci); free(flags);
    orc_genome_free(&g);
    if (summary_path && summary_path[0]) write_summary(summary_path, summary);
}

static char *slurp(const char *path, size_t *n) {
    FILE *f = strcmp(path, "-") == 0 ? stdin : fopen(path, "rb");
    if (!f) return NULL;
    sbuf b = {0}; char buf[65536]; size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) sb_putn(&b, buf, r);
    if (f != stdin) fclose(f);
    if (!b.s) { b.s = xstrdup(""); }
    *n = b.n; return b.s;
}

/* main.cpp:72-157 after argument parsing. */
int orc_run(const char *input_path, const char *output_path, const char *summary_path, int nthreads) {
    size_t n; char *text = slurp(input_path, &n);
    char err[512];
    orc_query q;
    if (!text) { fprintf(stderr, "Critical error! Cannot read input file.\n"); return 1; }
    if (orc_parse_input(text, n, &q, err, sizeof err)) {
        fprintf(stderr, "Critical error! %s\n", err); free(text); return 1;
    }
    free(text);
    FILE *fo = strcmp(output_path, "-") == 0 ? stdout : fopen(output_path, "w");
    if (!fo) { orc_query_free(&q); return 1; }
    fprintf(fo, "##Generated by Cas-OFFinder %s\n", ORC_VERSION);
    fputs("#Id\tBulge Type\tcrRNA\tDNA\tChromosome\tLocation\tDirection\tMismatches\tBulge Size\n", fo);
    stable summary = {0};
    int issum = summary_path && summary_path[0];
    DIR *dir = opendir(q.chrdir);
    int rc = 0;
    if (!dir) {
        struct stat st;
        if (stat(q.chrdir, &st) == 0) run_file(q.chrdir, &q, fo, issum ? &summary : NULL, summary_path, nthreads);
        else { fprintf(stderr, "An error has occured while opening directory: %s\n", q.chrdir); rc = 1; }
    } else {
        struct dirent *ent;
        while ((ent = readdir(dir)) != NULL) {
            if (ent->d_type == DT_REG || ent->d_type == DT_LNK) {
                size_t pl = strlen(q.chrdir) + strlen(ent->d_name) + 2;
                char *p = (char *)xrealloc(NULL, pl);
                snprintf(p, pl, "%s/%s", q.chrdir, ent->d_name);
                run_file(p, &q, fo, issum ? &summary : NULL, summary_path, nthreads);
                free(p);
            }
        }
        closedir(dir);
    }
    if (fo != stdout) fclose(fo); else fflush(fo);
    for (size_t i = 0; i < summary.n; i++) free(summary.v[i].k);
    free(summary.v);
    orc_query_free(&q);
    return rc;
}

double orc_time_search(const char *input_path, int nthreads, uint64_t *genome_len,
                       uint64_t *ncand, uint64_t *npairs, uint64_t *nhits) {
    size_t n; char *text = slurp(input_path, &n);
    char err[512]; orc_query q; orc_genome g = {0};
    *genome_len = *ncand = *npairs = *nhits = 0;
    if (!text) return -1.0;
    if (orc_parse_input(text, n, &q, err, sizeof err)) { free(text); return -1.0; }
    free(text);
    if (orc_read_fasta(q.chrdir, &g) && orc_read_twobit(q.chrdir, &g)) { orc_query_free(&q); return -1.0; }
    double t0 = now_s();
    uint64_t *loci; char *flags;
    size_t nl = orc_finder(&g, &q, &loci, &flags, nthreads);
    uint64_t nh = 0, np = 0;
    for (size_t ci = 0; ci < q.ncmp; ci++) {
        orc_hit *hits;
        nh += orc_comparer(&g, &q, q.cmp[ci].key, q.cmp[ci].threshold, loci, flags, nl, &hits, nthreads);
        for (size_t i = 0; i < nl; i++) np += (flags[i] == 0) ? 2 : 1;
        free(hits);
    }
    double t1 = now_s();
    *genome_len = g.len; *ncand = nl; *npairs = np; *nhits = nh;
    free(loci); free(flags); orc_genome_free(&g); orc_query_free(&q);
    return t1 - t0;
}

/* ----------------------------------------------------------- oracle CLI */
#ifdef ORC_MAIN
int main(int argc, char **argv) {
    const char *summary = "";
    const char *pos[8]; int np = 0;
    for (int i = 1; i < argc; i++) {
        if (strncmp(argv[i], "--", 2) == 0 && strlen(argv[i]) > 2) {
            if (strcmp(argv[i] + 2, "summary") == 0 && i + 1 < argc) summary = argv[++i];
            else { printf("usage: cas-offinder-oracle input C output [--summary file]\n"); return 0; }
        } else if (np < 8) pos[np++] = argv[i];
    }
    if (np < 3) { printf("usage: cas-offinder-oracle input C output [--summary file]\n"); return 0; }
    const char *nt = getenv("ORC_THREADS");
    return orc_run(pos[0], pos[2], summary, nt ? atoi(nt) : 0);
}
#endif
