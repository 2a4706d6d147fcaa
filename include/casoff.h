/*
 * casoff.h -- C ABI of libcasoff.so, the MI355X (gfx950) Cas-OFFinder search engine.
 *
 * The reference (Cas-OFFinder 3.0.0) has no library API: its search path sits
 * behind the Cas_OFFinder class (cas-offinder.h:84-106) and the fail-fast OpenCL
 * wrappers (oclfunctions.h:20-133).  Each entry point below names the reference
 * member or call sequence it replaces.  Conventions:
 *   - every function returns int: 0 ok, < 0 error (COF_E_*); nothing calls exit();
 *   - the library owns device memory and results; callers release with *_free;
 *   - a context is single-threaded (one HIP stream); cof_last_error() holds the
 *     message of the last failing call on it;
 *   - positions are 0-based; strands are '+' / '-'.
 */
#ifndef CASOFF_H
#define CASOFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COF_OK 0
#define COF_E_INVALID (-1)   /* malformed input, bad argument */
#define COF_E_NOMEM (-2)     /* host or device allocation failed, admission refused */
#define COF_E_DEVICE (-3)    /* HIP runtime error, no device */
#define COF_E_IO (-4)        /* genome path unreadable */
#define COF_E_UNSUPPORTED (-5)

#define COF_MODE_PLAIN 0     /* finder launch, then one comparer launch per compare key */
#define COF_MODE_BATCHED 1   /* fused finder+comparer over all keys, one launch per query */

typedef struct cof_ctx cof_ctx;
typedef struct cof_query cof_query;
typedef struct cof_result cof_result;
typedef struct cof_server cof_server;

/* One device-side hit: the window start in the context's global genome
 * coordinate, the compare key index (std::map order of the query's keys) and
 * the mismatch count.  cof_result_records() returns these sorted by
 * (query, key, pos, strand). */
typedef struct {
    uint64_t pos;
    uint32_t key;
    uint16_t mm;
    uint8_t strand;   /* 0 '+', 1 '-' */
    uint8_t query;    /* index of the query in the cof_search batch */
} cof_hit;

/* Per-search timing and work counters of the last cof_search on a context. */
typedef struct {
    double t_total_ms;      /* submission to compact records in host memory */
    double t_finder_ms;     /* plain mode: finder kernels (HIP events) */
    double t_comparer_ms;   /* plain mode: all comparer launches; batched: fused kernels */
    double t_fused_ms;      /* batched mode: fused kernels (HIP events) */
    uint64_t n_candidates;  /* (locus, strand) candidates that passed the PAM test */
    uint64_t n_pairs;       /* candidate x compare-key evaluations */
    uint64_t n_hits;
    uint64_t n_launches;
    uint64_t genome_len;    /* positions scanned (global coordinate incl. separators) */
    uint64_t n_base_cmp;    /* sum over pairs of the key's comparable (non-N) positions */
    uint64_t n_model_ops;   /* int32 VALU ops of the comparer model: fused kernel 3 per (pair,
                               32-position word) for A/C/G/T keys, 6 with IUPAC codes; bulge kernel
                               per (candidate, guide, word) 2 per window alignment (5 with IUPAC
                               codes) + 1 + 3 per bulge family */
    double t_collect_ms;    /* device sort of the hits + copy to host + decode */
    uint64_t n_loci;        /* plain mode: (locus, flag) entries the finder wrote */
    uint64_t n_finder_positions; /* plain mode: window starts the finder launches scanned */
    double t_tune_ms;       /* batched mode: autotuner trials run by this search (first use of a
                               launch signature on the context; COF_AUTOTUNE=0 disables them) */
    uint64_t graph;         /* batched mode: 0 launches issued one by one, 1 captured as HIP
                               graphs and launched, 2 the previous search's graphs replayed */
    double t_sort_ms;       /* collection: device sort and packing of the hits (HIP events) */
    double t_copy_ms;       /* collection: device-to-host copies of the hits (HIP events) */
    double t_decode_ms;     /* collection: host decoding into cof_hit records (overlaps the copies) */
    double t_upload_ms;     /* the context's last genome upload: the three bit planes through pinned
                               staging buffers, plus the poison and tile tables (not per search) */
    double t_plan_ms;       /* batched mode: host time building the search's key / bulge tables and
                               seed indexes (0 when the lane reused the plan of the same queries) */
} cof_stats;

/* Synthetic genome specification (seeded, deterministic; no reference
 * counterpart -- inputs for tests and benchmarks). */
typedef struct {
    uint64_t seed;
    uint32_t n_contigs;
    uint32_t length_mode;    /* 0: weights U(0.5,1.5) scaled to total_len; 1: log-uniform in [min_len,max_len] scaled to total_len */
    uint64_t total_len;
    uint64_t min_len, max_len;
    double gc;               /* fraction of G+C */
    double n_frac;           /* fraction of positions covered by N runs */
    uint32_t n_run_min, n_run_max;
    double lower_frac;       /* fraction in soft-masked (lowercase) runs, FASTA output only */
    uint32_t line_width;     /* FASTA line width */
} cof_synth_spec;

/* ---- context: replaces Cas_OFFinder(devtype, devarg) + initOpenCLDevices
 *      (cas-offinder.cpp:80-156) and ~Cas_OFFinder (:158-173).  device -1
 *      opens a host-only context (genome loading, parsing, formatting; no
 *      search). */
int cof_open(int device, cof_ctx **out);
int cof_close(cof_ctx *ctx);
const char *cof_last_error(const cof_ctx *ctx);
const char *cof_version(void);

/* ---- device ids: the text after the device letter of the CLI's
 *      {C|G|A}[ids] argument, read as Cas_OFFinder(devtype, devarg) reads it
 *      (cas-offinder.cpp:131-153) and applied as initOpenCLDevices applies it
 *      (:80-104): comma-separated ids (atoi) and start:end[:step] ranges; an
 *      argument that lists no id selects every device.  selected[d] = 1 for
 *      each device index d < n_devices that it selects, else 0; *n_bad counts
 *      the pieces that are neither an id nor a range (the reference warns
 *      "Something wrong with the device ID argument..." and skips them).  A
 *      negative step counts down from start to 0; a step of 0 over a
 *      non-empty range is COF_E_INVALID (the reference loops without end).
 *      Host-only: no device access.  cof_device_count gives the number of
 *      HIP devices (0 when there is none). */
int cof_device_ids(const char *devarg, int n_devices, uint8_t *selected, int *n_bad);
int cof_device_count(int *n);

/* ---- genome: replaces read_fasta (read_fasta.cpp:11-57) / read_twobit
 *      (read_twobit.cpp:44-131) as driven by main.cpp:43-58 and the per-chunk
 *      upload setChrData/loadNextChunk (cas-offinder.cpp:175-252).  `path` is a
 *      directory (every regular file or symlink in readdir order, non-FASTA /
 *      non-2bit files skipped) or one file.  The whole genome is packed as
 *      2-bit bases + a non-ACGT bit plane and stays resident on the device
 *      until the next load or cof_close.  Returns the number of files loaded
 *      in *nfiles (may be NULL).  A load that fails leaves the previous
 *      genome in place; one that succeeds replaces it, and results searched
 *      on the previous genome no longer format (COF_E_INVALID).  Loads,
 *      cof_genome_synth and cof_close return COF_E_INVALID while a server
 *      (cof_server_start) is attached to the context. */
int cof_genome_load(cof_ctx *ctx, const char *path, int *nfiles);
/* Generates the genome directly into the packed resident form. */
int cof_genome_synth(cof_ctx *ctx, const cof_synth_spec *spec);
/* Total positions (global coordinate, separators included), file and contig counts. */
int cof_genome_info(const cof_ctx *ctx, uint64_t *total_len, uint64_t *n_files, uint64_t *n_contigs);
/* Contig i (0-based over all files): name, global position of its first
 * base, length, and the file it came from. */
int cof_genome_contig(const cof_ctx *ctx, uint64_t i, const char **name, uint64_t *pos, uint64_t *len,
                      uint32_t *file);
/* Copies the genome text of [pos, pos+n) as the reference holds it
 * (uppercase, ';' between contigs). */
int cof_genome_text(const cof_ctx *ctx, uint64_t pos, uint64_t n, char *out);
/* Writes a synthetic genome as FASTA (format 0) or UCSC 2bit (format 1). */
int cof_synth_write(const cof_synth_spec *spec, const char *path, int format);

/* ---- query: replaces Cas_OFFinder::readInputFile / parseInput
 *      (cas-offinder.cpp:590-751): pattern line with DNA/RNA bulge sizes,
 *      guides with thresholds and ids, expanded into compare keys in std::map
 *      order with first-(id,threshold)-wins.  On error returns COF_E_INVALID
 *      and the reference's message ("Malformed input file." ...) in
 *      cof_last_error. */
int cof_query_parse(cof_ctx *ctx, const char *text, size_t len, cof_query **out);
int cof_query_free(cof_query *q);
/* Diagnostic text of the parsed query (no reference counterpart): line
 * "P D R pattern", then "key<TAB>id<TAB>threshold<TAB>first:second,..." per
 * compare key in map order.  Caller frees *out with cof_free. */
int cof_query_dump(const cof_query *q, char **out, size_t *len);
/* Genome path (first input line), padded pattern length, key count. */
int cof_query_info(const cof_query *q, const char **chrdir, uint32_t *pattern_len, uint32_t *n_keys);

/* ---- search: replaces findPattern (cas-offinder.cpp:254-297) and the launch
 *      part of compareAll (:356-393) for every chunk (main.cpp:58-65).
 *      `queries` run over the resident genome; mode COF_MODE_PLAIN keeps the
 *      reference's launch model (finder, then one comparer launch per key),
 *      COF_MODE_BATCHED runs one fused finder+comparer launch per pattern
 *      group of the queries (queries with DNA/RNA bulges: the variants are
 *      expanded on the device). */
int cof_search(cof_ctx *ctx, cof_query *const *queries, int n_queries, int mode, cof_result **out);
int cof_stats_get(const cof_ctx *ctx, cof_stats *out);
/* Batched search that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  As
 * cof_search in COF_MODE_BATCHED, but no device or pinned host memory is
 * allocated (no hipMalloc / hipHostMalloc / hipFree, no library
 * temporaries, no graphs, events or autotuner trials: launch variants come
 * from the context's tuner decisions); only host memory for the queries'
 * tables.  The key tables and launch counters live in the library's
 * scratch array of 8 MiB per GPU (COF_E_NOMEM when a search needs more);
 * every call on one GPU shares it, so calls that overlap in time on one GPU
 * race.  hit_keys / hit_mm are caller-owned DEVICE buffers of cap entries;
 * the hits are written unsorted, hit_keys[i] = global key index << 40 |
 * window start << 1 | strand (0 '+', 1 '-'), where a query's global key
 * indices start at the sum of the key counts of the queries before it
 * (cof_query_info), and hit_mm[i] = its mismatches.  *n_hits = hits found;
 * more than cap: only cap are written and COF_E_NOMEM is returned.  The
 * call returns after the hits are written. */
int cof_search_noalloc(cof_ctx *ctx, cof_query *const *queries, int n_queries, uint64_t *hit_keys,
                       uint16_t *hit_mm, uint64_t cap, uint64_t *n_hits);
/* Autotuner decisions of the context so far, one "signature<TAB>variant" line
 * each (no reference counterpart: the reference's launch shape is fixed).
 * Caller frees *out with cof_free. */
int cof_tuner_info(cof_ctx *ctx, char **out, size_t *len);
/* Tuner file (no reference counterpart): a JSON object {"signature":
 * variant, ...} of autotuner decisions.  A context opens with the decisions
 * of $COF_TUNER_FILE ("none": no file), else of tuned_gfx950.json beside
 * libcasoff.so when it exists (cof_open fails with COF_E_INVALID when that
 * file is malformed); launches whose signature it holds run no trials.
 * cof_tuner_load merges a file's decisions into the context's; cof_tuner_save
 * writes all of them. */
int cof_tuner_load(cof_ctx *ctx, const char *path);
int cof_tuner_save(cof_ctx *ctx, const char *path);

/* Batched launches of the last cof_search on the context, one line each:
 * "kernel<TAB>signature<TAB>variant<TAB>ms<TAB>candidates": the kernel
 * instance as profilers name it, the autotuner's launch signature and
 * variant, the launch's duration (HIP events around it on the context's
 * stream) and the (locus, strand) candidates it scanned.  Empty after a
 * plain search.  No reference counterpart (its kernels are launched without
 * profiling, cas-offinder.cpp:123).  Caller frees *out with cof_free. */
int cof_launch_info(const cof_ctx *ctx, char **out, size_t *len);

/* ---- results */
int64_t cof_result_count(const cof_result *r);
int cof_result_records(const cof_result *r, cof_hit *out, size_t cap);
/* Writes the records of query `qi` as the reference's compareAll does
 * (cas-offinder.cpp:394-456): one TSV line per (hit, bulge info).  `path` "-"
 * is stdout; `append` 0 truncates and writes the two header lines of
 * writeHeaders (:319-328) first.  When `summary_path` is non-empty the
 * --summary table (writeSummaryTable, :471-500) is written there. */
int cof_result_format(cof_ctx *ctx, const cof_result *r, int qi, const char *path, int append,
                      const char *summary_path);
/* Formatting reads the context's genome: a result formats only on the
 * context and genome it was searched on (else COF_E_INVALID). */
/* Same records into a malloc-ed buffer (caller frees with cof_free). */
int cof_result_format_mem(cof_ctx *ctx, const cof_result *r, int qi, char **out, size_t *len);
/* Releases a result; one whose hit array holds 64 MiB or more is released
 * on a short-lived thread of the library (the call returns at once). */
int cof_result_free(cof_result *r);
/* Formats caller-supplied hits of one query (hit.query ignored) against the
 * context's genome, exactly as cof_result_format_mem; no header lines. */
int cof_format_hits(cof_ctx *ctx, const cof_query *q, const cof_hit *hits, size_t n, char **out, size_t *len);
void cof_free(void *p);

/* ---- serving (no reference counterpart: the reference runs one input per
 *      process).  A server owns the context on its own threads (two lanes,
 *      each with hit buffers of its share of mem_budget_bytes), admits queued
 *      requests while their estimated key-table and hit bytes fit a lane and
 *      the budget, and merges admitted queries into batched passes as they
 *      come. */
typedef struct {
    uint64_t mem_budget_bytes;   /* key tables + hit buffers of in-flight queries */
    uint32_t max_batch_queries;  /* queries merged into one pass */
} cof_server_cfg;
int cof_server_start(cof_ctx *ctx, const cof_server_cfg *cfg, cof_server **out);
/* One request of n_queries queries (for example one per pattern group); the
 * server keeps its own references, the caller may free the queries.
 * *ticket identifies the request. */
int cof_server_submit(cof_server *s, cof_query *const *queries, int n_queries, uint64_t *ticket);
/* 0: done (*out set, one result per request with its queries indexed 0..n-1,
 * latency from submit to records in host memory in *latency_ms),
 * 1: pending, < 0: error.  A done or failed ticket is forgotten. */
int cof_server_poll(cof_server *s, uint64_t ticket, int wait, cof_result **out, double *latency_ms);
int cof_server_stop(cof_server *s);
/* What the server did so far: each lane's hit capacity (its share of
 * mem_budget_bytes at 20 device bytes per hit, 1 M to 64 M hits), served
 * batches / requests / queries / hits, batches whose hits outgrew the lane
 * (grown and rerun), admissions that waited for running batches, the peak
 * estimate of in-flight bytes, the running average of hits per compare
 * key that sizes the admission estimate, and, summed over the batches, their
 * launches and the milliseconds of the batched search (wall time), of its
 * host table building, of its kernels (HIP events) and of splitting the
 * merged records per request. */
typedef struct {
    uint64_t lane_hits;
    uint64_t batches, requests, queries, hits;
    uint64_t lane_grows;
    uint64_t admission_waits;
    uint64_t peak_inflight_bytes;
    double hits_per_key;
    uint64_t launches;
    double t_search_ms, t_plan_ms, t_kernel_ms, t_split_ms;
} cof_server_stats;
int cof_server_stats_get(cof_server *s, cof_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
