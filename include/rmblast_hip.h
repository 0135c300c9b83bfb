/*
 * rmblast_hip.h — C ABI of librmblast_hip.so, the MI355X (gfx950) search path
 * of rmblastn.
 *
 * The library replaces the per-(query, subject) search of the reference's
 * rmblast-lib search API (rmblast-lib/src/search/mod.rs:9-21):
 *   build_query_lookup       (rmblast-lib/src/search/engine.rs:606)
 *   search_with_query_lookup (rmblast-lib/src/search/engine.rs:715-729)
 * driven over every subject by rmblastn's search_db_parallel
 * (rmblastn/src/main.rs:1133-1234).  One call searches a batch of query
 * records against the loaded library (both strands) and returns, per
 * (query, subject) pair, the final HSPs of the reference's traceback phase
 * (engine.rs:1438-1762: after complexity adjustment, the common-endpoint purge
 * and the final containment pass) with their edit scripts.  The host tail
 * after it (E-value reap, mask level, output order, alignment statistics and
 * the outfmt-6 text; main.rs:1224-1234, engine.rs:863-1017, stats.rs,
 * output.rs) is offered by the rmb_report_* and rmb_mask_level entry points,
 * which run on the library's host threads; a caller may keep its own.
 *
 * Conventions
 *  - Sequences are BLASTNA codes (A0 C1 G2 T3 R4 Y5 M6 K7 W8 S9 B10 D11 H12
 *    V13 N14; encoding.rs:19-54) WITHOUT the sentinel bytes; the library adds
 *    them.  Library subjects are given with their original ambiguity codes;
 *    the library resolves them for seeding exactly as FastaDb does
 *    (seq/fasta_db.rs:86-105) and keeps the originals for traceback.
 *  - Coordinates are 0-based, half-open; subject coordinates of minus-strand
 *    HSPs are on the plus strand (hits.rs:62-98).
 *  - Every function returns RMB_OK (0) or an error code; the message of the
 *    last error is available from rmb_last_error().  No exception or abort
 *    crosses this boundary (the convention of the reference's only C ABI,
 *    rmblast-stats/rmstats/csrc/alp_shim.cpp:44-117).
 *  - One context per GPU; calls on one context must be serialised by the
 *    caller.  The caller owns all host buffers; the library owns device memory
 *    for the life of the context.
 */
#ifndef RMBLAST_HIP_H
#define RMBLAST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMB_OK 0
#define RMB_ERR_ARG 1       /* invalid argument (message says which) */
#define RMB_ERR_DEVICE 2    /* HIP runtime or kernel failure */
#define RMB_ERR_STATE 3     /* call out of order (e.g. search before library load) */
#define RMB_ERR_CAPACITY 4  /* a device work limit was exceeded; message names it */

/* Search parameters: the rmblastn options in scope (rmblast-lib/src/options.rs:12-104,
 * rmblastn/src/main.rs:128-281).  dust = 1 applies symmetric DUST (window 64,
 * level 20, linker 1) to each query on the host, as build_query_lookup does
 * (engine.rs:606-619): the lookup table and the exact word extension read the
 * masked query, ungapped and gapped extension the query itself. */
typedef struct rmb_params {
  int32_t gap_open;
  int32_t gap_extend;
  int32_t word_size;
  int32_t xdrop_ungap;
  int32_t xdrop_gap;
  int32_t xdrop_gap_final;
  int32_t min_raw_gapped_score;
  /* Ungapped keep threshold.  >= 0: this value for every query record.
   * < 0: resolved per query record as rmblastn resolves it before each
   * search (rmblastn/src/main.rs:1154-1170 calling ungapped_cutoff,
   * ka_cutoff.rs:126-161): from ka_lambda / ka_k below when both are > 0,
   * with the record's full length (a split query's whole length) and the
   * loaded library's average length; else NCBI's cutoff_score_max
   * (options.rs:80-82). */
  int32_t ungapped_cutoff;
  int32_t complexity_adjust;
  int32_t dust;
  /* Gapped Karlin-Altschul lambda and K of the matrix + gap costs, as
   * rmb_ka_lookup returns them; 0 when the combination has none. */
  double ka_lambda;
  double ka_k;
} rmb_params;

/* Substitution matrix in BLASTNA order, as ScoreMatrix holds it after parsing
 * (matrix.rs:62-72): scores[q][s], base frequencies (A,C,G,T at 0..3) and the
 * lambda used by complexity adjustment (matrix.rs:306-354). */
typedef struct rmb_matrix {
  int32_t scores[16][16];
  double freqs[16];
  double lambda;
} rmb_matrix;

/* One final HSP (hits.rs:62-98 Hsp without the aligned strings). */
typedef struct rmb_hsp {
  uint32_t query;     /* index into the query batch */
  uint32_t subject;   /* index into the loaded library */
  int32_t score;
  uint32_t q_start, q_end;
  uint32_t s_start, s_end;
  uint32_t strand;    /* 0 plus, 1 minus */
  uint32_t ops_offset; /* first edit op in the ops array */
  uint32_t ops_count;
} rmb_hsp;

/* Edit op: low 2 bits = kind (0 Sub, 1 GapInQuery, 2 GapInSubject; hits.rs:21-30),
 * upper 30 bits = run length. Ops are in forward query order. */
typedef uint32_t rmb_op;

typedef struct rmb_ctx rmb_ctx;

/* Parse an RMBlast matrix file (FREQS comment, header row, one row per
 * symbol) into BLASTNA order with NCBI's defaults for absent entries and the
 * sentinel row/column, and estimate lambda.  Replaces ScoreMatrix::from_file
 * (matrix.rs:76-216, lambda :306-354).  err receives a message on failure. */
int rmb_matrix_load(const char* path, rmb_matrix* out, char* err, size_t err_len);

/* Gapped Karlin-Altschul (lambda, K) of a RepeatMasker matrix at its gap
 * costs, from NCBI's ALP table (blast_stat.c rmblast_*_values, as restated in
 * rmblast-lib/src/search/ka_cutoff.rs:40-87).  The matrix file's base name
 * (the part after the last '/') is compared case-insensitively and both gap
 * costs must match the entry.  Replaces lookup_ka (ka_cutoff.rs:105-114).
 * *found = 1 and out = {lambda, K} on a match, else *found = 0. */
int rmb_ka_lookup(const char* matrix_name, int32_t gap_open, int32_t gap_extend, double out[2], int32_t* found);

/* The ungapped pre-filter cutoff for one query (ka_cutoff.rs:126-161):
 * es = ceil(ln(K * min(avg_subj, 2 * query_length) * avg_subj / 0.05) / lambda)
 * clamped to [1, cutoff_score_max] when rmb_ka_lookup matches and
 * avg_subj_length > 0, else cutoff_score_max (min_raw_gapped_score / 2, or
 * 1 when that is 0).  query_length is the whole query record's length.
 * info (may be NULL) receives {found, es, lambda, K, searchsp} for the run log
 * rmblastn prints (main.rs:1098-1131). */
typedef struct rmb_ka_cutoff_info {
  int32_t found;
  int32_t es;          /* the KA score before the cap */
  double lambda, k;
  uint64_t searchsp;
} rmb_ka_cutoff_info;
int rmb_ungapped_cutoff(const char* matrix_name, int32_t gap_open, int32_t gap_extend, int32_t min_raw_gapped_score,
                        uint64_t query_length, uint64_t avg_subj_length, int32_t* cutoff, rmb_ka_cutoff_info* info);

/* ---- E-value statistics (host only) ---------------------------------------
 * The gapped Karlin-Altschul block behind the evalue / bitscore fields and
 * the --evalue reap, resolved as rmblastn resolves it (rmblast-lib/src/
 * ka_stats.rs:153-300): the table of RepeatMasker matrices by file base name
 * and gap costs (rmblast-stats/rmstats/src/tables_data.rs), then the
 * -matrix_lambda/_k/_alpha overrides given together (with -matrix_beta), then
 * a "# KARLIN" comment in the matrix file, else none: E-values print 1.0 and
 * bit scores 0.0, and nothing is reaped.  ncbi_compat = 1 takes NCBI 2.17.1's
 * sources only (no comparison.matrix entry, its 30p53g placeholder, no
 * comment).  The reference's optional ALP fit of other matrices is not part
 * of this library. */
#define RMB_KA_SENTINEL 0
#define RMB_KA_TABLE 1
#define RMB_KA_CLI 2
#define RMB_KA_COMMENT 3
typedef struct rmb_ka_stats_options {
  double matrix_lambda, matrix_k, matrix_alpha, matrix_beta;  /* 0 when not given */
  int32_t ncbi_compat;
  int32_t pad;
} rmb_ka_stats_options;
typedef struct rmb_ka_stats {
  int32_t source;              /* RMB_KA_* */
  int32_t pad;
  double lambda, k, log_k, h;  /* -1 each without statistics */
  double alpha, beta;          /* the length adjustment's slope (alpha / lambda) and intercept */
} rmb_ka_stats;
/* Resolve the statistics for a matrix file at its gap costs; a "# KARLIN"
 * line that does not parse is RMB_ERR_ARG with its message in err. */
int rmb_ka_stats_resolve(const char* matrix_path, int32_t gap_open, int32_t gap_extend,
                         const rmb_ka_stats_options* opt, rmb_ka_stats* out, char* err, size_t err_len);
/* Effective search space and length adjustment of a query of query_length
 * bases against db_num_seqs sequences of db_length bases in all
 * (BLAST_CalcEffLengths, rmstats/src/eff_lengths.rs:52-128, with
 * BLAST_ComputeLengthAdjustment, length_adjust.rs:7-93); 0 without statistics. */
int rmb_ka_search_space(const rmb_ka_stats* ka, int32_t query_length, int64_t db_length, int32_t db_num_seqs,
                        int64_t* eff_searchsp, int32_t* length_adjustment);
/* E-value and bit score of a raw score (rmstats/src/evalue.rs:96-126). */
int rmb_ka_evalue(const rmb_ka_stats* ka, int64_t eff_searchsp, int32_t score, double* evalue, double* bit_score);

/* Create a context on HIP device `device`.  Replaces the per-search state the
 * reference builds in rmblastn main (main.rs:516-700): parameters + matrix. */
int rmb_ctx_create(int device, const rmb_params* params, const rmb_matrix* matrix, rmb_ctx** out);
int rmb_ctx_destroy(rmb_ctx* ctx);

/* Schedule of a context's kernels (no counterpart in the reference: the
 * choices the autotuner in rmblast_amd/autotune.py sweeps and stores per
 * matrix and word size).  0 in a field keeps the library's default.  Every
 * setting gives the same results; only the time changes.
 *  batches          pipeline batches a search is cut into (1..64; default 8)
 *  prelim_ring      preliminary-DP LDS band ring, 16 or 64 cells (default by
 *                   X-drop: 64 past 12 band cells); wider bands rerun with
 *                   global rings either way
 *  prelim_wg_per_cu preliminary-DP workgroups per CU (1..16; default 8 of
 *                   128 lanes with 16-cell rings, 9 of 64 lanes with 64)
 *  tb_ring          traceback LDS band ring, 64 or 256 cells (default by
 *                   X-drop: 256 past 48 band cells)
 *  tb_wg_per_cu     traceback workgroups (64 lanes) per CU (1..16, capped at
 *                   9, or 2 with the 256-cell rings, by LDS and the lane
 *                   arenas' memory; default 9, or 2)
 *  tb_wave_halves   a traceback launch of at most this many halves runs
 *                   every half one per wave, a band row per vector step,
 *                   instead of one per lane (1..2^30; -1 never; default 112
 *                   per CU): with few halves the lanes' serial walks along
 *                   the longest ones set the time, with many the lanes'
 *                   throughput wins; only with the 64-cell rings */
typedef struct rmb_tuning {
  int32_t batches;
  int32_t prelim_ring;
  int32_t prelim_wg_per_cu;
  int32_t tb_ring;
  int32_t tb_wg_per_cu;
  int32_t tb_wave_halves;
} rmb_tuning;
int rmb_ctx_set_tuning(rmb_ctx* ctx, const rmb_tuning* tuning);

/* Copy of the last error message (NUL-terminated, truncated to len). */
int rmb_last_error(const rmb_ctx* ctx, char* buf, size_t len);

/* Load the subject library (the TE families): n sequences with their original
 * ambiguity codes.  Replaces SubjectDb::open + get_full_sequence_blastna +
 * get_prepared (seq/mod.rs:155-290, seq/fasta_db.rs:86-132). */
int rmb_library_load(rmb_ctx* ctx, uint32_t n, const uint8_t* const* seqs, const uint32_t* lens);

/* Search a batch of query records against every loaded subject, both strands.
 * Replaces build_query_lookup (engine.rs:606) + search_with_query_lookup over
 * all subjects (main.rs:1205-1223).  A query of 1,999,800 bases or more is
 * searched as NCBI's overlapping chunks with the cross-chunk merge, as
 * search_db_parallel does (main.rs:1150-1300).  Queries longer than 2^31-1
 * are refused.  Results are held by the context until the next search. */
int rmb_search(rmb_ctx* ctx, uint32_t n_queries, const uint8_t* const* queries, const uint32_t* lens);

/* Search only the listed (query, subject) pairs of a batch: the reference's
 * search_with_query_lookup for one query and one subject (engine.rs:715-729),
 * over many pairs at once.  Duplicate pairs are searched once; results as for
 * rmb_search, in (query, subject) order. */
int rmb_search_pairs(rmb_ctx* ctx, uint32_t n_queries, const uint8_t* const* queries, const uint32_t* lens,
                     uint64_t n_pairs, const uint32_t* pair_query, const uint32_t* pair_subject);

/* DUST intervals of BLASTNA bases (no sentinels): [first, last] pairs, 0-based
 * inclusive, into out (2 x *n_out entries on input); *n_out receives the
 * count.  Pass out = NULL to get the count.  Replaces dust_mask / the BLAST
 * filter path of rmblast-lib/src/filter/dust.rs:339-484.  Host only. */
int rmb_dust_intervals(const uint8_t* bases, uint64_t n, uint64_t* out, uint64_t* n_out);

/* Two-call result retrieval: sizes, then copy into caller buffers. */
int rmb_results_size(const rmb_ctx* ctx, uint64_t* n_hsps, uint64_t* n_ops);
int rmb_results_copy(const rmb_ctx* ctx, rmb_hsp* hsps, rmb_op* ops);

/* ---- the host tail after a search (rmblastn/src/main.rs:1224-1234) ------- */

/* Alignment statistics of one HSP (rmblast-lib/src/stats.rs:52-243,
 * compute_align_stats) from its aligned IUPAC strings of length n ('-' for a
 * gap), the subject taken as the ancestral sequence. */
typedef struct rmb_align_stats {
  uint32_t identities, mismatches, query_gaps, subject_gaps;
  uint32_t transitions, transversions, cpg_sites, pad;
  double kdiv, cpg_kdiv;
  float perc_sub, perc_query_gap, perc_db_gap, pad2;
} rmb_align_stats;
int rmb_align_stats_of(const char* query_aligned, const char* subject_aligned, uint64_t n, rmb_align_stats* out);

/* Mask level over one query's HSPs from all subjects (apply_mask_level in
 * its cross-subject form, engine.rs:863-974): sorted by score, then subject
 * index (the OID) descending, then ScoreCompareHSPs (hits.rs:104-112), an HSP
 * is dropped when an accepted HSP of at least its score covers
 * (int)(100 * overlap / span) >= mask_level percent of its query range
 * (minus-strand ranges shifted by one; at level 0 touching counts).  q_len is
 * the query's length; keep[i] receives 1 for a survivor.  Host only. */
int rmb_mask_level(const rmb_hsp* hsps, uint64_t n, uint32_t q_len, int32_t mask_level, uint8_t* keep);

/* What rmblastn reports of the last search.  mask_level < 0 or >= 100 keeps
 * every HSP.  ka (may be NULL: no statistics) gives the evalue / bitscore
 * fields, per query from its whole length and the loaded library's residues
 * and sequences; with evalue_max > 0 and statistics, HSPs whose E-value
 * exceeds it are dropped first (the --evalue reap, main.rs:1224-1226 and
 * 1418-1420, before mask level).  fields, query_ids and subject_ids are used
 * by rmb_report_tabular only. */
typedef struct rmb_report_spec {
  int32_t mask_level;
  int32_t pad;
  const char* fields;             /* field names separated by spaces */
  const char* const* query_ids;   /* qseqid of each query of the last search */
  const char* const* subject_ids; /* sseqid of each loaded subject */
  const rmb_ka_stats* ka;
  double evalue_max;
} rmb_report_spec;

/* The HSPs of the last search that rmblastn reports, in its output order:
 * per query, the reap, then mask level across all subjects (apply_mask_level,
 * engine.rs:863-974), then NCBI's order (sort_hit_list_order_by_oid,
 * engine.rs:998-1017, after main.rs:1230-1234).  rows receives indices into
 * the results of rmb_results_copy, queries in batch order.  Two-call:
 * rows = NULL gives the count in *n. */
int rmb_report_order(rmb_ctx* ctx, const rmb_report_spec* spec, uint64_t* rows, uint64_t* n);

/* outfmt 6 of the last search (output.rs:294-363): the rows of
 * rmb_report_order with the tab-separated fields, each line ending in '\n'.
 * Fields: score perc_sub perc_query_gap perc_db_gap qseqid qstart qend qlen
 * sstrand sseqid sstart send slen kdiv cpg_kdiv transi transv cpg_sites qseq
 * sseq evalue bitscore (with NCBI's column quirks, output.rs:338-340, and its
 * number formats, output.rs:256-288).  Queries are formatted on the
 * library's host threads.  The text is held by the context until the next
 * search or report; rmb_report_copy copies it out. */
int rmb_report_tabular(rmb_ctx* ctx, const rmb_report_spec* spec, uint64_t* n_bytes);

/* rmb_search, then rmb_report_tabular's text of that search, in one call:
 * rmblastn's search then report (main.rs:1133-1234), with the report written
 * while the search runs.  The search takes its queries in batches of whole
 * queries; as soon as a batch's HSPs are final, its queries' rows (reap,
 * mask level, order, statistics, outfmt 6 text) are written on a host worker
 * thread while the GPU and the host go on with the next batches, so only the
 * last batch's rows remain after the search.  The text is the bytes
 * rmb_report_tabular gives with the same spec (in place through
 * rmb_report_text, *n_bytes its length); the results are rmb_search's.
 * spec, its strings and the ids must stay valid until the call returns.
 * No counterpart call in the reference: the CLI's two steps in one. */
int rmb_search_report(rmb_ctx* ctx, uint32_t n_queries, const uint8_t* const* queries, const uint32_t* lens,
                      const rmb_report_spec* spec, uint64_t* n_bytes);
int rmb_report_copy(const rmb_ctx* ctx, char* out);
/* The report text in place (not NUL-terminated), valid until the context's
 * next search, report or destruction. */
int rmb_report_text(const rmb_ctx* ctx, const char** text, uint64_t* n_bytes);

/* Stage counters of the last search (the reference's diagnostics counters,
 * engine.rs:28-37 and gapped.rs:25-36): [0] seeds that reached the diagonal
 * hash, [1] ungapped hits kept, [2] preliminary DPs run, [3] traceback DPs run,
 * [4] DP cells (both phases, each band cell counted once), [5] of those, the
 * traceback phase's, [6] final HSPs,
 * [7] ungapped extension steps.  Stage wall times in microseconds:
 * [8] lookup build, [9] seeding + ungapped, [10] preliminary gapped,
 * [11] traceback, [12] total; [13] time inside the seeding kernel in
 * nanoseconds (HIP events on the stream it is launched on), [14] time inside
 * the gapped kernels (preliminary, score and traceback DP) in nanoseconds,
 * [15] launches of the seeding kernel, [16] seeds staged by the seeding
 * kernel, [17..22] seeding-kernel stage time (scan, output-slot reservation,
 * seed fill + extension, bucket sort, hash replay, emit) in 100 MHz ticks
 * summed over workgroups,
 * [23] time inside the preliminary DP kernel in nanoseconds (part of [14]),
 * [24] (rmb_search_report) microseconds of report writing left after the
 * last batch's HSPs were final.  Entries past `n` are not written;
 * RMB_NSTATS is the full length. */
#define RMB_NSTATS 25
int rmb_stats(const rmb_ctx* ctx, uint64_t* out, int n);

/* ---- parity taps (one stage at a time; same inputs as rmb_search) -------- */

/* Seeding + diagonal dedup + ungapped extension for query 0 of a one-query
 * batch against subject `subject`: the sorted ungapped hits the reference's
 * search_with_query_lookup_ungapped collects (engine.rs:817-832, order of
 * sort_ungapped_like_ncbi :1280-1309) in the record frame: per hit
 * {q_seed, s_seed, q_start, q_end, s_start, s_end, score, strand}.
 * Two-call: pass out = NULL to get the count in *n. */
int rmb_tap_ungapped(rmb_ctx* ctx, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out,
                     uint64_t* n);

/* Subject `subject` as the search scans it, from the device: the plus strand
 * with its ambiguity codes resolved by makeblastdb's generator (seeded with
 * the length, one draw per ambiguous base 5' to 3'; seq/twobit.rs:50-134,
 * seq/mod.rs:155-290 get_full_sequence_blastna) and its reverse complement,
 * each *n bases (no sentinels).  Two-call: with plus and rc NULL, or *n too
 * small, only the length is set. */
int rmb_tap_subject(const rmb_ctx* ctx, uint32_t subject, uint8_t* plus, uint8_t* rc, uint32_t* n);

/* The query's lookup table, as build_query_lookup builds it (engine.rs:606,
 * blast/nalookup.rs:128-181, blast/mblookup.rs:106-245) for a one-query
 * batch: *lut_width, then CSR cells[4^lut_width + 1] (cell w's offsets are
 * entries[cells[w] .. cells[w + 1])) and entries (offsets into the two-strand
 * query buffer: plus-strand word starts below q_len + 1, minus-strand ones
 * from there, in the order the scan visits them).  Two-call: with cells or
 * entries NULL, or *n_cells / *n_entries too small, only the sizes are set. */
int rmb_tap_lookup(rmb_ctx* ctx, const uint8_t* query, uint32_t q_len, int32_t* lut_width, uint32_t* cells,
                   uint64_t* n_cells, uint32_t* entries, uint64_t* n_entries);

/* The seeds the diagonal hash lets through for the query against subject
 * `subject`, as search_with_query_lookup_seeds collects them
 * (engine.rs:796-816; pushed at :1165 and :1237): per seed {q_off, s_off,
 * strand (0 plus, 1 minus)}, the reference's SeedRecord, in its scan order.
 * Two-call as rmb_tap_ungapped. */
int rmb_tap_seeds(rmb_ctx* ctx, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out,
                  uint64_t* n);

/* Phase 2a's preliminary HSPs for the query against subject `subject` (the
 * PrelimHsp list run_phase2a returns, engine.rs:1317-1411 with the common-
 * endpoint purge :1969-2083; the reference's search_phase2a,
 * engine.rs:1775-1784, per query and subject): per prelim {q_seed, s_seed
 * (the traceback seed), q_start, q_end, s_start, s_end, score, strand}, record
 * frame, in Phase 2a's order.  Runs the whole search of that pair.  Two-call
 * as rmb_tap_ungapped. */
int rmb_tap_prelims(rmb_ctx* ctx, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out,
                    uint64_t* n);

/* Score-only X-drop half alignment (gapped.rs:124-337, align_ex_score_only):
 * out = {score, a_len, b_len}. */
int rmb_tap_align_score_only(rmb_ctx* ctx, const uint8_t* a, uint32_t m, const uint8_t* b, uint32_t n, int32_t xdrop,
                             int32_t reverse, int64_t out[3]);

/* X-drop half alignment with traceback (gapped.rs:450-757, align_ex):
 * out = {score, a_len, b_len, n_ops}; ops receives up to max_ops ops. */
int rmb_tap_align_ex(rmb_ctx* ctx, const uint8_t* a, uint32_t m, const uint8_t* b, uint32_t n, int32_t xdrop,
                     int32_t reverse, int64_t out[4], rmb_op* ops, uint32_t max_ops);

/* ---- score-only X-drop DPs that allocate nothing ------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the score-only half alignments of rmb_tap_align_score_only for
 * n_halves halves whose sequences, descriptors and outputs are already in
 * device memory.  halves[i].a / .b are device pointers to BLASTNA codes (no
 * sentinels); out receives 3 int32 per half {score, a_len, b_len}.  The
 * library allocates nothing in this call: it works in one fixed scratch area
 * of about 19 MB per GPU, held until the process ends.  Limits:
 * n_halves <= RMB_NOALLOC_MAX_HALVES (else RMB_ERR_ARG); a band wider than
 * 1,024 cells gives RMB_ERR_CAPACITY (rmb_tap_align_score_only takes bands up
 * to 2^17 cells).  The call runs on the context's stream and returns when the
 * results are in `out`.  Calls that overlap on one GPU (from two threads or
 * two contexts) race on the scratch area. */
#define RMB_NOALLOC_MAX_HALVES 65536u
typedef struct rmb_dp_half {
  const uint8_t* a;   /* outer sequence (rows), device memory */
  const uint8_t* b;   /* inner sequence (columns), device memory */
  uint32_t m, n;      /* their lengths */
  int32_t xdrop;
  uint32_t reverse;   /* 1: align from the sequences' ends backwards */
} rmb_dp_half;
int rmb_align_score_only_noalloc(rmb_ctx* ctx, const rmb_dp_half* halves, uint32_t n_halves, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* RMBLAST_HIP_H */
