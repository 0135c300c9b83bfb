// CPU ORACLE — test infrastructure only.
//
// A plain C++ restatement of the rmblastn search path of Dfam-consortium/RMBlast
// (Rust port of NCBI RMBlast 2.17.1), written by reading the reference as text.
// It is the checker for the HIP product path and the timed CPU baseline; it is
// never linked into the product library.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg use it.
//
// Each function cites the reference file:line it follows (paths relative to the
// reference repository root).  Coordinates follow the reference: 0-based,
// half-open, sequences sentinel-wrapped BLASTNA ([15, bases..., 15]).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <utility>

namespace orc {

// ---- encoding (rmblast-lib/src/encoding.rs:19-141) --------------------------
extern const char kBlastnaToIupac[17];
extern const uint8_t kComplement[16];
std::vector<uint8_t> encode_iupac(const std::string& s);          // sentinel-wrapped
std::vector<uint8_t> revcomp(const std::vector<uint8_t>& s);      // whole buffer reversed+complemented

// ---- matrix (rmblast-lib/src/matrix.rs:62-354) ------------------------------
struct Matrix {
  int32_t s[16][16];
  double freqs[16];
  double lambda;
  std::string name;
  int32_t score(uint8_t q, uint8_t b) const { return s[q & 15][b & 15]; }
};
bool matrix_from_text(const std::string& name, const std::string& text, Matrix& m, std::string& err);
bool matrix_from_file(const std::string& path, Matrix& m, std::string& err);

// ---- parameters (rmblast-lib/src/options.rs:12-104) -------------------------
struct KaStats;
struct Params {
  int gap_open = 20, gap_extend = 5;
  int word_size = 11;
  int xdrop_ungap = 20, xdrop_gap = 30, xdrop_gap_final = 100;
  int min_raw_gapped_score = 0;
  int ungapped_cutoff = -1;     // <0: unset -> cutoff_score_max()
  bool complexity_adjust = false;
  bool dust = false;
  int mask_level = 101;
  // --evalue reap before mask level (main.rs:1224-1226): statistics, the
  // query's effective search space and the threshold; null = off
  const KaStats* reap_ka = nullptr;
  int64_t reap_searchsp = 0;
  double reap_evalue = 0;
  int gapped_cutoff() const { return min_raw_gapped_score > 1 ? min_raw_gapped_score : 1; }
  int cutoff_score_max() const { return min_raw_gapped_score > 0 ? min_raw_gapped_score / 2 : 1; }
  int ungapped_keep() const { return ungapped_cutoff >= 0 ? ungapped_cutoff : cutoff_score_max(); }
};

// ---- NcbiRandom (rmblast-lib/src/seq/twobit.rs:50-134) ----------------------
struct NcbiRandom {
  uint32_t state[33];
  int rj, rk;
  explicit NcbiRandom(uint32_t seed);
  uint32_t rand32();
  uint32_t get_rand() { return rand32() >> 1; }
  uint8_t next_base_iupac(uint8_t code);
};

// ---- FASTA + subject db (rmblast-lib/src/seq/fasta.rs, fasta_db.rs:86-132) --
struct FastaRecord { std::string id, defline; std::vector<uint8_t> seq; };  // seq sentinel-wrapped
bool read_fasta(const std::string& path, std::vector<FastaRecord>& out, std::string& err);
std::vector<uint8_t> resolve_ambiguity(const std::vector<uint8_t>& raw);   // fasta_db.rs:86-105
std::vector<uint8_t> n_mask_of(const std::vector<uint8_t>& raw);           // fasta_db.rs:115-132

struct PreparedSubject {          // seq/mod.rs:38-100
  std::vector<uint8_t> plus;      // ambiguity-resolved, sentinel-wrapped
  std::vector<uint8_t> rc;        // revcomp(plus)
  std::vector<uint8_t> align;     // plus with original ambiguity codes restored
  std::vector<uint8_t> align_rc;  // revcomp(align)
  std::vector<uint8_t> n_mask;    // empty if unambiguous
  std::vector<uint8_t> n_mask_rc;
};
PreparedSubject prepare_subject(const std::vector<uint8_t>& raw_sentinel_wrapped);

// ---- lookup tables (blast/nalookup.rs, blast/mblookup.rs) -------------------
// Seeds come out of a table as the reference's scanners deliver them: subject
// positions in scan order, and for each position the cell's query offsets in
// cell order (ascending for the small table, LIFO/descending for the megablast
// table).  The table is stored as CSR: cell_start[4^lw + 1], entries[].
struct Lookup {
  int lut_width = 0, word_length = 0, scan_step = 1;
  bool mega = false;
  int longest_chain = 0;
  std::vector<uint32_t> cell_start;
  std::vector<uint32_t> entries;
};
int choose_lut_width(int word_size, size_t approx_entries);   // nalookup.rs:254-291
struct QueryLookup {                                           // engine.rs:579-690
  Lookup lut;
  uint32_t fwd_n = 0, c1_start = 0;
  std::vector<uint8_t> masked;  // DUST-masked query (sentinel-wrapped); empty with dust off
};
// engine.rs:606-619: with dust on, the table is built from the DUST-masked
// query and exact word extension reads it; extensions use the query itself.
QueryLookup build_query_lookup(const std::vector<uint8_t>& query, const Params& p);

// ---- core kernels ------------------------------------------------------------
struct Ungapped { int32_t score; uint32_t q_start, q_end, s_start, s_end; };
Ungapped extend_ungapped(const uint8_t* query, size_t qlen_with_sent, const uint8_t* subject,
                         size_t slen_with_sent, uint32_t q_off, uint32_t s_off,
                         const Matrix& m, int xdrop);          // ungapped.rs:38-121

struct DiagHash {                                              // diag_hash.rs:15-111
  std::vector<uint32_t> backbone;
  struct Cell { int32_t diag; uint32_t level; uint32_t next; };
  std::vector<Cell> chain;
  uint32_t occupancy = 1;
  DiagHash();
  void clear();
  static uint32_t bucket(int32_t diag) { return ((uint32_t)diag * 0x9E370001u) % 512u; }
  uint32_t get(int32_t diag) const;
  void insert(int32_t diag, uint32_t s_end, uint32_t s_off, int32_t window);
};

enum Op : uint8_t { OP_SUB = 0, OP_GAP_IN_QUERY = 1, OP_GAP_IN_SUBJECT = 2 };
using EditOps = std::vector<std::pair<Op, uint32_t>>;
void push_op(EditOps& e, Op op, uint32_t n);

struct DpCounters { uint64_t score_only_cells = 0, traceback_cells = 0; };
// gapped.rs:124-337 (score-only X-drop) -> (score, a_len, b_len)
void align_score_only(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int gap_open, int gap_extend,
                      int xdrop, const Matrix& mat, bool reverse, int32_t& score, size_t& a_len, size_t& b_len,
                      DpCounters* ctr);
// gapped.rs:360-418
bool gapped_score_only(const std::vector<uint8_t>& query, const std::vector<uint8_t>& subject, uint32_t q_seed,
                       uint32_t s_seed, int gap_open, int gap_extend, int xdrop, const Matrix& mat, int32_t& score,
                       uint32_t& q0, uint32_t& q1, uint32_t& s0, uint32_t& s1, DpCounters* ctr);
// gapped.rs:450-757 (ALIGN_EX with traceback)
struct GapResult { int32_t score; size_t a_len, b_len; EditOps ops; };
GapResult align_ex(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int gap_open, int gap_extend, int xdrop,
                   const Matrix& mat, bool reverse, DpCounters* ctr);
// engine.rs:2953-3039
bool gapped_extend_bidirectional(const std::vector<uint8_t>& query, const std::vector<uint8_t>& subject,
                                 uint32_t q_seed, uint32_t s_seed, int gap_open, int gap_extend, int xdrop,
                                 const Matrix& mat, int32_t& score, uint32_t& q0, uint32_t& q1, uint32_t& s0,
                                 uint32_t& s1, EditOps& ops, DpCounters* ctr);
// gapped.rs:773-820
void extract_aligned(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen, size_t a_start, size_t b_start,
                     const EditOps& ops, const std::vector<uint8_t>& n_mask, std::vector<uint8_t>& qa,
                     std::vector<uint8_t>& sa);

// ---- statistics (stats.rs:52-243), complexity adjust (complexity_adjust.rs) --
struct AlignStats {
  uint32_t matches = 0, mismatches = 0, query_gaps = 0, subject_gaps = 0;
  uint32_t transitions = 0, transversions = 0, cpg_sites = 0;
  double kdiv = 0, cpg_kdiv = 0;
  float perc_sub = 0, perc_query_gap = 0, perc_db_gap = 0;
};
AlignStats compute_align_stats(const std::string& q, const std::string& s);
std::string to_iupac(const std::vector<uint8_t>& v);
double compute_adjusted_score(int32_t raw, const std::vector<uint8_t>& bases, const Matrix& m);
bool apply_complexity_adjust(int32_t raw, const std::vector<uint8_t>& query, uint32_t q_start, const EditOps& ops,
                             const Matrix& m, int cutoff, bool minus, int32_t& out);

// ---- results ------------------------------------------------------------------
enum Strand : uint8_t { PLUS = 0, MINUS = 1 };
struct Hsp {
  int32_t score;
  uint32_t q_start, q_end, q_len, s_start, s_end, s_len;
  Strand strand;
  EditOps ops;
  std::vector<uint8_t> q_seq, s_seq;
};
struct AlignResult { Hsp hsp; std::string query_id, subject_id; AlignStats stats; };

struct SeedRecord { uint32_t q_off, s_off; Strand strand; };
struct UngappedHit { uint32_t q_start, q_end, s_start, s_end; int32_t score; Strand strand; };
struct UngappedRecord {
  uint32_t q_seed, s_seed, q_start, q_end, s_start, s_end; int32_t score; Strand strand;
};
struct PrelimHsp {
  uint32_t q_seed, s_seed, q_start, q_end, s_start, s_end; int32_t score; Strand strand;
};

struct Taps {                       // optional stage outputs for parity tests
  std::vector<SeedRecord>* seeds = nullptr;
  std::vector<UngappedRecord>* ungapped = nullptr;   // sorted, record frame
  std::vector<PrelimHsp>* prelims = nullptr;          // after Phase 2a purge
  DpCounters* counters = nullptr;
  uint64_t* n_seed_evals = nullptr;                   // seeds that reached the diag hash
  uint64_t* n_ungapped_steps = nullptr;
};

// engine.rs:734-790 single-chunk search of one query against one subject
struct ChunkPlan { size_t size = 0, count = 1, step = 0; };   // rmblastn/src/main.rs:1150-1190
ChunkPlan chunk_plan(size_t query_len);
// rmblastn/src/main.rs:1016-1085
void merge_chunk_prelims(std::vector<PrelimHsp>& acc, std::vector<PrelimHsp> nw, uint32_t split);
std::vector<AlignResult> search_one_subject(const QueryLookup& ql, const std::vector<uint8_t>& query,
                                            const std::vector<uint8_t>& query_rc, const std::string& query_id,
                                            const PreparedSubject& subj, const std::string& subject_id,
                                            const Params& p, const Matrix& m, Taps* taps);
// engine.rs:863-974
void apply_mask_level(std::vector<AlignResult>& r, int mask_level, const std::vector<std::string>& subject_names);
// engine.rs:998-1017
void sort_hit_list_order(std::vector<AlignResult>& r, const std::vector<std::string>& subject_names);
// main.rs:1133-1234 single-chunk form (mask level + ordering)
std::vector<AlignResult> search_query(const std::vector<uint8_t>& query, const std::string& query_id,
                                      const std::vector<PreparedSubject>& subjects,
                                      const std::vector<std::string>& subject_names, const Params& p,
                                      const Matrix& m, int threads);
// output.rs:294-363 outfmt 6
std::string format_tabular(const AlignResult& r, const std::vector<std::string>& fields, const KaStats* ka = nullptr,
                           int64_t searchsp = 0);

// engine.rs:383-572 (query/subject without sentinels)
void improve_seed(const uint8_t* query, size_t qlen, const uint8_t* subject, size_t slen, uint32_t q_seed,
                  uint32_t s_seed, uint32_t pqs, uint32_t pqe, uint32_t pss, uint32_t pse, bool minus, uint32_t& oq,
                  uint32_t& os);

// ---- Karlin-Altschul ungapped cutoff (rmblast-lib/src/search/ka_cutoff.rs) ----
struct KaCutoffInfo { bool found; double lambda, k; uint64_t searchsp; int es; };
bool ka_lookup(const std::string& matrix_name, int gap_open, int gap_extend, double& lambda, double& k);  // :105-114
// :126-153; query_length = the whole query record, avg_subj_length = db residues / sequences
int ungapped_cutoff(const std::string& matrix_name, int gap_open, int gap_extend, int min_raw_gapped_score,
                    uint64_t query_length, uint64_t avg_subj_length, KaCutoffInfo* info);

// ---- E-value statistics (rmblast-lib/src/ka_stats.rs:153-300, rmstats) ----------
enum { KA_SENTINEL = 0, KA_TABLE = 1, KA_CLI = 2, KA_COMMENT = 3 };
struct KaStats { int source; double lambda, k, log_k, h, alpha, beta; };
struct KaOptions { double lambda = 0, k = 0, alpha = 0, beta = 0; bool ncbi_compat = false; };
bool karlin_comment(const std::string& path, bool& found, double kv[5], std::string& err);
bool resolve_ka_stats(const std::string& matrix_path, int go, int ge, const KaOptions& o, KaStats& out,
                      std::string& err);
int compute_length_adjustment(double K, double logK, double a_d_l, double beta, int qlen, int64_t dblen, int nseq);
int64_t effective_search_space(const KaStats& s, int qlen, int64_t dblen, int nseq, int* la_out);
double evalue_of(const KaStats& s, int64_t searchsp, int score);
double bit_score_of(const KaStats& s, int score);
std::string fmt_evalue(double e);
std::string fmt_bits(double v);

// unit-level entry points used by tests
void purge_prelim_common_endpoints(std::vector<PrelimHsp>& v, uint32_t sa_len);
int purge_hsps_with_common_endpoints(std::vector<AlignResult*>& arr, bool purge);

}  // namespace orc
