// Host tail of librmblast_hip: what rmblastn does with a query's HSPs after
// the search (rmblastn/src/main.rs:1224-1234): mask level across subjects
// (rmblast-lib/src/search/engine.rs:863-974), NCBI's output order
// (engine.rs:998-1017), alignment statistics (rmblast-lib/src/stats.rs:52-243)
// and the outfmt 6 writer (rmblast-lib/src/output.rs:294-363).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rmblast_hip.h"

namespace rmb {

// The HSPs of a search and the sequences they refer to (all host memory).
struct ReportData {
  const rmb_hsp* hsps = nullptr;
  uint64_t n_hsps = 0;
  const rmb_op* ops = nullptr;
  uint32_t n_queries = 0;
  const uint8_t* const* queries = nullptr;  // sentinel-wrapped BLASTNA, n_queries
  const uint32_t* query_lens = nullptr;     // bases, without sentinels
  uint32_t n_subjects = 0;
  const uint8_t* const* subjects = nullptr;     // sentinel-wrapped, original ambiguity codes
  const uint8_t* const* subjects_rc = nullptr;  // their reverse complements
  const uint32_t* subject_lens = nullptr;
  // E-value statistics (null: none) with the database's residues and
  // sequences, and the reap threshold (<= 0: no reap)
  const rmb_ka_stats* ka = nullptr;
  int64_t db_length = 0;
  int32_t db_num_seqs = 0;
  double evalue_max = 0;
};

// NCBI's tabular number formats (output.rs:256-288).
std::string format_evalue(double evalue);
std::string format_bit_score(double bits);

// Per query, the indices of its HSPs that survive mask level, in output
// order.  mask_level >= 100 (or < 0) keeps every HSP.  Queries run on
// `threads` host threads.  With a range [q0, q1), only those queries (entry
// q - q0), whose HSPs are all in d.
std::vector<std::vector<uint64_t>> report_order(const ReportData& d, int32_t mask_level, unsigned threads,
                                                uint32_t q0 = 0, uint32_t q1 = UINT32_MAX);

// Mask level over one query's HSPs (indices into d.hsps), in place.
void mask_level_filter(const ReportData& d, std::vector<uint64_t>& rows, int32_t level, uint32_t q_len);

// A report's text in one buffer.
struct ReportText {
  std::shared_ptr<char> data;
  uint64_t size = 0;
};

// The outfmt 6 field list: field codes, and whether the rows need the
// aligned strings.  Throws std::invalid_argument naming an unknown field.
struct ReportFields {
  std::vector<int> codes;
  bool need_aln = false, need_stats = false;
};
ReportFields parse_report_fields(const std::vector<std::string>& fields);

// The outfmt 6 rows of `order` (one line per HSP, tab-separated `fields`).
// Throws std::invalid_argument naming an unknown field.
ReportText report_tabular(const ReportData& d, const std::vector<std::vector<uint64_t>>& order,
                           const std::vector<std::string>& fields, const char* const* query_ids,
                           const char* const* subject_ids, unsigned threads);

// A report written while the search runs: the search hands over the final
// HSPs of consecutive query ranges as it completes them (add), and each
// range's rows (reap, mask level, order, statistics, outfmt 6 text) are
// appended to one buffer in query order, reserved up front so that it never
// moves; finish() gives the whole text, the same bytes report_tabular writes
// over the finished search.
class ReportStream {
 public:
  ReportStream(const std::vector<std::string>& fields, const char* const* query_ids, const char* const* subject_ids,
               int32_t mask_level);
  ~ReportStream();
  ReportStream(const ReportStream&) = delete;
  ReportStream& operator=(const ReportStream&) = delete;
  // d holds every HSP of queries [q0, q1) (and no other query's)
  void add(const ReportData& d, uint32_t q0, uint32_t q1, unsigned threads);
  ReportText finish();

 private:
  void reserve(uint64_t need);
  ReportFields f_;
  const char* const* qids_;
  const char* const* sids_;
  int32_t mask_level_;
  char* buf_ = nullptr;
  uint64_t cap_ = 0, used_ = 0;
  bool mapped_ = false;
};

// The aligned IUPAC strings of one HSP (forward query frame; a minus-strand
// subject reads reverse-complemented).
void aligned_strings(const ReportData& d, const rmb_hsp& h, std::string& qa, std::string& sa);

// stats.rs:52-243 on aligned IUPAC strings, the subject as the ancestral
// sequence.
rmb_align_stats align_stats(const char* q, const char* s, uint64_t n);

}  // namespace rmb
