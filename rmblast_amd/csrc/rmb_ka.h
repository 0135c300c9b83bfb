// Karlin-Altschul ungapped cutoff (host code): the table lookup and the
// per-query cutoff of rmblast-lib/src/search/ka_cutoff.rs:105-161.
#pragma once
#include <cstdint>
#include <string>

#include "../../include/rmblast_hip.h"

namespace rmb {

// The part of a matrix path after its last '/'.
const char* matrix_base_name(const char* path);

// (lambda, K) of a table matrix at its gap costs; false when not in the table.
bool ka_lookup(const char* matrix_name, int32_t gap_open, int32_t gap_extend, double& lambda, double& k);

// min_raw_gapped_score / 2, or 1 when that is 0 (options.rs:80-82).
int32_t cutoff_score_max(int32_t min_raw_gapped_score);

// The KA-derived cutoff for one query, capped at cutoff_score_max, or -1 when
// no statistics apply (lambda or K not positive, or no subjects).  es_out and
// searchsp_out (may be null) receive the uncapped score and the search space.
int32_t ka_ungapped_cutoff(double lambda, double k, int32_t min_raw_gapped_score, uint64_t query_length,
                           uint64_t avg_subj_length, int32_t* es_out, uint64_t* searchsp_out);

// The "# KARLIN" comment of a matrix file (matrix.rs:248-266): v = {lambda,
// k, h, alpha, beta}, 0 where not given; false with a message when a key or
// value does not parse.
bool read_karlin_comment(const char* path, double v[5], bool& found, std::string& err);

// E-value statistics of a matrix + gap costs (ka_stats.rs:153-300); false
// with a message when the matrix file's "# KARLIN" comment does not parse.
bool ka_stats_resolve(const char* matrix_path, int32_t gap_open, int32_t gap_extend,
                      const rmb_ka_stats_options& opt, rmb_ka_stats& out, std::string& err);

// BLAST_ComputeLengthAdjustment (rmstats/src/length_adjust.rs:7-93).
int32_t length_adjustment(double k, double log_k, double alpha_d_lambda, double beta, int32_t query_length,
                          int64_t db_length, int32_t db_num_seqs);

// Effective search space of one query (rmstats/src/eff_lengths.rs:52-128).
void ka_search_space(const rmb_ka_stats& s, int32_t query_length, int64_t db_length, int32_t db_num_seqs,
                     int64_t& eff_searchsp, int32_t& length_adjustment);

// E-value and bit score of a raw score (rmstats/src/evalue.rs:96-126).
double ka_evalue(const rmb_ka_stats& s, int64_t eff_searchsp, int32_t score);
double ka_bit_score(const rmb_ka_stats& s, int32_t score);

}  // namespace rmb
