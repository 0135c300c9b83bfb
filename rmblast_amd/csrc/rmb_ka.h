// Karlin-Altschul ungapped cutoff (host code): the table lookup and the
// per-query cutoff of rmblast-lib/src/search/ka_cutoff.rs:105-161.
#pragma once
#include <cstdint>

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

}  // namespace rmb
