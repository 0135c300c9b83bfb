// CPU ORACLE (test infrastructure): 2-D interval tree for HSP containment.
// Restates rmblast-lib/src/search/itree.rs:60-545 (NCBI blast_itree.c,
// eQueryAndSubject): query-axis centred tree whose straddling nodes own a
// subject-axis tree; add() does the endpoint-sharing dedup, add_simple() not.
#pragma once
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace orc {

struct ITreeHsp { uint32_t q_start, q_end, s_start, s_end; int32_t score; };

class IntervalTree {
 public:
  IntervalTree(uint32_t q_start, uint32_t q_end, uint32_t s_min, uint32_t s_max, bool rc_query = false);
  bool add(const ITreeHsp& h, bool plus_strand);
  void add_simple(const ITreeHsp& h);
  bool contains(uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse, int32_t score, bool plus,
                int64_t min_diag_sep) const;

 private:
  static constexpr uint32_t INTERNAL = 0xFFFFFFFFu, NIL = 0;
  struct Node { uint32_t leftend, rightend, leftptr, midptr, rightptr, hsp_idx; };
  std::vector<Node> nodes_;
  std::vector<ITreeHsp> hsps_;
  std::vector<uint8_t> deleted_;
  std::map<std::pair<uint32_t, uint32_t>, std::pair<int32_t, uint32_t>> left_ep_, right_ep_;
  uint32_t s_min_, s_max_;
  bool rc_query_;

  uint32_t alloc_internal(uint32_t le, uint32_t re);
  uint32_t alloc_leaf(const ITreeHsp& h);
  void remove_endpoints(uint32_t idx, bool plus);
  void insert_structural(uint32_t new_idx);
  bool midpoint_contains(uint32_t root, uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse, int32_t score,
                         bool plus, int64_t sep) const;
  bool hsp_is_contained(uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse, int32_t score, bool plus,
                        int64_t sep, uint32_t idx) const;
};

}  // namespace orc
