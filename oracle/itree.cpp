// CPU ORACLE (test infrastructure): interval tree, restating
// rmblast-lib/src/search/itree.rs:60-545.
#include "itree.hpp"

#include <cstdlib>

namespace orc {

IntervalTree::IntervalTree(uint32_t q_start, uint32_t q_end, uint32_t s_min, uint32_t s_max, bool rc_query)
    : s_min_(s_min), s_max_(s_max), rc_query_(rc_query) {
  nodes_.push_back(Node{q_start, q_end, NIL, NIL, NIL, INTERNAL});
}

uint32_t IntervalTree::alloc_internal(uint32_t le, uint32_t re) {
  nodes_.push_back(Node{le, re, NIL, NIL, NIL, INTERNAL});
  return (uint32_t)nodes_.size() - 1;
}

uint32_t IntervalTree::alloc_leaf(const ITreeHsp& h) {
  uint32_t hi = (uint32_t)hsps_.size();
  hsps_.push_back(h);
  deleted_.push_back(0);
  nodes_.push_back(Node{0, 0, NIL, NIL, NIL, hi});
  return (uint32_t)nodes_.size() - 1;
}

static bool new_wins(int32_t ns, uint32_t nq, uint32_t nsl, int32_t es, uint32_t eq, uint32_t esl) {
  // itree.rs:246-255
  if (ns != es) return ns > es;
  if (nq != eq) return nq < eq;
  if (nsl != esl) return nsl < esl;
  return false;
}

void IntervalTree::remove_endpoints(uint32_t idx, bool plus) {  // itree.rs:222-237
  deleted_[idx] = 1;
  const ITreeHsp& e = hsps_[idx];
  std::pair<uint32_t, uint32_t> l = plus ? std::make_pair(e.q_start, e.s_start) : std::make_pair(e.q_end, e.s_start);
  std::pair<uint32_t, uint32_t> r = plus ? std::make_pair(e.q_end, e.s_end) : std::make_pair(e.q_start, e.s_end);
  auto il = left_ep_.find(l);
  if (il != left_ep_.end() && il->second.second == idx) left_ep_.erase(il);
  auto ir = right_ep_.find(r);
  if (ir != right_ep_.end() && ir->second.second == idx) right_ep_.erase(ir);
}

bool IntervalTree::add(const ITreeHsp& h, bool plus) {  // itree.rs:164-216
  std::pair<uint32_t, uint32_t> l = plus ? std::make_pair(h.q_start, h.s_start) : std::make_pair(h.q_end, h.s_start);
  std::pair<uint32_t, uint32_t> r = plus ? std::make_pair(h.q_end, h.s_end) : std::make_pair(h.q_start, h.s_end);
  uint32_t nq = h.q_end - h.q_start, ns = h.s_end - h.s_start;
  auto il = left_ep_.find(l);
  if (il != left_ep_.end()) {
    uint32_t ei = il->second.second;
    const ITreeHsp& e = hsps_[ei];
    if (!new_wins(h.score, nq, ns, il->second.first, e.q_end - e.q_start, e.s_end - e.s_start)) return false;
    remove_endpoints(ei, plus);
  }
  auto ir = right_ep_.find(r);
  if (ir != right_ep_.end()) {
    uint32_t ei = ir->second.second;
    const ITreeHsp& e = hsps_[ei];
    if (!new_wins(h.score, nq, ns, ir->second.first, e.q_end - e.q_start, e.s_end - e.s_start)) return false;
    remove_endpoints(ei, plus);
  }
  uint32_t ni = alloc_leaf(h);
  uint32_t hi = nodes_[ni].hsp_idx;
  left_ep_[l] = {h.score, hi};
  right_ep_[r] = {h.score, hi};
  insert_structural(ni);
  return true;
}

void IntervalTree::add_simple(const ITreeHsp& h) { insert_structural(alloc_leaf(h)); }

void IntervalTree::insert_structural(uint32_t new_idx) {  // itree.rs:270-373
  const ITreeHsp& h = hsps_[nodes_[new_idx].hsp_idx];
  const uint32_t s_off = h.s_start, s_end = h.s_end;
  uint32_t rs = h.q_start, re = h.q_end;
  bool subj = false;
  uint32_t root = 0;
  for (;;) {
    const uint32_t rle = nodes_[root].leftend, rre = nodes_[root].rightend;
    const uint64_t middle = ((uint64_t)rle + rre) / 2;
    bool left_half;
    uint32_t old_index;
    if ((uint64_t)re < middle) {
      uint32_t lp = nodes_[root].leftptr;
      if (lp == NIL) { nodes_[root].leftptr = new_idx; return; }
      if (nodes_[lp].hsp_idx == INTERNAL) { root = lp; continue; }
      old_index = lp;
      left_half = true;
    } else if ((uint64_t)rs > middle) {
      uint32_t rp = nodes_[root].rightptr;
      if (rp == NIL) { nodes_[root].rightptr = new_idx; return; }
      if (nodes_[rp].hsp_idx == INTERNAL) { root = rp; continue; }
      old_index = rp;
      left_half = false;
    } else {
      if (subj) {
        nodes_[new_idx].midptr = nodes_[root].midptr;
        nodes_[root].midptr = new_idx;
        return;
      }
      subj = true;
      if (nodes_[root].midptr == NIL) {
        uint32_t mid = alloc_internal(s_min_, s_max_);
        nodes_[root].midptr = mid;
      }
      root = nodes_[root].midptr;
      rs = s_off;
      re = s_end;
      continue;
    }
    uint32_t mle = left_half ? rle : (uint32_t)middle + 1;
    uint32_t mre = left_half ? (uint32_t)middle : rre;
    uint32_t mi = alloc_internal(mle, mre);
    if (left_half) nodes_[root].leftptr = mi; else nodes_[root].rightptr = mi;
    const ITreeHsp& o = hsps_[nodes_[old_index].hsp_idx];
    uint32_t ors = subj ? o.s_start : o.q_start, ore = subj ? o.s_end : o.q_end;
    root = mi;
    const uint64_t mmid = ((uint64_t)mle + mre) / 2;
    if ((uint64_t)ore < mmid) {
      nodes_[mi].leftptr = old_index;
    } else if ((uint64_t)ors > mmid) {
      nodes_[mi].rightptr = old_index;
    } else if (subj) {
      nodes_[mi].midptr = old_index;
    } else {
      uint32_t os = o.s_start, oe = o.s_end;
      uint32_t mid2 = alloc_internal(s_min_, s_max_);
      nodes_[mi].midptr = mid2;
      const uint64_t m2 = ((uint64_t)s_min_ + s_max_) / 2;
      if ((uint64_t)oe < m2) nodes_[mid2].leftptr = old_index;
      else if ((uint64_t)os > m2) nodes_[mid2].rightptr = old_index;
      else nodes_[mid2].midptr = old_index;
    }
  }
}

bool IntervalTree::contains(uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse, int32_t score, bool plus,
                            int64_t sep) const {  // itree.rs:379-429
  uint32_t node = 0;
  for (;;) {
    uint32_t mid = nodes_[node].midptr;
    if (mid != NIL && midpoint_contains(mid, cqs, cqe, css, cse, score, plus, sep)) return true;
    const uint64_t middle = ((uint64_t)nodes_[node].leftend + nodes_[node].rightend) / 2;
    uint32_t next;
    if ((uint64_t)cqe < middle) next = nodes_[node].leftptr;
    else if ((uint64_t)cqs > middle) next = nodes_[node].rightptr;
    else return false;
    if (next == NIL) return false;
    if (nodes_[next].hsp_idx != INTERNAL)
      return hsp_is_contained(cqs, cqe, css, cse, score, plus, sep, nodes_[next].hsp_idx);
    node = next;
  }
}

bool IntervalTree::midpoint_contains(uint32_t root, uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse,
                                     int32_t score, bool plus, int64_t sep) const {  // itree.rs:437-485
  uint32_t node = root;
  for (;;) {
    if (nodes_[node].hsp_idx != INTERNAL)
      return hsp_is_contained(cqs, cqe, css, cse, score, plus, sep, nodes_[node].hsp_idx);
    for (uint32_t m = nodes_[node].midptr; m != NIL; m = nodes_[m].midptr)
      if (hsp_is_contained(cqs, cqe, css, cse, score, plus, sep, nodes_[m].hsp_idx)) return true;
    const uint64_t middle = ((uint64_t)nodes_[node].leftend + nodes_[node].rightend) / 2;
    uint32_t next;
    if ((uint64_t)cse < middle) next = nodes_[node].leftptr;
    else if ((uint64_t)css > middle) next = nodes_[node].rightptr;
    else return false;
    if (next == NIL) return false;
    node = next;
  }
}

bool IntervalTree::hsp_is_contained(uint32_t cqs, uint32_t cqe, uint32_t css, uint32_t cse, int32_t score,
                                    bool plus, int64_t sep, uint32_t idx) const {  // itree.rs:496-544
  if (deleted_[idx]) return false;
  const ITreeHsp& h = hsps_[idx];
  if (score > h.score) return false;
  if (!(h.q_start <= cqs && h.q_end >= cqe && h.s_start <= css && h.s_end >= cse)) return false;
  int64_t d0, d1;
  if (plus || rc_query_) {
    d0 = std::llabs(((int64_t)cqs - (int64_t)css) - ((int64_t)h.q_start - (int64_t)h.s_start));
    d1 = std::llabs(((int64_t)cqe - (int64_t)cse) - ((int64_t)h.q_end - (int64_t)h.s_end));
  } else {
    d0 = std::llabs(((int64_t)cqe + (int64_t)css) - ((int64_t)h.q_end + (int64_t)h.s_start));
    d1 = std::llabs(((int64_t)cqs + (int64_t)cse) - ((int64_t)h.q_start + (int64_t)h.s_end));
  }
  return d0 < sep || d1 < sep;
}

}  // namespace orc
