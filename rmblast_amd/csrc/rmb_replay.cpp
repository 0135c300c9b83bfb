// Ordered host replay of the reference's per-(query, subject) decisions; see
// rmb_replay.h for the reference sections each routine follows.
#include "rmb_replay.h"

#include <algorithm>
#include <cmath>
#include <optional>
#include <tuple>

namespace rmb {

const uint8_t kBlastnaComplement[16] = {3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15};

void script_push(Script& s, uint8_t kind, uint32_t len) {
  if (!len) return;
  if (!s.empty() && s.back().kind == kind) s.back().len += len;
  else s.push_back(Run{kind, len});
}

static inline uint32_t sub_sat(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

// ------------------------------------------------------------ ordering ------
static inline std::tuple<int32_t, uint32_t, uint32_t, uint64_t, uint32_t> ung_key(const UngRec& a, uint32_t q_len,
                                                                                  uint32_t s_len) {
  // score DESC, fwd s_start ASC, length DESC, combined-query key ASC, seed key ASC
  const uint32_t ss = a.minus ? s_len - a.s_end : a.s_start;
  const uint64_t qk = a.minus ? ((uint64_t)q_len + 1) + (q_len > a.q_end ? (uint64_t)(q_len - a.q_end) : 0)
                              : a.q_start;
  const uint32_t sk = a.minus ? s_len - a.s_seed : a.s_seed;
  return std::make_tuple(-a.score, ss, ~(a.s_end - a.s_start), qk, sk);
}

bool ung_before(const UngRec& a, const UngRec& b, uint32_t q_len, uint32_t s_len) {
  return ung_key(a, q_len, s_len) < ung_key(b, q_len, s_len);
}

void order_ungapped(std::vector<UngRec>& v, uint32_t q_len, uint32_t s_len) {
  std::sort(v.begin(), v.end(), [&](const UngRec& a, const UngRec& b) { return ung_before(a, b, q_len, s_len); });
}

void pivot_of(const UngRec& u, uint32_t qa_len, uint32_t sa_len, uint32_t lw, uint32_t& pq, uint32_t& ps,
              uint32_t& tq, uint32_t& ts) {
  if (!u.minus) {
    const bool shift = u.s_end >= u.s_seed + 8;
    tq = u.q_seed + (shift ? 3 : 0);
    ts = u.s_seed + (shift ? 3 : 0);
    const uint32_t up = 4 - (ts & 3);  // next 4-base subject boundary
    ps = ts + up;
    pq = tq + up;
    if (ps >= sa_len || pq >= qa_len) { ps -= 4; pq -= 4; }
    return;
  }
  // minus: the +3 rule is decided in forward-subject space; the pivot is
  // computed in the (RC query, forward subject) frame the DP runs in.
  const uint32_t saved_fwd = sa_len - u.s_seed - lw;
  const bool shift = (sa_len - u.s_start) >= saved_fwd + 8;
  tq = shift ? u.q_seed + lw - 4 : u.q_seed;
  ts = shift ? u.s_seed + lw - 4 : u.s_seed;
  const uint32_t qn = (qa_len - u.q_seed - lw) + (shift ? 3 : 0);
  const uint32_t sn = (sa_len - u.s_seed - lw) + (shift ? 3 : 0);
  const uint32_t up = 4 - (sn & 3);
  pq = qn + up;
  ps = sn + up;
  if (pq > qa_len || ps > sa_len) { pq -= 4; ps -= 4; }
}

void purge_prelims(std::vector<PreRec>& v, uint32_t sa_len) {
  const size_t n = v.size();
  if (n < 2) return;
  // NCBI context-space keys: (strand, query.offset, subject.offset) and (strand, query.end, subject.end)
  struct K { uint32_t ctx, qo, so, qe, se; int32_t sc; };
  std::vector<K> k(n);
  for (size_t i = 0; i < n; ++i) {
    const PreRec& h = v[i];
    k[i].ctx = h.minus;
    k[i].qo = h.minus ? h.q_end : h.q_start;
    k[i].qe = h.minus ? h.q_start : h.q_end;
    k[i].so = h.minus ? sa_len - h.s_end : h.s_start;
    k[i].se = h.minus ? sa_len - h.s_start : h.s_end;
    k[i].sc = h.score;
  }
  std::vector<uint8_t> drop(n, 0);
  std::vector<uint32_t> ord(n);
  for (int pass = 0; pass < 2; ++pass) {
    for (size_t i = 0; i < n; ++i) ord[i] = (uint32_t)i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
      const K &x = k[a], &y = k[b];
      const uint32_t xa = pass ? x.qe : x.qo, ya = pass ? y.qe : y.qo;
      const uint32_t xb = pass ? x.se : x.so, yb = pass ? y.se : y.so;
      const uint32_t xc = pass ? x.qo : x.qe, yc = pass ? y.qo : y.qe;
      const uint32_t xd = pass ? x.so : x.se, yd = pass ? y.so : y.se;
      if (x.ctx != y.ctx) return x.ctx < y.ctx;
      if (xa != ya) return xa < ya;
      if (xb != yb) return xb < yb;
      if (x.sc != y.sc) return x.sc > y.sc;
      if (xc != yc) return x.ctx ? xc < yc : xc > yc;  // strand-aware tie-breaker
      if (xd != yd) return xd > yd;
      return a < b;
    });
    for (size_t i = 0; i < n;) {
      if (drop[ord[i]]) { ++i; continue; }
      const K& h = k[ord[i]];
      size_t j = i + 1;
      for (; j < n; ++j) {
        const K& g = k[ord[j]];
        const bool same = g.ctx == h.ctx && (pass ? (g.qe == h.qe && g.se == h.se) : (g.qo == h.qo && g.so == h.so));
        if (!same) break;
        drop[ord[j]] = 1;
      }
      i = j;
    }
  }
  size_t w = 0;
  for (size_t i = 0; i < n; ++i)
    if (!drop[i]) v[w++] = v[i];
  v.resize(w);
}

// ------------------------------------------------------------ seed improve ---
void improve_seed(const uint8_t* q, size_t qlen, const uint8_t* s, size_t slen, uint32_t q_seed, uint32_t s_seed,
                  uint32_t pqs, uint32_t pqe, uint32_t pss, uint32_t pse, bool minus, uint32_t& oq, uint32_t& os) {
  constexpr int32_t kRun = 10;
  constexpr int32_t kLong = 15;  // (int)(10 * 1.5)
  oq = q_seed;
  os = s_seed;
  // Is the seed already inside an identity run longer than kRun?
  int32_t run = -1;
  for (size_t qi = q_seed, si = s_seed; qi < qlen && si < slen && (minus || qi < pqe) && q[qi] == s[si]; ++qi, ++si)
    if (++run > kRun) return;
  for (size_t qi = q_seed, si = s_seed;; --qi, --si) {
    if (q[qi] != s[si]) break;
    if (++run > kRun) return;
    if (minus ? (qi <= pqs || si == 0) : (qi == 0 || si == 0)) break;
  }
  if (!minus) {
    const size_t back = std::min<size_t>(sub_sat(q_seed, pqs), sub_sat(s_seed, pss));
    const size_t q0 = q_seed - back, s0 = s_seed - back;
    const size_t len = std::min<size_t>(sub_sat(pse, (uint32_t)s0), sub_sat(pqe, (uint32_t)q0));
    int32_t best = 0, cur = 0;
    size_t best_at = q0;
    bool m = false, prev = false;
    for (size_t i = 0; i < len; ++i) {
      const size_t qi = q0 + i, si = s0 + i;
      if (qi >= qlen || si >= slen) break;
      m = q[qi] == s[si];
      if (m != prev) {
        prev = m;
        if (m) cur = 1;
        else if (cur > best) { best = cur; best_at = qi - (size_t)(cur / 2); }
      } else if (m && ++cur > kLong) {
        oq = (uint32_t)(qi - kLong / 2);
        os = (uint32_t)(s0 + (oq - q0));
        return;
      }
    }
    if (m && cur > best) { best = cur; best_at = q0 + len - (size_t)(cur / 2); }
    if (best > 0) { oq = (uint32_t)best_at; os = (uint32_t)(s0 + (best_at - q0)); }
    return;
  }
  // minus: scan downward in forward-query coordinates from the prelim's top
  const size_t up = std::min<size_t>(sub_sat(sub_sat(pqe, 1), q_seed), sub_sat(sub_sat(pse, 1), s_seed));
  const size_t qt = q_seed + up, st = s_seed + up;
  const size_t len = std::min<size_t>(sub_sat((uint32_t)qt, pqs) + 1, sub_sat((uint32_t)st, pss) + 1);
  int32_t best = 0, cur = 0;
  int64_t best_at = q_seed;
  bool m = false, prev = false;
  for (size_t i = 0; i < len; ++i) {
    const size_t qi = qt - i, si = st - i;
    if (qi >= qlen || si >= slen) continue;
    m = q[qi] == s[si];
    if (m != prev) {
      prev = m;
      if (m) cur = 1;
      else if (cur > best) { best = cur; best_at = (int64_t)qi + cur / 2; }
    } else if (m && ++cur > kLong) {
      const size_t nq = qi + kLong / 2;
      oq = (uint32_t)nq;
      os = (uint32_t)(st - (qt - nq));
      return;
    }
  }
  if (m && cur > best) { best = cur; best_at = ((int64_t)qt - (int64_t)len) + cur / 2; }
  if (best > 0 && best_at >= 0) {
    const int64_t ns = (int64_t)st - ((int64_t)qt - best_at);
    oq = (uint32_t)best_at;
    os = (uint32_t)(ns > 0 ? ns : 0);
  }
}

// ------------------------------------------------------ complexity adjust ---
bool complexity_adjust(int32_t raw, const uint8_t* qw, size_t qw_len, uint32_t q_start, const Script& ops,
                       const Scoring& m, int32_t cutoff, bool minus, int32_t& out) {
  uint32_t counts[16] = {0};
  size_t qpos = q_start, n = 0;
  for (const Run& r : ops) {
    if (r.kind == K_SUB) {
      for (uint32_t i = 0; i < r.len; ++i) {
        const size_t idx = 1 + qpos + i;
        if (idx < qw_len && qw[idx] != 15) {
          const uint8_t b = qw[idx] & 15;
          counts[minus ? kBlastnaComplement[b] : b]++;
          ++n;
        }
      }
      qpos += r.len;
    } else if (r.kind == K_GAP_S) {
      qpos += r.len;
    }
  }
  double adj;
  if (m.lambda <= 0.0 || n == 0) {
    adj = (double)raw;
  } else {
    double tf = 0, ts = 0, tc = 0;
    for (int i = 0; i < 16; ++i) {
      const double f = m.freqs[i];
      if (counts[i] > 0 && f > 0.0 && std::log(f) != 0.0) {
        const double c = (double)counts[i];
        tf += c * std::log(c);
        ts += c * std::log(f);
        tc += c;
      }
    }
    if (tc == 0.0) return false;  // NCBI's 0*log(0) NaN: the hit is dropped
    tf -= tc * std::log(tc);
    ts -= tf;
    adj = (double)raw + ts / m.lambda + 0.999;
    if (adj < 0.0) adj = 0.0;
  }
  const int32_t r = (int32_t)adj;
  if (r < cutoff) return false;
  out = r;
  return true;
}

// ---------------------------------------------------------------- tree ------
HspTree::HspTree(uint32_t q_hi, uint32_t s_hi, bool rc_query) : s_hi_(s_hi), rcq_(rc_query) {
  nodes_.push_back(Node{0, q_hi, 0, 0, 0, kInner});
}
uint32_t HspTree::new_inner(uint32_t lo, uint32_t hi) {
  nodes_.push_back(Node{lo, hi, 0, 0, 0, kInner});
  return (uint32_t)nodes_.size() - 1;
}
static inline uint64_t ep(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

void HspTree::retire(uint32_t idx, bool plus) {
  dead_[idx] = 1;
  const Box& e = boxes_[idx];
  const uint64_t l = plus ? ep(e.qs, e.ss) : ep(e.qe, e.ss), r = plus ? ep(e.qe, e.se) : ep(e.qs, e.se);
  auto a = lend_.find(l);
  if (a != lend_.end() && a->second.second == idx) lend_.erase(a);
  auto b = rend_.find(r);
  if (b != rend_.end() && b->second.second == idx) rend_.erase(b);
}

bool HspTree::add(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus) {
  const uint64_t l = plus ? ep(qs, ss) : ep(qe, ss), r = plus ? ep(qe, se) : ep(qs, se);
  auto wins = [&](const std::pair<int32_t, uint32_t>& ex) {
    const Box& e = boxes_[ex.second];
    if (sc != ex.first) return sc > ex.first;
    if (qe - qs != e.qe - e.qs) return qe - qs < e.qe - e.qs;
    if (se - ss != e.se - e.ss) return se - ss < e.se - e.ss;
    return false;
  };
  auto a = lend_.find(l);
  if (a != lend_.end()) {
    if (!wins(a->second)) return false;
    retire(a->second.second, plus);
  }
  auto b = rend_.find(r);
  if (b != rend_.end()) {
    if (!wins(b->second)) return false;
    retire(b->second.second, plus);
  }
  add_plain(qs, qe, ss, se, sc);
  const uint32_t bi = (uint32_t)boxes_.size() - 1;
  lend_[l] = {sc, bi};
  rend_[r] = {sc, bi};
  return true;
}

void HspTree::add_plain(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc) {
  boxes_.push_back(Box{qs, qe, ss, se, sc});
  dead_.push_back(0);
  nodes_.push_back(Node{0, 0, 0, 0, 0, (uint32_t)boxes_.size() - 1});
  place((uint32_t)nodes_.size() - 1);
}

void HspTree::place(uint32_t leaf) {
  const Box h = boxes_[nodes_[leaf].box];
  uint32_t lo = h.qs, hi = h.qe, at = 0;
  bool on_subject = false;
  for (;;) {
    const uint64_t mid = ((uint64_t)nodes_[at].lo + nodes_[at].hi) / 2;
    uint32_t* slot;
    bool left;
    if ((uint64_t)hi < mid) { slot = &nodes_[at].left; left = true; }
    else if ((uint64_t)lo > mid) { slot = &nodes_[at].right; left = false; }
    else {
      if (on_subject) {  // straddles a subject centre: LIFO list on the node
        nodes_[leaf].mid = nodes_[at].mid;
        nodes_[at].mid = leaf;
        return;
      }
      on_subject = true;
      if (!nodes_[at].mid) {
        const uint32_t t = new_inner(0, s_hi_);
        nodes_[at].mid = t;
      }
      at = nodes_[at].mid;
      lo = h.ss;
      hi = h.se;
      continue;
    }
    const uint32_t child = *slot;
    if (!child) { *slot = leaf; return; }
    if (nodes_[child].box == kInner) { at = child; continue; }
    // leaf collision: split the half into a new inner node holding the old leaf
    const uint32_t nlo = left ? nodes_[at].lo : (uint32_t)mid + 1, nhi = left ? (uint32_t)mid : nodes_[at].hi;
    const uint32_t inner = new_inner(nlo, nhi);
    if (left) nodes_[at].left = inner; else nodes_[at].right = inner;
    const Box& o = boxes_[nodes_[child].box];
    const uint32_t olo = on_subject ? o.ss : o.qs, ohi = on_subject ? o.se : o.qe;
    const uint64_t m2 = ((uint64_t)nlo + nhi) / 2;
    if ((uint64_t)ohi < m2) nodes_[inner].left = child;
    else if ((uint64_t)olo > m2) nodes_[inner].right = child;
    else if (on_subject) nodes_[inner].mid = child;
    else {
      const uint32_t st = new_inner(0, s_hi_);
      nodes_[inner].mid = st;
      const uint64_t m3 = (uint64_t)s_hi_ / 2;
      if ((uint64_t)o.se < m3) nodes_[st].left = child;
      else if ((uint64_t)o.ss > m3) nodes_[st].right = child;
      else nodes_[st].mid = child;
    }
    at = inner;
  }
}

bool HspTree::test(uint32_t idx, uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus) const {
  if (dead_[idx]) return false;
  const Box& h = boxes_[idx];
  if (sc > h.sc) return false;
  if (h.qs > qs || h.qe < qe || h.ss > ss || h.se < se) return false;
  int64_t d0, d1;
  if (plus || rcq_) {
    d0 = ((int64_t)qs - ss) - ((int64_t)h.qs - h.ss);
    d1 = ((int64_t)qe - se) - ((int64_t)h.qe - h.se);
  } else {
    d0 = ((int64_t)qe + ss) - ((int64_t)h.qe + h.ss);
    d1 = ((int64_t)qs + se) - ((int64_t)h.qs + h.se);
  }
  return (d0 < 0 ? -d0 : d0) < 50 || (d1 < 0 ? -d1 : d1) < 50;  // MB_HSP_CLOSE, engine.rs:173
}

bool HspTree::subject_tree_covers(uint32_t at, uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc,
                                  bool plus) const {
  for (;;) {
    if (nodes_[at].box != kInner) return test(nodes_[at].box, qs, qe, ss, se, sc, plus);
    for (uint32_t x = nodes_[at].mid; x; x = nodes_[x].mid)
      if (test(nodes_[x].box, qs, qe, ss, se, sc, plus)) return true;
    const uint64_t mid = ((uint64_t)nodes_[at].lo + nodes_[at].hi) / 2;
    uint32_t next;
    if ((uint64_t)se < mid) next = nodes_[at].left;
    else if ((uint64_t)ss > mid) next = nodes_[at].right;
    else return false;
    if (!next) return false;
    at = next;
  }
}

bool HspTree::covers(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus) const {
  uint32_t at = 0;
  for (;;) {
    if (nodes_[at].mid && subject_tree_covers(nodes_[at].mid, qs, qe, ss, se, sc, plus)) return true;
    const uint64_t mid = ((uint64_t)nodes_[at].lo + nodes_[at].hi) / 2;
    uint32_t next;
    if ((uint64_t)qe < mid) next = nodes_[at].left;
    else if ((uint64_t)qs > mid) next = nodes_[at].right;
    else return false;
    if (!next) return false;
    if (nodes_[next].box != kInner) return test(nodes_[next].box, qs, qe, ss, se, sc, plus);
    at = next;
  }
}

// --------------------------------------------- common-endpoint cut/purge ----
namespace {

struct Piece {  // a final HSP in the purge arrays
  FinalRec h;
};

// Walk ops from one end until `nq` query and `ns` subject bases are consumed.
bool find_cut(const Script& ops, bool from_back, uint32_t nq, uint32_t ns, size_t& at, uint32_t& part, uint32_t& gq,
              uint32_t& gs) {
  gq = gs = 0;
  for (size_t c = 0; c < ops.size(); ++c) {
    const size_t i = from_back ? ops.size() - 1 - c : c;
    const Run& r = ops[i];
    if (r.kind == K_SUB) {
      if (gq + r.len >= nq && gs + r.len >= ns) {  // first unit where both are reached
        const uint32_t need = std::max(nq > gq ? nq - gq : 0, ns > gs ? ns - gs : 0);
        const uint32_t take = need ? need : 1;
        gq += take;
        gs += take;
        at = i;
        part = take;
        return true;
      }
      gq += r.len;
      gs += r.len;
    } else {
      if (r.kind == K_GAP_Q) gs += r.len; else gq += r.len;
      if (gq >= nq && gs >= ns) { at = i; part = r.len; return true; }
    }
  }
  return false;
}

bool cut_start(const FinalRec& p, FinalRec& s) {
  const bool extends = p.minus ? s.q_start < p.q_start : s.q_end > p.q_end;
  if (!extends) return false;
  size_t at; uint32_t part, gq, gs;
  if (!p.minus) {
    const uint32_t nq = sub_sat(p.q_end, s.q_start), ns = sub_sat(p.s_end, s.s_start);
    if (nq == 0 && ns == 0) return !s.ops.empty();
    if (find_cut(s.ops, false, nq, ns, at, part, gq, gs)) {
      const Run r = s.ops[at];
      s.ops.erase(s.ops.begin(), s.ops.begin() + at + 1);
      if (r.len > part) s.ops.insert(s.ops.begin(), Run{r.kind, r.len - part});
      s.q_start += gq;
      s.s_start += gs;
    }
  } else {
    const uint32_t nq = sub_sat(s.q_end, p.q_start), ns = sub_sat(p.s_end, s.s_start);
    if (nq == 0 && ns == 0) return !s.ops.empty();
    if (find_cut(s.ops, true, nq, ns, at, part, gq, gs)) {
      const Run r = s.ops[at];
      s.ops.resize(at);
      if (r.len > part) s.ops.push_back(Run{r.kind, r.len - part});
      s.q_end -= gq;
      s.s_start += gs;
    }
  }
  return !s.ops.empty();
}

bool cut_end(const FinalRec& p, FinalRec& s) {
  const bool extends = p.minus ? s.q_end > p.q_end : s.q_start < p.q_start;
  if (!extends) return false;
  size_t at; uint32_t part, gq, gs;
  if (!p.minus) {
    const uint32_t nq = sub_sat(p.q_start, s.q_start), ns = sub_sat(p.s_start, s.s_start);
    if (nq == 0 && ns == 0) return !s.ops.empty();
    if (find_cut(s.ops, false, nq, ns, at, part, gq, gs)) {
      s.ops.resize(at + 1);
      s.ops[at].len = std::min(s.ops[at].len, part);
      s.q_end = s.q_start + gq;
      s.s_end = s.s_start + gs;
    }
  } else {
    const uint32_t nq = sub_sat(s.q_end, p.q_end), ns = sub_sat(p.s_start, s.s_start);
    if (nq == 0 && ns == 0) return !s.ops.empty();
    if (find_cut(s.ops, true, nq, ns, at, part, gq, gs)) {
      s.ops.erase(s.ops.begin(), s.ops.begin() + at);
      if (part < s.ops[0].len) s.ops[0].len = part;
      s.q_start = sub_sat(s.q_end, gq);
      s.s_end = s.s_start + gs;
    }
  }
  return !s.ops.empty();
}

bool before_by_start(const FinalRec& a, const FinalRec& b) {
  if (a.minus != b.minus) return a.minus < b.minus;
  const uint32_t ka = a.minus ? a.q_end : a.q_start, kb = b.minus ? b.q_end : b.q_start;
  if (ka != kb) return ka < kb;
  if (a.s_start != b.s_start) return a.s_start < b.s_start;
  if (a.score != b.score) return a.score > b.score;
  if (a.minus) { if (a.q_start != b.q_start) return a.q_start < b.q_start; }
  else if (a.q_end != b.q_end) return a.q_end > b.q_end;
  return a.s_end > b.s_end;
}
bool before_by_end(const FinalRec& a, const FinalRec& b) {
  if (a.minus != b.minus) return a.minus < b.minus;
  const uint32_t ka = a.minus ? a.q_start : a.q_end, kb = b.minus ? b.q_start : b.q_end;
  if (ka != kb) return ka < kb;
  if (a.s_end != b.s_end) return a.s_end < b.s_end;
  if (a.score != b.score) return a.score > b.score;
  if (a.minus) { if (a.q_end != b.q_end) return a.q_end < b.q_end; }
  else if (a.q_start != b.q_start) return a.q_start > b.q_start;
  return a.s_start > b.s_start;
}
bool share_start(const FinalRec& a, const FinalRec& b) {
  return a.minus == b.minus && (a.minus ? a.q_end == b.q_end : a.q_start == b.q_start) && a.s_start == b.s_start;
}
bool share_end(const FinalRec& a, const FinalRec& b) {
  return a.minus == b.minus && (a.minus ? a.q_start == b.q_start : a.q_end == b.q_end) && a.s_end == b.s_end;
}

using Slots = std::vector<std::optional<FinalRec>>;

size_t collapse(Slots& arr, size_t count, bool purge, bool (*same)(const FinalRec&, const FinalRec&),
                bool (*cut)(const FinalRec&, FinalRec&)) {
  Slots removed;
  size_t live = 0;
  for (size_t r = 0; r < count; ++r) {
    if (live > 0 && same(*arr[live - 1], *arr[r])) {
      FinalRec sec = std::move(*arr[r]);
      arr[r].reset();
      if (!purge && cut(*arr[live - 1], sec)) removed.emplace_back(std::move(sec));
      else removed.emplace_back(std::nullopt);
    } else {
      if (live != r) std::swap(arr[live], arr[r]);
      ++live;
    }
  }
  for (size_t k = 0; k < removed.size(); ++k) arr[live + k] = std::move(removed[removed.size() - 1 - k]);
  return live;
}

size_t purge_endpoints(Slots& arr, bool purge) {
  size_t n = arr.size();
  if (n < 2) return n;
  std::stable_sort(arr.begin(), arr.begin() + n, [](const auto& a, const auto& b) { return before_by_start(*a, *b); });
  n = collapse(arr, n, purge, share_start, cut_start);
  std::stable_sort(arr.begin(), arr.begin() + n, [](const auto& a, const auto& b) { return before_by_end(*a, *b); });
  return collapse(arr, n, purge, share_end, cut_end);
}

void columns(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen, size_t a0, size_t b0, const Script& ops,
             const std::vector<uint8_t>& mask, std::vector<uint8_t>& qa, std::vector<uint8_t>& sa) {
  qa.clear();
  sa.clear();
  auto sb = [&](size_t i) -> uint8_t {
    if (i < mask.size() && mask[i]) return mask[i];
    return i < blen ? (b[i] & 15) : 14;
  };
  for (const Run& r : ops)
    for (uint32_t k = 0; k < r.len; ++k) {
      if (r.kind != K_GAP_Q) { qa.push_back(a0 < alen ? a[a0] & 15 : 14); ++a0; } else qa.push_back(15);
      if (r.kind != K_GAP_S) { sa.push_back(sb(b0)); ++b0; } else sa.push_back(15);
    }
}

// Blast_HSPReevaluateWithAmbiguitiesGapped (engine.rs:2268-2567); true = delete.
bool reevaluate(FinalRec& h, const std::vector<uint8_t>& qc, const std::vector<uint8_t>& sc, const Scoring& m, int32_t go,
                int32_t ge, int32_t min_score, const uint8_t* rq, size_t rq_len, size_t q0, const uint8_t* rs,
                size_t rs_len, size_t s0) {
  if (h.ops.empty()) return true;
  const bool minus = h.minus;
  const Script fwd = h.ops;
  const size_t L = fwd.size();
  // scan in subject-forward order: reversed for minus
  auto op_at = [&](size_t i) -> const Run& { return minus ? fwd[L - 1 - i] : fwd[i]; };
  const size_t ncol = qc.size();
  auto col = [&](size_t i, bool query) -> uint8_t {
    const size_t j = minus ? ncol - 1 - i : i;
    const std::vector<uint8_t>& v = query ? qc : sc;
    uint8_t x = (j < v.size() ? v[j] : 14) & 0xf;
    return minus ? (kBlastnaComplement[x] & 0xf) : x;
  };
  int32_t sum = 0, score = 0;
  size_t ci = 0;
  uint32_t qp = 0, sp = 0, bqs = 0, bss = 0, bqe = 0, bse = 0, cqs = 0, css = 0;
  size_t bsi = 0, bei = 0, csi = 0;
  uint32_t ben = 0, csk = 0, bsk = 0;
  for (size_t e = 0; e < L; ++e) {
    const Run& r = op_at(e);
    if (r.kind == K_SUB) {
      for (uint32_t k = 0; k < r.len; ++k) {
        sum += m.s[col(ci, true)][col(ci, false)];
        ++ci; ++qp; ++sp;
        const uint32_t used = k + 1;
        if (sum < 0) {
          if (used < r.len) { csi = e; csk = used; } else { csi = e + 1; csk = 0; }
          sum = 0; cqs = qp; css = sp;
          if (score < min_score) { bqs = qp; bss = sp; score = 0; bsi = csi; bsk = csk; bei = csi; ben = 0; }
        } else if (sum > score) {
          score = sum; bqs = cqs; bss = css; bqe = qp; bse = sp; bsi = csi; bsk = csk; bei = e; ben = used;
        }
      }
    } else {
      sum -= go + ge * (int32_t)r.len;
      ci += r.len;
      if (r.kind == K_GAP_Q) sp += r.len; else qp += r.len;
      if (sum < 0) {
        csi = e + 1; csk = 0; sum = 0; cqs = qp; css = sp;
        if (score < min_score) { bqs = qp; bss = sp; score = 0; bsi = csi; bsk = 0; bei = csi; ben = 0; }
      } else if (sum > score) {
        score = sum; bqs = cqs; bss = css; bqe = qp; bse = sp; bsi = csi; bsk = csk; bei = e; ben = r.len;
      }
    }
  }
  if (score < min_score) return true;
  if (minus) {  // back to forward indexing
    uint32_t qt = 0, st = 0;
    for (const Run& r : fwd) {
      if (r.kind != K_GAP_Q) qt += r.len;
      if (r.kind != K_GAP_S) st += r.len;
    }
    const size_t fs = L - 1 - bei, fe = L - 1 - bsi;
    const uint32_t fen = fwd[fe].len - bsk, fsk = fwd[fs].len - ben;
    const uint32_t a = qt - bqe, b = qt - bqs, c = st - bse, d = st - bss;
    bsi = fs; bei = fe; bsk = fsk; ben = fen; bqs = a; bqe = b; bss = c; bse = d;
  }
  Script& ops = h.ops;
  if (bei < (L ? L - 1 : 0) || ben < fwd[bei].len) {
    ops.resize(bei + 1);
    if (ben > 0) ops[bei].len = ben; else ops.resize(bei);
  }
  if (bsi > 0) ops.erase(ops.begin(), ops.begin() + std::min(bsi, ops.size()));
  if (bsk > 0 && !ops.empty()) {
    if (bsk < ops[0].len) ops[0].len -= bsk; else ops.erase(ops.begin());
  }
  int64_t eqs = bqs, ess = bss, eqe = bqe, ese = bse;
  for (;;) {  // exact-match extension on the raw sequences (score unchanged)
    const int64_t qi = (int64_t)q0 + eqs, si = (int64_t)s0 + ess;
    if (qi <= 0 || si <= 0) break;
    const uint8_t a = rq[qi - 1], b = rs[si - 1];
    if (!(a < 4 && a == b)) break;
    --eqs; --ess;
    if (!ops.empty() && ops.front().kind == K_SUB) ops.front().len++;
    else ops.insert(ops.begin(), Run{K_SUB, 1});
  }
  for (;;) {
    const int64_t qi = (int64_t)q0 + eqe, si = (int64_t)s0 + ese;
    if (qi >= (int64_t)rq_len || si >= (int64_t)rs_len) break;
    const uint8_t a = rq[qi], b = rs[si];
    if (!(a < 4 && a == b)) break;
    ++eqe; ++ese;
    if (!ops.empty() && ops.back().kind == K_SUB) ops.back().len++;
    else ops.push_back(Run{K_SUB, 1});
  }
  h.score = score;
  auto nn = [](int64_t v) { return (uint32_t)(v > 0 ? v : 0); };
  if (minus) {
    const int64_t oq = h.q_start, ose = h.s_end;
    h.q_start = nn(oq + eqs);
    h.q_end = nn(oq + eqe);
    h.s_end = nn(ose - ess);
    h.s_start = nn(ose - ese);
  } else {
    const uint32_t nq = nn((int64_t)h.q_start + eqs), ns = nn((int64_t)h.s_start + ess);
    h.q_start = nq;
    h.q_end = nq + (uint32_t)(eqe - eqs);
    h.s_start = ns;
    h.s_end = ns + (uint32_t)(ese - ess);
  }
  return false;
}

}  // namespace

void finish_subject(std::vector<FinalRec>& hsps, const uint8_t* qw, uint32_t qa_len, const uint8_t* sf,
                    const uint8_t* sr, uint32_t sa_len, const std::vector<uint8_t>& n_mask,
                    const std::vector<uint8_t>& n_mask_rc, const Scoring& m, int32_t go, int32_t ge,
                    int32_t cutoff) {
  Slots arr;
  arr.reserve(hsps.size());
  for (auto& h : hsps) arr.emplace_back(std::move(h));
  hsps.clear();
  const size_t extra = purge_endpoints(arr, false);
  std::vector<uint8_t> qc, sc;
  for (size_t i = extra; i < arr.size(); ++i) {
    if (!arr[i]) continue;
    FinalRec& h = *arr[i];
    const bool minus = h.minus;
    const uint8_t* b = minus ? sr + 1 : sf + 1;
    const std::vector<uint8_t>& mk = minus ? n_mask_rc : n_mask;
    const size_t b0 = minus ? sa_len - h.s_end : h.s_start;
    columns(qw + 1, qa_len, b, sa_len, h.q_start, b0, h.ops, mk, qc, sc);
    if (reevaluate(h, qc, sc, m, go, ge, cutoff, qw + 1, qa_len, h.q_start, b, sa_len, b0)) arr[i].reset();
  }
  Slots live;
  for (auto& s : arr)
    if (s) live.emplace_back(std::move(s));
  purge_endpoints(live, true);
  std::vector<FinalRec> v;
  for (auto& s : live)
    if (s) v.push_back(std::move(*s));
  // final containment pass, ScoreCompareHSPs order (engine.rs:1731-1759)
  std::stable_sort(v.begin(), v.end(), [&](const FinalRec& a, const FinalRec& b) {
    if (a.score != b.score) return a.score > b.score;
    if (a.s_start != b.s_start) return a.s_start < b.s_start;
    if (a.s_end != b.s_end) return a.s_end > b.s_end;
    const uint32_t qa = a.minus ? qa_len - a.q_end : a.q_start, qb = b.minus ? qa_len - b.q_end : b.q_start;
    if (qa != qb) return qa < qb;
    const uint32_t ea = a.minus ? qa_len - a.q_start : a.q_end, eb = b.minus ? qa_len - b.q_start : b.q_end;
    return ea > eb;
  });
  HspTree tp(qa_len + 1, sa_len + 1), tm(qa_len + 1, sa_len + 1);
  for (auto& h : v) {
    HspTree& t = h.minus ? tm : tp;
    if (t.covers(h.q_start, h.q_end, h.s_start, h.s_end, h.score, !h.minus)) continue;
    t.add_plain(h.q_start, h.q_end, h.s_start, h.s_end, h.score);
    hsps.push_back(std::move(h));
  }
}

}  // namespace rmb

namespace rmb {

QuerySplit query_split(uint32_t n) {
  constexpr uint32_t kChunk = 1000000, kOverlap = 100;
  const uint32_t count = n / (kChunk - kOverlap);
  if (count <= 1) return QuerySplit{n, 1, n > kOverlap ? n - kOverlap : 0};
  uint64_t size = ((uint64_t)n + (uint64_t)(count - 1) * kOverlap) / count;
  if (count + kOverlap < size) ++size;  // NCBI: num_chunks < chunk_size - overlap
  return QuerySplit{(uint32_t)size, count, (uint32_t)size - kOverlap};
}

void merge_chunk_prelims(std::vector<PreRec>& acc, std::vector<PreRec>&& later, uint32_t split) {
  // Only plus-strand pieces that reach into the 100-base overlap take part:
  // accumulated ones ending past the split, later ones starting before its end.
  // Each accumulated piece, best first, absorbs every later piece on a
  // diagonal within 10 of its end whose start or end lies in its box; the box
  // grows as it absorbs, and the score becomes the larger of the better piece
  // and the pair's score density over the grown query span.
  auto ranks_before = [](const PreRec& a, const PreRec& b) {
    return std::make_tuple(-(int64_t)a.score, a.s_start, ~a.s_end, a.q_start, ~a.q_end) <
           std::make_tuple(-(int64_t)b.score, b.s_start, ~b.s_end, b.q_start, ~b.q_end);
  };
  std::vector<uint32_t> lhs, rhs;
  for (uint32_t i = 0; i < acc.size(); ++i)
    if (!acc[i].minus && acc[i].q_end > split) lhs.push_back(i);
  for (uint32_t j = 0; j < later.size(); ++j)
    if (!later[j].minus && later[j].q_start < split + 100u) rhs.push_back(j);
  std::stable_sort(lhs.begin(), lhs.end(), [&](uint32_t x, uint32_t y) { return ranks_before(acc[x], acc[y]); });
  std::stable_sort(rhs.begin(), rhs.end(), [&](uint32_t x, uint32_t y) { return ranks_before(later[x], later[y]); });
  std::vector<uint8_t> gone(later.size(), 0);
  auto inside = [](const PreRec& box, uint32_t q, uint32_t s) {
    return q >= box.q_start && q <= box.q_end && s >= box.s_start && s <= box.s_end;
  };
  for (uint32_t i : lhs) {
    PreRec& a = acc[i];
    for (uint32_t j : rhs) {
      if (gone[j]) continue;
      const PreRec& b = later[j];
      const int64_t d = ((int64_t)a.q_end - a.s_end) - ((int64_t)b.q_start - b.s_start);
      if (d <= -10 || d >= 10) continue;
      if (!inside(a, b.q_start, b.s_start) && !inside(a, b.q_end, b.s_end)) continue;
      const double density =
          (double)(a.score + b.score) / (double)((a.q_end - a.q_start) + (b.q_end - b.q_start));
      a.q_start = std::min(a.q_start, b.q_start);
      a.q_end = std::max(a.q_end, b.q_end);
      a.s_start = std::min(a.s_start, b.s_start);
      a.s_end = std::max(a.s_end, b.s_end);
      if (b.score > a.score) {
        a.score = b.score;
        a.q_seed = b.q_seed;
        a.s_seed = b.s_seed;
      }
      const int32_t spread = (int32_t)(density * (double)(a.q_end - a.q_start));
      if (spread > a.score) a.score = spread;
      gone[j] = 1;
    }
  }
  for (uint32_t j = 0; j < later.size(); ++j)
    if (!gone[j]) acc.push_back(later[j]);
}

}  // namespace rmb
