// CPU ORACLE (test infrastructure): extension and alignment kernels.
// Restates rmblast-lib/src/search/ungapped.rs:38-121, diag_hash.rs:30-111,
// gapped.rs:124-820, engine.rs:2953-3039 (bidirectional traceback alignment),
// complexity_adjust.rs:19-132 and stats.rs:52-251.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace orc {

static const int32_t MININT = INT32_MIN / 2;  // gapped.rs:47

// ------------------------------------------------------------ ungapped -------
Ungapped extend_ungapped(const uint8_t* query, size_t qn, const uint8_t* subject, size_t sn, uint32_t q_off,
                         uint32_t s_off, const Matrix& m, int xdrop) {
  const int32_t x = -xdrop;
  const int64_t q_real_len = (int64_t)qn - 1, s_real_len = (int64_t)sn - 1;  // ungapped.rs:235-236
  const int64_t qo = q_off, so = s_off;
  int32_t score = 0, sum = 0;
  int64_t q_beg = qo;
  {  // left, fixed X (ungapped.rs:248-267)
    int64_t q = qo - 1, s = so - 1;
    for (;;) {
      if (q < 0 || s < 0) break;
      uint8_t qb = query[1 + q], sb = subject[1 + s];
      if (qb == 15 || sb == 15) break;
      sum += m.score(qb, sb);
      if (sum > 0) {
        q_beg = q;
        score += sum;
        sum = 0;
      } else if (sum < x) {
        break;
      }
      --q;
      --s;
    }
  }
  Ungapped r;
  r.q_start = (uint32_t)q_beg;
  r.s_start = (uint32_t)(so - (qo - q_beg));
  int64_t q_end = qo;
  int32_t x_cur = x;
  sum = 0;
  {  // right, adaptive X (ungapped.rs:277-297)
    int64_t q = qo, s = so;
    for (;;) {
      if (q >= q_real_len || s >= s_real_len) break;
      uint8_t qb = query[1 + q], sb = subject[1 + s];
      if (qb == 15 || sb == 15) break;
      sum += m.score(qb, sb);
      if (sum > 0) {
        q_end = q + 1;
        score += sum;
        x_cur = (-score > x) ? -score : x;
        sum = 0;
      } else if (sum < x_cur) {
        break;
      }
      ++q;
      ++s;
    }
  }
  r.q_end = (uint32_t)q_end;
  r.s_end = (uint32_t)(so + (q_end - qo));
  r.score = score;
  return r;
}

// ----------------------------------------------------------- diag hash -------
DiagHash::DiagHash() : backbone(512, 0), chain(257, Cell{0, 0, 0}) {}
void DiagHash::clear() {
  std::fill(backbone.begin(), backbone.end(), 0u);
  occupancy = 1;
}
uint32_t DiagHash::get(int32_t diag) const {  // diag_hash.rs:57-67
  uint32_t idx = backbone[bucket(diag)];
  while (idx != 0) {
    const Cell& c = chain[idx];
    if (c.diag == diag) return c.level;
    idx = c.next;
  }
  return 0;
}
void DiagHash::insert(int32_t diag, uint32_t s_end, uint32_t s_off, int32_t window) {  // diag_hash.rs:78-111
  uint32_t b = bucket(diag);
  uint32_t idx = backbone[b];
  while (idx != 0) {
    Cell& c = chain[idx];
    if (c.diag == diag) { c.level = s_end; return; }
    if ((int32_t)(s_off - c.level) > window) {  // wrapping sub, then signed compare
      c.diag = diag;
      c.level = s_end;
      return;
    }
    idx = c.next;
  }
  uint32_t ni = occupancy;
  if (ni >= chain.size()) chain.push_back(Cell{0, 0, 0});
  chain[ni] = Cell{diag, s_end, backbone[b]};
  backbone[b] = ni;
  ++occupancy;
}

// ----------------------------------------------------------- edit ops --------
void push_op(EditOps& e, Op op, uint32_t n) {  // hits.rs:43-54
  if (n == 0) return;
  if (!e.empty() && e.back().first == op) { e.back().second += n; return; }
  e.emplace_back(op, n);
}

// --------------------------------------------------- score-only X-drop -------
struct Cell2 { int32_t best, best_gap; };

void align_score_only(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int gap_open, int gap_extend,
                      int x_dropoff, const Matrix& mat, bool reverse, int32_t& out_score, size_t& out_a,
                      size_t& out_b, DpCounters* ctr) {
  // gapped.rs:124-337.  dp is indexed by absolute b column; every cell read in
  // a call has been written earlier in the same call.
  if (m == 0 || n == 0) { out_score = 0; out_a = 0; out_b = 0; return; }
  thread_local std::vector<Cell2> dp;
  if (dp.size() < n + 2) dp.resize(n + 2);
  const int32_t oe = gap_open + gap_extend;
  const int32_t xdrop = std::max(x_dropoff, oe);
  dp[0].best = 0;
  dp[0].best_gap = -oe;
  int32_t score = -oe;
  size_t b_size = 1;
  for (size_t j = 1; j <= n; ++j) {
    if (score < -xdrop) break;
    dp[j].best = score;
    dp[j].best_gap = score - oe;
    score -= gap_extend;
    b_size = j + 1;
  }
  int32_t best_score = 0;
  size_t best_a = 0, best_b = 0, first_b = 0;
  uint64_t cells = 0;
  for (size_t ai = 0; ai < m; ++ai) {
    const size_t ap = reverse ? m - 1 - ai : ai;
    const int32_t* row = mat.s[a[ap] & 15];
    int32_t sc = MININT, sgr = MININT;
    size_t last_b = first_b;
    const int32_t prev_best = best_score;
    cells += b_size - first_b;
    for (size_t bi = first_b; bi < b_size; ++bi) {
      int32_t next_sc = 0;
      int32_t sgc = dp[bi].best_gap;
      if (bi < n) {
        const size_t bp = reverse ? n - 1 - bi : bi;
        next_sc = dp[bi].best + row[b[bp]];
      }
      int32_t c = sc;
      if (sgc > c) c = sgc;
      if (sgr > c) c = sgr;
      if (best_score - c > xdrop) {
        if (bi == first_b) ++first_b;
        else dp[bi].best = MININT;
      } else {
        last_b = bi;
        if (c > best_score) { best_b = bi; best_score = c; }
        sgr -= gap_extend;
        const int32_t open_gap = c - oe;
        const int32_t sgc_ext = sgc - gap_extend;
        dp[bi].best_gap = sgc_ext > open_gap ? sgc_ext : open_gap;
        dp[bi].best = c;
        sgr = std::max(sgr, open_gap);
      }
      sc = next_sc;
    }
    if (best_score > prev_best) best_a = ai + 1;
    if (first_b >= b_size) break;
    if (last_b + 1 < b_size) {
      b_size = last_b + 1;
    } else {
      while (sgr >= best_score - xdrop && b_size <= n) {
        dp[b_size].best = sgr;
        dp[b_size].best_gap = sgr - oe;
        sgr -= gap_extend;
        ++b_size;
      }
    }
    if (b_size <= n) {
      dp[b_size].best = MININT;
      dp[b_size].best_gap = MININT;
      ++b_size;
    }
  }
  if (ctr) ctr->score_only_cells += cells;
  out_score = best_score;
  out_a = best_a;
  out_b = best_b;
}

bool gapped_score_only(const std::vector<uint8_t>& query, const std::vector<uint8_t>& subject, uint32_t q_seed,
                       uint32_t s_seed, int go, int ge, int xdrop, const Matrix& mat, int32_t& score, uint32_t& q0,
                       uint32_t& q1, uint32_t& s0, uint32_t& s1, DpCounters* ctr) {
  // gapped.rs:360-418: subject is the outer (a) sequence; the pivot starts the right half.
  const uint8_t* qa = query.data() + 1;
  const uint8_t* sa = subject.data() + 1;
  const size_t qlen = query.size() - 2, slen = subject.size() - 2;
  int32_t ls;
  size_t la, lb;
  align_score_only(sa, qa, s_seed, q_seed, go, ge, xdrop, mat, true, ls, la, lb, ctr);
  s0 = s_seed - (uint32_t)la;
  q0 = q_seed - (uint32_t)lb;
  size_t rm = slen > s_seed ? slen - s_seed : 0;
  size_t rn = qlen > q_seed ? qlen - q_seed : 0;
  int32_t rs = 0;
  size_t ra = 0, rb = 0;
  if (rm > 0 && rn > 0) align_score_only(sa + s_seed, qa + q_seed, rm, rn, go, ge, xdrop, mat, false, rs, ra, rb, ctr);
  score = ls + rs;
  if (score < 0) return false;
  s1 = s_seed + (uint32_t)ra;
  q1 = q_seed + (uint32_t)rb;
  return true;
}

// ------------------------------------------------ ALIGN_EX + traceback -------
static const uint8_t SCRIPT_GAP_IN_A = 0, SCRIPT_SUB = 3, SCRIPT_GAP_IN_B = 6, SCRIPT_OP_MASK = 7,
                     SCRIPT_EXTEND_GAP_A = 0x10, SCRIPT_EXTEND_GAP_B = 0x40;

static EditOps traceback_flat(const std::vector<uint8_t>& flat, const std::vector<std::pair<size_t, size_t>>& rows,
                              size_t end_a, size_t end_b) {
  // gapped.rs:692-757
  EditOps rev;
  if (end_a == 0 && end_b == 0) return rev;
  size_t a = end_a, b = end_b;
  uint8_t prev = SCRIPT_SUB;
  for (size_t step = 0; step < end_a + end_b; ++step) {
    size_t off = rows[a].first, fb = rows[a].second;
    size_t row_end = a + 1 < rows.size() ? rows[a + 1].first : flat.size();
    uint8_t raw;
    if (b >= fb) {
      size_t idx = off + (b - fb);
      raw = idx < row_end ? flat[idx] : SCRIPT_GAP_IN_A;
    } else {
      raw = SCRIPT_GAP_IN_A;
    }
    uint8_t op;
    switch (prev & SCRIPT_OP_MASK) {
      case SCRIPT_GAP_IN_A: op = (raw & SCRIPT_EXTEND_GAP_A) ? SCRIPT_GAP_IN_A : (raw & SCRIPT_OP_MASK); break;
      case SCRIPT_GAP_IN_B: op = (raw & SCRIPT_EXTEND_GAP_B) ? SCRIPT_GAP_IN_B : (raw & SCRIPT_OP_MASK); break;
      default: op = raw & SCRIPT_OP_MASK;
    }
    prev = op;
    if (op == SCRIPT_GAP_IN_A) { push_op(rev, OP_GAP_IN_QUERY, 1); if (b) --b; }
    else if (op == SCRIPT_GAP_IN_B) { push_op(rev, OP_GAP_IN_SUBJECT, 1); if (a) --a; }
    else { push_op(rev, OP_SUB, 1); if (a) --a; if (b) --b; }
    if (a == 0 && b == 0) break;
  }
  if (a != 0 || b != 0) { std::abort(); }
  std::reverse(rev.begin(), rev.end());
  EditOps out;
  for (auto& o : rev) push_op(out, o.first, o.second);
  return out;
}

GapResult align_ex(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int gap_open, int gap_extend,
                   int x_dropoff, const Matrix& mat, bool reverse, DpCounters* ctr) {
  // gapped.rs:450-689
  GapResult res{0, 0, 0, {}};
  if (m == 0 || n == 0) return res;
  thread_local std::vector<Cell2> dp;
  thread_local std::vector<uint8_t> flat;
  thread_local std::vector<std::pair<size_t, size_t>> rows;
  if (dp.size() < n + 2) dp.resize(n + 2);
  flat.clear();
  rows.clear();
  const int32_t oe = gap_open + gap_extend;
  const int32_t xdrop = std::max(x_dropoff, oe);
  dp[0].best = 0;
  dp[0].best_gap = -oe;
  rows.emplace_back(0, 0);
  flat.push_back(SCRIPT_GAP_IN_B);
  int32_t score = -oe;
  size_t b_size = 1;
  for (size_t j = 1; j <= n; ++j) {
    if (score < -xdrop) break;
    dp[j].best = score;
    dp[j].best_gap = score - oe;
    flat.push_back(SCRIPT_GAP_IN_A);
    score -= gap_extend;
    b_size = j + 1;
  }
  int32_t best_score = 0;
  size_t best_a = 0, best_b = 0, first_b = 0;
  uint64_t cells = 0;
  for (size_t ai = 0; ai < m; ++ai) {
    const size_t ap = reverse ? m - 1 - ai : ai;
    const int32_t* row = mat.s[a[ap] & 15];
    rows.emplace_back(flat.size(), first_b);
    int32_t sc = MININT, sgr = MININT;
    size_t last_b = first_b;
    cells += b_size - first_b;
    for (size_t bi = first_b; bi < b_size; ++bi) {
      uint8_t bb = 0;
      if (bi < n) bb = b[reverse ? n - 1 - bi : bi] & 15;
      int32_t sgc = dp[bi].best_gap;
      const int32_t next_sc = dp[bi].best + row[bb];
      uint8_t op = SCRIPT_SUB;
      int32_t c = sc;
      if (sgc > c) { c = sgc; op = SCRIPT_GAP_IN_B; }
      if (sgr > c) { c = sgr; op = SCRIPT_GAP_IN_A; }
      if (best_score - c > xdrop) {
        if (bi == first_b) ++first_b;
        else dp[bi].best = MININT;
      } else {
        last_b = bi;
        if (c > best_score) { best_score = c; best_a = ai + 1; best_b = bi; }
        sgr -= gap_extend;
        sgc -= gap_extend;
        const int32_t open_gap = c - oe;
        if (sgc < open_gap) dp[bi].best_gap = open_gap;
        else { dp[bi].best_gap = sgc; op |= SCRIPT_EXTEND_GAP_B; }
        dp[bi].best = c;
        if (sgr < open_gap) sgr = open_gap;
        else op |= SCRIPT_EXTEND_GAP_A;
      }
      sc = next_sc;
      flat.push_back(op);
    }
    if (first_b >= b_size) break;
    if (last_b + 1 < b_size) {
      b_size = last_b + 1;
    } else {
      while (sgr >= best_score - xdrop && b_size <= n) {
        dp[b_size].best = sgr;
        dp[b_size].best_gap = sgr - oe;
        sgr -= gap_extend;
        flat.push_back(SCRIPT_GAP_IN_A);
        ++b_size;
      }
    }
    if (b_size <= n) {
      dp[b_size].best = MININT;
      dp[b_size].best_gap = MININT;
      ++b_size;
    }
  }
  if (ctr) ctr->traceback_cells += cells;
  res.score = best_score;
  res.a_len = best_a;
  res.b_len = best_b;
  res.ops = traceback_flat(flat, rows, best_a, best_b);
  return res;
}

bool gapped_extend_bidirectional(const std::vector<uint8_t>& query, const std::vector<uint8_t>& subject,
                                 uint32_t q_seed, uint32_t s_seed, int go, int ge, int xdrop, const Matrix& mat,
                                 int32_t& score, uint32_t& q0, uint32_t& q1, uint32_t& s0, uint32_t& s1,
                                 EditOps& ops, DpCounters* ctr) {
  // engine.rs:2953-3039: left half includes the seed, right half excludes it.
  const uint8_t* qa = query.data() + 1;
  const uint8_t* sa = subject.data() + 1;
  const size_t qlen = query.size() - 2, slen = subject.size() - 2;
  GapResult left = align_ex(qa, sa, (size_t)q_seed + 1, (size_t)s_seed + 1, go, ge, xdrop, mat, true, ctr);
  q0 = q_seed + 1 - (uint32_t)left.a_len;
  s0 = s_seed + 1 - (uint32_t)left.b_len;
  size_t rq = (size_t)q_seed + 1, rs = (size_t)s_seed + 1;
  size_t rm = qlen > rq ? qlen - rq : 0, rn = slen > rs ? slen - rs : 0;
  GapResult right{0, 0, 0, {}};
  if (rm > 0 && rn > 0) right = align_ex(qa + rq, sa + rs, rm, rn, go, ge, xdrop, mat, false, ctr);
  int32_t total = left.score + right.score;
  q1 = q_seed + 1 + (uint32_t)right.a_len;
  s1 = s_seed + 1 + (uint32_t)right.b_len;
  ops.clear();
  for (auto it = left.ops.rbegin(); it != left.ops.rend(); ++it) push_op(ops, it->first, it->second);
  for (auto& o : right.ops) push_op(ops, o.first, o.second);
  while (!ops.empty() && ops.front().first != OP_SUB) {  // engine.rs:3016-3024
    auto [op, cnt] = ops.front();
    total += go + (int32_t)cnt * ge;
    if (op == OP_GAP_IN_QUERY) s0 += cnt; else q0 += cnt;
    ops.erase(ops.begin());
  }
  while (!ops.empty() && ops.back().first != OP_SUB) {  // engine.rs:3025-3033
    auto [op, cnt] = ops.back();
    total += go + (int32_t)cnt * ge;
    if (op == OP_GAP_IN_QUERY) s1 -= cnt; else q1 -= cnt;
    ops.pop_back();
  }
  if (total < 0) return false;
  score = total;
  return true;
}

void extract_aligned(const uint8_t* a, size_t alen, const uint8_t* b, size_t blen, size_t a_start, size_t b_start,
                     const EditOps& ops, const std::vector<uint8_t>& n_mask, std::vector<uint8_t>& qa,
                     std::vector<uint8_t>& sa) {
  // gapped.rs:773-820
  auto sb = [&](size_t si) -> uint8_t {
    uint8_t v = si < blen ? (b[si] & 15) : 14;
    if (si < n_mask.size() && n_mask[si] != 0) return n_mask[si];
    return v;
  };
  qa.clear();
  sa.clear();
  size_t qi = a_start, si = b_start;
  for (auto& [op, cnt] : ops) {
    for (uint32_t k = 0; k < cnt; ++k) {
      if (op == OP_SUB) { qa.push_back(qi < alen ? a[qi] & 15 : 14); sa.push_back(sb(si)); ++qi; ++si; }
      else if (op == OP_GAP_IN_SUBJECT) { qa.push_back(qi < alen ? a[qi] & 15 : 14); sa.push_back(15); ++qi; }
      else { qa.push_back(15); sa.push_back(sb(si)); ++si; }
    }
  }
}

// ---------------------------------------------------- complexity adjust ------
double compute_adjusted_score(int32_t raw, const std::vector<uint8_t>& bases, const Matrix& m) {
  // complexity_adjust.rs:57-102
  if (m.lambda <= 0.0 || bases.empty()) return (double)raw;
  uint32_t counts[16] = {0};
  for (uint8_t b : bases) counts[b & 15]++;
  uint32_t total = 0;
  for (int i = 0; i < 16; ++i) total += counts[i];
  if (total == 0) return (double)raw;
  double t_factor = 0, t_sum = 0, t_counts = 0;
  for (int i = 0; i < 16; ++i) {
    uint32_t c = counts[i];
    double f = m.freqs[i];
    if (c > 0 && f > 0.0 && std::log(f) != 0.0) {
      double cf = (double)c;
      t_factor += cf * std::log(cf);
      t_sum += cf * std::log(f);
      t_counts += cf;
    }
  }
  if (t_counts == 0.0) return NAN;
  t_factor -= t_counts * std::log(t_counts);
  t_sum -= t_factor;
  double adj = (double)raw + t_sum / m.lambda + 0.999;
  return adj < 0.0 ? 0.0 : adj;
}

bool apply_complexity_adjust(int32_t raw, const std::vector<uint8_t>& query, uint32_t q_start, const EditOps& ops,
                             const Matrix& m, int cutoff, bool minus, int32_t& out) {
  // complexity_adjust.rs:19-50, 111-132
  std::vector<uint8_t> bases;
  size_t q = q_start;
  for (auto& [op, cnt] : ops) {
    if (op == OP_SUB) {
      for (uint32_t i = 0; i < cnt; ++i) {
        size_t idx = 1 + q + i;
        if (idx < query.size() && query[idx] != 15) bases.push_back(query[idx]);
      }
      q += cnt;
    } else if (op == OP_GAP_IN_SUBJECT) {
      q += cnt;
    }
  }
  if (minus)
    for (auto& b : bases) b = kComplement[b & 15];
  double adj = compute_adjusted_score(raw, bases, m);
  if (std::isnan(adj)) return false;
  int32_t r = (int32_t)adj;
  if (r < cutoff) return false;
  out = r;
  return true;
}

// ------------------------------------------------------------- statistics ----
std::string to_iupac(const std::vector<uint8_t>& v) {  // stats.rs:247-251
  std::string s(v.size(), '-');
  for (size_t i = 0; i < v.size(); ++i) s[i] = v[i] == 15 ? '-' : kBlastnaToIupac[v[i] & 15];
  return s;
}

enum PairType { P_ID, P_TRANSI, P_TRANSV, P_IUBMATCH, P_IUBSUB, P_GAP, P_UNKNOWN };
static bool unambig(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }
static uint8_t iub_mask(char c) {
  switch (c) {
    case 'A': return 1; case 'C': return 2; case 'G': return 4; case 'T': case 'U': return 8;
    case 'R': return 5; case 'Y': return 10; case 'M': return 3; case 'K': return 12; case 'W': return 9;
    case 'S': return 6; case 'B': return 14; case 'D': return 13; case 'H': return 11; case 'V': return 7;
    default: return 15;
  }
}
static PairType classify(char a, char b) {  // stats.rs:162-191
  if (a == '-' || b == '-') return P_GAP;
  if (a == b) return P_ID;
  if (!unambig(a) || !unambig(b)) {
    bool one_pure = unambig(a) ^ unambig(b);
    return (one_pure && (iub_mask(a) & iub_mask(b))) ? P_IUBMATCH : P_IUBSUB;
  }
  bool ap = a == 'A' || a == 'G', bp = b == 'A' || b == 'G';
  return ap == bp ? P_TRANSI : P_TRANSV;
}
static double kimura(double p, double q) {  // stats.rs:236-243
  double op = (1.0 - 2.0 * p - q) * std::sqrt(1.0 - 2.0 * q);
  if (op > 0.0) return std::fabs(-0.5 * std::log(op)) * 100.0;
  return 100.0;
}

AlignStats compute_align_stats(const std::string& qs, const std::string& ss) {
  // stats.rs:52-160 with query_is_ancestral = false: ancestral = subject.
  const std::string& anc = ss;
  const std::string& der = qs;
  AlignStats st;
  uint32_t unambig_pairs = 0, q_len = 0, d_len = 0;
  double cpg_transi = 0.0;
  char prev_anc = ' ';
  PairType prev_pair = P_UNKNOWN;
  for (size_t i = 0; i < anc.size(); ++i) {
    char ac = (char)std::toupper((unsigned char)anc[i]);
    char dc = (char)std::toupper((unsigned char)der[i]);
    if (qs[i] != '-') ++q_len;
    if (ss[i] != '-') ++d_len;
    if (prev_anc == 'C' && ac == 'G') {
      st.cpg_sites++;
      PairType cur = classify(ac, dc);
      if (prev_pair == P_TRANSI && cur == P_TRANSI) cpg_transi -= 1.0;
      else if (prev_pair == P_TRANSI || cur == P_TRANSI) cpg_transi -= 0.9;
    }
    if (ac == '-' || dc == '-') {
      if (qs[i] == '-') st.query_gaps++; else st.subject_gaps++;
      if (ac != '-') { prev_anc = ac; prev_pair = P_GAP; }
    } else {
      PairType pt = classify(ac, dc);
      switch (pt) {
        case P_ID: st.matches++; if (unambig(ac) && unambig(dc)) unambig_pairs++; break;
        case P_TRANSI:
          st.mismatches++;
          if (unambig(ac) && unambig(dc)) { unambig_pairs++; st.transitions++; cpg_transi += 1.0; }
          break;
        case P_TRANSV:
          st.mismatches++;
          if (unambig(ac) && unambig(dc)) { unambig_pairs++; st.transversions++; }
          break;
        case P_IUBSUB: st.mismatches++; break;
        default: break;
      }
      prev_anc = ac;
      prev_pair = pt;
    }
  }
  if (unambig_pairs < 1) {
    st.kdiv = 100.0;
    st.cpg_kdiv = 100.0;
  } else {
    double n = (double)unambig_pairs;
    double p = (double)st.transitions / n, cp = cpg_transi / n, q = (double)st.transversions / n;
    st.kdiv = kimura(p, q);
    st.cpg_kdiv = kimura(cp, q);
  }
  volatile float fq = (float)q_len, fd = (float)d_len;
  st.perc_sub = q_len > 0 ? ((float)st.mismatches / fq) * 100.0f : 0.0f;
  st.perc_query_gap = q_len > 0 ? ((float)st.query_gaps / fq) * 100.0f : 0.0f;
  st.perc_db_gap = d_len > 0 ? ((float)st.subject_gaps / fd) * 100.0f : 0.0f;
  return st;
}

}  // namespace orc
