// Host tail of librmblast_hip: mask level, output order, alignment statistics
// and the outfmt 6 writer, per query on the host threads.
#include "rmb_report.h"

#include "rmb_ka.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <thread>

#include <sys/mman.h>

namespace rmb {
namespace {

const char kIupac[] = "ACGTRYMKWSBDHVN-";

template <class F>
void for_each_query(uint32_t n, unsigned threads, F&& fn) {
  threads = std::max(1u, std::min<unsigned>(threads, n));
  if (threads == 1) {
    for (uint32_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (uint32_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}

// Query offset and end in NCBI's strand frame (hits.rs:104-112): a minus HSP
// counts from the query's far end.
inline uint32_t q_off(const rmb_hsp& h, uint32_t q_len) { return h.strand ? q_len - h.q_end : h.q_start; }
inline uint32_t q_fin(const rmb_hsp& h, uint32_t q_len) { return h.strand ? q_len - h.q_start : h.q_end; }

// ScoreCompareHSPs (hits.rs:104-112): < 0 when a comes first
int score_compare(const rmb_hsp& a, const rmb_hsp& b, uint32_t q_len) {
  if (a.score != b.score) return a.score > b.score ? -1 : 1;
  if (a.s_start != b.s_start) return a.s_start < b.s_start ? -1 : 1;
  if (a.s_end != b.s_end) return a.s_end > b.s_end ? -1 : 1;
  const uint32_t oa = q_off(a, q_len), ob = q_off(b, q_len);
  if (oa != ob) return oa < ob ? -1 : 1;
  const uint32_t ea = q_fin(a, q_len), eb = q_fin(b, q_len);
  if (ea != eb) return ea > eb ? -1 : 1;
  return 0;
}

// Accepted query ranges of the mask-level sweep: one ordered map per span
// class (spans in [2^(k-1), 2^k)), each searched back only as far as the
// longest span it holds, so a check costs O(classes * log n) instead of a
// scan over every range within the longest span (the reference's SpanIndex,
// engine.rs:211-246, has the same answer).
class SpanIndex {
 public:
  void insert(uint32_t s, uint32_t e, int32_t score) {
    const uint32_t span = e - s;
    Class& c = cls_[32 - __builtin_clz(span)];
    c.longest = std::max(c.longest, span);
    c.ranges.emplace(s, std::make_pair(e, score));
  }
  template <class T>
  bool any(uint32_t s, uint32_t hi, T&& test) const {
    for (const Class& c : cls_) {
      if (c.ranges.empty()) continue;
      const uint32_t lo = s > c.longest ? s - c.longest : 0;
      if (lo > hi) continue;
      for (auto it = c.ranges.lower_bound(lo); it != c.ranges.end() && it->first <= hi; ++it)
        if (test(it->first, it->second.first, it->second.second)) return true;
    }
    return false;
  }

 private:
  struct Class {
    uint32_t longest = 0;
    std::multimap<uint32_t, std::pair<uint32_t, int32_t>> ranges;  // start -> (end, score)
  };
  Class cls_[33];
};

}  // namespace

// engine.rs:863-974 in its cross-subject form: sort by score, then subject
// (oid) descending, then ScoreCompareHSPs; keep an HSP unless an accepted HSP
// of at least its score covers (int)(100 * overlap / span) >= level percent of
// its query range.  Minus-strand ranges are shifted by one (NCBI's interval
// tree coordinates); at level 0 a range that only touches masks too.
void mask_level_filter(const ReportData& d, std::vector<uint64_t>& rows, int32_t level, uint32_t q_len) {
  std::sort(rows.begin(), rows.end(), [&](uint64_t x, uint64_t y) {
    const rmb_hsp &a = d.hsps[x], &b = d.hsps[y];
    if (a.score != b.score) return a.score > b.score;
    if (a.subject != b.subject) return a.subject > b.subject;
    return score_compare(a, b, q_len) < 0;
  });
  SpanIndex kept;
  std::vector<uint64_t> out;
  out.reserve(rows.size());
  const bool touching = level == 0;
  for (uint64_t r : rows) {
    const rmb_hsp& h = d.hsps[r];
    const uint32_t s = h.q_start + (h.strand ? 1 : 0), e = h.q_end + (h.strand ? 1 : 0);
    const int64_t span = (int64_t)e - (int64_t)s;
    if (span == 0) continue;
    const uint32_t hi = touching ? e : e - 1;
    const bool masked = kept.any(s, hi, [&](uint32_t sj, uint32_t ej, int32_t scj) {
      if (ej < s || (ej == s && !touching)) return false;
      if (scj < h.score) return false;
      const int64_t ov = (int64_t)std::min(ej, e) - (int64_t)std::max(sj, s);
      return (int64_t)(100.0 * ((double)ov / (double)span)) >= (int64_t)level;
    });
    if (masked) continue;
    kept.insert(s, e, h.score);
    out.push_back(r);
  }
  rows.swap(out);
}

namespace {

// main.rs:1230-1234 then sort_hit_list_order_by_oid (engine.rs:998-1017):
// subjects by their best score, then oid, descending; a subject's HSPs by
// ScoreCompareHSPs, ties in the order of (strand, query start).
void output_order(const ReportData& d, std::vector<uint64_t>& rows, uint32_t q_len) {
  std::sort(rows.begin(), rows.end(), [&](uint64_t x, uint64_t y) {
    const rmb_hsp &a = d.hsps[x], &b = d.hsps[y];
    if (a.subject != b.subject) return a.subject < b.subject;
    if (a.strand != b.strand) return a.strand < b.strand;
    return a.q_start < b.q_start;
  });
  std::map<uint32_t, int32_t> top;
  for (uint64_t r : rows) {
    auto it = top.emplace(d.hsps[r].subject, INT32_MIN).first;
    it->second = std::max(it->second, d.hsps[r].score);
  }
  std::stable_sort(rows.begin(), rows.end(), [&](uint64_t x, uint64_t y) {
    const rmb_hsp &a = d.hsps[x], &b = d.hsps[y];
    const int32_t ta = top[a.subject], tb = top[b.subject];
    if (ta != tb) return ta > tb;
    if (a.subject != b.subject) return a.subject > b.subject;
    return score_compare(a, b, q_len) < 0;
  });
}

void append_f2(std::string& o, double v) {
  char b[48];
  std::snprintf(b, sizeof b, "%.2f", v);
  o += b;
}

}  // namespace

rmb_align_stats align_stats(const char* q, const char* s, uint64_t n) {
  auto pure = [](char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; };
  auto bits = [](char c) -> int {  // IUPAC code as a set of bases
    switch (c) {
      case 'A': return 1; case 'C': return 2; case 'G': return 4; case 'T': return 8; case 'R': return 5;
      case 'Y': return 10; case 'M': return 3; case 'K': return 12; case 'W': return 9; case 'S': return 6;
      case 'B': return 14; case 'D': return 13; case 'H': return 11; case 'V': return 7; default: return 15;
    }
  };
  enum { ID, TI, TV, IUBM, IUBS, GAP, NONE };
  auto kind = [&](char a, char b) {
    if (a == '-' || b == '-') return (int)GAP;
    if (a == b) return (int)ID;
    if (!pure(a) || !pure(b)) return ((pure(a) != pure(b)) && (bits(a) & bits(b))) ? (int)IUBM : (int)IUBS;
    return ((a == 'A' || a == 'G') == (b == 'A' || b == 'G')) ? (int)TI : (int)TV;
  };
  rmb_align_stats st{};
  uint32_t pairs = 0, qlen = 0, dlen = 0;
  double cpg_ti = 0;
  char prev = ' ';
  int prevk = NONE;
  for (uint64_t i = 0; i < n; ++i) {
    const char a = s[i], b = q[i];  // a: ancestral (subject), b: query
    if (b != '-') ++qlen;
    if (a != '-') ++dlen;
    if (prev == 'C' && a == 'G') {  // a CpG site of the ancestral sequence
      ++st.cpg_sites;
      const int k = kind(a, b);
      if (prevk == TI && k == TI) cpg_ti -= 1.0;
      else if (prevk == TI || k == TI) cpg_ti -= 0.9;
    }
    if (a == '-' || b == '-') {
      if (b == '-') ++st.query_gaps; else ++st.subject_gaps;
      if (a != '-') { prev = a; prevk = GAP; }
      continue;
    }
    const int k = kind(a, b);
    const bool both = pure(a) && pure(b);
    if (k == ID) { ++st.identities; if (both) ++pairs; }
    else if (k == TI) { ++st.mismatches; if (both) { ++pairs; ++st.transitions; cpg_ti += 1.0; } }
    else if (k == TV) { ++st.mismatches; if (both) { ++pairs; ++st.transversions; } }
    else if (k == IUBS) ++st.mismatches;
    prev = a;
    prevk = k;
  }
  auto kimura = [](double p, double qq) {
    const double o = (1.0 - 2.0 * p - qq) * std::sqrt(1.0 - 2.0 * qq);
    return o > 0.0 ? std::fabs(-0.5 * std::log(o)) * 100.0 : 100.0;
  };
  if (pairs < 1) {
    st.kdiv = st.cpg_kdiv = 100.0;
  } else {
    const double np = pairs;
    st.kdiv = kimura(st.transitions / np, st.transversions / np);
    st.cpg_kdiv = kimura(cpg_ti / np, st.transversions / np);
  }
  st.perc_sub = qlen ? ((float)st.mismatches / (float)qlen) * 100.0f : 0.0f;
  st.perc_query_gap = qlen ? ((float)st.query_gaps / (float)qlen) * 100.0f : 0.0f;
  st.perc_db_gap = dlen ? ((float)st.subject_gaps / (float)dlen) * 100.0f : 0.0f;
  return st;
}

void aligned_strings(const ReportData& d, const rmb_hsp& h, std::string& qa, std::string& sa) {
  const uint8_t* Q = d.queries[h.query] + 1;
  const uint8_t* S = (h.strand ? d.subjects_rc : d.subjects)[h.subject] + 1;
  uint64_t qi = h.q_start, si = h.strand ? d.subject_lens[h.subject] - h.s_end : h.s_start;
  qa.clear();
  sa.clear();
  for (uint32_t k = 0; k < h.ops_count; ++k) {
    const uint32_t op = d.ops[h.ops_offset + k], kind = op & 3, n = op >> 2;
    for (uint32_t j = 0; j < n; ++j) {
      qa += kind != 1 ? kIupac[Q[qi++] & 15] : '-';
      sa += kind != 2 ? kIupac[S[si++] & 15] : '-';
    }
  }
}

// The effective search space of query q (0 without statistics).
static int64_t query_searchsp(const ReportData& d, uint32_t q) {
  if (!d.ka) return 0;
  int64_t sp = 0;
  int32_t la = 0;
  ka_search_space(*d.ka, (int32_t)d.query_lens[q], d.db_length, d.db_num_seqs, sp, la);
  return sp;
}

std::string format_evalue(double e) {
  char b[64];
  if (e < 1.0e-180) return "0.0";
  if (e < 0.0009) std::snprintf(b, sizeof b, "%.2e", e);
  else if (e < 0.1) std::snprintf(b, sizeof b, "%.3f", e);
  else if (e < 1.0) std::snprintf(b, sizeof b, "%.2f", e);
  else if (e < 10.0) std::snprintf(b, sizeof b, "%.1f", e);
  else std::snprintf(b, sizeof b, "%.0f", e);
  return b;
}

std::string format_bit_score(double bits) {
  char b[64];
  if (bits > 99999.0) std::snprintf(b, sizeof b, "%.3e", bits);
  else if (bits > 99.9) std::snprintf(b, sizeof b, "%3lld", (long long)bits);  // truncated toward zero
  else std::snprintf(b, sizeof b, "%4.1f", bits);
  return b;
}

std::vector<std::vector<uint64_t>> report_order(const ReportData& d, int32_t mask_level, unsigned threads,
                                                uint32_t q0, uint32_t q1) {
  q1 = std::min(q1, d.n_queries);
  q0 = std::min(q0, q1);
  std::vector<std::vector<uint64_t>> rows(q1 - q0);
  for (uint64_t i = 0; i < d.n_hsps; ++i) {
    const uint32_t q = d.hsps[i].query;
    if (q >= q0 && q < q1) rows[q - q0].push_back(i);
  }
  const bool masking = mask_level >= 0 && mask_level < 100;
  const bool reap = d.ka && d.ka->lambda > 0.0 && d.evalue_max > 0.0;
  for_each_query(q1 - q0, threads, [&](uint32_t k) {
    const uint32_t q = q0 + k;
    std::vector<uint64_t>& rq = rows[k];
    if (reap) {  // the --evalue reap, before mask level
      const int64_t sp = query_searchsp(d, q);
      std::vector<uint64_t> kept;
      kept.reserve(rq.size());
      for (uint64_t r : rq)
        if (ka_evalue(*d.ka, sp, d.hsps[r].score) <= d.evalue_max) kept.push_back(r);
      rq.swap(kept);
    }
    if (masking) mask_level_filter(d, rq, mask_level, d.query_lens[q]);
    output_order(d, rq, d.query_lens[q]);
  });
  return rows;
}

namespace {
enum FieldCode {
  SCORE, PERC_SUB, PERC_QGAP, PERC_DGAP, QSEQID, QSTART, QEND, QLEN, SSTRAND, SSEQID, SSTART, SEND, SLEN, KDIV,
  CPG_KDIV, TRANSI, TRANSV, CPG_SITES, QSEQ, SSEQ, EVALUE, BITSCORE
};
const char* const kFieldNames[] = {"score", "perc_sub", "perc_query_gap", "perc_db_gap", "qseqid", "qstart", "qend",
                                   "qlen", "sstrand", "sseqid", "sstart", "send", "slen", "kdiv", "cpg_kdiv",
                                   "transi", "transv", "cpg_sites", "qseq", "sseq", "evalue", "bitscore"};

// The rows of queries [q0, q0 + order.size()) into text[k] (query q0 + k).
void report_rows(const ReportData& d, const std::vector<std::vector<uint64_t>>& order, uint32_t q0,
                 const ReportFields& rf, const char* const* query_ids, const char* const* subject_ids,
                 unsigned threads, std::vector<std::string>& text) {
  const std::vector<int>& f = rf.codes;
  text.assign(order.size(), std::string());
  for_each_query((uint32_t)order.size(), threads, [&](uint32_t k) {
    std::string& o = text[k];
    std::string qa, sa;
    const int64_t sp = query_searchsp(d, q0 + k);
    for (uint64_t r : order[k]) {
      const rmb_hsp& h = d.hsps[r];
      if (rf.need_aln) aligned_strings(d, h, qa, sa);
      const rmb_align_stats st = rf.need_stats ? align_stats(qa.data(), sa.data(), qa.size()) : rmb_align_stats{};
      for (size_t j = 0; j < f.size(); ++j) {
        if (j) o += '\t';
        switch (f[j]) {
          case SCORE: o += std::to_string(h.score); break;
          case PERC_SUB: append_f2(o, (double)st.perc_sub); break;
          case PERC_QGAP: append_f2(o, (double)st.perc_query_gap); break;
          case PERC_DGAP: append_f2(o, (double)st.perc_db_gap); break;
          case QSEQID: o += query_ids[h.query]; break;
          case QSTART: o += std::to_string(h.q_start + 1); break;
          case QEND: o += std::to_string(h.q_end); break;
          case QLEN: o += std::to_string(d.query_lens[h.query]); break;
          case SSTRAND: o += h.strand ? "minus" : "plus"; break;
          case SSEQID: o += subject_ids[h.subject]; break;
          case SSTART: o += std::to_string(h.strand ? h.s_end : h.s_start + 1); break;
          case SEND: o += std::to_string(h.strand ? h.s_start + 1 : h.s_end); break;
          case SLEN: o += std::to_string(d.subject_lens[h.subject]); break;
          case KDIV: append_f2(o, st.kdiv); break;
          case CPG_KDIV: append_f2(o, st.cpg_kdiv); break;
          // NCBI's column quirks (output.rs:338-340): transi prints the
          // transversions, transv the CpG sites, cpg_sites the transitions
          case TRANSI: o += std::to_string(st.transversions); break;
          case TRANSV: o += std::to_string(st.cpg_sites); break;
          case CPG_SITES: o += std::to_string(st.transitions); break;
          case QSEQ: o += qa; break;
          case SSEQ: o += sa; break;
          case EVALUE: o += format_evalue(d.ka ? ka_evalue(*d.ka, sp, h.score) : 1.0); break;
          case BITSCORE: o += format_bit_score(d.ka ? ka_bit_score(*d.ka, h.score) : 0.0); break;
        }
      }
      o += '\n';
    }
  });
}
}  // namespace

ReportFields parse_report_fields(const std::vector<std::string>& fields) {
  ReportFields rf;
  for (const std::string& name : fields) {
    const auto it = std::find_if(std::begin(kFieldNames), std::end(kFieldNames),
                                 [&](const char* k) { return name == k; });
    if (it == std::end(kFieldNames)) throw std::invalid_argument("unknown outfmt 6 field '" + name + "'");
    const int k = (int)(it - std::begin(kFieldNames));
    rf.codes.push_back(k);
    if (k == QSEQ || k == SSEQ) rf.need_aln = true;
    if (k == PERC_SUB || k == PERC_QGAP || k == PERC_DGAP || (k >= KDIV && k <= CPG_SITES))
      rf.need_aln = rf.need_stats = true;
  }
  return rf;
}

ReportText report_tabular(const ReportData& d, const std::vector<std::vector<uint64_t>>& order,
                           const std::vector<std::string>& fields, const char* const* query_ids,
                           const char* const* subject_ids, unsigned threads) {
  const ReportFields rf = parse_report_fields(fields);
  std::vector<std::string> text;
  report_rows(d, order, 0, rf, query_ids, subject_ids, threads, text);
  // one buffer, the queries' texts copied in parallel (not zero-filled first)
  const uint32_t nq = (uint32_t)text.size();
  std::vector<uint64_t> at(nq + 1, 0);
  for (uint32_t q = 0; q < nq; ++q) at[q + 1] = at[q] + text[q].size();
  ReportText out;
  out.size = at[nq];
  out.data = std::shared_ptr<char>(new char[out.size ? out.size : 1], std::default_delete<char[]>());
  for_each_query(nq, threads, [&](uint32_t q) {
    if (!text[q].empty()) std::memcpy(out.data.get() + at[q], text[q].data(), text[q].size());
    std::string().swap(text[q]);
  });
  return out;
}

// The stream's buffer is an address range reserved once (pages are backed as
// the text reaches them), so appends never move what is written; should the
// text outgrow it, a larger buffer takes a copy.
static constexpr uint64_t kStreamReserve = 1ull << 36;  // 64 GiB of address space

ReportStream::ReportStream(const std::vector<std::string>& fields, const char* const* query_ids,
                           const char* const* subject_ids, int32_t mask_level)
    : f_(parse_report_fields(fields)), qids_(query_ids), sids_(subject_ids), mask_level_(mask_level) {
  void* p = mmap(nullptr, kStreamReserve, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (p != MAP_FAILED) {
    buf_ = static_cast<char*>(p);
    cap_ = kStreamReserve;
    mapped_ = true;
  }
}

ReportStream::~ReportStream() {
  if (!buf_) return;
  if (mapped_) munmap(buf_, cap_);
  else delete[] buf_;
}

void ReportStream::reserve(uint64_t need) {
  if (need <= cap_) return;
  const uint64_t cap = std::max<uint64_t>(need, 2 * cap_ + (1u << 20));
  char* nb = new char[cap];
  if (used_) std::memcpy(nb, buf_, used_);
  if (buf_) {
    if (mapped_) munmap(buf_, cap_);
    else delete[] buf_;
  }
  buf_ = nb;
  cap_ = cap;
  mapped_ = false;
}

void ReportStream::add(const ReportData& d, uint32_t q0, uint32_t q1, unsigned threads) {
  if (q1 <= q0) return;
  const auto order = report_order(d, mask_level_, threads, q0, q1);
  std::vector<std::string> text;
  report_rows(d, order, q0, f_, qids_, sids_, threads, text);
  std::vector<uint64_t> at(text.size() + 1, used_);
  for (size_t k = 0; k < text.size(); ++k) at[k + 1] = at[k] + text[k].size();
  reserve(at.back());
  for_each_query((uint32_t)text.size(), threads, [&](uint32_t k) {
    if (!text[k].empty()) std::memcpy(buf_ + at[k], text[k].data(), text[k].size());
    std::string().swap(text[k]);
  });
  used_ = at.back();
}

ReportText ReportStream::finish() {
  ReportText out;
  out.size = used_;
  char* b = buf_;
  const uint64_t cap = cap_;
  if (!b) {
    out.data = std::shared_ptr<char>(new char[1], std::default_delete<char[]>());
  } else if (mapped_) {
    out.data = std::shared_ptr<char>(b, [cap](char* p) { munmap(p, cap); });
  } else {
    out.data = std::shared_ptr<char>(b, std::default_delete<char[]>());
  }
  buf_ = nullptr;
  cap_ = used_ = 0;
  return out;
}

}  // namespace rmb
