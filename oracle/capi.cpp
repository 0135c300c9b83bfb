// CPU ORACLE (test infrastructure): C entry points for the unit-level tests.
#include <cstring>

#include "oracle.hpp"

using namespace orc;

static Matrix mat_from(const int32_t* s) {
  Matrix m{};
  std::memcpy(m.s, s, sizeof(m.s));
  return m;
}

extern "C" {

// ungapped.rs:38-121; out = {score, q_start, q_end, s_start, s_end}
void orc_extend_ungapped(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, uint32_t q_off, uint32_t s_off,
                         const int32_t* matrix, int xdrop, int32_t* out) {
  Matrix m = mat_from(matrix);
  Ungapped u = extend_ungapped(q, qn, s, sn, q_off, s_off, m, xdrop);
  out[0] = u.score; out[1] = (int32_t)u.q_start; out[2] = (int32_t)u.q_end; out[3] = (int32_t)u.s_start; out[4] = (int32_t)u.s_end;
}

// gapped.rs:124-337; out = {score, a_len, b_len}
void orc_align_score_only(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int go, int ge, int x,
                          const int32_t* matrix, int reverse, int64_t* out) {
  Matrix mm = mat_from(matrix);
  int32_t sc; size_t al, bl;
  align_score_only(a, b, m, n, go, ge, x, mm, reverse != 0, sc, al, bl, nullptr);
  out[0] = sc; out[1] = (int64_t)al; out[2] = (int64_t)bl;
}

// gapped.rs:450-757; out = {score, a_len, b_len, n_ops}; ops = (op, count) pairs
int orc_align_ex(const uint8_t* a, const uint8_t* b, size_t m, size_t n, int go, int ge, int x, const int32_t* matrix,
                 int reverse, int64_t* out, uint32_t* ops, int max_ops) {
  Matrix mm = mat_from(matrix);
  GapResult r = align_ex(a, b, m, n, go, ge, x, mm, reverse != 0, nullptr);
  out[0] = r.score; out[1] = (int64_t)r.a_len; out[2] = (int64_t)r.b_len; out[3] = (int64_t)r.ops.size();
  for (size_t i = 0; i < r.ops.size() && (int)i < max_ops; ++i) { ops[2 * i] = r.ops[i].first; ops[2 * i + 1] = r.ops[i].second; }
  return (int)r.ops.size();
}

// diag_hash.rs: run a list of (kind, diag, a, b) operations; kind 0 = get -> out[i], 1 = insert(diag, s_end=a, s_off=b)
void orc_diag_hash_run(const int32_t* ops, int n, int window, uint32_t* out) {
  DiagHash h;
  for (int i = 0; i < n; ++i) {
    const int32_t* o = ops + 4 * i;
    if (o[0] == 0) out[i] = h.get(o[1]);
    else if (o[0] == 1) { h.insert(o[1], (uint32_t)o[2], (uint32_t)o[3], window); out[i] = 0; }
    else { h.clear(); out[i] = 0; }
  }
}

// matrix.rs lambda for a parsed matrix file; returns 0 on success
int orc_matrix_load(const char* path, int32_t* scores, double* freqs, double* lambda) {
  Matrix m; std::string err;
  if (!matrix_from_file(path, m, err)) return 1;
  std::memcpy(scores, m.s, sizeof(m.s));
  std::memcpy(freqs, m.freqs, sizeof(m.freqs));
  *lambda = m.lambda;
  return 0;
}

// twobit.rs NcbiRandom: n draws of next_base_iupac(code) after seeding
void orc_ncbi_random(uint32_t seed, const uint8_t* codes, int n, uint8_t* out) {
  NcbiRandom r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.next_base_iupac(codes[i]);
}

}

extern "C" {

// engine.rs:796-832 search_with_query_lookup_ungapped for one pair; q, s are raw
// BLASTNA (no sentinels).  out: 8 u32 per hit {q_seed, s_seed, q_start, q_end,
// s_start, s_end, score, strand} in the record frame, sorted like NCBI.
int orc_ungapped_hits(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, int word_size, int xdrop_ungap,
                      int min_raw, const int32_t* matrix, uint32_t* out, int cap) {
  Matrix m = mat_from(matrix);
  Params p;
  p.word_size = word_size;
  p.xdrop_ungap = xdrop_ungap;
  p.min_raw_gapped_score = min_raw;
  std::vector<uint8_t> qv(q, q + qn), sv(s, s + sn);
  qv.insert(qv.begin(), 15); qv.push_back(15);
  sv.insert(sv.begin(), 15); sv.push_back(15);
  QueryLookup ql = build_query_lookup(qv, p);
  PreparedSubject ps = prepare_subject(sv);
  std::vector<UngappedRecord> ung;
  Taps t;
  t.ungapped = &ung;
  search_one_subject(ql, qv, revcomp(qv), "q", ps, "s", p, m, &t);
  for (int i = 0; i < (int)ung.size() && i < cap; ++i) {
    const UngappedRecord& u = ung[i];
    uint32_t* o = out + 8 * i;
    o[0] = u.q_seed; o[1] = u.s_seed; o[2] = u.q_start; o[3] = u.q_end; o[4] = u.s_start; o[5] = u.s_end;
    o[6] = (uint32_t)u.score; o[7] = u.strand;
  }
  return (int)ung.size();
}

// complexity_adjust.rs:57-102 with the matrix file's FREQS and lambda
int orc_adjusted_score(const char* matrix_path, int32_t raw, const uint8_t* bases, int n, double* out) {
  Matrix m; std::string err;
  if (!matrix_from_file(matrix_path, m, err)) return 1;
  *out = compute_adjusted_score(raw, std::vector<uint8_t>(bases, bases + n), m);
  return 0;
}

// engine.rs:863-974 per-subject form; hsps: 6 u32 per HSP {score, q_start, q_end, s_start, s_end, strand}
int orc_mask_level(const uint32_t* h, int n, uint32_t q_len, int mask_level, uint8_t* keep) {
  std::vector<AlignResult> r(n);
  for (int i = 0; i < n; ++i) {
    const uint32_t* x = h + 6 * i;
    r[i].hsp = Hsp{(int32_t)x[0], x[1], x[2], q_len, x[3], x[4], 0, x[5] ? MINUS : PLUS, {}, {}, {}};
    r[i].subject_id = std::to_string(i);
  }
  apply_mask_level(r, mask_level, {});
  for (int i = 0; i < n; ++i) keep[i] = 0;
  for (auto& x : r) keep[std::stoi(x.subject_id)] = 1;
  return (int)r.size();
}

// engine.rs:863-974, cross-subject form: 7 u32 per HSP {score, q_start, q_end,
// s_start, s_end, strand, subject}; subject k is the k-th of n_subjects names
int orc_mask_level_xs(const uint32_t* h, int n, uint32_t q_len, int n_subjects, int mask_level, uint8_t* keep) {
  std::vector<AlignResult> r(n);
  std::vector<std::string> names;
  for (int k = 0; k < n_subjects; ++k) names.push_back("s" + std::to_string(k));
  for (int i = 0; i < n; ++i) {
    const uint32_t* x = h + 7 * i;
    r[i].hsp = Hsp{(int32_t)x[0], x[1], x[2], q_len, x[3], x[4], 0, x[5] ? MINUS : PLUS, {}, {}, {}};
    r[i].subject_id = names[x[6]];
    r[i].query_id = std::to_string(i);
  }
  apply_mask_level(r, mask_level, names);
  for (int i = 0; i < n; ++i) keep[i] = 0;
  for (auto& x : r) keep[std::stoi(x.query_id)] = 1;
  return (int)r.size();
}

// engine.rs:1969-2083; prelims: 6 u32 {score, q_start, q_end, s_start, s_end, strand}
int orc_purge_prelims(const uint32_t* h, int n, uint32_t sa_len, uint8_t* keep) {
  std::vector<PrelimHsp> v(n);
  for (int i = 0; i < n; ++i) {
    const uint32_t* x = h + 6 * i;
    v[i] = PrelimHsp{(uint32_t)i, 0, x[1], x[2], x[3], x[4], (int32_t)x[0], x[5] ? MINUS : PLUS};
  }
  purge_prelim_common_endpoints(v, sa_len);
  for (int i = 0; i < n; ++i) keep[i] = 0;
  for (auto& p : v) keep[p.q_seed] = 1;
  return (int)v.size();
}

// engine.rs:2953-3039 on sentinel-free query/subject; out = {score, q0, q1, s0, s1, ok}
void orc_gapped_bidir(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, uint32_t q_seed, uint32_t s_seed,
                      int go, int ge, int x, const int32_t* matrix, int64_t* out) {
  Matrix m = mat_from(matrix);
  std::vector<uint8_t> qv(q, q + qn), sv(s, s + sn);
  qv.insert(qv.begin(), 15); qv.push_back(15);
  sv.insert(sv.begin(), 15); sv.push_back(15);
  int32_t sc = 0; uint32_t q0 = 0, q1 = 0, s0 = 0, s1 = 0; EditOps ops;
  bool ok = gapped_extend_bidirectional(qv, sv, q_seed, s_seed, go, ge, x, m, sc, q0, q1, s0, s1, ops, nullptr);
  out[0] = sc; out[1] = q0; out[2] = q1; out[3] = s0; out[4] = s1; out[5] = ok;
}

}

extern "C" void orc_improve_seed(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, const uint32_t* in6,
                                 int minus, uint32_t* out2) {
  // in6 = {q_seed, s_seed, prelim q_start, q_end, s_start, s_end}
  improve_seed(q, qn, s, sn, in6[0], in6[1], in6[2], in6[3], in6[4], in6[5], minus != 0, out2[0], out2[1]);
}

// main.rs:1150-1193: {chunk size, chunk count, step}
extern "C" void orc_chunk_plan(uint64_t n, uint64_t* out3) {
  const ChunkPlan c = chunk_plan((size_t)n);
  out3[0] = c.size;
  out3[1] = c.count;
  out3[2] = c.step;
}

// main.rs:1016-1085; prelims as 8 u32 {q_seed, s_seed, q_start, q_end, s_start, s_end, score, strand};
// out holds n_acc + n_new entries; returns the merged count
extern "C" int orc_merge_chunk_prelims(const uint32_t* acc, int n_acc, const uint32_t* nw, int n_new, uint32_t split,
                                       uint32_t* out) {
  auto unpack = [](const uint32_t* x, int n) {
    std::vector<PrelimHsp> v(n);
    for (int i = 0; i < n; ++i) {
      const uint32_t* r = x + 8 * i;
      v[i] = PrelimHsp{r[0], r[1], r[2], r[3], r[4], r[5], (int32_t)r[6], r[7] ? MINUS : PLUS};
    }
    return v;
  };
  std::vector<PrelimHsp> a = unpack(acc, n_acc);
  merge_chunk_prelims(a, unpack(nw, n_new), split);
  for (size_t i = 0; i < a.size(); ++i) {
    const PrelimHsp& h = a[i];
    const uint32_t r[8] = {h.q_seed, h.s_seed, h.q_start, h.q_end, h.s_start, h.s_end, (uint32_t)h.score,
                           h.strand == MINUS ? 1u : 0u};
    std::copy(r, r + 8, out + 8 * i);
  }
  return (int)a.size();
}

extern "C" {
// ka_cutoff.rs:126-153; out = {found, es}, out_sp = searchsp; returns the cutoff
int orc_ungapped_cutoff(const char* name, int go, int ge, int min_raw, uint64_t qlen, uint64_t avg, int* out,
                        uint64_t* out_sp) {
  KaCutoffInfo info{};
  const int c = ungapped_cutoff(name, go, ge, min_raw, qlen, avg, &info);
  out[0] = info.found ? 1 : 0;
  out[1] = info.es;
  *out_sp = info.searchsp;
  return c;
}
}

extern "C" {
// ka_stats.rs resolve: opt = {lambda, k, alpha, beta}; out = {lambda, k, log_k, h, alpha, beta}; returns the
// source (0 sentinel, 1 table, 2 CLI, 3 comment) or -1 on a bad KARLIN line
int orc_ka_stats(const char* path, int go, int ge, const double* opt, int ncbi_compat, double* out) {
  KaOptions o;
  o.lambda = opt[0]; o.k = opt[1]; o.alpha = opt[2]; o.beta = opt[3]; o.ncbi_compat = ncbi_compat != 0;
  KaStats s{};
  std::string err;
  if (!resolve_ka_stats(path, go, ge, o, s, err)) return -1;
  const double v[6] = {s.lambda, s.k, s.log_k, s.h, s.alpha, s.beta};
  for (int i = 0; i < 6; ++i) out[i] = v[i];
  return s.source;
}
int orc_length_adjustment(double K, double logK, double a_d_l, double beta, int qlen, int64_t dblen, int nseq) {
  return compute_length_adjustment(K, logK, a_d_l, beta, qlen, dblen, nseq);
}
// writes the E-value and bit-score strings (NUL-terminated) into e and b (64 bytes each)
void orc_format_scores(double evalue, double bits, char* e, char* b) {
  const std::string x = fmt_evalue(evalue), y = fmt_bits(bits);
  std::snprintf(e, 64, "%s", x.c_str());
  std::snprintf(b, 64, "%s", y.c_str());
}
}

extern "C" {

static QueryLookup lookup_of(const uint8_t* q, size_t qn, const Params& p, std::vector<uint8_t>& qv) {
  qv.assign(q, q + qn);
  qv.insert(qv.begin(), 15);
  qv.push_back(15);
  return build_query_lookup(qv, p);
}

// engine.rs:606 build_query_lookup of a raw BLASTNA query (no sentinels), dust
// off: *lw, CSR cells (4^lw + 1, written when cells_cap suffices) and entries
// (written when ent_cap suffices); returns the entry count.
int64_t orc_lookup(const uint8_t* q, size_t qn, int word_size, int* lw, uint32_t* cells, size_t cells_cap,
                   uint32_t* entries, size_t ent_cap) {
  Params p;
  p.word_size = word_size;
  std::vector<uint8_t> qv;
  QueryLookup ql = lookup_of(q, qn, p, qv);
  *lw = ql.lut.lut_width;
  if (cells && cells_cap >= ql.lut.cell_start.size())
    std::memcpy(cells, ql.lut.cell_start.data(), ql.lut.cell_start.size() * sizeof(uint32_t));
  if (entries && ent_cap >= ql.lut.entries.size())
    std::memcpy(entries, ql.lut.entries.data(), ql.lut.entries.size() * sizeof(uint32_t));
  return (int64_t)ql.lut.entries.size();
}

// engine.rs:796-816 search_with_query_lookup_seeds for one pair: 3 u32 per seed
// {q_off, s_off, strand} in push order; returns the count.
int orc_seeds(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, int word_size, int xdrop_ungap, int min_raw,
              const int32_t* matrix, uint32_t* out, int cap) {
  Matrix m = mat_from(matrix);
  Params p;
  p.word_size = word_size;
  p.xdrop_ungap = xdrop_ungap;
  p.min_raw_gapped_score = min_raw;
  std::vector<uint8_t> qv, sv(s, s + sn);
  QueryLookup ql = lookup_of(q, qn, p, qv);
  sv.insert(sv.begin(), 15);
  sv.push_back(15);
  PreparedSubject ps = prepare_subject(sv);
  std::vector<SeedRecord> seeds;
  Taps t;
  t.seeds = &seeds;
  search_one_subject(ql, qv, revcomp(qv), "q", ps, "s", p, m, &t);
  for (int i = 0; i < (int)seeds.size() && i < cap; ++i) {
    out[3 * i] = seeds[i].q_off;
    out[3 * i + 1] = seeds[i].s_off;
    out[3 * i + 2] = seeds[i].strand;
  }
  return (int)seeds.size();
}

// Phase 2a's prelims for one pair (engine.rs:1317-1411 + purge :1969-2083): 8 u32
// per prelim {q_seed, s_seed, q_start, q_end, s_start, s_end, score, strand};
// returns the count.
int orc_prelims(const uint8_t* q, size_t qn, const uint8_t* s, size_t sn, int word_size, int xdrop_ungap,
                int xdrop_gap, int gap_open, int gap_extend, int min_raw, const int32_t* matrix, uint32_t* out,
                int cap) {
  Matrix m = mat_from(matrix);
  Params p;
  p.word_size = word_size;
  p.xdrop_ungap = xdrop_ungap;
  p.xdrop_gap = xdrop_gap;
  p.gap_open = gap_open;
  p.gap_extend = gap_extend;
  p.min_raw_gapped_score = min_raw;
  std::vector<uint8_t> qv, sv(s, s + sn);
  QueryLookup ql = lookup_of(q, qn, p, qv);
  sv.insert(sv.begin(), 15);
  sv.push_back(15);
  PreparedSubject ps = prepare_subject(sv);
  std::vector<PrelimHsp> pre;
  Taps t;
  t.prelims = &pre;
  search_one_subject(ql, qv, revcomp(qv), "q", ps, "s", p, m, &t);
  for (int i = 0; i < (int)pre.size() && i < cap; ++i) {
    const PrelimHsp& h = pre[i];
    uint32_t* o = out + 8 * i;
    o[0] = h.q_seed; o[1] = h.s_seed; o[2] = h.q_start; o[3] = h.q_end; o[4] = h.s_start; o[5] = h.s_end;
    o[6] = (uint32_t)h.score; o[7] = h.strand;
  }
  return (int)pre.size();
}

}  // extern "C"
