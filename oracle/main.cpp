// CPU ORACLE (test infrastructure): command-line driver with the rmblastn
// flags in scope (rmblastn/src/main.rs:128-281; NCBI single-dash spellings are
// normalised as in main.rs:64-101) plus stage dumps used by the parity tests.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>

#include "oracle.hpp"

using namespace orc;

static std::string norm_flag(std::string a) {
  if (a.size() > 1 && a[0] == '-' && a[1] != '-') a = "-" + a;
  for (auto& c : a) if (c == '_') c = '-';
  return a;
}

int main(int argc, char** argv) {
  std::map<std::string, std::string> kv;
  bool complexity = false, ncbi_compat = false;
  for (int i = 1; i < argc; ++i) {
    std::string f = norm_flag(argv[i]);
    if (f == "--complexity-adjust") { complexity = true; continue; }
    if (f == "--ncbi-compat") { ncbi_compat = true; continue; }
    if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", argv[i]); return 2; }
    kv[f] = argv[++i];
  }
  auto get = [&](const char* k, const char* d) { auto it = kv.find(k); return it == kv.end() ? std::string(d) : it->second; };
  if (!kv.count("--db") || !kv.count("--query") || !kv.count("--gapopen") || !kv.count("--gapextend")) {
    fprintf(stderr, "usage: rmblastn_oracle --db F --query F --matrix F --gapopen N --gapextend N [options]\n");
    return 2;
  }
  Params p;
  p.gap_open = std::stoi(get("--gapopen", "0"));
  p.gap_extend = std::stoi(get("--gapextend", "1"));
  p.word_size = std::stoi(get("--word-size", "11"));
  p.xdrop_ungap = std::stoi(get("--xdrop-ungap", "20"));
  p.xdrop_gap = std::stoi(get("--xdrop-gap", "30"));
  p.xdrop_gap_final = std::stoi(get("--xdrop-gap-final", "100"));
  p.min_raw_gapped_score = std::stoi(get("--min-raw-gapped-score", "0"));
  p.complexity_adjust = complexity;
  int ml = std::stoi(get("--mask-level", "-1"));
  p.mask_level = ml < 0 ? 101 : ml;
  std::string dust = get("--dust", "yes");
  p.dust = (dust == "yes" || dust == "true");
  int threads = std::stoi(get("--num-threads", "1"));
  std::string dump = get("--dump", "");
  std::vector<std::string> fields;
  {
    std::istringstream is(get("--outfmt", "6 score qseqid qstart qend sstrand sseqid sstart send"));
    std::string w;
    is >> w;
    if (w != "6") { fprintf(stderr, "only -outfmt 6 is supported\n"); return 1; }
    while (is >> w) fields.push_back(w);
  }
  Matrix m;
  std::string err;
  if (!matrix_from_file(get("--matrix", ""), m, err)) { fprintf(stderr, "matrix: %s\n", err.c_str()); return 1; }
  std::vector<FastaRecord> db, qs;
  if (!read_fasta(kv["--db"], db, err) || !read_fasta(kv["--query"], qs, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  std::vector<PreparedSubject> subjects;
  std::vector<std::string> names;
  for (auto& r : db) { subjects.push_back(prepare_subject(r.seq)); names.push_back(r.id); }
  // average subject length for the KA ungapped cutoff (main.rs:603-610)
  uint64_t db_letters = 0;
  for (auto& r : db) db_letters += r.seq.size() - 2;
  const uint64_t avg_subj = db.empty() ? 0 : db_letters / db.size();
  const std::string matrix_name = get("--matrix", "");
  // E-value statistics when the fields print them or a reap threshold is set
  // (main.rs:612-686): --evalue, or NCBI's 10 under --ncbi-compat
  KaOptions kopt;
  kopt.lambda = std::stod(get("--matrix-lambda", "0"));
  kopt.k = std::stod(get("--matrix-k", "0"));
  kopt.alpha = std::stod(get("--matrix-alpha", "0"));
  kopt.beta = std::stod(get("--matrix-beta", "0"));
  kopt.ncbi_compat = ncbi_compat;
  const bool evalue_given = kv.count("--evalue") > 0;
  const double threshold = evalue_given ? std::stod(kv["--evalue"]) : (ncbi_compat ? 10.0 : 0.0);
  bool stats_fields = false;
  for (auto& f : fields) stats_fields |= f == "evalue" || f == "bitscore";
  KaStats ka{};
  const bool have_ka = stats_fields || threshold > 0;
  if (have_ka && !resolve_ka_stats(matrix_name, p.gap_open, p.gap_extend, kopt, ka, err)) {
    fprintf(stderr, "matrix: %s\n", err.c_str());
    return 1;
  }
  const bool reap = threshold > 0 && ka.lambda > 0;
  auto t0 = std::chrono::steady_clock::now();
  for (auto& q : qs) {
    // the cutoff of this query, from its whole length (main.rs:1154-1170)
    p.ungapped_cutoff = ungapped_cutoff(matrix_name, p.gap_open, p.gap_extend, p.min_raw_gapped_score,
                                        q.seq.size() - 2, avg_subj, nullptr);
    if (!dump.empty()) {  // stage dumps: seeds | ungapped | prelims, per subject
      QueryLookup ql = build_query_lookup(q.seq, p);
      std::vector<uint8_t> qrc = revcomp(q.seq);
      for (size_t i = 0; i < subjects.size(); ++i) {
        std::vector<SeedRecord> seeds;
        std::vector<UngappedRecord> ung;
        std::vector<PrelimHsp> pre;
        Taps t;
        t.seeds = &seeds; t.ungapped = &ung; t.prelims = &pre;
        search_one_subject(ql, q.seq, qrc, q.id, subjects[i], names[i], p, m, &t);
        if (dump == "seeds")
          for (auto& s : seeds) printf("%s\t%s\t%u\t%u\t%c\n", q.id.c_str(), names[i].c_str(), s.q_off, s.s_off, s.strand == PLUS ? '+' : '-');
        if (dump == "ungapped")
          for (auto& u : ung)
            printf("%s\t%s\t%c\t%u\t%u\t%u\t%u\t%d\t%u\t%u\n", q.id.c_str(), names[i].c_str(), u.strand == PLUS ? '+' : '-',
                   u.q_start, u.q_end, u.s_start, u.s_end, u.score, u.q_seed, u.s_seed);
        if (dump == "prelims")
          for (auto& u : pre)
            printf("%s\t%s\t%c\t%u\t%u\t%u\t%u\t%d\t%u\t%u\n", q.id.c_str(), names[i].c_str(), u.strand == PLUS ? '+' : '-',
                   u.q_start, u.q_end, u.s_start, u.s_end, u.score, u.q_seed, u.s_seed);
      }
      continue;
    }
    const int64_t searchsp = have_ka ? effective_search_space(ka, (int)(q.seq.size() - 2), (int64_t)db_letters,
                                                               (int)db.size(), nullptr)
                                     : 0;
    p.reap_ka = reap ? &ka : nullptr;
    p.reap_searchsp = searchsp;
    p.reap_evalue = threshold;
    std::vector<AlignResult> res = search_query(q.seq, q.id, subjects, names, p, m, threads);
    sort_hit_list_order(res, names);
    std::string out;
    for (auto& r : res) { out += format_tabular(r, fields, have_ka ? &ka : nullptr, searchsp); out += '\n'; }
    fwrite(out.data(), 1, out.size(), stdout);
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (getenv("RMB_ORACLE_TIME")) fprintf(stderr, "oracle_search_seconds %.6f\n", secs);
  return 0;
}
