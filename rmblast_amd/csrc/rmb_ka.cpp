// Karlin-Altschul ungapped cutoff of librmblast_hip (host code).
//
// rmblastn keeps an ungapped hit when its score reaches a cutoff that NCBI
// derives from gapped Karlin-Altschul statistics whenever the matrix + gap
// costs have them (BlastInitialWordParametersNew in NCBI's
// blast_parameters.c; restated by the reference in
// rmblast-lib/src/search/ka_cutoff.rs:1-161), and otherwise from
// min_raw_gapped_score / 2.  The statistics come from NCBI's ALP table of the
// RepeatMasker matrices, keyed by the matrix file's base name and one gap
// pair per matrix.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "rmb_ka.h"

namespace rmb {
namespace {

struct KaEntry {
  const char* name;
  int32_t gap_open, gap_extend;
  double lambda, k;
};

// NCBI blast_stat.c rmblast_*_values (2.17), one row per RepeatMasker matrix:
// the table of ka_cutoff.rs:40-87.  The numbers are NCBI's fitted values and
// must be these exactly; 30p53g carries NCBI's placeholder (0.1, 0.1).
constexpr KaEntry kKaTable[] = {
    {"14p35g.matrix", 29, 6, 0.1241860392, 0.2565118357}, {"14p37g.matrix", 29, 6, 0.1243509013, 0.2538327145},
    {"14p39g.matrix", 29, 6, 0.1181081970, 0.2381799563}, {"14p41g.matrix", 29, 6, 0.1209164962, 0.2516961980},
    {"14p43g.matrix", 29, 6, 0.1204338701, 0.2484920137}, {"14p45g.matrix", 29, 6, 0.1281012415, 0.2651519027},
    {"14p47g.matrix", 29, 6, 0.1169571490, 0.2334091021}, {"14p49g.matrix", 29, 6, 0.1167697593, 0.2289233061},
    {"14p51g.matrix", 29, 6, 0.1238779017, 0.2467119379}, {"14p53g.matrix", 29, 6, 0.1211022467, 0.2307304663},
    {"18p35g.matrix", 28, 5, 0.1204776630, 0.2111410357}, {"18p37g.matrix", 28, 5, 0.1213800379, 0.2057492194},
    {"18p39g.matrix", 28, 5, 0.1228853168, 0.2111281814}, {"18p41g.matrix", 28, 5, 0.1096830298, 0.1715966796},
    {"18p43g.matrix", 28, 5, 0.1179924113, 0.1957371364}, {"18p45g.matrix", 28, 5, 0.1184247414, 0.2068045787},
    {"18p47g.matrix", 28, 5, 0.1158783320, 0.1842046578}, {"18p49g.matrix", 28, 5, 0.1248964218, 0.2141853759},
    {"18p51g.matrix", 28, 5, 0.1106793836, 0.1621592148}, {"18p53g.matrix", 28, 5, 0.1082347513, 0.1408677318},
    {"20p35g.matrix", 25, 5, 0.1112832877, 0.1466667286}, {"20p37g.matrix", 25, 5, 0.1134511409, 0.1501236494},
    {"20p39g.matrix", 25, 5, 0.1233401001, 0.1853070864}, {"20p41g.matrix", 25, 5, 0.1088138918, 0.1409300313},
    {"20p43g.matrix", 25, 5, 0.1097495426, 0.1408421343}, {"20p45g.matrix", 25, 5, 0.1101110565, 0.1402262486},
    {"20p47g.matrix", 25, 5, 0.1198508511, 0.1752301830}, {"20p49g.matrix", 25, 5, 0.1150022018, 0.1500562415},
    {"20p51g.matrix", 25, 5, 0.1101356126, 0.1310353805}, {"20p53g.matrix", 25, 5, 0.1239625969, 0.1752022586},
    {"25p35g.matrix", 22, 5, 0.0965155465, 0.0677355511}, {"25p37g.matrix", 22, 5, 0.0989264764, 0.0763719670},
    {"25p39g.matrix", 22, 5, 0.1005843197, 0.0805691511}, {"25p41g.matrix", 22, 5, 0.1030661339, 0.0840485952},
    {"25p43g.matrix", 22, 5, 0.1155434360, 0.1226554129}, {"25p45g.matrix", 22, 5, 0.0973402625, 0.0742020190},
    {"25p47g.matrix", 22, 5, 0.0967346256, 0.0743885008}, {"25p49g.matrix", 22, 5, 0.1070577141, 0.0941947942},
    {"25p51g.matrix", 22, 5, 0.1044972455, 0.0789036639}, {"25p53g.matrix", 22, 5, 0.0977828435, 0.0607500307},
    {"30p53g.matrix", 22, 5, 0.1, 0.1},
};

constexpr double kCutoffE = 0.05;  // NCBI CUTOFF_E_BLASTN

bool same_name(const char* a, const char* b) {
  for (; *a && *b; ++a, ++b)
    if (std::tolower((unsigned char)*a) != std::tolower((unsigned char)*b)) return false;
  return *a == *b;
}

// f64 -> i32 as Rust's `as` converts: truncation, NaN to 0, saturation at the range ends
int32_t to_i32_saturating(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
  return (int32_t)v;
}

uint64_t mul_saturating(uint64_t a, uint64_t b) {
  uint64_t r;
  return __builtin_mul_overflow(a, b, &r) ? ~0ull : r;
}

}  // namespace

const char* matrix_base_name(const char* path) {
  const char* slash = std::strrchr(path, '/');
  return slash && slash[1] ? slash + 1 : path;
}

bool ka_lookup(const char* matrix_name, int32_t gap_open, int32_t gap_extend, double& lambda, double& k) {
  const char* base = matrix_base_name(matrix_name);
  for (const KaEntry& e : kKaTable)
    if (e.gap_open == gap_open && e.gap_extend == gap_extend && same_name(e.name, base)) {
      lambda = e.lambda;
      k = e.k;
      return true;
    }
  return false;
}

int32_t cutoff_score_max(int32_t min_raw_gapped_score) {
  return min_raw_gapped_score > 0 ? min_raw_gapped_score / 2 : 1;
}

int32_t ka_ungapped_cutoff(double lambda, double k, int32_t min_raw_gapped_score, uint64_t query_length,
                           uint64_t avg_subj_length, int32_t* es_out, uint64_t* searchsp_out) {
  const int32_t fallback = cutoff_score_max(min_raw_gapped_score);
  if (!(avg_subj_length > 0 && lambda > 0.0 && k > 0.0)) return -1;
  // searchsp = min(avg subject length, both query strands) * avg subject length
  const uint64_t q2 = mul_saturating(query_length, 2);
  const uint64_t searchsp = mul_saturating(avg_subj_length < q2 ? avg_subj_length : q2, avg_subj_length);
  // BlastKarlinEtoS_simple: the smallest score whose E-value is at most 0.05
  const int32_t es = to_i32_saturating(std::ceil(std::log(k * (double)searchsp / kCutoffE) / lambda));
  if (es_out) *es_out = es;
  if (searchsp_out) *searchsp_out = searchsp;
  const int32_t lo = es > 1 ? es : 1;
  return lo < fallback ? lo : fallback;
}

// ---- E-value statistics ------------------------------------------------------
//
// The gapped Karlin-Altschul block of a custom matrix for the evalue /
// bitscore fields and the --evalue reap, resolved as rmblastn does
// (rmblast-lib/src/ka_stats.rs:153-300): the table of RepeatMasker matrices
// (NCBI's verbatim entries with ALP fits for 30p53g and comparison.matrix,
// rmblast-stats/rmstats/src/tables_data.rs), then -matrix_lambda/k/alpha
// (/beta) given together, then a "# KARLIN" comment in the matrix file, else
// none (E-value 1.0, bit score 0.0).  The ALP fit of unknown matrices that the
// reference can build in is not part of this library.
namespace {

struct StatsEntry {
  const char* name;
  int32_t gap_open, gap_extend;
  double lambda, k, h;
};

// rmstats' table of the RepeatMasker matrices: name, gap costs, lambda, K, H.
constexpr StatsEntry kStatsTable[] = {
    {"14p35g.matrix", 29, 6, 0.1241860392, 0.2565118357, 0.3349210581},
    {"14p37g.matrix", 29, 6, 0.1243509013, 0.2538327145, 0.3484386549},
    {"14p39g.matrix", 29, 6, 0.118108197, 0.2381799563, 0.3209988017},
    {"14p41g.matrix", 29, 6, 0.1209164962, 0.251696198, 0.3474870514},
    {"14p43g.matrix", 29, 6, 0.1204338701, 0.2484920137, 0.3397421893},
    {"14p45g.matrix", 29, 6, 0.1281012415, 0.2651519027, 0.3664022073},
    {"14p47g.matrix", 29, 6, 0.116957149, 0.2334091021, 0.3226817428},
    {"14p49g.matrix", 29, 6, 0.1167697593, 0.2289233061, 0.3136272703},
    {"14p51g.matrix", 29, 6, 0.1238779017, 0.2467119379, 0.32413383},
    {"14p53g.matrix", 29, 6, 0.1211022467, 0.2307304663, 0.2949314272},
    {"18p35g.matrix", 28, 5, 0.120477663, 0.2111410357, 0.2484902185},
    {"18p37g.matrix", 28, 5, 0.1213800379, 0.2057492194, 0.2637651575},
    {"18p39g.matrix", 28, 5, 0.1228853168, 0.2111281814, 0.2802003038},
    {"18p41g.matrix", 28, 5, 0.1096830298, 0.1715966796, 0.2171487218},
    {"18p43g.matrix", 28, 5, 0.1179924113, 0.1957371364, 0.2544060456},
    {"18p45g.matrix", 28, 5, 0.1184247414, 0.2068045787, 0.2641082233},
    {"18p47g.matrix", 28, 5, 0.115878332, 0.1842046578, 0.2407688057},
    {"18p49g.matrix", 28, 5, 0.1248964218, 0.2141853759, 0.2675489103},
    {"18p51g.matrix", 28, 5, 0.1106793836, 0.1621592148, 0.1979815532},
    {"18p53g.matrix", 28, 5, 0.1082347513, 0.1408677318, 0.1771724651},
    {"20p35g.matrix", 25, 5, 0.1112832877, 0.1466667286, 0.1741696491},
    {"20p37g.matrix", 25, 5, 0.1134511409, 0.1501236494, 0.193180875},
    {"20p39g.matrix", 25, 5, 0.1233401001, 0.1853070864, 0.2324952723},
    {"20p41g.matrix", 25, 5, 0.1088138918, 0.1409300313, 0.1778018484},
    {"20p43g.matrix", 25, 5, 0.1097495426, 0.1408421343, 0.1813610542},
    {"20p45g.matrix", 25, 5, 0.1101110565, 0.1402262486, 0.1803401526},
    {"20p47g.matrix", 25, 5, 0.1198508511, 0.175230183, 0.2232474228},
    {"20p49g.matrix", 25, 5, 0.1150022018, 0.1500562415, 0.1891763921},
    {"20p51g.matrix", 25, 5, 0.1101356126, 0.1310353805, 0.1457676446},
    {"20p53g.matrix", 25, 5, 0.1239625969, 0.1752022586, 0.210665934},
    {"25p35g.matrix", 22, 5, 0.0965155465, 0.0677355511, 0.0964040436},
    {"25p37g.matrix", 22, 5, 0.0989264764, 0.076371967, 0.1025772304},
    {"25p39g.matrix", 22, 5, 0.1005843197, 0.0805691511, 0.1108913867},
    {"25p41g.matrix", 22, 5, 0.1030661339, 0.0840485952, 0.1237017673},
    {"25p43g.matrix", 22, 5, 0.115543436, 0.1226554129, 0.1495157932},
    {"25p45g.matrix", 22, 5, 0.0973402625, 0.074202019, 0.1029029688},
    {"25p47g.matrix", 22, 5, 0.0967346256, 0.0743885008, 0.0995494785},
    {"25p49g.matrix", 22, 5, 0.1070577141, 0.0941947942, 0.1181023643},
    {"25p51g.matrix", 22, 5, 0.1044972455, 0.0789036639, 0.1081508568},
    {"25p53g.matrix", 22, 5, 0.0977828435, 0.0607500307, 0.0834178835},
    {"30p53g.matrix", 22, 5, 0.1204393447, 0.1065198197, 0.1243912966},
    {"comparison.matrix", 20, 5, 0.09836223, 0.0789918084, 0.1116129577},
};

}  // namespace

bool read_karlin_comment(const char* path, double v[5], bool& found, std::string& err) {
  // "# KARLIN key value ..." (keys lambda k h alpha beta, any case and order;
  // matrix.rs:248-266); FREQS lines are the other comment the format knows
  found = false;
  FILE* f = std::fopen(path, "r");
  if (!f) {
    err = std::string("cannot open ") + path;
    return false;
  }
  char line[4096];
  bool ok = true;
  while (ok && std::fgets(line, sizeof line, f)) {
    const char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p != '#' || std::strstr(p, "FREQS")) continue;
    const char* rest = nullptr;
    if (!std::strncmp(p, "# KARLIN", 8)) rest = p + 8;
    else if (!std::strncmp(p, "#KARLIN", 7)) rest = p + 7;
    if (!rest) continue;
    for (int i = 0; i < 5; ++i) v[i] = 0.0;
    found = true;
    std::vector<std::string> tok;
    {
      std::string cur;
      for (const char* q = rest;; ++q) {
        if (*q == 0 || std::isspace((unsigned char)*q)) {
          if (!cur.empty()) tok.push_back(cur);
          cur.clear();
          if (!*q) break;
        } else {
          cur += *q;
        }
      }
    }
    for (size_t i = 0; i + 1 < tok.size(); i += 2) {
      char* end = nullptr;
      const double val = std::strtod(tok[i + 1].c_str(), &end);
      if (end == tok[i + 1].c_str() || *end) {
        err = "bad KARLIN value '" + tok[i + 1] + "' for key '" + tok[i] + "'";
        ok = false;
        break;
      }
      std::string key = tok[i];
      for (auto& ch : key) ch = (char)std::tolower((unsigned char)ch);
      const char* keys[5] = {"lambda", "k", "h", "alpha", "beta"};
      int idx = -1;
      for (int k = 0; k < 5; ++k)
        if (key == keys[k]) idx = k;
      if (idx < 0) {
        err = "unknown KARLIN key '" + key + "'";
        ok = false;
        break;
      }
      v[idx] = val;
    }
  }
  std::fclose(f);
  return ok;
}

namespace {

rmb_ka_stats sentinel_stats() {
  rmb_ka_stats s{};
  s.source = RMB_KA_SENTINEL;
  s.lambda = s.k = s.log_k = s.h = -1.0;
  return s;
}

// Mode-2 semantics (blast_setup.c's -matrix_* path, ka_stats.rs:305-312):
// H = lambda / alpha, alpha and beta feed the length adjustment
rmb_ka_stats cli_like(int32_t source, double lambda, double k, double alpha, double beta) {
  rmb_ka_stats s{};
  s.source = source;
  s.lambda = lambda;
  s.k = k;
  s.log_k = std::log(k);
  s.h = lambda / alpha;
  s.alpha = alpha;
  s.beta = beta;
  return s;
}

}  // namespace

bool ka_stats_resolve(const char* matrix_path, int32_t gap_open, int32_t gap_extend,
                      const rmb_ka_stats_options& opt, rmb_ka_stats& out, std::string& err) {
  const char* base = matrix_base_name(matrix_path);
  const StatsEntry* hit = nullptr;
  for (const StatsEntry& e : kStatsTable)
    if (same_name(e.name, base)) {
      if (e.gap_open == gap_open && e.gap_extend == gap_extend) hit = &e;
      break;
    }
  const bool cli_set = opt.matrix_lambda > 0.0 && opt.matrix_k > 0.0 && opt.matrix_alpha > 0.0;
  auto table = [&](const StatsEntry& e) {
    rmb_ka_stats s{};
    s.source = RMB_KA_TABLE;
    s.lambda = e.lambda;
    s.k = e.k;
    s.log_k = std::log(e.k);
    s.h = e.h;
    // the -matrix_* alpha/beta take part even on a table hit (blast_setup.c's
    // read_in_matrix arm; eff_lengths.rs:68-79), else alpha = lambda / H, beta 0
    s.alpha = cli_set ? opt.matrix_alpha : e.lambda / e.h;
    s.beta = cli_set ? opt.matrix_beta : 0.0;
    return s;
  };
  if (opt.ncbi_compat) {
    // NCBI 2.17.1's own sources only: its C table lacks comparison.matrix and
    // holds a placeholder for 30p53g (lambda = K = H = 0.1, alpha 1, beta 0)
    if (hit && !same_name(hit->name, "comparison.matrix")) {
      out = same_name(hit->name, "30p53g.matrix") ? cli_like(RMB_KA_TABLE, 0.1, 0.1, 1.0, 0.0) : table(*hit);
      return true;
    }
    out = cli_set ? cli_like(RMB_KA_CLI, opt.matrix_lambda, opt.matrix_k, opt.matrix_alpha, opt.matrix_beta)
                  : sentinel_stats();
    return true;
  }
  if (hit) {
    out = table(*hit);
    return true;
  }
  if (cli_set) {
    out = cli_like(RMB_KA_CLI, opt.matrix_lambda, opt.matrix_k, opt.matrix_alpha, opt.matrix_beta);
    return true;
  }
  double v[5] = {0, 0, 0, 0, 0};  // lambda k h alpha beta
  bool found = false;
  if (!read_karlin_comment(matrix_path, v, found, err)) return false;
  if (found) {
    const double alpha = v[3] > 0.0 ? v[3] : (v[2] > 0.0 ? v[0] / v[2] : 0.0);
    if (v[0] > 0.0 && v[1] > 0.0 && alpha > 0.0) {
      out = cli_like(RMB_KA_COMMENT, v[0], v[1], alpha, v[4]);
      return true;
    }
  }
  out = sentinel_stats();
  return true;
}

// BLAST_ComputeLengthAdjustment (blast_stat.c, restated in
// rmstats/src/length_adjust.rs:7-93): the largest integer ell with
// alpha/lambda * (ln K + ln((m - ell)(n - N ell))) + beta >= ell, by at most
// 20 steps of fixed-point iteration and bisection from [0, ell_max].
int32_t length_adjustment(double k, double log_k, double alpha_d_lambda, double beta, int32_t query_length,
                          int64_t db_length, int32_t db_num_seqs) {
  const double m = query_length, n = (double)db_length, nn = db_num_seqs;
  double ell_min = 0, ell_max, ell_next = 0;
  bool converged = false;
  {
    // the largest ell with K (m - ell)(n - N ell) > max(m, n): a root of a
    // quadratic, in the cancellation-free form 2c / (-b + sqrt(b^2 - 4ac))
    const double a = nn, mb = m * nn + n, c = n * m - std::max(m, n) / k;
    if (c < 0) return 0;
    ell_max = 2.0 * c / (mb + std::sqrt(mb * mb - 4.0 * a * c));
  }
  for (int i = 1; i <= 20; ++i) {
    const double ell = ell_next;
    const double ss = (m - ell) * (n - nn * ell);
    const double ell_bar = alpha_d_lambda * (log_k + std::log(ss)) + beta;
    if (ell_bar >= ell) {
      ell_min = ell;
      if (ell_bar - ell_min <= 1.0) {
        converged = true;
        break;
      }
      if (ell_min == ell_max) break;
    } else {
      ell_max = ell;
    }
    if (ell_min <= ell_bar && ell_bar <= ell_max) ell_next = ell_bar;
    else ell_next = i == 1 ? ell_max : (ell_min + ell_max) / 2.0;
  }
  int32_t la = (int32_t)ell_min;
  if (converged) {
    const double ell = std::ceil(ell_min);
    if (ell <= ell_max) {
      const double ss = (m - ell) * (n - nn * ell);
      if (alpha_d_lambda * (log_k + std::log(ss)) + beta >= ell) la = (int32_t)ell;
    }
  }
  return la;
}

// BLAST_CalcEffLengths' read_in_matrix arm for one query (eff_lengths.rs:52-128)
void ka_search_space(const rmb_ka_stats& s, int32_t query_length, int64_t db_length, int32_t db_num_seqs,
                     int64_t& eff_searchsp, int32_t& la) {
  eff_searchsp = 0;
  la = 0;
  if (query_length <= 0 || !(s.lambda > 0.0)) return;
  la = length_adjustment(s.k, s.log_k, s.alpha / s.lambda, s.beta, query_length, db_length, db_num_seqs);
  int64_t eff_db = db_length - (int64_t)db_num_seqs * la;
  if (eff_db <= 0) eff_db = 1;
  eff_searchsp = eff_db * (int64_t)(query_length - la);
}

// Blast_HSPListGetEvalues / GetBitScores with the sentinel guard
// (rmstats/src/evalue.rs:96-126): 1.0 and 0.0 without statistics
double ka_evalue(const rmb_ka_stats& s, int64_t eff_searchsp, int32_t score) {
  if (!(s.lambda > 0.0)) return 1.0;
  if (s.k < 0.0 || s.h < 0.0) return -1.0;
  return (double)eff_searchsp * std::exp(-s.lambda * (double)score + s.log_k);
}

double ka_bit_score(const rmb_ka_stats& s, int32_t score) {
  if (!(s.lambda > 0.0)) return 0.0;
  return ((double)score * s.lambda - s.log_k) / 0.69314718055994530942;
}

}  // namespace rmb
