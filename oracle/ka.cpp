// CPU ORACLE (test infrastructure): the Karlin-Altschul ungapped cutoff of
// rmblast-lib/src/search/ka_cutoff.rs, restated.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle.hpp"

namespace orc {

namespace {
// ka_cutoff.rs:40-87, grouped as the reference groups it: each RepeatMasker
// family of matrices ("14p", "18p", ...) has one gap pair, and per matrix
// ("14p35g" ...) a (lambda, K) pair.
struct Family { const char* prefix; int go, ge; struct { int pct; double lambda, k; } m[10]; int n; };
const Family kFamilies[] = {
    {"14p", 29, 6, {{35, 0.1241860392, 0.2565118357}, {37, 0.1243509013, 0.2538327145},
                    {39, 0.1181081970, 0.2381799563}, {41, 0.1209164962, 0.2516961980},
                    {43, 0.1204338701, 0.2484920137}, {45, 0.1281012415, 0.2651519027},
                    {47, 0.1169571490, 0.2334091021}, {49, 0.1167697593, 0.2289233061},
                    {51, 0.1238779017, 0.2467119379}, {53, 0.1211022467, 0.2307304663}}, 10},
    {"18p", 28, 5, {{35, 0.1204776630, 0.2111410357}, {37, 0.1213800379, 0.2057492194},
                    {39, 0.1228853168, 0.2111281814}, {41, 0.1096830298, 0.1715966796},
                    {43, 0.1179924113, 0.1957371364}, {45, 0.1184247414, 0.2068045787},
                    {47, 0.1158783320, 0.1842046578}, {49, 0.1248964218, 0.2141853759},
                    {51, 0.1106793836, 0.1621592148}, {53, 0.1082347513, 0.1408677318}}, 10},
    {"20p", 25, 5, {{35, 0.1112832877, 0.1466667286}, {37, 0.1134511409, 0.1501236494},
                    {39, 0.1233401001, 0.1853070864}, {41, 0.1088138918, 0.1409300313},
                    {43, 0.1097495426, 0.1408421343}, {45, 0.1101110565, 0.1402262486},
                    {47, 0.1198508511, 0.1752301830}, {49, 0.1150022018, 0.1500562415},
                    {51, 0.1101356126, 0.1310353805}, {53, 0.1239625969, 0.1752022586}}, 10},
    {"25p", 22, 5, {{35, 0.0965155465, 0.0677355511}, {37, 0.0989264764, 0.0763719670},
                    {39, 0.1005843197, 0.0805691511}, {41, 0.1030661339, 0.0840485952},
                    {43, 0.1155434360, 0.1226554129}, {45, 0.0973402625, 0.0742020190},
                    {47, 0.0967346256, 0.0743885008}, {49, 0.1070577141, 0.0941947942},
                    {51, 0.1044972455, 0.0789036639}, {53, 0.0977828435, 0.0607500307}}, 10},
    {"30p", 22, 5, {{53, 0.1, 0.1}}, 1},
};

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}
}  // namespace

// ka_cutoff.rs:105-114: base name of the path, case-insensitive, exact gaps
bool ka_lookup(const std::string& matrix_name, int gap_open, int gap_extend, double& lambda, double& k) {
  const size_t slash = matrix_name.rfind('/');
  const std::string base = lower(slash == std::string::npos ? matrix_name : matrix_name.substr(slash + 1));
  for (const Family& f : kFamilies)
    for (int i = 0; i < f.n; ++i) {
      const std::string name = std::string(f.prefix) + std::to_string(f.m[i].pct) + "g.matrix";
      if (f.go == gap_open && f.ge == gap_extend && name == base) {
        lambda = f.m[i].lambda;
        k = f.m[i].k;
        return true;
      }
    }
  return false;
}

// ka_cutoff.rs:126-153
int ungapped_cutoff(const std::string& matrix_name, int gap_open, int gap_extend, int min_raw_gapped_score,
                    uint64_t query_length, uint64_t avg_subj_length, KaCutoffInfo* info) {
  const int fallback = min_raw_gapped_score > 0 ? min_raw_gapped_score / 2 : 1;
  double lambda = 0, k = 0;
  if (info) info->found = false;
  if (!ka_lookup(matrix_name, gap_open, gap_extend, lambda, k) || avg_subj_length == 0 || !(lambda > 0) || !(k > 0))
    return fallback;
  auto sat_mul = [](uint64_t a, uint64_t b) {
    unsigned __int128 r = (unsigned __int128)a * b;
    return r > ~0ull ? ~0ull : (uint64_t)r;
  };
  const uint64_t qlen2 = sat_mul(query_length, 2);
  const uint64_t searchsp = sat_mul(std::min(avg_subj_length, qlen2), avg_subj_length);
  const double v = std::ceil(std::log(k * (double)searchsp / 0.05) / lambda);
  // Rust `as i32`: NaN -> 0, saturating at both ends, else truncation
  const int es = std::isnan(v) ? 0 : v >= 2147483647.0 ? 2147483647 : v <= -2147483648.0 ? (-2147483647 - 1) : (int)v;
  if (info) *info = KaCutoffInfo{true, lambda, k, searchsp, es};
  return std::min(std::max(es, 1), fallback);
}

// ---- E-value statistics (rmblast-lib/src/ka_stats.rs, rmblast-stats/rmstats) ----
namespace {
struct StatRow { const char* stem; int go, ge; double lkh[3]; };  // lambda, K, H
// rmstats/src/tables_data.rs RMBLAST_MATRIX_KA: "<stem>.matrix" at its gap pair
const StatRow kStatRows[] = {
    {"14p35g", 29, 6, {0.1241860392, 0.2565118357, 0.3349210581}},
    {"14p37g", 29, 6, {0.1243509013, 0.2538327145, 0.3484386549}},
    {"14p39g", 29, 6, {0.118108197, 0.2381799563, 0.3209988017}},
    {"14p41g", 29, 6, {0.1209164962, 0.251696198, 0.3474870514}},
    {"14p43g", 29, 6, {0.1204338701, 0.2484920137, 0.3397421893}},
    {"14p45g", 29, 6, {0.1281012415, 0.2651519027, 0.3664022073}},
    {"14p47g", 29, 6, {0.116957149, 0.2334091021, 0.3226817428}},
    {"14p49g", 29, 6, {0.1167697593, 0.2289233061, 0.3136272703}},
    {"14p51g", 29, 6, {0.1238779017, 0.2467119379, 0.32413383}},
    {"14p53g", 29, 6, {0.1211022467, 0.2307304663, 0.2949314272}},
    {"18p35g", 28, 5, {0.120477663, 0.2111410357, 0.2484902185}},
    {"18p37g", 28, 5, {0.1213800379, 0.2057492194, 0.2637651575}},
    {"18p39g", 28, 5, {0.1228853168, 0.2111281814, 0.2802003038}},
    {"18p41g", 28, 5, {0.1096830298, 0.1715966796, 0.2171487218}},
    {"18p43g", 28, 5, {0.1179924113, 0.1957371364, 0.2544060456}},
    {"18p45g", 28, 5, {0.1184247414, 0.2068045787, 0.2641082233}},
    {"18p47g", 28, 5, {0.115878332, 0.1842046578, 0.2407688057}},
    {"18p49g", 28, 5, {0.1248964218, 0.2141853759, 0.2675489103}},
    {"18p51g", 28, 5, {0.1106793836, 0.1621592148, 0.1979815532}},
    {"18p53g", 28, 5, {0.1082347513, 0.1408677318, 0.1771724651}},
    {"20p35g", 25, 5, {0.1112832877, 0.1466667286, 0.1741696491}},
    {"20p37g", 25, 5, {0.1134511409, 0.1501236494, 0.193180875}},
    {"20p39g", 25, 5, {0.1233401001, 0.1853070864, 0.2324952723}},
    {"20p41g", 25, 5, {0.1088138918, 0.1409300313, 0.1778018484}},
    {"20p43g", 25, 5, {0.1097495426, 0.1408421343, 0.1813610542}},
    {"20p45g", 25, 5, {0.1101110565, 0.1402262486, 0.1803401526}},
    {"20p47g", 25, 5, {0.1198508511, 0.175230183, 0.2232474228}},
    {"20p49g", 25, 5, {0.1150022018, 0.1500562415, 0.1891763921}},
    {"20p51g", 25, 5, {0.1101356126, 0.1310353805, 0.1457676446}},
    {"20p53g", 25, 5, {0.1239625969, 0.1752022586, 0.210665934}},
    {"25p35g", 22, 5, {0.0965155465, 0.0677355511, 0.0964040436}},
    {"25p37g", 22, 5, {0.0989264764, 0.076371967, 0.1025772304}},
    {"25p39g", 22, 5, {0.1005843197, 0.0805691511, 0.1108913867}},
    {"25p41g", 22, 5, {0.1030661339, 0.0840485952, 0.1237017673}},
    {"25p43g", 22, 5, {0.115543436, 0.1226554129, 0.1495157932}},
    {"25p45g", 22, 5, {0.0973402625, 0.074202019, 0.1029029688}},
    {"25p47g", 22, 5, {0.0967346256, 0.0743885008, 0.0995494785}},
    {"25p49g", 22, 5, {0.1070577141, 0.0941947942, 0.1181023643}},
    {"25p51g", 22, 5, {0.1044972455, 0.0789036639, 0.1081508568}},
    {"25p53g", 22, 5, {0.0977828435, 0.0607500307, 0.0834178835}},
    {"30p53g", 22, 5, {0.1204393447, 0.1065198197, 0.1243912966}},
    {"comparison", 20, 5, {0.09836223, 0.0789918084, 0.1116129577}},
};
bool ieq(const std::string& a, const std::string& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
  return true;
}
KaStats sentinel() { return KaStats{KA_SENTINEL, -1, -1, -1, -1, 0, 0}; }
KaStats mode2(int src, double lambda, double k, double alpha, double beta) {  // ka_stats.rs:300-312
  return KaStats{src, lambda, k, std::log(k), lambda / alpha, alpha, beta};
}
}  // namespace

// matrix.rs:100-104, 248-266
bool karlin_comment(const std::string& path, bool& found, double kv[5], std::string& err) {
  found = false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos || line[a] != '#') continue;
    std::string t = line.substr(a);
    if (t.find("FREQS") != std::string::npos) continue;
    std::string rest;
    if (t.rfind("# KARLIN", 0) == 0) rest = t.substr(8);
    else if (t.rfind("#KARLIN", 0) == 0) rest = t.substr(7);
    else continue;
    found = true;
    for (int i = 0; i < 5; ++i) kv[i] = 0;
    std::istringstream is(rest);
    std::vector<std::string> tok;
    for (std::string w; is >> w;) tok.push_back(w);
    for (size_t i = 0; i + 1 < tok.size(); i += 2) {
      double v;
      try {
        size_t used = 0;
        v = std::stod(tok[i + 1], &used);
        if (used != tok[i + 1].size()) throw std::invalid_argument("x");
      } catch (...) {
        err = "bad KARLIN value '" + tok[i + 1] + "' for key '" + tok[i] + "'";
        return false;
      }
      std::string key = tok[i];
      for (auto& ch : key) ch = (char)std::tolower((unsigned char)ch);
      if (key == "lambda") kv[0] = v;
      else if (key == "k") kv[1] = v;
      else if (key == "h") kv[2] = v;
      else if (key == "alpha") kv[3] = v;
      else if (key == "beta") kv[4] = v;
      else { err = "unknown KARLIN key '" + key + "'"; return false; }
    }
  }
  return true;
}

// ka_stats.rs:153-300 resolve (without the ALP fit) + rmstats kbp_for_recipe
bool resolve_ka_stats(const std::string& matrix_path, int go, int ge, const KaOptions& o, KaStats& out,
                      std::string& err) {
  const size_t slash = matrix_path.rfind('/');
  const std::string base = slash == std::string::npos ? matrix_path : matrix_path.substr(slash + 1);
  const StatRow* row = nullptr;
  for (const StatRow& r : kStatRows)
    if (ieq(std::string(r.stem) + ".matrix", base)) { row = &r; break; }
  const bool hit = row && row->go == go && row->ge == ge;
  const bool cli = o.lambda > 0 && o.k > 0 && o.alpha > 0;
  auto from_table = [&]() {
    KaStats s{KA_TABLE, row->lkh[0], row->lkh[1], std::log(row->lkh[1]), row->lkh[2], 0, 0};
    // eff_lengths.rs:68-79: user alpha/beta win, else alpha = lambda/H, beta 0
    s.alpha = cli ? o.alpha : s.lambda / s.h;
    s.beta = cli ? o.beta : 0.0;
    return s;
  };
  if (o.ncbi_compat) {
    if (hit && !ieq(base, "comparison.matrix")) {
      out = ieq(base, "30p53g.matrix") ? mode2(KA_TABLE, 0.1, 0.1, 1.0, 0.0) : from_table();
      return true;
    }
    out = cli ? mode2(KA_CLI, o.lambda, o.k, o.alpha, o.beta) : sentinel();
    return true;
  }
  if (hit) { out = from_table(); return true; }
  if (cli) { out = mode2(KA_CLI, o.lambda, o.k, o.alpha, o.beta); return true; }
  bool found = false;
  double kv[5];
  if (!karlin_comment(matrix_path, found, kv, err)) return false;
  if (found) {
    double alpha = kv[3] > 0 ? kv[3] : kv[2] > 0 ? kv[0] / kv[2] : 0;
    if (kv[0] > 0 && kv[1] > 0 && alpha > 0) { out = mode2(KA_COMMENT, kv[0], kv[1], alpha, kv[4]); return true; }
  }
  out = sentinel();
  return true;
}

// rmstats/src/length_adjust.rs:7-93 (BLAST_ComputeLengthAdjustment)
int compute_length_adjustment(double K, double logK, double a_d_l, double beta, int qlen, int64_t dblen, int nseq) {
  const double m = qlen, n = (double)dblen, N = nseq;
  double lo = 0, hi, next = 0;
  bool conv = false;
  const double A = N, B = m * N + n, C = n * m - (m > n ? m : n) / K;
  if (C < 0) return 0;
  hi = 2 * C / (B + std::sqrt(B * B - 4 * A * C));
  for (int it = 1; it <= 20; ++it) {
    const double ell = next;
    const double bar = a_d_l * (logK + std::log((m - ell) * (n - N * ell))) + beta;
    if (bar >= ell) {
      lo = ell;
      if (bar - lo <= 1.0) { conv = true; break; }
      if (lo == hi) break;
    } else {
      hi = ell;
    }
    next = (lo <= bar && bar <= hi) ? bar : (it == 1 ? hi : (lo + hi) / 2);
  }
  int la = (int)lo;
  if (conv) {
    const double c = std::ceil(lo);
    if (c <= hi && a_d_l * (logK + std::log((m - c) * (n - N * c))) + beta >= c) la = (int)c;
  }
  return la;
}

// rmstats/src/eff_lengths.rs:52-128
int64_t effective_search_space(const KaStats& s, int qlen, int64_t dblen, int nseq, int* la_out) {
  int la = 0;
  int64_t sp = 0;
  if (qlen > 0 && s.lambda > 0) {
    la = compute_length_adjustment(s.k, s.log_k, s.alpha / s.lambda, s.beta, qlen, dblen, nseq);
    int64_t eff_db = dblen - (int64_t)nseq * la;
    if (eff_db <= 0) eff_db = 1;
    sp = eff_db * (int64_t)(qlen - la);
  }
  if (la_out) *la_out = la;
  return sp;
}

// rmstats/src/evalue.rs:96-126
double evalue_of(const KaStats& s, int64_t searchsp, int score) {
  if (s.lambda <= 0) return 1.0;
  if (s.k < 0 || s.h < 0) return -1.0;
  return (double)searchsp * std::exp(-s.lambda * score + s.log_k);
}
double bit_score_of(const KaStats& s, int score) {
  if (s.lambda <= 0) return 0.0;
  return (score * s.lambda - s.log_k) / std::log(2.0);
}

// output.rs:256-288
std::string fmt_evalue(double e) {
  char b[64];
  if (e < 1.0e-180) return "0.0";
  if (e < 0.0009) snprintf(b, sizeof b, "%.2e", e);
  else if (e < 0.1) snprintf(b, sizeof b, "%.3f", e);
  else if (e < 1.0) snprintf(b, sizeof b, "%.2f", e);
  else if (e < 10.0) snprintf(b, sizeof b, "%.1f", e);
  else snprintf(b, sizeof b, "%.0f", e);
  return b;
}
std::string fmt_bits(double v) {
  char b[64];
  if (v > 99999.0) snprintf(b, sizeof b, "%.3e", v);
  else if (v > 99.9) snprintf(b, sizeof b, "%3ld", (long)v);
  else snprintf(b, sizeof b, "%4.1f", v);
  return b;
}

}  // namespace orc
