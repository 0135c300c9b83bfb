// CPU ORACLE (test infrastructure): the Karlin-Altschul ungapped cutoff of
// rmblast-lib/src/search/ka_cutoff.rs, restated.
#include <cctype>
#include <cmath>
#include <cstring>
#include <string>

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

}  // namespace orc
