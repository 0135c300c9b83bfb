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
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

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

}  // namespace rmb
