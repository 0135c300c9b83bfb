// rmblastn front end over librmblast_hip (include/rmblast_hip.h).
//
// Keeps the reference CLI's flags and NCBI spellings (rmblastn/src/main.rs:64-281)
// its outfmt 6 fields (rmblast-lib/src/output.rs:142-363) and its pairwise
// output, outfmt 0, the default (output.rs:365-720).  The GPU library
// returns each (query, subject) pair's final HSPs; this program does what the
// reference does after the search: mask level across subjects
// (engine.rs:863-974), NCBI output order (main.rs:1230-1234, engine.rs:998-1017),
// alignment statistics (stats.rs:52-243) and the writers.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iterator>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/rmblast_hip.h"

namespace {

const char kIupac[] = "ACGTRYMKWSBDHVN-";
const uint8_t kComp[16] = {3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15};

uint8_t code_of(char ch) {
  switch (std::toupper((unsigned char)ch)) {
    case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': case 'U': return 3;
    case 'R': return 4; case 'Y': return 5; case 'M': return 6; case 'K': return 7; case 'W': return 8;
    case 'S': return 9; case 'B': return 10; case 'D': return 11; case 'H': return 12; case 'V': return 13;
    default: return 14;  // N, X, '-' and anything else read as N (encoding.rs:19-52)
  }
}

// defline: the header line after '>' with its line ending removed
struct Record { std::string id, defline; std::vector<uint8_t> bases; };

bool load_fasta(const std::string& path, std::vector<Record>& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { err = "cannot open " + path; return false; }
  std::string line;
  Record cur;
  bool open = false;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (!line.empty() && line[0] == '>') {
      if (open) out.push_back(std::move(cur));
      cur = Record{};
      cur.defline = line.substr(1);
      std::istringstream is(cur.defline);
      is >> cur.id;
      if (cur.id.empty()) { err = "'>' with no sequence id"; return false; }
      open = true;
      continue;
    }
    size_t a = line.find_first_not_of(" \t"), b = line.find_last_not_of(" \t");
    if (a == std::string::npos) continue;
    if (!open) { err = "sequence data before the first header"; return false; }
    for (size_t i = a; i <= b; ++i) cur.bases.push_back(code_of(line[i]));
  }
  if (open) out.push_back(std::move(cur));
  return true;
}

// UCSC .2bit database (seq/twobit.rs:187-300): header, index of names and
// record offsets, then per record dnaSize, N blocks, mask blocks (ignored: case
// does not matter), a reserved word and the bases packed four to a byte, T C A
// G = 0 1 2 3 from the high bits.  N-block bases read as N, which the library
// resolves exactly as the reference resolves them (one NcbiRandom draw per
// position, seeded by the length, 5' to 3': twobit.rs:346-384); blocks must
// therefore be ascending and disjoint, as faToTwoBit writes them.
bool load_2bit(const std::string& path, std::vector<Record>& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { err = "cannot open " + path; return false; }
  const std::string d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  bool swap = false;
  auto r32 = [&](size_t pos, uint32_t& v) {
    if (pos > d.size() || d.size() - pos < 4) return false;
    uint32_t x;
    memcpy(&x, d.data() + pos, 4);
    v = swap ? __builtin_bswap32(x) : x;
    return true;
  };
  uint32_t magic = 0, version = 0, count = 0;
  if (!r32(0, magic)) { err = path + ": shorter than the 16-byte 2bit header"; return false; }
  if (magic == 0x4327411Au) swap = true;
  else if (magic != 0x1A412743u) { err = path + ": not a 2bit file"; return false; }
  if (!r32(4, version) || !r32(8, count) || d.size() < 16) { err = path + ": truncated 2bit header"; return false; }
  if (version != 0) { err = path + ": unsupported 2bit version"; return false; }
  if (count > (d.size() - 16) / 5) { err = path + ": 2bit index larger than the file"; return false; }
  static const uint8_t kFromUcsc[4] = {3, 1, 0, 2};
  std::vector<std::pair<std::string, uint32_t>> index;
  std::set<std::string> seen;
  size_t pos = 16;
  for (uint32_t i = 0; i < count; ++i) {
    if (pos >= d.size()) { err = path + ": truncated 2bit index"; return false; }
    const size_t nl = (uint8_t)d[pos++];
    uint32_t off = 0;
    if (d.size() - pos < nl || !r32(pos + nl, off)) { err = path + ": truncated 2bit index"; return false; }
    index.emplace_back(d.substr(pos, nl), off);
    if (!seen.insert(index.back().first).second) { err = path + ": duplicate name " + index.back().first; return false; }
    pos += nl + 4;
  }
  for (const auto& [name, off] : index) {
    uint32_t n = 0, nb = 0, mb = 0, reserved = 0;
    if (!r32(off, n) || !r32((size_t)off + 4, nb)) { err = path + ": truncated record " + name; return false; }
    const size_t ns = (size_t)off + 8, nz = ns + 4ull * nb;
    if (!r32(nz + 4ull * nb, mb) || !r32(nz + 4ull * nb + 4 + 8ull * mb, reserved)) {
      err = path + ": truncated record " + name;
      return false;
    }
    const size_t dna = nz + 4ull * nb + 8 + 8ull * mb;
    if (n > 0x7FFFFFFFu || d.size() < dna || d.size() - dna < (n + 3ull) / 4) {
      err = path + ": record " + name + " runs past the end of the file";
      return false;
    }
    Record r;
    r.id = r.defline = name;
    r.bases.resize(n);
    for (uint32_t p = 0; p < n; ++p)
      r.bases[p] = kFromUcsc[((uint8_t)d[dna + p / 4] >> (6 - 2 * (p % 4))) & 3];
    uint32_t prev_end = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      uint32_t s = 0, z = 0;
      r32(ns + 4ull * b, s);
      r32(nz + 4ull * b, z);
      const uint32_t e = (uint32_t)std::min<uint64_t>((uint64_t)s + z, n), s0 = std::min(s, e);
      if (s0 < prev_end) { err = path + ": N blocks of " + name + " overlap or are out of order"; return false; }
      std::fill(r.bases.begin() + s0, r.bases.begin() + e, (uint8_t)14);
      prev_end = e;
    }
    out.push_back(std::move(r));
  }
  return true;
}

bool load_db(const std::string& path, std::vector<Record>& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  char m[4] = {0};
  if (f && f.read(m, 4) && ((uint8_t)m[0] == 0x43 && (uint8_t)m[1] == 0x27 && (uint8_t)m[2] == 0x41 && (uint8_t)m[3] == 0x1A))
    return load_2bit(path, out, err);
  if (f && (uint8_t)m[0] == 0x1A && (uint8_t)m[1] == 0x41 && (uint8_t)m[2] == 0x27 && (uint8_t)m[3] == 0x43)
    return load_2bit(path, out, err);
  return load_fasta(path, out, err);
}

struct Row {
  rmb_hsp h;
  uint32_t q_len, s_len;
  std::string qa, sa;  // aligned IUPAC strings
};

std::string norm(std::string a) {
  if (a.size() > 1 && a[0] == '-' && a[1] != '-') a = "-" + a;
  for (auto& c : a) if (c == '_') c = '-';
  return a;
}

// ---------------------------------------------------------------- outfmt 0 --

// 11207 -> "11,207" (output.rs:28-37)
std::string commas(uint64_t n) {
  const std::string d = std::to_string(n);
  std::string o;
  for (size_t i = 0; i < d.size(); ++i) {
    if (i && (d.size() - i) % 3 == 0) o += ',';
    o += d[i];
  }
  return o;
}

bool ascii_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\f' || c == '\r'; }

// NCBI's flat-file wrap of the database title (output.rs:44-102): each line
// ends at the best break within `width` columns -- a newline, else the last
// space, else just after the last comma, else just after the last hyphen,
// else wherever the width runs out; spaces after a space break are dropped.
std::vector<std::string> wrap_flat(const std::string& s, size_t width) {
  enum { kForced, kHyphen, kComma, kSpace, kNewline };
  std::vector<std::string> lines;
  const size_t len = s.size();
  size_t pos = 0;
  while (pos < len) {
    size_t nl = s.find('\n', pos);
    if (nl == std::string::npos) nl = len;
    const size_t start = nl - pos <= width ? nl : pos;
    size_t col = 0, at = SIZE_MAX, p = start;
    int best = kForced;
    for (; p < len && col <= width; ++p, ++col) {
      const char c = s[p];
      if (c == '\n') { at = p; best = kNewline; break; }
      int kind = kForced;
      size_t cut = p;
      if (ascii_space(c)) kind = kSpace;
      else if (c == ',' && col < width && p + 1 < len) { kind = kComma; ++cut; }
      else if (c == '-' && p + 1 < len && col < width) { kind = kHyphen; ++cut; }
      if (kind >= best && cut > start) { at = cut; best = kind; }
    }
    if (best != kNewline && col <= width) at = len;
    at = std::max(std::min(at, len), pos + 1);
    lines.push_back(s.substr(pos, at - pos));
    pos = at;
    if (best == kSpace) {
      while (pos < len && s[pos] == ' ') ++pos;
      if (pos < len && s[pos] == '\n') ++pos;
    } else if (best == kNewline) {
      ++pos;
    }
  }
  if (lines.empty()) lines.emplace_back();
  return lines;
}

// Subject title wrap (output.rs:108-126): once a line passes 60 characters,
// break after the next whitespace, which stays on the line.
std::string wrap_title(const std::string& s) {
  if (s.size() <= 60) return s;
  std::string o;
  bool due = false;
  for (size_t i = 0; i < s.size(); ++i) {
    if (i && i % 60 == 0) due = true;
    o += s[i];
    if (due && ascii_space(s[i])) { o += '\n'; due = false; }
  }
  return o;
}

// NCBI's GetPercentMatch (output.rs:369-374): 100 only when n == d
unsigned percent_match(uint32_t n, uint32_t d) {
  if (d == 0) return 0;
  if (n == d) return 100;
  return std::min(99u, (unsigned)(0.5 + 100.0 * n / d));
}

std::string header_block(const std::string& db_name, size_t n_seqs, uint64_t letters) {  // output.rs:431-452
  std::string o = "RMBLASTN 2.17.1+\n\n\n"
                  "Reference: Robert M. Hubley, Arian Smit\n"
                  "RMBlast - RepeatMasker Search Engine\n"
                  "2010 <http://www.repeatmasker.org>\n\n"
                  "Reference: Stephen F. Altschul, Thomas L. Madden, Alejandro A.\n"
                  "Schaffer, Jinghui Zhang, Zheng Zhang, Webb Miller, and David J.\n"
                  "Lipman (1997), \"Gapped BLAST and PSI-BLAST: a new generation of\n"
                  "protein database search programs\", Nucleic Acids Res. 25:3389-3402.\n\n\n\n"
                  "Database: ";
  for (const auto& l : wrap_flat(db_name, 68)) o += l + "\n";
  o += "           " + commas(n_seqs) + " sequences; " + commas(letters) + " total letters\n\n\n\n";
  return o;
}

// "Posted date" line value from the database file's modification time, in
// UTC: "Mon D, YYYY  H:MM AM/PM" (output.rs:383-420)
std::string posted_date(time_t t) {
  static const char* kMon[12] = {"Jan", "Feb", "Mar", "Apr", "May", "Jun", "Jul", "Aug", "Sep", "Oct", "Nov", "Dec"};
  struct tm u;
  gmtime_r(&t, &u);
  const int h12 = u.tm_hour % 12 == 0 ? 12 : u.tm_hour % 12;
  char b[64];
  snprintf(b, sizeof b, "%s %d, %d  %d:%02d %s", kMon[u.tm_mon], u.tm_mday, u.tm_year + 1900, h12, u.tm_min,
           u.tm_hour < 12 ? "AM" : "PM");
  return b;
}

std::string footer_block(const std::string& db_name, size_t n_seqs, uint64_t letters, const time_t* mtime,
                         int gap_open, int gap_extend) {  // output.rs:652-680
  std::string o = "  Database: ";
  for (const auto& l : wrap_flat(db_name, 68)) o += l + "\n";
  if (mtime) o += "    Posted date:  " + posted_date(*mtime) + "\n";
  o += "  Number of letters in database: " + commas(letters) + "\n";
  o += "  Number of sequences in database:  " + commas(n_seqs) + "\n\n\n\n";
  char b[96];
  snprintf(b, sizeof b, "Matrix: blastn matrix 0 0\nGap Penalties: Existence: %d, Extension: %.2f\n", gap_open,
           (double)gap_extend);
  return o + b;
}

// One HSP (output.rs:510-649): the statistics block, then 60-column rows of
// query, match line and subject.  Query bases inside the DUST intervals
// `mask` ([first, last] pairs, sorted) print in lower case.
void hsp_block(std::string& o, const Row& r, const rmb_align_stats& st, const std::vector<uint64_t>& mask) {
  const rmb_hsp& h = r.h;
  const uint32_t n = (uint32_t)r.qa.size(), gaps = st.query_gaps + st.subject_gaps;
  char b[256];
  snprintf(b, sizeof b,
           " Score = %d\n Substitutions = %.2f%%, Query Gaps = %.2f%%, DB Gaps = %.2f%%\n"
           " TransI/TransV = %u/%u,  CpG_sites = %u\n Kimura Div = %.2f%%, Kimura Div CpG Adjusted = %.2f%%\n",
           h.score, (double)st.perc_sub, (double)st.perc_query_gap, (double)st.perc_db_gap, st.transitions,
           st.transversions, st.cpg_sites, st.kdiv, st.cpg_kdiv);
  o += b;
  snprintf(b, sizeof b, " Identities = %u/%u (%u%%), Gaps = %u/%u (%u%%)\n Strand=%s\n\n", st.identities, n,
           percent_match(st.identities, n), gaps, n, percent_match(gaps, n), h.strand ? "Plus/Minus" : "Plus/Plus");
  o += b;
  std::string qa = r.qa;
  if (!mask.empty()) {
    const size_t nm = mask.size() / 2;
    size_t m = 0;
    uint64_t q = h.q_start;
    while (m < nm && mask[2 * m + 1] < q) ++m;
    for (char& c : qa) {
      if (c == '-') continue;
      while (m < nm && mask[2 * m + 1] < q) ++m;
      if (m < nm && mask[2 * m] <= q) c = (char)std::tolower((unsigned char)c);
      ++q;
    }
  }
  const uint32_t top = std::max(h.q_end, h.s_end);
  const int w = (int)std::to_string(top).size();
  uint32_t qp = h.q_start + 1, sp = h.strand ? h.s_end : h.s_start + 1;
  for (uint32_t i = 0; i < n; i += 60) {
    const uint32_t e = std::min(n, i + 60);
    uint32_t qn = 0, sn = 0;
    std::string mid;
    for (uint32_t k = i; k < e; ++k) {
      const char q = qa[k], s = r.sa[k];
      qn += q != '-';
      sn += s != '-';
      mid += (q != '-' && s != '-' && std::toupper((unsigned char)q) == std::toupper((unsigned char)s)) ? '|' : ' ';
    }
    const uint32_t qe = qn ? qp + qn - 1 : qp;
    const uint32_t se = sn ? (h.strand ? sp - (sn - 1) : sp + sn - 1) : sp;
    snprintf(b, sizeof b, "Query  %-*u  ", w, qp);
    o += b;
    o.append(qa, i, e - i);
    o += "  " + std::to_string(qe) + "\n" + std::string(9 + w, ' ') + mid + "\n";
    snprintf(b, sizeof b, "Sbjct  %-*u  ", w, sp);
    o += b;
    o.append(r.sa, i, e - i);
    o += "  " + std::to_string(se) + "\n\n";
    qp += qn;
    sp = h.strand ? (sp >= sn ? sp - sn : 0) : sp + sn;
  }
  o += '\n';
}

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, std::string> kv;
  bool complexity = false, ncbi_compat = false;
  for (int i = 1; i < argc; ++i) {
    const std::string f = norm(argv[i]);
    if (f == "--complexity-adjust") { complexity = true; continue; }
    if (f == "--ncbi-compat") { ncbi_compat = true; continue; }
    if (i + 1 >= argc) { fprintf(stderr, "error: %s needs a value\n", argv[i]); return 2; }
    kv[f] = argv[++i];
  }
  auto opt = [&](const char* k, const char* d) { auto it = kv.find(k); return it == kv.end() ? std::string(d) : it->second; };
  for (const char* req : {"--db", "--query", "--gapopen", "--gapextend"})
    if (!kv.count(req)) { fprintf(stderr, "error: %s is required\n", req); return 2; }
  rmb_params P{};
  P.gap_open = std::stoi(kv["--gapopen"]);
  P.gap_extend = std::stoi(kv["--gapextend"]);
  P.word_size = std::stoi(opt("--word-size", "11"));
  P.xdrop_ungap = std::stoi(opt("--xdrop-ungap", "20"));
  P.xdrop_gap = std::stoi(opt("--xdrop-gap", "30"));
  P.xdrop_gap_final = std::stoi(opt("--xdrop-gap-final", "100"));
  P.min_raw_gapped_score = std::stoi(opt("--min-raw-gapped-score", "0"));
  P.ungapped_cutoff = -1;  // per query record (main.rs:1154-1170)
  P.complexity_adjust = complexity;
  const std::string matrix_path = opt("--matrix", "");
  {
    double ka[2] = {0, 0};
    int32_t found = 0;
    rmb_ka_lookup(matrix_path.c_str(), P.gap_open, P.gap_extend, ka, &found);
    if (found) { P.ka_lambda = ka[0]; P.ka_k = ka[1]; }
  }
  // --dust yes|true|no|false (rmblastn/src/main.rs:330-342); NCBI's
  // 'level window linker' form is rejected
  const std::string dust = opt("--dust", "yes");
  int dust_on = 0;
  if (dust == "yes" || dust == "true") {
    dust_on = 1;
  } else if (dust != "no" && dust != "false") {
    fprintf(stderr, "error: --dust '%s': expected yes, no, true or false\n", dust.c_str());
    return 1;
  }
  P.dust = dust_on;
  int mask_level = std::stoi(opt("--mask-level", "-1"));
  if (mask_level < 0) mask_level = 101;
  // --outfmt: "0" pairwise (the default) or "6 <fields>" (main.rs:201-206,380-391)
  std::vector<std::string> fields;
  bool pairwise = false;
  {
    const std::string spec = opt("--outfmt", "0");
    std::istringstream is(spec);
    std::string w;
    is >> w;
    static const char* kFields[] = {"score", "perc_sub", "perc_query_gap", "perc_db_gap", "qseqid", "qstart",
                                    "qend", "qlen", "sstrand", "sseqid", "sstart", "send", "slen", "kdiv",
                                    "cpg_kdiv", "transi", "transv", "cpg_sites", "qseq", "sseq", "evalue",
                                    "bitscore"};
    std::string names;
    for (const char* f : kFields) names += std::string(names.empty() ? "" : " ") + f;
    if (w == "0") {
      pairwise = true;
      if (is >> w) { fprintf(stderr, "error: --outfmt '%s': format 0 (pairwise) takes no fields\n", spec.c_str()); return 1; }
    } else if (w != "6") {  // output.rs:223-242
      fprintf(stderr, "error: --outfmt '%s': output format '%s' is not supported; use 0 (pairwise) or 6 <fields> (tabular)\n",
              spec.c_str(), w.c_str());
      return 1;
    }
    while (is >> w) {
      if (std::find_if(std::begin(kFields), std::end(kFields), [&](const char* f) { return w == f; }) == std::end(kFields)) {
        fprintf(stderr, "error: --outfmt '%s': unknown -outfmt field '%s'. Supported fields: %s\n", spec.c_str(), w.c_str(),
                names.c_str());
        return 1;
      }
      fields.push_back(w);
    }
    if (!pairwise && fields.empty()) {
      fprintf(stderr, "error: --outfmt '%s': -outfmt 6 needs a field list; NCBI's default columns are not supported. "
              "Fields: %s\n", spec.c_str(), names.c_str());
      return 1;
    }
  }
  // E-value statistics options (main.rs:243-281): --evalue > 0, the -matrix_*
  // overrides finite (lambda, k and alpha not negative), --ncbi-compat
  rmb_ka_stats_options kopt{};
  kopt.ncbi_compat = ncbi_compat ? 1 : 0;
  double evalue = 0;
  bool evalue_given = false;
  {
    auto num = [&](const char* flag, double& v, bool nonneg) {
      if (!kv.count(flag)) return true;
      const std::string t = kv[flag];
      char* end = nullptr;
      v = std::strtod(t.c_str(), &end);
      if (t.empty() || *end || !std::isfinite(v)) {
        fprintf(stderr, "error: invalid value '%s' for '%s': not a finite number\n", t.c_str(), flag);
        return false;
      }
      if (nonneg && v < 0) {
        fprintf(stderr, "error: invalid value '%s' for '%s': '%s' must not be negative\n", t.c_str(), flag, t.c_str());
        return false;
      }
      return true;
    };
    if (!num("--matrix-lambda", kopt.matrix_lambda, true) || !num("--matrix-k", kopt.matrix_k, true) ||
        !num("--matrix-alpha", kopt.matrix_alpha, true) || !num("--matrix-beta", kopt.matrix_beta, false) ||
        !num("--evalue", evalue, false))
      return 2;
    evalue_given = kv.count("--evalue") > 0;
    if (evalue_given && !(evalue > 0)) {
      fprintf(stderr, "error: invalid value '%s' for '--evalue': '%s': the expect value must be greater than zero\n",
              kv["--evalue"].c_str(), kv["--evalue"].c_str());
      return 2;
    }
  }
  std::string err;
  rmb_matrix M;
  {
    char mbuf[256] = {0};
    if (rmb_matrix_load(matrix_path.c_str(), &M, mbuf, sizeof mbuf)) { fprintf(stderr, "error: matrix: %s\n", mbuf); return 1; }
  }
  std::vector<Record> db, qs;
  if (!load_db(kv["--db"], db, err) || !load_fasta(kv["--query"], qs, err)) {
    fprintf(stderr, "error: %s\n", err.c_str());
    return 1;
  }
  if (db.empty()) { fprintf(stderr, "error: database '%s' contains no sequences\n", kv["--db"].c_str()); return 1; }
  if (kv.count("--gilist")) {  // main.rs:578-600, 1444-1470: subjects named in the file, one per line
    std::ifstream f(kv["--gilist"]);
    if (!f) { fprintf(stderr, "error: opening gilist '%s'\n", kv["--gilist"].c_str()); return 1; }
    std::set<std::string> keep;
    std::string line;
    while (std::getline(f, line)) {
      const size_t a = line.find_first_not_of(" \t\r\n\f\v"), b = line.find_last_not_of(" \t\r\n\f\v");
      if (a == std::string::npos) continue;
      const std::string e = line.substr(a, b - a + 1);
      // a bare GI number also names the sequence "gi|<n>"
      if (std::all_of(e.begin(), e.end(), [](char c) { return c >= '0' && c <= '9'; })) keep.insert("gi|" + e);
      keep.insert(e);
    }
    std::vector<Record> sel;
    for (Record& r : db)
      if (keep.count(r.id)) sel.push_back(std::move(r));
    db.swap(sel);
    if (db.empty())
      fprintf(stderr, "Warning: [rmblastn] the gilist '%s' matches no sequence in the database; nothing to search\n",
              kv["--gilist"].c_str());
  }
  std::map<std::string, size_t> oid;
  for (size_t i = 0; i < db.size(); ++i)
    if (!oid.emplace(db[i].id, i).second) { fprintf(stderr, "error: duplicate subject id %s\n", db[i].id.c_str()); return 1; }
  rmb_ctx* ctx = nullptr;
  int rc = rmb_ctx_create(std::stoi(opt("--device", "0")), &P, &M, &ctx);
  auto fail = [&](const char* what) {
    char buf[512] = {0};
    if (ctx) rmb_last_error(ctx, buf, sizeof buf);
    fprintf(stderr, "error: %s: %s\n", what, buf);
    rmb_ctx_destroy(ctx);
    return 1;
  };
  if (rc) return fail("device");
  if (kv.count("--tuning-file")) {
    // the autotuner's table (rmblast_amd/autotune.py): the row for this
    // matrix file name, word size and X-drops, if any, sets the schedule
    std::ifstream f(kv["--tuning-file"]);
    if (!f) { fprintf(stderr, "error: cannot open tuning file '%s'\n", kv["--tuning-file"].c_str()); return fail("tuning"); }
    const size_t slash = matrix_path.rfind('/');
    const std::string base = slash == std::string::npos ? matrix_path : matrix_path.substr(slash + 1);
    std::string line;
    std::vector<std::string> head;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::vector<std::string> cols;
      std::stringstream ss(line);
      for (std::string c; std::getline(ss, c, '\t');) cols.push_back(c);
      if (head.empty()) { head = cols; continue; }
      std::map<std::string, std::string> row;
      for (size_t k = 0; k < head.size() && k < cols.size(); ++k) row[head[k]] = cols[k];
      auto num = [&](const char* k) { return row.count(k) ? std::atoi(row[k].c_str()) : 0; };
      if (row["matrix"] != base || num("word_size") != P.word_size || num("xdrop_gap") != P.xdrop_gap ||
          num("xdrop_gap_final") != P.xdrop_gap_final)
        continue;
      rmb_tuning t{num("batches"), num("prelim_ring"), num("prelim_wg_per_cu"), num("tb_ring"), num("tb_wg_per_cu"),
                    num("tb_wave_halves")};
      if (rmb_ctx_set_tuning(ctx, &t)) return fail("tuning");
      break;
    }
  }
  std::string out;
  uint64_t db_letters = 0;
  for (const Record& r : db) db_letters += r.bases.size();
  const std::string& db_name = kv["--db"];
  // Karlin-Altschul statistics, resolved when they are used: evalue/bitscore
  // fields or a reap threshold (--evalue, or NCBI's 10 under --ncbi-compat)
  // (main.rs:612-686)
  const bool stats_requested =
      !pairwise && std::any_of(fields.begin(), fields.end(), [](const std::string& f) { return f == "evalue" || f == "bitscore"; });
  const double reap_threshold = ncbi_compat ? (evalue_given ? evalue : 10.0) : (evalue_given ? evalue : 0.0);
  rmb_ka_stats ka{};
  bool have_ka = false, reap_on = false;
  if (stats_requested || reap_threshold > 0) {
    char kbuf[256] = {0};
    if (rmb_ka_stats_resolve(matrix_path.c_str(), P.gap_open, P.gap_extend, &kopt, &ka, kbuf, sizeof kbuf)) {
      fprintf(stderr, "error: matrix: %s\n", kbuf);
      return fail("statistics");
    }
    have_ka = true;
    reap_on = reap_threshold > 0 && ka.source != RMB_KA_SENTINEL;
    static const char* kDescribe[] = {"none (sentinel: evalue=1.0, bitscore=0.0)", "baked RMBlast table (NCBI Mode 1)",
                                      "-matrix_lambda/k/alpha CLI overrides (NCBI Mode 2)",
                                      "# KARLIN comment in matrix file"};
    fprintf(stderr, "# rmblastn: E-value statistics%s: %s\n", ncbi_compat ? " (--ncbi-compat 2.17.1 emulation)" : "",
            kDescribe[ka.source]);
    if (ka.source == RMB_KA_TABLE) {
      const size_t slash = matrix_path.rfind('/');
      fprintf(stderr, "#   Karlin-Altschul parameters for '%s' at gap costs %d/%d taken from the built-in table by "
              "matrix file name; the matrix contents are not checked\n",
              (slash == std::string::npos ? matrix_path : matrix_path.substr(slash + 1)).c_str(), P.gap_open, P.gap_extend);
    }
    if (ka.lambda > 0) fprintf(stderr, "#   lambda=%.10f K=%.10f H=%.10f\n", ka.lambda, ka.k, ka.h);
    const std::string t = reap_threshold > 0 ? (evalue_given ? kv["--evalue"] : std::string("10")) : "";
    if (reap_on) fprintf(stderr, "#   -evalue cutoff ACTIVE: threshold %s\n", t.c_str());
    else if (reap_threshold > 0)
      fprintf(stderr, "#   -evalue cutoff requested (threshold %s) but no statistics available \xe2\x80\x94 nothing culled\n",
              t.c_str());
    else fprintf(stderr, "#   -evalue cutoff: off (no --evalue given)\n");
  }
  std::vector<const char*> qids, sids;
  for (const Record& r : qs) qids.push_back(r.id.c_str());
  for (const Record& r : db) sids.push_back(r.id.c_str());
  // outfmt 6: the library writes the rows while the search runs (mask level,
  // output order, statistics, text: its host tail)
  std::string fl;
  for (const std::string& f : fields) fl += (fl.empty() ? "" : " ") + f;
  rmb_report_spec spec{};
  spec.mask_level = mask_level;
  spec.ka = have_ka ? &ka : nullptr;
  spec.evalue_max = reap_on ? reap_threshold : 0.0;
  spec.fields = fl.c_str();
  spec.query_ids = qids.data();
  spec.subject_ids = sids.data();
  uint64_t nh = 0, no = 0;
  std::vector<rmb_hsp> H;
  std::vector<rmb_op> O;
  if (!db.empty() && !qs.empty()) {  // a gilist may select no subject: a search with no hits
    std::vector<const uint8_t*> ptr;
    std::vector<uint32_t> len;
    for (auto& r : db) { ptr.push_back(r.bases.data()); len.push_back((uint32_t)r.bases.size()); }
    if (rmb_library_load(ctx, (uint32_t)db.size(), ptr.data(), len.data())) return fail("library");
    std::vector<const uint8_t*> qp;
    std::vector<uint32_t> ql;
    for (auto& r : qs) { qp.push_back(r.bases.data()); ql.push_back((uint32_t)r.bases.size()); }
    // The run-log note of a KA-derived ungapped cutoff, once, with the first
    // query's numbers (main.rs:1098-1131)
    if (P.ka_lambda > 0 && P.ka_k > 0) {
      uint64_t letters = 0;
      for (auto& r : db) letters += r.bases.size();
      int32_t cut = 0;
      rmb_ka_cutoff_info info{};
      rmb_ungapped_cutoff(matrix_path.c_str(), P.gap_open, P.gap_extend, P.min_raw_gapped_score, ql[0],
                          letters / db.size(), &cut, &info);
      if (info.found) {
        const size_t slash = matrix_path.rfind('/');
        const std::string base = slash == std::string::npos ? matrix_path : matrix_path.substr(slash + 1);
        fprintf(stderr, "# rmblastn: Karlin-Altschul ungapped cutoff IN EFFECT (matrix=%s gapopen=%d gapextend=%d)\n",
                base.c_str(), P.gap_open, P.gap_extend);
        fprintf(stderr, "#   lambda=%.10f K=%.10f E=0.05 avg_db_seq_len=%llu searchsp=%llu -> KA score=%d\n",
                info.lambda, info.k, (unsigned long long)(letters / db.size()), (unsigned long long)info.searchsp,
                info.es);
        fprintf(stderr, "#   ungapped_cutoff = %d (overrides default min_raw_gapped_score/2 = %d)\n", cut,
                P.min_raw_gapped_score / 2);
      }
    }
    if (!pairwise) {
      uint64_t nb = 0;
      if (rmb_search_report(ctx, (uint32_t)qs.size(), qp.data(), ql.data(), &spec, &nb)) return fail("search");
      out.resize(nb);
      if (rmb_report_copy(ctx, out.data())) return fail("report");
      rmb_results_size(ctx, &nh, &no);
    } else {
      if (rmb_search(ctx, (uint32_t)qs.size(), qp.data(), ql.data())) return fail("search");
      rmb_results_size(ctx, &nh, &no);
      H.resize(nh);
      O.resize(no);
      rmb_results_copy(ctx, H.data(), O.data());
    }
  }
  if (pairwise) {
    std::vector<uint64_t> order;
    uint64_t n_rows = 0;
    if (nh) {
      rmb_report_spec spec{};
      spec.mask_level = mask_level;
      spec.ka = have_ka ? &ka : nullptr;
      spec.evalue_max = reap_on ? reap_threshold : 0.0;
      if (rmb_report_order(ctx, &spec, nullptr, &n_rows)) return fail("report");
      order.resize(n_rows);
      if (rmb_report_order(ctx, &spec, order.data(), &n_rows)) return fail("report");
    }
    std::vector<std::vector<Row>> per_query(qs.size());
    for (uint64_t k : order) {
      const rmb_hsp& h = H[k];
      const Record& Q = qs[h.query];
      const Record& S = db[h.subject];
      Row r;
      r.h = h;
      r.q_len = (uint32_t)Q.bases.size();
      r.s_len = (uint32_t)S.bases.size();
      // aligned columns in the forward query frame; the minus subject is read reverse-complemented
      size_t qi = h.q_start, si = h.strand ? S.bases.size() - h.s_end : h.s_start;
      for (uint32_t j = 0; j < h.ops_count; ++j) {
        const uint32_t kindv = O[h.ops_offset + j] & 3, n = O[h.ops_offset + j] >> 2;
        for (uint32_t t = 0; t < n; ++t) {
          if (kindv != 1) r.qa += kIupac[Q.bases[qi++] & 15]; else r.qa += '-';
          if (kindv != 2) {
            const uint8_t b = h.strand ? kComp[S.bases[S.bases.size() - 1 - si] & 15] : (S.bases[si] & 15);
            ++si;
            r.sa += kIupac[b];
          } else {
            r.sa += '-';
          }
        }
      }
      per_query[h.query].push_back(std::move(r));
    }
    out += header_block(db_name, db.size(), db_letters);
    for (size_t qi = 0; qi < qs.size(); ++qi) {  // output.rs:685-720; rows are grouped by subject already
      const std::vector<Row>& rows = per_query[qi];
      std::string ql = qs[qi].defline;
      while (!ql.empty() && ascii_space(ql.back())) ql.pop_back();
      out += "Query= " + ql + "\n\nLength=" + std::to_string(qs[qi].bases.size()) + "\n\n";
      if (rows.empty()) {
        out += "\n***** No hits found *****\n\n\n\n\n";
        continue;
      }
      std::vector<uint64_t> mask;
      if (dust_on) {
        const Record& Q = qs[qi];
        uint64_t nm = 0;
        if (rmb_dust_intervals(Q.bases.data(), Q.bases.size(), nullptr, &nm)) return fail("dust");
        mask.resize(2 * nm);
        if (nm && rmb_dust_intervals(Q.bases.data(), Q.bases.size(), mask.data(), &nm)) return fail("dust");
      }
      for (size_t k = 0; k < rows.size(); ++k) {
        const Row& r = rows[k];
        if (k == 0 || r.h.subject != rows[k - 1].h.subject) {
          // ">id title": the defline past the id and one character (output.rs:487-508)
          const Record& S = db[r.h.subject];
          std::string title;
          if (S.defline.compare(0, S.id.size(), S.id) == 0 && S.defline.size() > S.id.size())
            title = S.defline.substr(S.id.size() + 1);
          out += ">" + S.id + " " + wrap_title(title) + "\nLength=" + std::to_string(r.s_len) + "\n\n";
        }
        rmb_align_stats st{};
        rmb_align_stats_of(r.qa.data(), r.sa.data(), r.qa.size(), &st);
        hsp_block(out, r, st, mask);
      }
      out += "\n\n";
    }
  }
  if (pairwise) {
    struct stat sb;
    const bool have = stat(db_name.c_str(), &sb) == 0;
    out += footer_block(db_name, db.size(), db_letters, have ? &sb.st_mtime : nullptr, P.gap_open, P.gap_extend);
  }
  fwrite(out.data(), 1, out.size(), stdout);
  if (std::getenv("RMB_STATS")) {
    uint64_t s[16] = {0};
    rmb_stats(ctx, s, 16);
    fprintf(stderr, "seed_evals %llu ungapped_hits %llu prelim_dps %llu tb_dps %llu dp_cells %llu hsps %llu "
            "ung_steps %llu t_seed_us %llu t_prelim_us %llu t_tb_us %llu t_total_us %llu\n",
            (unsigned long long)s[0], (unsigned long long)s[1], (unsigned long long)s[2], (unsigned long long)s[3],
            (unsigned long long)s[4], (unsigned long long)s[6], (unsigned long long)s[7], (unsigned long long)s[9],
            (unsigned long long)s[10], (unsigned long long)s[11], (unsigned long long)s[12]);
  }
  rmb_ctx_destroy(ctx);
  return 0;
}
