// CPU ORACLE (test infrastructure): encoding, matrix parsing, NcbiRandom,
// FASTA and subject preparation.  Restates rmblast-lib/src/encoding.rs,
// matrix.rs, seq/twobit.rs:50-134, seq/fasta.rs, seq/fasta_db.rs:86-132 and
// seq/mod.rs:38-100.
#include "oracle.hpp"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace orc {

// encoding.rs:19-52 — unknown -> N (14); '-' and 'X' in sequences are N.
static uint8_t iupac_code(unsigned char c) {
  switch (std::toupper(c)) {
    case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': return 3; case 'U': return 3;
    case 'R': return 4; case 'Y': return 5; case 'M': return 6; case 'K': return 7; case 'W': return 8;
    case 'S': return 9; case 'B': return 10; case 'D': return 11; case 'H': return 12; case 'V': return 13;
    default: return 14;
  }
}
const char kBlastnaToIupac[17] = "ACGTRYMKWSBDHVN-";
// encoding.rs:68-85
const uint8_t kComplement[16] = {3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15};

std::vector<uint8_t> encode_iupac(const std::string& s) {  // encoding.rs:125-133
  std::vector<uint8_t> out;
  out.reserve(s.size() + 2);
  out.push_back(15);
  for (unsigned char c : s) out.push_back(c < 128 ? iupac_code(c) : 14);
  out.push_back(15);
  return out;
}

std::vector<uint8_t> revcomp(const std::vector<uint8_t>& s) {  // encoding.rs:136-141
  std::vector<uint8_t> out(s.size());
  for (size_t i = 0; i < s.size(); ++i) out[i] = kComplement[s[s.size() - 1 - i] & 15];
  return out;
}

// ---------------------------------------------------------------- matrix ----
static uint8_t matrix_symbol(uint8_t sym) {  // matrix.rs:239-246
  switch (sym) {
    case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': return 3;
    case 'R': return 4; case 'Y': return 5; case 'M': return 6; case 'K': return 7; case 'W': return 8;
    case 'S': return 9; case 'B': return 10; case 'D': return 11; case 'H': return 12; case 'V': return 13;
    case 'N': return 14; default: return 15;
  }
}

static std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> t;
  std::istringstream is(s);
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

static bool parse_int(const std::string& t, int32_t& v) {
  if (t.empty()) return false;
  size_t i = (t[0] == '-' || t[0] == '+') ? 1 : 0;
  if (i >= t.size()) return false;
  for (size_t k = i; k < t.size(); ++k)
    if (!std::isdigit((unsigned char)t[k])) return false;
  v = (int32_t)std::stol(t);
  return true;
}

// matrix.rs:306-354 — doubling then bisection to 1e-5; returns last midpoint.
static double estimate_lambda(const int32_t s[16][16], const double f[16]) {
  double tot = f[0] + f[1] + f[2] + f[3];
  if (tot < 1e-9) return 0.0;
  auto eval = [&](double lam) {
    double sum = 0;
    for (int i = 0; i < 16; ++i)
      for (int j = 0; j < 16; ++j)
        if (f[i] > 0.0 && f[j] > 0.0) sum += f[i] * f[j] * std::exp(lam * (double)s[i][j]);
    return sum;
  };
  double lo = 0.0, lam = 0.5;
  for (;;) {
    double v = eval(lam);
    if (v < 1.0) {
      lo = lam;
      lam *= 2.0;
      if (lam > 1000.0) return 0.0;
    } else {
      break;
    }
  }
  double hi = lam;
  while (hi - lo > 0.00001) {
    lam = (lo + hi) / 2.0;
    if (eval(lam) >= 1.0) hi = lam; else lo = lam;
  }
  return lam;
}

bool matrix_from_text(const std::string& name, const std::string& text, Matrix& m, std::string& err) {
  // matrix.rs:76-213
  double freqs[16] = {0};
  bool have_freqs = false;
  std::vector<uint8_t> header;
  std::vector<std::vector<int32_t>> rows;
  std::istringstream is(text);
  std::string line;
  int lineno = 0;
  while (std::getline(is, line)) {
    ++lineno;
    size_t a = line.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) continue;
    std::string t = line.substr(a);
    while (!t.empty() && std::isspace((unsigned char)t.back())) t.pop_back();
    if (t[0] == '#') {
      size_t at = t.find("FREQS");
      if (at != std::string::npos) {
        if (have_freqs) { err = "second FREQS line"; return false; }
        auto tok = split_ws(t.substr(at + 5));
        if (tok.size() != 8) { err = "FREQS needs 4 pairs"; return false; }
        bool seen[4] = {false, false, false, false};
        for (int k = 0; k < 4; ++k) {
          int idx = -1;
          char c = (char)std::toupper((unsigned char)tok[2 * k][0]);
          if (tok[2 * k].size() == 1) idx = (c == 'A') ? 0 : (c == 'C') ? 1 : (c == 'G') ? 2 : (c == 'T') ? 3 : -1;
          if (idx < 0 || seen[idx]) { err = "bad FREQS base"; return false; }
          seen[idx] = true;
          char* endp = nullptr;
          double v = std::strtod(tok[2 * k + 1].c_str(), &endp);
          if (*endp != 0 || !std::isfinite(v) || v < 0) { err = "bad FREQS value"; return false; }
          freqs[idx] = v;
        }
        double sum = freqs[0] + freqs[1] + freqs[2] + freqs[3];
        if (!(sum >= 0.999 && sum <= 1.001)) { err = "FREQS do not sum to 1"; return false; }
        have_freqs = true;
      }
      continue;
    }
    auto tok = split_ws(t);
    if (header.empty()) {
      for (auto& w : tok) {
        if (w.size() != 1 || !(std::isalpha((unsigned char)w[0]) || w[0] == '-')) {
          err = "bad matrix header at line " + std::to_string(lineno);
          return false;
        }
        uint8_t sym = (uint8_t)std::toupper((unsigned char)w[0]);
        for (uint8_t h : header) if (h == sym) { err = "repeated header symbol"; return false; }
        header.push_back(sym);
      }
      continue;
    }
    if (rows.size() == header.size()) { err = "more rows than columns"; return false; }
    size_t first = 0;
    int32_t tmp;
    if (!parse_int(tok[0], tmp)) {
      if (tok[0].size() != 1 || std::toupper((unsigned char)tok[0][0]) != header[rows.size()]) {
        err = "row label does not match header";
        return false;
      }
      first = 1;
    }
    std::vector<int32_t> r;
    for (size_t k = first; k < tok.size(); ++k) {
      if (!parse_int(tok[k], tmp)) { err = "score is not an integer at line " + std::to_string(lineno); return false; }
      r.push_back(tmp);
    }
    if (r.size() != header.size()) { err = "row length mismatch"; return false; }
    rows.push_back(r);
  }
  if (header.empty()) { err = "no column header"; return false; }
  if (rows.size() != header.size()) { err = "matrix must be square"; return false; }
  if (!have_freqs) { err = "no FREQS line"; return false; }
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) m.s[i][j] = -32768;   // BLAST_SCORE_MIN (matrix.rs:175-183)
  for (size_t ri = 0; ri < rows.size(); ++ri) {
    uint8_t r = matrix_symbol(header[ri]);
    for (size_t ci = 0; ci < rows[ri].size(); ++ci) m.s[r][matrix_symbol(header[ci])] = rows[ri][ci];
  }
  for (int i = 0; i < 16; ++i) {                      // matrix.rs:187-190
    m.s[15][i] = INT32_MIN / 2;
    m.s[i][15] = INT32_MIN / 2;
  }
  double expected = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) expected += freqs[i] * freqs[j] * (double)m.s[i][j];
  if (expected >= 0.0) { err = "expected score must be negative"; return false; }
  for (int i = 0; i < 16; ++i) m.freqs[i] = freqs[i];
  m.lambda = estimate_lambda(m.s, m.freqs);
  m.name = name;
  return true;
}

bool matrix_from_file(const std::string& path, Matrix& m, std::string& err) {
  std::ifstream f(path);
  if (!f) { err = "cannot open matrix " + path; return false; }
  std::stringstream ss;
  ss << f.rdbuf();
  return matrix_from_text(path, ss.str(), m, err);
}

// ----------------------------------------------------------- NcbiRandom -----
NcbiRandom::NcbiRandom(uint32_t seed) {  // twobit.rs:80-95
  state[0] = seed;
  for (int i = 1; i < 33; ++i) state[i] = 1103515245u * state[i - 1] + 12345u;
  rj = 12;
  rk = 32;
  for (int i = 0; i < 330; ++i) rand32();
}

uint32_t NcbiRandom::rand32() {  // twobit.rs:97-109
  uint32_t r = state[rk] + state[rj];
  state[rk] = r;
  --rk;
  --rj;
  if (rk < 0) rk = 32;
  else if (rj < 0) rj = 32;
  return r;
}

uint8_t NcbiRandom::next_base_iupac(uint8_t code) {  // twobit.rs:121-130, encoding.rs:97-114
  static const uint8_t allowed[16][4] = {
      {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {2, 0, 2, 0}, {2, 1, 3, 0},
      {2, 0, 1, 0}, {2, 2, 3, 0}, {2, 0, 3, 0}, {2, 1, 2, 0}, {3, 1, 2, 3}, {3, 0, 2, 3},
      {3, 0, 1, 3}, {3, 0, 1, 2}, {4, 0, 0, 0}, {0, 0, 0, 0}};
  uint32_t r = get_rand();
  if (code == 14) return (uint8_t)(r & 3);
  uint32_t count = allowed[code][0];
  return allowed[code][1 + r % count];
}

// ---------------------------------------------------------------- FASTA -----
bool read_fasta(const std::string& path, std::vector<FastaRecord>& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { err = "cannot open " + path; return false; }
  std::string line, defline, seq;
  bool have = false;
  auto flush = [&]() {
    if (!have) return;
    FastaRecord r;
    r.defline = defline;
    std::istringstream is(defline);
    is >> r.id;
    r.seq = encode_iupac(seq);
    out.push_back(std::move(r));
  };
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (!line.empty() && line[0] == '>') {
      flush();
      defline = line.substr(1);
      seq.clear();
      have = true;
      continue;
    }
    size_t a = line.find_first_not_of(" \t");
    if (a == std::string::npos) continue;
    size_t b = line.find_last_not_of(" \t");
    if (!have) { err = "sequence before header"; return false; }
    seq.append(line, a, b - a + 1);
  }
  flush();
  return true;
}

std::vector<uint8_t> resolve_ambiguity(const std::vector<uint8_t>& raw) {  // fasta_db.rs:86-105
  std::vector<uint8_t> out = raw;
  if (out.size() < 2) return out;
  uint32_t dna_size = (uint32_t)(out.size() - 2);
  NcbiRandom* rng = nullptr;
  NcbiRandom storage(0);
  for (size_t i = 1; i + 1 < out.size(); ++i) {
    uint8_t c = out[i];
    if (c >= 4 && c <= 14) {
      if (!rng) { storage = NcbiRandom(dna_size); rng = &storage; }
      out[i] = rng->next_base_iupac(c);
    }
  }
  return out;
}

std::vector<uint8_t> n_mask_of(const std::vector<uint8_t>& raw) {  // fasta_db.rs:115-132
  std::vector<uint8_t> m;
  if (raw.size() < 2) return m;
  bool any = false;
  for (size_t i = 1; i + 1 < raw.size(); ++i) any |= (raw[i] >= 4 && raw[i] <= 14);
  if (!any) return m;
  m.assign(raw.size() - 2, 0);
  for (size_t i = 1; i + 1 < raw.size(); ++i)
    if (raw[i] >= 4 && raw[i] <= 14) m[i - 1] = raw[i];
  return m;
}

PreparedSubject prepare_subject(const std::vector<uint8_t>& raw) {  // seq/mod.rs:64-100
  PreparedSubject p;
  p.plus = resolve_ambiguity(raw);
  p.rc = revcomp(p.plus);
  p.n_mask = n_mask_of(raw);
  if (p.n_mask.empty()) {
    p.align = p.plus;
    p.align_rc = p.rc;
  } else {
    p.align = p.plus;
    for (size_t i = 0; i < p.n_mask.size(); ++i)
      if (p.n_mask[i] != 0) p.align[i + 1] = p.n_mask[i];
    p.align_rc = revcomp(p.align);
    p.n_mask_rc.resize(p.n_mask.size());
    for (size_t i = 0; i < p.n_mask.size(); ++i) {
      uint8_t c = p.n_mask[p.n_mask.size() - 1 - i];
      p.n_mask_rc[i] = c == 0 ? 0 : kComplement[c];
    }
  }
  return p;
}

}  // namespace orc
