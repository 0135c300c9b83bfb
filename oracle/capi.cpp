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

// ungapped.rs:223-306; out = {score, q_start, q_end, s_start, s_end}
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
