// CPU ORACLE (test infrastructure): query lookup tables.
// Restates rmblast-lib/src/blast/nalookup.rs:128-335 (small table, entries in
// ascending query offset), blast/mblookup.rs:106-245 (megablast table, LIFO
// chains = descending query offset) and engine.rs:630-690 (combined two-context
// buffer [fwd | N | rc]).  Both tables index every lut_width window of a
// location that contains no code >= 4; a location shorter than word_length is
// skipped (nalookup.rs:134-141, mblookup.rs:137-147).
#include "../rmblast_amd/csrc/rmb_dust.h"
#include "oracle.hpp"

#include <algorithm>

namespace orc {

int choose_lut_width(int ws, size_t e) {  // nalookup.rs:254-291
  switch (ws) {
    case 4: case 5: case 6: return ws;
    case 7: return e < 250 ? 6 : 7;
    case 8: return e < 8500 ? 7 : 8;
    case 9: return e < 1250 ? 7 : (e < 21000 ? 8 : 9);
    case 10: return e < 1250 ? 7 : (e < 8500 ? 8 : (e < 18000 ? 9 : 10));
    case 11: return e < 12000 ? 8 : (e < 180000 ? 10 : 11);
    case 12: return e < 8500 ? 8 : (e < 18000 ? 9 : (e < 60000 ? 10 : (e < 900000 ? 11 : 12)));
    default: return e < 8500 ? 8 : (e < 300000 ? 11 : 12);  // 13..
  }
}

static Lookup build_table(const std::vector<uint8_t>& seq, const std::vector<std::pair<int, int>>& locs, int word_length,
                          int lw, bool mega) {
  Lookup L;
  L.lut_width = lw;
  L.word_length = word_length;
  L.scan_step = word_length - lw + 1;
  L.mega = mega;
  const size_t cells = (size_t)1 << (2 * lw);
  const uint32_t mask = (uint32_t)(cells - 1);
  std::vector<uint32_t> count(cells + 1, 0);
  // every (index, offset) pair in position order
  std::vector<std::pair<uint32_t, uint32_t>> words;
  for (auto [from, to] : locs) {
    if (word_length > to - from + 1) continue;
    uint32_t code = 0;
    int valid = 0;
    for (int pos = from; pos <= to; ++pos) {
      uint8_t b = seq[pos];
      if (b & 0xFC) { valid = 0; code = 0; continue; }
      code = ((code << 2) | b) & mask;
      if (++valid >= lw) words.emplace_back(code, (uint32_t)(pos - lw + 1));
    }
  }
  for (auto& w : words) count[w.first + 1]++;
  L.cell_start.assign(cells + 1, 0);
  for (size_t i = 0; i < cells; ++i) L.cell_start[i + 1] = L.cell_start[i] + count[i + 1];
  L.entries.assign(words.size(), 0);
  std::vector<uint32_t> fill(L.cell_start.begin(), L.cell_start.end() - 1);
  for (auto& w : words) L.entries[fill[w.first]++] = w.second;
  int longest = 0;
  for (size_t i = 0; i < cells; ++i) {
    uint32_t a = L.cell_start[i], b = L.cell_start[i + 1];
    longest = std::max<int>(longest, (int)(b - a));
    // megablast chains are LIFO (mblookup.rs:170 prepends): descending offsets
    if (mega) std::reverse(L.entries.begin() + a, L.entries.begin() + b);
  }
  L.longest_chain = longest;
  return L;
}

QueryLookup build_query_lookup(const std::vector<uint8_t>& query, const Params& p) {  // engine.rs:606-662
  QueryLookup ql;
  if (p.dust && query.size() > 2) {  // filter/dust.rs:462-484, mask value 14
    ql.masked = query;
    rmb::dust_mask(ql.masked.data() + 1, ql.masked.size() - 2);
  }
  const std::vector<uint8_t>& q = ql.masked.empty() ? query : ql.masked;
  size_t n = q.size() >= 2 ? q.size() - 2 : 0;
  ql.fwd_n = (uint32_t)n;
  ql.c1_start = (uint32_t)n + 1;
  std::vector<uint8_t> combined;
  combined.reserve(2 * n + 1);
  for (size_t i = 0; i < n; ++i) combined.push_back(q[1 + i]);
  combined.push_back(14);
  for (size_t i = 0; i < n; ++i) combined.push_back(kComplement[q[n - i] & 15]);
  int lw = choose_lut_width(p.word_size, 2 * n);
  std::vector<std::pair<int, int>> locs = {{0, (int)n - 1}, {(int)n + 1, 2 * (int)n}};
  ql.lut = build_table(combined, locs, p.word_size, lw, lw > 8);
  return ql;
}

}  // namespace orc
