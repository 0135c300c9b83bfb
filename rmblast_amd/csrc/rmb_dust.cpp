// Symmetric DUST (BLAST filter path), restated from
// rmblast-lib/src/filter/dust.rs:1-484, which ports NCBI's CSymDustMasker
// (symdust.cpp).  The window slides over base triplets; a suffix of the window
// whose triplet-repeat score exceeds level x length is a "perfect" interval,
// and the best perfect interval leaving the window is reported, merged with
// the previous one when they are within `linker` bases.
#include "rmb_dust.h"

#include <algorithm>
#include <deque>

namespace rmb {
namespace {

// NCBI CRandom lagged-Fibonacci generator with its fixed initial state
// (dust.rs:23-67); N bases draw from one stream per query.
class DustRandom {
 public:
  uint32_t next() {
    const uint32_t r = s_[k_] + s_[j_];
    s_[k_] = r;
    if (--k_ < 0) k_ = 32;
    else if (--j_ < 0) j_ = 32;
    return r >> 1;
  }

 private:
  uint32_t s_[33] = {0xd53f1852, 0xdfc78b83, 0x4f256096, 0x0e643df7, 0x82c359bf, 0xc7794dfa, 0xd5e9ffaa,
                     0x2c8cb64a, 0x2f07b334, 0xad5a7eb5, 0x96dc0cde, 0x6fc24589, 0xa5853646, 0xe71576e2,
                     0x0dae30df, 0xb09ce711, 0x5e56ef87, 0x4b4b0082, 0x6f4f340e, 0xc5bb17e8, 0xd788d765,
                     0x67498087, 0x9d7aba26, 0x261351d4, 0x411ee7ea, 0x0393a263, 0x2c5a5835, 0xc115fcd8,
                     0x25e9132c, 0xd0c6e906, 0xc2bc5b2d, 0x6c065c98, 0x6e37bd55};
  int k_ = 32, j_ = 12;
};

// BLASTNA -> 2-bit code for the triplets: C 1, G 2, T 3, N random, else 0
// (dust.rs:71-79).
inline uint8_t dust_code(uint8_t b, DustRandom& rng) {
  switch (b) {
    case 1: return 1;
    case 2: return 2;
    case 3: return 3;
    case 14: return (uint8_t)(rng.next() & 3u);
    default: return 0;
  }
}

struct Perfect {
  size_t start, stop;  // [start, stop] inclusive once reported
  uint32_t score, len;
};

// The sliding window of triplets (dust.rs:96-278).  `list` holds the window,
// newest at the front; c_w / r_w count triplet pairs over the whole window,
// c_v / r_v over its suffix starting at `l`.
struct Window {
  std::deque<uint8_t> list;
  uint32_t c_w[64] = {0}, c_v[64] = {0};
  uint32_t r_w = 0, r_v = 0, num_diff = 0;
  size_t max_size;
  uint32_t low_k;
  size_t start, stop, l;

  Window(size_t window, uint32_t lk, size_t at) : max_size(window >= 2 ? window - 2 : 0), low_k(lk), start(at),
                                                  stop(at), l(at) {}

  void add_w(uint8_t t) {
    if (c_w[t] == 0) ++num_diff;
    r_w += c_w[t]++;
  }
  void rem_w(uint8_t t) {
    if (--c_w[t] == 0) --num_diff;
    r_w -= c_w[t];
  }
  void add_v(uint8_t t) { r_v += c_v[t]++; }
  void rem_v(uint8_t t) { r_v -= --c_v[t]; }

  bool shift_high(uint8_t t, std::vector<Perfect>& perfect) {
    rem_w(list.back());
    list.pop_back();
    ++start;
    list.push_front(t);
    add_w(t);
    ++stop;
    if (num_diff <= 1) {
      perfect.insert(perfect.begin(), Perfect{start, stop + 1, 0, 0});
      return false;
    }
    return true;
  }

  // false when the whole window holds one triplet
  bool shift(uint8_t t, std::vector<Perfect>& perfect) {
    if (list.size() >= max_size) {
      if (num_diff <= 1) return shift_high(t, perfect);
      const uint8_t s = list.back();
      list.pop_back();
      rem_w(s);
      if (l == start) {
        ++l;
        rem_v(s);
      }
      ++start;
    }
    list.push_front(t);
    add_w(t);
    add_v(t);
    if (c_v[t] > low_k) {  // drop suffix triplets up to the oldest copy of t
      size_t idx = list.size() - (l - start) - 1;
      for (;;) {
        const uint8_t s = list[idx];
        rem_v(s);
        ++l;
        if (s == t || idx == 0) break;
        --idx;
      }
    }
    ++stop;
    if (list.size() >= max_size && num_diff <= 1) {
      perfect.clear();
      perfect.insert(perfect.begin(), Perfect{start, stop + 1, 0, 0});
      return false;
    }
    return true;
  }

  bool needs_processing(const std::vector<uint32_t>& thr) const {
    const size_t count = stop - l;
    return count < list.size() && r_w * 10 > thr[count];
  }

  // extend the suffix back through the window, recording better intervals
  void find_perfect(std::vector<Perfect>& perfect, const std::vector<uint32_t>& thr) const {
    uint32_t counts[64];
    std::copy(c_v, c_v + 64, counts);
    uint32_t score = r_v;
    uint32_t count = (uint32_t)(stop - l);
    size_t it = 0;
    uint32_t max_score = 0, max_len = 0;
    for (size_t idx = stop - l; idx < list.size(); ++idx, ++count) {
      const uint8_t ti = list[idx];
      const uint32_t cnt = counts[ti];
      score += counts[ti]++;
      if (cnt > 0 && score * 10 > thr[count]) {
        const size_t pos = (stop - idx) - 1;  // wraps like NCBI's size_type at 0
        while (it < perfect.size() && pos <= perfect[it].start) {
          const Perfect& p = perfect[it];
          if (max_score == 0 || max_len * p.score > max_score * p.len) {
            max_score = p.score;
            max_len = p.len;
          }
          ++it;
        }
        if (max_score == 0 || score * max_len >= max_score * count) {
          max_score = score;
          max_len = count;
          perfect.insert(perfect.begin() + (std::ptrdiff_t)it, Perfect{pos, stop + 1, score, count});
        }
      }
    }
  }
};

// report the oldest perfect interval once it has left the window
// (dust.rs:284-312)
void save_masked(std::vector<Perfect>& perfect, std::vector<std::pair<size_t, size_t>>& res, size_t wstart,
                 size_t linker) {
  if (perfect.empty()) return;
  const Perfect b = perfect.back();
  if (b.start >= wstart) return;
  if (!res.empty() && res.back().second + linker >= b.start) res.back().second = std::max(res.back().second, b.stop);
  else res.emplace_back(b.start, b.stop);
  while (!perfect.empty() && perfect.back().start < wstart) perfect.pop_back();
}

}  // namespace

std::vector<std::pair<size_t, size_t>> dust_intervals(const uint8_t* bases, size_t n, int window, int level,
                                                      int linker) {
  // out-of-range parameters fall back to the defaults (dust.rs:318-324)
  if (window < 8 || window > 64) window = kDustWindow;
  if (level < 2 || level > 64) level = kDustLevel;
  if (linker < 1 || linker > 32) linker = kDustLinker;
  std::vector<std::pair<size_t, size_t>> res;
  if (n < 3) return res;
  DustRandom rng;
  std::vector<uint8_t> conv(n, 0);
  std::vector<uint32_t> thr((size_t)window - 2, 1u);
  for (size_t i = 1; i < thr.size(); ++i) thr[i] = (uint32_t)i * (uint32_t)level;
  const uint32_t low_k = (uint32_t)level / 5;
  std::vector<Perfect> perfect;
  size_t j_start = 0;
  // The window restarts where it stood after a run of one repeated triplet;
  // bases from the restart are converted again, so N positions draw fresh
  // random values on each visit (dust.rs:350-366).
  while (n >= j_start + 4) {
    perfect.clear();
    Window w((size_t)window, low_k, j_start);
    const size_t j_end = n - 2;
    size_t j = j_start, conv_from = j_start;
    auto triplet = [&](size_t at) {
      while (conv_from <= at + 2) {
        conv[conv_from] = dust_code(bases[conv_from], rng);
        ++conv_from;
      }
      return (uint8_t)((conv[at] << 4) | (conv[at + 1] << 2) | conv[at + 2]);
    };
    while (j < j_end) {
      save_masked(perfect, res, w.start, (size_t)linker);
      const uint8_t t = triplet(j);
      ++j;
      if (w.shift(t, perfect)) {
        if (w.needs_processing(thr)) w.find_perfect(perfect, thr);
        continue;
      }
      // the window holds one triplet: slide without looking for intervals
      // until it is diverse again
      while (j < j_end) {
        save_masked(perfect, res, w.start, (size_t)linker);
        if (w.shift(triplet(j), perfect)) break;
        ++j;
      }
      break;
    }
    for (size_t ws = w.start; !perfect.empty(); ++ws) save_masked(perfect, res, ws, (size_t)linker);
    if (w.start > j_start) j_start = w.start;
    else break;
  }
  return res;
}

std::vector<std::pair<size_t, size_t>> dust_mask(uint8_t* bases, size_t n) {
  std::vector<std::pair<size_t, size_t>> iv = dust_intervals(bases, n);
  for (const auto& p : iv)
    for (size_t i = p.first; i <= std::min(p.second, n - 1); ++i) bases[i] = 14;
  return iv;
}

}  // namespace rmb
