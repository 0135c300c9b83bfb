// Dimensions and flat parameter layout of the learned triage model.
#pragma once
#include <cstddef>
#include <cstdint>

namespace rmb {
namespace tri {

constexpr int kKmerDim = 2080;  // canonical 6-mers: (4^6 + 4^3) / 2
constexpr int kH1 = 512;        // tower hidden width
constexpr int kEmb = 256;       // tower output width
constexpr int kHead = 64;       // head hidden width

// Flat FP32 parameter vector: both towers, then the head.
struct Layout {
  static constexpr size_t W1 = (size_t)kH1 * kKmerDim, B1 = kH1, W2 = (size_t)kEmb * kH1, B2 = kEmb;
  static constexpr size_t TOWER = W1 + B1 + W2 + B2;
  static constexpr size_t tower(int t) { return (size_t)t * TOWER; }  // t = 0 window tower, 1 family tower
  static constexpr size_t w1(int t) { return tower(t); }
  static constexpr size_t b1(int t) { return tower(t) + W1; }
  static constexpr size_t w2(int t) { return tower(t) + W1 + B1; }
  static constexpr size_t b2(int t) { return tower(t) + W1 + B1 + W2; }
  static constexpr size_t HEAD = 2 * TOWER;
  static constexpr size_t wh1() { return HEAD; }                         // [kHead][kEmb]
  static constexpr size_t hp() { return HEAD + (size_t)kHead * kEmb; }   // bh1[64], wh2[64], bh2, a, c
  static constexpr size_t HP = 2 * kHead + 3;
  static constexpr size_t TOTAL = HEAD + (size_t)kHead * kEmb + HP;
};

}  // namespace tri
}  // namespace rmb
