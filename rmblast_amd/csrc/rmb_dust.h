// Symmetric DUST of a query (the BLAST filter path), restated from
// rmblast-lib/src/filter/dust.rs (a port of NCBI's CSymDustMasker): window 64,
// level 20, linker 1; N bases take random values from NCBI's CRandom stream.
#pragma once
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

namespace rmb {

constexpr int kDustWindow = 64;
constexpr int kDustLevel = 20;
constexpr int kDustLinker = 1;

// Masked intervals [first, last] (0-based, inclusive) of BLASTNA bases
// (no sentinels), in order (dust.rs:339-448, dust_intervals).
std::vector<std::pair<size_t, size_t>> dust_intervals(const uint8_t* bases, size_t n, int window = kDustWindow,
                                                      int level = kDustLevel, int linker = kDustLinker);

// Mask bases[0, n) in place with code 14 over the DUST intervals
// (dust.rs:462-484, dust_mask); returns the intervals.
std::vector<std::pair<size_t, size_t>> dust_mask(uint8_t* bases, size_t n);

}  // namespace rmb
