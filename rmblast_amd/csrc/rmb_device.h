// Shared device/host record layouts of librmblast_hip.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace rmb {

// Phase-1 launch geometry.  One workgroup owns one (record, subject) work item
// at a time; kSeedCap seeds are staged in LDS per pass; the 512 diagonal-hash
// buckets are owned two per thread.
constexpr int kP1Threads = 256;
constexpr int kP1PerCu = 5;     // resident workgroups per CU (LDS and at most 96 VGPRs)
constexpr int kSeedCap = 1152;  // with it a workgroup needs 31 KB of LDS: five fit a CU
constexpr int kHashLds = 3;      // cells per bucket held in the owner thread's registers
constexpr int kHashGlobal = 253; // further cells per bucket in the workgroup's global spill area
constexpr int kBuckets = 512;
constexpr int kXcds = 8;         // MI355X: 8 XCDs, workgroups dispatched round-robin over them

struct RecDesc {        // one query record (a chunk of the genome)
  uint64_t q_base;      // offset of its sentinel-wrapped arrays in the query pools
  uint64_t cell_base;   // offset of its (4^lw + 1) cell starts in the cell pool
  uint32_t q_len;       // bases, without sentinels
  int32_t cutoff;       // ungapped keep threshold of the record's query (ka_cutoff.rs:126-161)
};

struct SubjDesc {       // one library sequence
  uint64_t s_base;      // offset of its sentinel-wrapped arrays in the subject pools
  uint32_t s_len;
  uint32_t pad;
};

struct P1Params {
  int32_t word_size, lut_width, scan_step, xdrop_ungap;
  uint32_t n_subj;
  uint64_t n_items;          // records x subjects, or the length of item_list
  uint64_t out_cap;
  const uint32_t* item_list; // null: every (record, subject) item; else these item ids
  int32_t known_word;        // whole-word table (ws >= 7, scan step 1) and positive ACGT match scores
  uint32_t part_log2;        // each item is searched by 2^part_log2 workgroups, each owning as many
                             // contiguous ranges of the 512 diagonal-hash buckets (0: one per item)
};

constexpr uint32_t kNoItem = 0xFFFFFFFFu;  // UngHit::item of an output slot reserved but left unused

struct UngHit {         // an ungapped hit at or above the cutoff, record frame
  uint32_t item;        // rec * n_subj + subj
  uint32_t q_seed, s_seed;
  uint32_t q_start, q_end, s_start, s_end;
  int32_t score;
  uint32_t strand;      // 0 plus, 1 minus (record frame: forward query, RC subject)
};

// Counters accumulated by the kernels (u64 atomics).
enum Counter {
  C_SEED_EVALS = 0, C_UNG_HITS = 1, C_UNG_STEPS = 2, C_DP_CELLS = 3, C_PRELIM_CELLS = 4, C_PRELIM_DPS = 5,
  C_SEEDS = 6,          // seeds staged by Phase 1 (valid or not)
  C_P1_STAGE = 8,       // 6 counters: Phase-1 stage A..F time, 100 MHz ticks summed over workgroups
  C_WAVE_ROWS = 14,     // rows of the wave-cooperative traceback DPs
  C_WAVE_ROUNDS = 15,   // their scan rounds (one per row when every first guess holds)
  C_TB_TICKS = 16,      // 4 maxima in 100 MHz ticks: a wave-run half's DP and walk, a lane-run wave of 64
                        // halves' DP and walk (the traceback kernel's critical paths)
  C_NCOUNTERS = 20
};

struct PreHit {         // an ungapped hit whose preliminary DP reached the gapped cutoff
  UngHit u;
  int32_t score;        // preliminary gapped score
  uint32_t q_start, q_end, s_start, s_end;  // record frame (RC subject for minus)
  uint32_t tq, ts;      // traceback seed (record frame)
};

struct PrelimParams {
  int32_t xdrop_gap, gap_open, gap_extend, cutoff, lut_width;
  uint32_t n_subj;
  uint64_t n_hits, out_cap, redo_cap;
  const unsigned long long* n_dev;  // when set, the task count is min(*n_dev, n_hits)
  const uint32_t* rec_ids;          // group-local record -> search record, for output item ids
  unsigned long long* next;         // work counter (zero at launch): waves take hits from it in runs
};

// Error bits written by kernels.
enum ErrBit : uint32_t {
  E_HASH_FULL = 1u, E_OUT_FULL = 2u, E_BAND_WIDE = 4u, E_TB_SPACE = 8u
};

// Per-lane DP band ring: 2^wlog2 cells plus one 128-byte line of padding, so
// the same ring cell of different lanes does not fall on the same L2 channel.
constexpr uint32_t kRingPad = 16;
inline constexpr uint64_t ring_stride(uint32_t wlog2) { return (1ull << wlog2) + kRingPad; }

struct DpTask {         // one half of an X-drop alignment, lane-serial
  const uint8_t* a;     // outer sequence (rows)
  const uint8_t* b;     // inner sequence (columns, the band)
  uint32_t m, n;
  int32_t xdrop;
  uint32_t reverse;
};

struct DpResult {
  int32_t score;
  uint32_t a_len, b_len;
  uint32_t flags;       // ErrBit E_BAND_WIDE when the band outgrew the ring
  uint64_t cells;       // DP cells evaluated (the reference's TOTAL_DP_CELLS unit)
  uint32_t tb_bytes;    // traceback bytes through the best row
  uint32_t tb_rows;     // traceback rows through the best row (row 0 included)
};

struct TbTask {         // traceback pass: a DpTask plus its storage
  DpTask t;
  uint64_t flat_off;    // byte offset of its traceback bytes
  uint64_t row_off;     // u32-pair offset of its row table
  uint64_t ops_off;     // u32 offset of its op buffer
  uint32_t stop_row;    // best_a from the score pass
  uint32_t best_b;
  uint32_t ops_cap;
  uint32_t pad;
};

struct TbResult { uint32_t n_ops; uint32_t flags; };

struct OnePassResult {  // a half alignment scored and traced back in one pass
  int32_t score;
  uint32_t a_len, b_len, n_ops;
  uint32_t flags;       // ErrBit: E_BAND_WIDE / E_TB_SPACE send the half to the two-pass path
  uint32_t pad;
  uint64_t ops_at;      // its ops in the packed pool
};

// Retry storage of the one-pass traceback: a half whose band outgrows the LDS
// ring, or whose traceback outgrows the lane's arena, is run again by its lane
// with one of `slots` global rings of 2^ring_log2 cells and a larger arena,
// slots handed out by *next (zero at launch).
struct OnePassBig {
  int2* ring;
  uint8_t* arena;
  unsigned int* next;
  uint32_t ring_log2, slots, flat_cap, rows_cap, ops_cap;
};

// Per-lane arena of the one-pass traceback: row table (8 B per row; none for
// the traceback stream with row trailers), then the traceback bytes (16-byte
// aligned), then the ops (4 B each).
inline constexpr uint64_t onepass_stride(uint32_t flat_cap, uint32_t rows_cap, uint32_t ops_cap) {
  return (uint64_t)rows_cap * 8 + ((flat_cap + 15u) & ~15u) + (((uint64_t)ops_cap * 4 + 15u) & ~15ull);
}

struct PackTask {       // copy n ops of one traceback from the op pool into the packed list
  uint64_t src, dst;
  uint32_t n, pad;
};

}  // namespace rmb
