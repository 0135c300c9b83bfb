// HIP kernels of librmblast_hip for gfx950 (wave64, 160 KB LDS per CU).
//
//  k_lut_keys      lookup-table build, stage 1: one (record, word, offset)
//                  key per window of the two-strand query buffer
//                  (blast/nalookup.rs:128-181, blast/mblookup.rs:106-245)
//  k_phase1        seeding + exact diagonal-hash dedup + ungapped X-drop for
//                  (record, subject) work items pulled from a device queue
//                  (search/engine.rs:1063-1277, diag_hash.rs, ungapped.rs)
//  k_prelim_fused  Phase 2a's score-only X-drop DP for every ungapped hit
//                  (search/engine.rs:1850-1895, search/gapped.rs:360-418)
//  k_dp_onepass    Phase 2b: score, traceback bytes and traceback walk of a
//                  half alignment in one lane-serial pass
//                  (search/gapped.rs:450-757)
//  k_dp_wave       the same for long halves, one per wave, a band row per
//                  vector step (prefix scans resolve the row's dependencies)
//  k_dp_score      lane-serial score-only X-drop half alignment
//                  (search/gapped.rs:124-337)
//  k_dp_traceback  the same recurrence with 1-byte traceback cells and the
//                  traceback walk, for halves beyond k_dp_onepass's storage
//                  (search/gapped.rs:450-757)
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "rmb_device.h"

namespace rmb {

static constexpr int32_t kMinInt = INT32_MIN / 2;  // gapped.rs:47
__constant__ uint8_t c_complement[16] = {3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15};

// --------------------------------------------------------------- helpers ----
__device__ __forceinline__ uint32_t block_incl_scan256(uint32_t v, uint32_t* wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  uint32_t add = 0;
  for (int w = 0; w < wid; ++w) add += wsum[w];
  __syncthreads();
  return v + add;
}

// Bytes [o, o + 16) of the 32-byte little-endian concatenation lo|hi, o < 16:
// a two-level word select and one v_alignbyte per output word.
__device__ __forceinline__ uint4 funnel16(const uint4 lo, const uint4 hi, uint32_t o) {
  const bool b8 = (o & 8u) != 0, b4 = (o & 4u) != 0;
  const uint32_t ob = o & 3u;
  const uint32_t t0 = b8 ? lo.z : lo.x, t1 = b8 ? lo.w : lo.y, t2 = b8 ? hi.x : lo.z, t3 = b8 ? hi.y : lo.w,
                 t4 = b8 ? hi.z : hi.x, t5 = b8 ? hi.w : hi.y;
  const uint32_t u0 = b4 ? t1 : t0, u1 = b4 ? t2 : t1, u2 = b4 ? t3 : t2, u3 = b4 ? t4 : t3, u4 = b4 ? t5 : t4;
  return make_uint4(__builtin_amdgcn_alignbyte(u1, u0, ob), __builtin_amdgcn_alignbyte(u2, u1, ob),
                    __builtin_amdgcn_alignbyte(u3, u2, ob), __builtin_amdgcn_alignbyte(u4, u3, ob));
}

// Bytes [e, e + 16) of a 16-byte-aligned pool (e >= -15), from the two aligned
// 16-byte blocks that hold them.  A block before the pool start is replaced by
// block 0: its bytes sit before the pool's leading sentinel, which ends every
// extension first.  Pools carry 64 bytes of tail, so the blocks after any
// sentinel lie inside the allocation.
__device__ __forceinline__ uint4 pool16(const uint4* __restrict__ V, int64_t e) {
  const int64_t b = e >> 4;  // floor
  const uint4 lo = V[b < 0 ? 0 : b], hi = V[b + 1];
  return funnel16(lo, hi, (uint32_t)e & 15u);
}

struct UngOut { int32_t score; uint32_t q_beg, q_end, steps; };

// `smat` is the scoring matrix with kSentScore in row and column 15, so
// reaching a sentinel ends a side through the X-drop test; that step is then
// taken back, as the reference tests for the sentinel before counting a step.
static constexpr int32_t kSentScore = -(1 << 20);
static constexpr int kLongChunk = 8;  // steps per chunk of the full extension
static constexpr int kFlankSteps = 5;  // flank steps per side of the fast first tier

// Ungapped X-drop from (q_off, s_off) over sentinel-wrapped sequences: left
// side fixed X, right side adaptive X (ungapped.rs:38-121).  qa / sa are the
// pool offsets of the start position.  Each side runs in chunks of 16 steps:
// the chunk's 16 bytes of both sequences come in with four aligned loads, its
// 16 matrix lookups are issued together, and the X-drop state advances
// branch-free under a per-lane `run` flag until the lane's side ends.
template <int CH>
__device__ __forceinline__ UngOut ungapped_chunked(const uint8_t* __restrict__ qpool, int64_t qa,
                                                   const uint8_t* __restrict__ spool, int64_t sa, int32_t q_off,
                                                   const int* smat, int32_t xdrop) {
  const uint4* QV = reinterpret_cast<const uint4*>(qpool);
  const uint4* SV = reinterpret_cast<const uint4*>(spool);
  auto index4 = [](uint32_t q, uint32_t s) { return ((q << 4) & 0xF0F0F0F0u) | (s & 0x0F0F0F0Fu); };
  const int32_t x = -xdrop;
  int32_t score = 0, sum = 0, q_beg = q_off;
  uint32_t steps = 0;
  bool run = true;
  // left: step k reads offset qa - 1 - k; chunk c holds steps 16c .. 16c + 15
  // at bytes 15 .. 0 of [qa - 16(c + 1), qa - 16c)
  for (int32_t c = 0; run; ++c) {
    const int64_t e = 16 + CH * (int64_t)c;
    const uint4 qv = pool16(QV, qa - e), sv = pool16(SV, sa - e);
    const uint32_t iw[4] = {index4(qv.x, sv.x), index4(qv.y, sv.y), index4(qv.z, sv.z), index4(qv.w, sv.w)};
    int32_t v[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int p = 15 - j;
      v[j] = smat[(iw[p >> 2] >> (8 * (p & 3))) & 0xFFu];
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int32_t ns = sum + v[j];
      const bool up = run && ns > 0;
      const bool stop = run && ns < x;
      score = up ? score + ns : score;
      q_beg = up ? q_off - 1 - (CH * c + j) : q_beg;
      sum = run ? (ns > 0 ? 0 : ns) : sum;
      steps += run ? 1u : 0u;
      run = run && !stop;
    }
  }
  if (sum <= kSentScore / 2) --steps;  // ended on a sentinel: that step is not counted
  int32_t q_end = q_off, xc = x;
  sum = 0;
  run = true;
  // right: step k reads offset qa + k; chunk c holds steps 16c .. 16c + 15
  for (int32_t c = 0; run; ++c) {
    const int64_t e = CH * (int64_t)c;
    const uint4 qv = pool16(QV, qa + e), sv = pool16(SV, sa + e);
    const uint32_t iw[4] = {index4(qv.x, sv.x), index4(qv.y, sv.y), index4(qv.z, sv.z), index4(qv.w, sv.w)};
    int32_t v[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) v[j] = smat[(iw[j >> 2] >> (8 * (j & 3))) & 0xFFu];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int32_t ns = sum + v[j];
      const bool up = run && ns > 0;
      const bool stop = run && ns < xc;
      score = up ? score + ns : score;
      q_end = up ? q_off + CH * c + j + 1 : q_end;
      xc = up ? ((-score > x) ? -score : x) : xc;
      sum = run ? (ns > 0 ? 0 : ns) : sum;
      steps += run ? 1u : 0u;
      run = run && !stop;
    }
  }
  if (sum <= kSentScore / 2) --steps;
  return UngOut{score, (uint32_t)q_beg, (uint32_t)q_end, steps};
}

// Fast first tier of the extension, for whole-word tables (ws >= 7) with
// positive ACGT match scores.  Every base of the seed word matches, so each
// of its steps raises the score: the state after them is known, and only the
// flanks need stepping.  Plus-strand seeds start at the word's first base (the
// right side opens with 7 word steps), minus-strand seeds at its last base
// (6 word steps on the left, 1 on the right): only the sum of the word's
// scores enters the result, so a minus seed is stepped in a frame moved 6
// bases left, where its word sits where a plus seed's does and both strands
// run the same code.  Then kFlankSteps branch-free flank steps per side, from
// the 24 bytes [b - 8, b + 16) of each sequence around the frame start b (an
// unaligned 8-byte and 16-byte load each: gfx950 global loads take any byte
// address).  Five steps suit random-like flanks, where a side ends after
// about three: each step costs every lane of the wave, and seeds that need
// more go to the full extension over a compact list.
// True: both sides ended within those steps and `out`
// is the full extension's result; false: a side still runs (or the frame
// starts within 16 bytes of its pool's start), and the caller runs the full
// extension.
__device__ __forceinline__ bool ungapped_first(const uint8_t* __restrict__ qpool, int64_t qa,
                                               const uint8_t* __restrict__ spool, int64_t sa, int32_t q_off,
                                               bool minus, const int* smat, int32_t xdrop, UngOut& out) {
  const int32_t shift = minus ? 6 : 0;
  qa -= shift;
  sa -= shift;
  q_off -= shift;
  if (qa < 16 || sa < 16) return false;
  static_assert(kFlankSteps <= 8, "the flank windows hold 8 bytes per side");
  uint2 qa8, sa8;
  uint4 qb16, sb16;
  __builtin_memcpy(&qa8, qpool + qa - 8, 8);
  __builtin_memcpy(&qb16, qpool + qa, 16);
  __builtin_memcpy(&sa8, spool + sa - 8, 8);
  __builtin_memcpy(&sb16, spool + sa, 16);
  // index bytes (q << 4 | s) at byte positions 8..31; byte 16 is the frame start
  uint32_t ix[8];
  {
    const uint32_t qw[6] = {qa8.x, qa8.y, qb16.x, qb16.y, qb16.z, qb16.w};
    const uint32_t sw[6] = {sa8.x, sa8.y, sb16.x, sb16.y, sb16.z, sb16.w};
    ix[0] = ix[1] = 0;
#pragma unroll
    for (int w = 0; w < 6; ++w) ix[w + 2] = ((qw[w] << 4) & 0xF0F0F0F0u) | (sw[w] & 0x0F0F0F0Fu);
  }
  auto sc = [&](int p) { return smat[(ix[p >> 2] >> (8 * (p & 3))) & 0xFFu]; };
  // the word's steps: bytes 16..22
  int32_t known = 0;
#pragma unroll
  for (int k = 16; k < 23; ++k) known += sc(k);
  const int32_t x = -xdrop;
  // left: flank steps from byte 15 down
  int32_t score = 0, sum = 0, q_beg = q_off;
  uint32_t steps = 0;
  bool run = true;
#pragma unroll
  for (int j = 0; j < kFlankSteps; ++j) {
    const int32_t ns = sum + sc(15 - j);
    const bool up = run && ns > 0;
    const bool stop = run && ns < x;
    score = up ? score + ns : score;
    q_beg = up ? q_off - 1 - j : q_beg;
    sum = run ? (ns > 0 ? 0 : ns) : sum;
    steps += run ? 1u : 0u;
    run = run && !stop;
  }
  if (run) return false;
  if (sum <= kSentScore / 2) --steps;  // ended on a sentinel: that step is not counted
  // right: the word's steps, then flank steps from byte 23 up
  score += known;
  steps += 7u;
  const int32_t qe0 = q_off + 7;
  int32_t q_end = qe0, xc = (-score > x) ? -score : x;
  sum = 0;
  run = true;
#pragma unroll
  for (int j = 0; j < kFlankSteps; ++j) {
    const int32_t ns = sum + sc(23 + j);
    const bool up = run && ns > 0;
    const bool stop = run && ns < xc;
    score = up ? score + ns : score;
    q_end = up ? qe0 + j + 1 : q_end;
    xc = up ? ((-score > x) ? -score : x) : xc;
    sum = run ? (ns > 0 ? 0 : ns) : sum;
    steps += run ? 1u : 0u;
    run = run && !stop;
  }
  if (run) return false;
  if (sum <= kSentScore / 2) --steps;
  out = UngOut{score, (uint32_t)q_beg, (uint32_t)q_end, steps};
  return true;
}

// --------------------------------------------------------- lookup build -----
// One key per window start p of record r's two-strand buffer [fwd | N | rc]
// (engine.rs:630-649).  Valid windows lie inside one strand, hold no code >= 4
// and belong to a strand of at least word_size bases; invalid ones get the
// sort-last key.  Small tables list offsets ascending per cell, megablast
// tables descending (mblookup.rs:170): the radix sort is stable, so the
// value order fed in decides it.
__global__ void k_lut_keys(const RecDesc* __restrict__ recs, uint32_t n_recs, const uint64_t* __restrict__ elem_base,
                           const uint8_t* __restrict__ qext, int lw, int word_size, int descending,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* __restrict__ counts,
                           uint64_t invalid_key, uint64_t total) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  // locate the record by binary search over elem_base (n_recs + 1 entries)
  uint32_t lo = 0, hi = n_recs;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (elem_base[mid] <= e) lo = mid; else hi = mid;
  }
  const uint32_t r = lo;
  const RecDesc R = recs[r];
  const int64_t n = R.q_len;
  const int64_t k = (int64_t)(e - elem_base[r]);  // 0 .. 2n
  const int64_t p = descending ? (2 * n - k) : k;
  keys[e] = invalid_key;
  vals[e] = (uint32_t)p;
  if (n < word_size) return;
  const int64_t last = p + lw - 1;
  const bool in_fwd = last <= n - 1;
  const bool in_rc = p >= n + 1 && last <= 2 * n;
  if (!in_fwd && !in_rc) return;
  const uint8_t* Q = qext + R.q_base;
  uint32_t code = 0;
  for (int i = 0; i < lw; ++i) {
    const int64_t pp = p + i;
    const uint8_t b = pp < n ? Q[1 + pp] : c_complement[Q[2 * n + 1 - pp] & 15];
    if (b & 0xFC) return;
    code = (code << 2) | b;
  }
  keys[e] = ((uint64_t)r << (2 * lw)) | code;
  atomicAdd(&counts[((uint64_t)r << (2 * lw)) + code], 1u);
}

// Ascending sort of eight values (Batcher's 19-comparator network).
__device__ __forceinline__ void sort8(uint32_t (&v)[8]) {
  constexpr int kNet[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                               {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    const uint32_t a = v[kNet[k][0]], b = v[kNet[k][1]];
    v[kNet[k][0]] = a < b ? a : b;
    v[kNet[k][1]] = a < b ? b : a;
  }
}

// Bit position of the n-th (from 0) set bit of m (n < popcount(m)).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t w = 32; w >= 1; w >>= 1) {
    const uint64_t low = m & ((1ull << w) - 1ull);
    const uint32_t c = (uint32_t)__popcll(low);
    const bool up = n >= c;
    n -= up ? c : 0u;
    m = up ? (m >> w) : low;
    pos += up ? w : 0u;
  }
  return pos;
}

// ------------------------------------------------------------ Phase 1 -------
// One workgroup per (record, subject) item at a time, items from a device
// queue.  Per pass it stages up to kSeedCap seeds of the reference scan order
// (subject offset ascending, then cell order) in LDS, extends every seed
// speculatively (the extension is a pure function of the seed), then replays
// the reference's 512-bucket diagonal hash exactly: buckets are independent,
// so each thread replays its two buckets' seeds in scan order.  Seeds the
// replay keeps and that reach the cutoff are emitted.
enum SeedFlag : uint8_t { F_VALID = 1, F_HIT = 2, F_EMIT = 4, F_MINUS = 8, F_LONG = 16 };
constexpr uint32_t kOutChunk = 256;  // output slots a workgroup reserves at a time

// SEEDS (the seeds tap): every seed the diagonal hash lets through is output
// instead of the hits, as {item, diagonal, hash check offset, strand, its
// ordinal in the item's scan}; the hash evolves exactly as in the search.
// WORD: whole-word tables (word size = lookup width, scan step 1, positive
// ACGT match scores; P.known_word).  A seed then needs no word extension and
// its frame follows from its scan position and lookup entry alone, so a hit's
// extension start and end are recomputed from the seed record at output time
// and only its left end and score are kept, in an LDS table of the pass
// (kHitLds entries, the rest in the workgroup's global area).
constexpr uint32_t kHitLds = 128;
constexpr uint32_t kHitSlotShift = 21;  // sd[].w: steps (16 bits) | flags << 16 (5 bits) | hit slot << 21
// SPLIT: items split over bucket ranges (P.part_log2 > 0).  With WORD, a
// pass's seeds are first materialised (B1) and each wave then extends only
// those on its part's diagonals (B2), handed to its lanes in scan order, so
// the fast tier's lanes stay full; otherwise the other parts' seeds are
// dropped where they are met.
template <bool SEEDS, bool WORD, bool SPLIT = false>
__global__ __launch_bounds__(kP1Threads, kP1PerCu) void k_phase1(
    P1Params P, const RecDesc* __restrict__ recs, const SubjDesc* __restrict__ subjs,
    const uint8_t* __restrict__ qfwd, const uint8_t* __restrict__ qrc, const uint8_t* __restrict__ qext,
    const uint8_t* __restrict__ splus, const uint8_t* __restrict__ src, const uint32_t* __restrict__ cells,
    const uint32_t* __restrict__ entries, const int32_t* __restrict__ gmat, UngHit* __restrict__ out,
    unsigned long long* __restrict__ out_count, uint2* __restrict__ spill, uint4* __restrict__ hitres,
    unsigned long long* __restrict__ queue, unsigned long long* __restrict__ counters, uint32_t* __restrict__ err) {
  __shared__ int mat[256];
  __shared__ uint32_t s_cum[kP1Threads], s_cst[kP1Threads], wsum[4];
  __shared__ uint4 sd[kSeedCap];  // per seed: {diagonal, hash check offset, level, steps | flags << 16}
  __shared__ uint16_t blist[kSeedCap];
  // seeds whose extension outruns the fast first tier: written in B, read in
  // C, so it shares blist's space (written from D on)
  uint16_t* const llist = blist;
  __shared__ uint8_t posmap[kSeedCap];  // pass seed -> its scan position in the window
  __shared__ uint32_t bcnt[kBuckets], bst[kBuckets];
  __shared__ unsigned long long s_item, s_out_next, s_pass_a, s_pass_b;
  __shared__ uint32_t s_part;
  __shared__ uint32_t s_out_left, s_pass_an, s_nemit, s_emit_i, s_nlong, s_nhit;
  __shared__ uint2 s_hit[WORD ? kHitLds : 1];  // WORD: {left end, score} of the pass's hits

  const int tid = threadIdx.x;
  const uint32_t xcd = blockIdx.x % kXcds;
  for (int i = tid; i < 256; i += kP1Threads) mat[i] = ((i >> 4) == 15 || (i & 15) == 15) ? kSentScore : gmat[i];
  if (tid == 0) {
    s_out_next = 0;
    s_out_left = 0;
    s_nlong = 0;
    s_nhit = 0;
  }
  uint2* my_spill = spill + (size_t)blockIdx.x * kBuckets * kHashGlobal;
  uint4* my_hitres = hitres + (size_t)blockIdx.x * kSeedCap;  // extension results of the pass's hits, by seed
  const int32_t ws = P.word_size, lw = P.lut_width, step = P.scan_step;
  const int32_t ext_needed = ws - lw;
  const int32_t window = 1 - ws;
  unsigned long long n_evals = 0, n_steps = 0, n_hits = 0;
  // per-stage time (thread 0, 100 MHz clock read after each barrier) and
  // seeds staged, kept in LDS: only thread 0 touches them
  __shared__ unsigned long long s_tick[7], s_nseeds;
  if (tid == 0) {
    for (int k = 0; k < 6; ++k) s_tick[k] = 0;
    s_tick[6] = wall_clock64();
    s_nseeds = 0;
  }
  auto tick = [&](int k) {
    if (tid == 0) {
      const unsigned long long t = wall_clock64();
      s_tick[k] += t - s_tick[6];
      s_tick[6] = t;
    }
  };
  // WORD: a slot of the pass's hit table for each lane with `hit` (one LDS
  // atomic per wave), its {left end, score} stored; returns the slot
  auto put_hit = [&](bool hit, uint32_t q_beg, int32_t score) -> uint32_t {
    const unsigned long long m = __ballot(hit);
    uint32_t slot = 0;
    if (m) {
      const uint32_t lane = (uint32_t)tid & 63u;
      const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(&s_nhit, (uint32_t)__popcll(m));
      base = __shfl(base, (int)leader, 64);
      slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (hit) {
        if (slot < kHitLds) s_hit[slot] = make_uint2(q_beg, (uint32_t)score);
        else my_hitres[slot - kHitLds] = make_uint4(q_beg, (uint32_t)score, 0u, 0u);
      }
    }
    return slot;
  };

  for (;;) {
    __syncthreads();
    if (tid == 0) {
      // XCD-aware queues: queue x hands out the items of records r = x (mod 8),
      // record-major, so one XCD's workgroups share a record's lookup table in
      // their L2; an XCD whose queue is empty takes work from the others.
      // An item split over 2^part_log2 workgroups is 2^part_log2 consecutive
      // work units, one per part.
      unsigned long long it = ~0ull;
      uint32_t part = 0;
      const uint32_t pl = SPLIT ? P.part_log2 : 0u;
      if (P.item_list) {  // listed items: the list is cut into kXcds contiguous slices
        for (int k = 0; k < kXcds && it == ~0ull; ++k) {
          const uint32_t x = (xcd + k) % kXcds;
          const unsigned long long lo = P.n_items * x / kXcds, hi = P.n_items * (x + 1) / kXcds;
          if (hi == lo) continue;
          const unsigned long long t = atomicAdd(&queue[x], 1ull);
          if (t < ((hi - lo) << pl)) {
            it = P.item_list[lo + (t >> pl)];
            part = (uint32_t)t & ((1u << pl) - 1u);
          }
        }
      } else {
        const uint32_t n_recs = (uint32_t)(P.n_items / P.n_subj);
        for (int k = 0; k < kXcds && it == ~0ull; ++k) {
          const uint32_t x = (xcd + k) % kXcds;
          const unsigned long long n_x =
              (unsigned long long)(n_recs > x ? (n_recs - x + kXcds - 1) / kXcds : 0) * P.n_subj;
          if (n_x == 0) continue;
          const unsigned long long tu = atomicAdd(&queue[x], 1ull);
          if (tu < (n_x << pl)) {
            const unsigned long long t = tu >> pl;
            it = (unsigned long long)(x + kXcds * (t / P.n_subj)) * P.n_subj + t % P.n_subj;
            part = (uint32_t)tu & ((1u << pl) - 1u);
          }
        }
      }
      s_item = it;
      s_part = part;
      s_nemit = 0;
      s_emit_i = 0;
    }
    for (int b = tid; b < kBuckets; b += kP1Threads) bcnt[b] = 0;
    __syncthreads();
    const unsigned long long item = s_item;
    if (item == ~0ull) break;
    // the buckets this workgroup owns: b >> own_shift == part (all when unsplit)
    const uint32_t part = SPLIT ? s_part : 0u, own_shift = 9u - (SPLIT ? P.part_log2 : 0u);
    const uint32_t r = (uint32_t)(item / P.n_subj), f = (uint32_t)(item % P.n_subj);
    const RecDesc R = recs[r];
    const SubjDesc S = subjs[f];
    const int32_t q_len = (int32_t)R.q_len, s_len = (int32_t)S.s_len;
    if (q_len < ws || s_len < ws) continue;
    const uint8_t* QE = qext + R.q_base;
    const uint8_t* SP = splus + S.s_base;
    const uint8_t* SR = src + S.s_base;
    const uint32_t* CS = cells + R.cell_base;
    const int32_t end_range = s_len - lw;
    const uint32_t n_slots = (uint32_t)(end_range / step) + 1;
    const int32_t c1 = q_len + 1;
    // diagonal hash: thread tid owns buckets tid and tid + kP1Threads for the
    // whole item; each bucket's first kHashLds (3) cells live in these
    // registers, the rest in the workgroup's spill area.  A register cell not
    // in use holds diagonal kNoDiag (no seed's) and level kFarLevel (never
    // stale while subject offsets stay below 2^30).
    constexpr int32_t kNoDiag = (int32_t)0x80000000u;
    constexpr uint32_t kFarLevel = 0x7FFFFFFFu;
    uint32_t hc[2] = {0u, 0u}, hl0[2] = {kFarLevel, kFarLevel}, hl1[2] = {kFarLevel, kFarLevel},
             hl2[2] = {kFarLevel, kFarLevel};
    int32_t hd0[2] = {kNoDiag, kNoDiag}, hd1[2] = {kNoDiag, kNoDiag}, hd2[2] = {kNoDiag, kNoDiag};
    // The scan runs over windows of kP1Threads subject positions; a window's
    // seeds (subject offset ascending, then cell order) are taken kSeedCap
    // per pass, so each window is scanned once.
    uint32_t w0 = 0, wk = 0, wtotal = 0;
    uint64_t wseen = 0;  // seeds of the item's earlier windows (the seeds tap's ordinals)
    bool scanned = false;
    for (;;) {
      if (!scanned) {
        if (w0 >= n_slots) break;
        // ---- A: hit counts of the window's scan positions
        uint32_t cnt = 0, cst = 0;
        const uint32_t slot = w0 + tid;
        if (slot < n_slots) {
          const int32_t s = (int32_t)slot * step;
          uint32_t idx = 0;
          for (int k = 0; k < lw; ++k) idx = (idx << 2) | (SP[1 + s + k] & 3);
          const uint32_t a = CS[idx], b = CS[idx + 1];
          cst = a;
          cnt = b - a;
        }
        s_cst[tid] = cst;
        const uint32_t inc = block_incl_scan256(cnt, wsum);
        s_cum[tid] = inc;
        __syncthreads();
        wtotal = s_cum[kP1Threads - 1];
        wk = 0;
        scanned = true;
        tick(0);
        if (wtotal == 0) {
          w0 += kP1Threads;
          scanned = false;
          continue;
        }
      }
      const uint32_t take = (wtotal - wk) < (uint32_t)kSeedCap ? (wtotal - wk) : (uint32_t)kSeedCap;
      if (tid == 0 && part == 0) s_nseeds += take;  // staged once per item
      // the pass's seed -> scan position map: each position's thread writes
      // the entries of its own seeds
      {
        const uint32_t hi = s_cum[tid], lo = tid ? s_cum[tid - 1] : 0u;
        const uint32_t a = lo > wk ? lo : wk, b = hi < wk + take ? hi : wk + take;
        for (uint32_t g = a; g < b; ++g) posmap[g - wk] = (uint8_t)tid;
      }
      __syncthreads();
      // ---- B+C: materialise each seed, exact word extension, diagonal,
      // speculative ungapped extension (the lane that fills a seed extends it)
      if constexpr (SPLIT && WORD) {
        // B1: every seed of the pass materialised; the wave's seeds on this
        // part's diagonals recorded per iteration as ballot masks
        constexpr int kIters = (kSeedCap + kP1Threads - 1) / kP1Threads;
        uint64_t own[kIters];
        uint32_t n_own = 0;
#pragma unroll
        for (int k = 0; k < kIters; ++k) {
          const uint32_t i = (uint32_t)tid + (uint32_t)(kP1Threads * k);
          bool owned = false;
          if (i < take) {
            const uint32_t g = wk + i;
            const int lo = posmap[i];
            const uint32_t prev = lo ? s_cum[lo - 1] : 0;
            const int32_t s_fwd = (int32_t)((w0 + (uint32_t)lo) * (uint32_t)step);
            const int32_t q0c = (int32_t)entries[s_cst[lo] + (g - prev)];
            const int32_t diag = (int32_t)((uint32_t)s_fwd - (uint32_t)q0c);
            owned = (((uint32_t)diag & (uint32_t)(kBuckets - 1)) >> own_shift) == part;
            const uint32_t fl = owned ? (F_VALID | (q0c >= c1 ? F_MINUS : 0)) : 0u;
            sd[i] = make_uint4(owned ? (uint32_t)diag : 0u, owned ? (uint32_t)s_fwd : 0u, 0u, fl << 16);
          }
          own[k] = __ballot(owned);
          n_own += (uint32_t)__popcll(own[k]);
        }
        // B2: the wave's owned seeds, its lanes taking them in scan order
        const uint32_t lane = (uint32_t)tid & 63u, wbase = (uint32_t)tid & ~63u;
        for (uint32_t j0 = 0; j0 < n_own; j0 += 64) {
          const uint32_t j = j0 + lane;
          const bool act = j < n_own;
          uint32_t i = 0, r = j;
          bool found = false;
#pragma unroll
          for (int k = 0; k < kIters; ++k) {
            const uint32_t ck = (uint32_t)__popcll(own[k]);
            const bool here = !found && r < ck;
            if (here) i = (uint32_t)(kP1Threads * k) + wbase + nth_set_bit(own[k], r);
            r -= (!found && !here) ? ck : 0u;
            found = found || here;
          }
          bool longer = false;
          if (act) {
            const uint4 e0 = sd[i];
            const bool minus = ((e0.w >> 16) & F_MINUS) != 0;
            const int32_t s_fwd = (int32_t)e0.y, diag = (int32_t)e0.x;
            const int32_t q0c = s_fwd - diag;
            const int32_t eq = minus ? q0c - c1 + lw - 1 : q0c, es = minus ? s_fwd + lw - 1 : s_fwd;
            const uint32_t chk = (uint32_t)s_fwd;
            const uint8_t* qpool = minus ? qrc : qfwd;
            const int64_t qa = (int64_t)R.q_base + 1 + eq, sa = (int64_t)S.s_base + 1 + es;
            uint32_t level, steps = 0, flag;
            UngOut u;
            if (ungapped_first(qpool, qa, splus, sa, eq, minus, mat, P.xdrop_ungap, u)) {
              const bool hit = u.score >= R.cutoff;
              const uint32_t slot = put_hit(hit, u.q_beg, u.score);
              level = hit ? (uint32_t)es + (u.q_end - (uint32_t)eq) : chk + (uint32_t)ws;
              steps = (u.steps < 65535u ? u.steps : 65535u) | (hit ? slot << kHitSlotShift : 0u);
              flag = F_VALID | (hit ? F_HIT : 0) | (minus ? F_MINUS : 0);
            } else {  // the full extension runs below; level holds eq meanwhile
              longer = true;
              level = (uint32_t)eq;
              flag = F_VALID | F_LONG | (minus ? F_MINUS : 0);
            }
            atomicAdd(&bcnt[(uint32_t)diag & (kBuckets - 1)], 1u);
            sd[i] = make_uint4((uint32_t)diag, chk, level, steps | (flag << 16));
          }
          const unsigned long long m = __ballot(longer);
          if (m) {
            const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_nlong, (uint32_t)__popcll(m));
            base = __shfl(base, (int)leader, 64);
            if (longer) llist[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)i;
          }
        }
      } else
      for (uint32_t i = tid; i < take; i += kP1Threads) {
        const uint32_t g = wk + i;  // seed index within the window
        const int lo = posmap[i];
        const uint32_t prev = lo ? s_cum[lo - 1] : 0;
        const int32_t s_fwd = (int32_t)((w0 + (uint32_t)lo) * (uint32_t)step);
        const int32_t q0c = (int32_t)entries[s_cst[lo] + (g - prev)];
        uint32_t flag = 0;
        int32_t diag = 0;
        uint32_t chk = 0, level = 0, steps = 0;
        // Strand-specific part: exact word extension, diagonal, hash check
        // offset and the extension frame (eq, es).  Plus strand: forward query
        // and subject; minus strand: reverse-complement query against the
        // forward subject, anchored at the word's last base.
        bool ok;
        int32_t eq = 0, es = 0;
        if constexpr (WORD) {
          // plus: forward query at q0c against the forward subject at s_fwd;
          // minus: the reverse-complement query at the word's last base
          // against the forward subject, the hash check offset s_fwd for both
          const bool minus = q0c >= c1;
          ok = true;
          diag = (int32_t)((uint32_t)s_fwd - (uint32_t)q0c);
          chk = (uint32_t)s_fwd;
          eq = minus ? q0c - c1 + lw - 1 : q0c;
          es = minus ? s_fwd + lw - 1 : s_fwd;
        } else if (q0c < c1) {
          int32_t q_off = q0c, s_off = s_fwd;
          ok = true;
          if (ext_needed > 0) {  // engine.rs:1130-1153
            int32_t el = 0, er = 0;
            while (el < ext_needed) {
              const int32_t qi = q0c - el - 1, si = s_fwd - el - 1;
              if (qi < 0 || si < 0) break;
              const uint8_t qb = QE[1 + qi], sb = SP[1 + si];
              if (qb >= 4 || sb >= 4 || qb != sb) break;
              ++el;
            }
            while (el + er < ext_needed) {
              const int32_t qi = q0c + lw + er, si = s_fwd + lw + er;
              if (qi >= q_len || si >= s_len) break;
              const uint8_t qb = QE[1 + qi], sb = SP[1 + si];
              if (qb >= 4 || sb >= 4 || qb != sb) break;
              ++er;
            }
            ok = el + er >= ext_needed;
            q_off = q0c - el;
            s_off = s_fwd - el;
          }
          diag = (int32_t)((uint32_t)s_off - (uint32_t)q_off);
          chk = (uint32_t)s_off;
          eq = q_off;
          es = s_off;
        } else {
          const int32_t rcq = q0c - c1;
          const int32_t q0 = q_len - rcq - lw, s0 = s_len - s_fwd - lw;
          int32_t q_off = q0, s_off = s0;
          ok = q0 >= 0 && s0 >= 0;
          if (ok && ext_needed > 0) {  // engine.rs:1201-1221
            int32_t el = 0, er = 0;
            while (er < ext_needed) {
              const int32_t qi = q0 + lw + er, si = s0 + lw + er;
              if (qi >= q_len || si >= s_len) break;
              const uint8_t qb = QE[1 + qi], sb = SR[1 + si];
              if (qb >= 4 || sb >= 4 || qb != sb) break;
              ++er;
            }
            while (el + er < ext_needed) {
              const int32_t qi = q0 - el - 1, si = s0 - el - 1;
              if (qi < 0 || si < 0) break;
              const uint8_t qb = QE[1 + qi], sb = SR[1 + si];
              if (qb >= 4 || sb >= 4 || qb != sb) break;
              ++el;
            }
            ok = el + er >= ext_needed;
            q_off = q0 + er;
            s_off = s0 + er;
          }
          diag = (int32_t)((uint32_t)s_fwd - (uint32_t)q0c);
          chk = (uint32_t)(s_len - s_off - lw);
          eq = q_len - 1 - q_off;
          es = s_len - 1 - s_off;
        }
        bool longer = false;
        // a seed on a diagonal of another part's buckets is that part's
        if (SPLIT) ok = ok && (((uint32_t)diag & (uint32_t)(kBuckets - 1)) >> own_shift) == part;
        if (ok) {
          // one extension call for both strands, so a wave holding seeds of
          // both runs it once
          const bool minus = q0c >= c1;
          const uint8_t* qpool = minus ? qrc : qfwd;
          const int64_t qa = (int64_t)R.q_base + 1 + eq, sa = (int64_t)S.s_base + 1 + es;
          UngOut u;
          // without the fast tier every seed takes the full extension below
          if (WORD && ungapped_first(qpool, qa, splus, sa, eq, minus, mat, P.xdrop_ungap, u)) {
            const bool hit = u.score >= R.cutoff;
            const uint32_t slot = put_hit(hit, u.q_beg, u.score);
            level = hit ? (uint32_t)es + (u.q_end - (uint32_t)eq) : chk + (uint32_t)ws;
            steps = (u.steps < 65535u ? u.steps : 65535u) | (hit ? slot << kHitSlotShift : 0u);
            flag = F_VALID | (hit ? F_HIT : 0) | (minus ? F_MINUS : 0);
          } else {  // the full extension runs below, over a compact list; level holds eq meanwhile
            longer = true;
            level = (uint32_t)eq;
            flag = F_VALID | F_LONG | (minus ? F_MINUS : 0);
          }
          atomicAdd(&bcnt[(uint32_t)diag & (kBuckets - 1)], 1u);
        } else {
          diag = 0;
          chk = 0;
        }
        sd[i] = make_uint4((uint32_t)diag, chk, level, (WORD ? steps : (steps < 65535u ? steps : 65535u)) | (flag << 16));
        // the list of longer extensions: one LDS atomic per wave
        const unsigned long long m = __ballot(longer);
        if (m) {
          const uint32_t lane = (uint32_t)tid & 63u;
          const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(&s_nlong, (uint32_t)__popcll(m));
          base = __shfl(base, (int)leader, 64);
          if (longer) llist[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)i;
        }
      }
      __syncthreads();
      // ---- C: full extensions of the listed seeds
      const uint32_t n_long = s_nlong;
      if (n_long) {
        for (uint32_t j = tid; j < n_long; j += kP1Threads) {
          const uint32_t i = llist[j];
          const uint4 e = sd[i];
          const uint32_t flag = (e.w >> 16) & (uint32_t)(F_VALID | F_MINUS);
          const bool minus = (flag & F_MINUS) != 0;
          const int32_t eq = (int32_t)e.z;
          const int32_t es = minus ? lw - 1 + (int32_t)e.y : (int32_t)e.y;
          const uint8_t* qpool = minus ? qrc : qfwd;
          const UngOut u = ungapped_chunked<kLongChunk>(qpool, (int64_t)R.q_base + 1 + eq, splus, (int64_t)S.s_base + 1 + es,
                                            eq, mat, P.xdrop_ungap);
          const bool hit = u.score >= R.cutoff;
          uint32_t steps = u.steps < 65535u ? u.steps : 65535u;
          if constexpr (WORD) {
            const uint32_t slot = put_hit(hit, u.q_beg, u.score);
            steps |= hit ? slot << kHitSlotShift : 0u;
          } else {
            if (hit) my_hitres[i] = make_uint4(u.q_beg, u.q_end, (uint32_t)u.score, (uint32_t)eq);
          }
          const uint32_t level = hit ? (uint32_t)es + (u.q_end - (uint32_t)eq) : e.y + (uint32_t)ws;
          sd[i] = make_uint4(e.x, e.y, level, steps | ((flag | (hit ? F_HIT : 0)) << 16));
        }
        __syncthreads();
      }
      tick(2);
      if (tid == 0) s_nlong = 0;  // every thread read it before the barrier above
      // ---- D: bucket lists (counting sort; each owner re-sorts by seed index)
      {
        const uint32_t v0 = bcnt[2 * tid], v1 = bcnt[2 * tid + 1];
        const uint32_t incl = block_incl_scan256(v0 + v1, wsum);
        bst[2 * tid] = incl - v0 - v1;
        bst[2 * tid + 1] = incl - v1;
      }
      __syncthreads();
      for (uint32_t i = tid; i < take; i += kP1Threads) {
        const uint4 e = sd[i];
        if ((e.w >> 16) & F_VALID) blist[atomicAdd(&bst[e.x & (kBuckets - 1)], 1u)] = (uint16_t)i;
      }
      __syncthreads();
      tick(3);
      // ---- E: exact replay of each bucket (diag_hash.rs:57-111).  Fast path:
      // a bucket with at most 8 seeds this pass whose cells fit the registers
      // sorts its seed list in registers and replays it branch-free; the
      // generic loop takes the rest (more seeds, or cells in the spill area).
      uint32_t nemit_local = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int bb = tid + h * kP1Threads;
        const uint32_t nb = bcnt[bb];
        const uint32_t st = bst[bb] - nb;
        uint32_t cnt_b = hc[h], l0 = hl0[h], l1 = hl1[h], l2 = hl2[h];
        int32_t d0 = hd0[h], d1 = hd1[h], d2 = hd2[h];
        bool slow = nb > 8 || cnt_b > (uint32_t)kHashLds;
        // seeds of the bucket the fast path replays: all of them, unless the
        // bucket needs the generic loop from the start or from a fourth cell
        uint32_t done_n = slow ? 0u : nb;
        if (__any(done_n > 0)) {
          uint32_t ix[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) ix[j] = (uint32_t)j < done_n ? (uint32_t)blist[st + j] : 0xFFFFu;
          sort8(ix);
          // unused register cells never match and are never stale, so the
          // tests below need no cell count; flags are written after the loop
          uint32_t emit_mask = 0, evals = 0, stepsum = 0;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const bool act = (uint32_t)j < done_n;
            if (!__any(act)) break;
            const uint4 e = sd[act ? ix[j] : 0u];
            const int32_t d = (int32_t)e.x;
            const uint32_t sc = e.y;
            // get(): newest cell first
            const bool m2 = d2 == d, m1 = d1 == d, m0 = d0 == d;
            const uint32_t lev = m2 ? l2 : (m1 ? l1 : (m0 ? l0 : 0u));
            // (bitwise & | on the tests: lane masks, no branches or selects)
            const bool eval = act & (sc >= lev);
            // insert(): the newest cell on the same diagonal or stale, else a
            // new cell; a fourth cell sends the bucket to the generic loop
            const bool s2 = (int32_t)(sc - l2) > window, s1 = (int32_t)(sc - l1) > window,
                       s0 = (int32_t)(sc - l0) > window;
            const bool u2 = m2 | s2;
            const bool u1 = !u2 & (m1 | s1);
            const bool u0 = !u2 & !u1 & (m0 | s0);
            const bool grow = !(u2 | u1 | u0);
            const bool to_slow = eval & grow & (cnt_b >= (uint32_t)kHashLds);
            done_n = to_slow ? (uint32_t)j : done_n;
            const bool apply = eval & !to_slow;
            evals += apply ? 1u : 0u;
            stepsum += apply ? (e.w & 0xFFFFu) : 0u;
            emit_mask |= (apply & (SEEDS || ((e.w >> 16) & F_HIT) != 0)) ? (1u << j) : 0u;
            const uint32_t L = e.z;
            const bool w2 = apply & (u2 | (grow & (cnt_b == 2)));
            const bool w1 = apply & (u1 | (grow & (cnt_b == 1)));
            const bool w0 = apply & (u0 | (grow & (cnt_b == 0)));
            d2 = w2 ? d : d2; l2 = w2 ? L : l2;
            d1 = w1 ? d : d1; l1 = w1 ? L : l1;
            d0 = w0 ? d : d0; l0 = w0 ? L : l0;
            cnt_b += (apply & grow) ? 1u : 0u;
          }
          slow = slow || done_n < nb;
          n_evals += evals;
          n_steps += stepsum;
          if (emit_mask) {
            nemit_local += (uint32_t)__popc(emit_mask);
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if ((emit_mask >> j) & 1u) sd[ix[j]].w |= (uint32_t)F_EMIT << 16;
          }
        }
        if (slow) {
          // generic replay from seed done_n: the list sorted in LDS, cells
          // beyond the registers in the workgroup's spill area
          for (uint32_t x = st + 1; x < st + nb; ++x) {
            const uint16_t v = blist[x];
            uint32_t y = x;
            while (y > st && blist[y - 1] > v) { blist[y] = blist[y - 1]; --y; }
            blist[y] = v;
          }
          uint2* GS = my_spill + (size_t)bb * kHashGlobal;
          for (uint32_t x = st + done_n; x < st + nb; ++x) {
            const uint32_t i = blist[x];
            const uint4 e = sd[i];
            const int32_t d = (int32_t)e.x;
            const uint32_t sc = e.y;
            // get(): newest cell first.  One pass over the spill cells, two
            // loads at a time (four cost the kernel register spills), finds
            // both the level (the newest cell on d) and insert()'s target (the
            // newest cell on d or stale): nothing is written between the two,
            // so they are the scans' answers.
            uint32_t lev = 0;
            bool found = false;
            int32_t ins = -1;
            for (int32_t k = (int32_t)cnt_b - 1; k >= kHashLds && !found; k -= 2) {
              uint2 gq[2];
#pragma unroll
              for (int q = 0; q < 2; ++q) gq[q] = GS[max(k - q, kHashLds) - kHashLds];
#pragma unroll
              for (int q = 0; q < 2; ++q) {
                if (k - q < kHashLds || found) continue;
                const bool same = (int32_t)gq[q].x == d;
                if (ins < 0 && (same || (int32_t)(sc - gq[q].y) > window)) ins = k - q;
                if (same) { lev = gq[q].y; found = true; }
              }
            }
            if (!found && cnt_b > 2 && d2 == d) { lev = l2; found = true; }
            if (!found && cnt_b > 1 && d1 == d) { lev = l1; found = true; }
            if (!found && cnt_b > 0 && d0 == d) lev = l0;
            if (sc < lev) continue;
            ++n_evals;
            n_steps += e.w & 0xFFFFu;
            if (SEEDS || ((e.w >> 16) & F_HIT)) {
              sd[i].w = e.w | ((uint32_t)F_EMIT << 16);
              ++nemit_local;
            }
            // insert(): same diagonal, else first stale cell, else a new head
            // cell; among the spill cells the scan above found the target
            const uint32_t L = e.z;
            bool done = false;
            if (ins >= 0) {
              GS[ins - kHashLds] = make_uint2((uint32_t)d, L);
              done = true;
            }
            if (!done && cnt_b > 2 && (d2 == d || (int32_t)(sc - l2) > window)) { d2 = d; l2 = L; done = true; }
            if (!done && cnt_b > 1 && (d1 == d || (int32_t)(sc - l1) > window)) { d1 = d; l1 = L; done = true; }
            if (!done && cnt_b > 0 && (d0 == d || (int32_t)(sc - l0) > window)) { d0 = d; l0 = L; done = true; }
            if (!done) {
              if (cnt_b >= (uint32_t)(kHashLds + kHashGlobal)) {
                atomicOr(err, (uint32_t)E_HASH_FULL);
              } else {
                if (cnt_b == 0) { d0 = d; l0 = L; }
                else if (cnt_b == 1) { d1 = d; l1 = L; }
                else if (cnt_b == 2) { d2 = d; l2 = L; }
                else GS[cnt_b - kHashLds] = make_uint2((uint32_t)d, L);
                ++cnt_b;
              }
            }
          }
        }
        hc[h] = cnt_b; hl0[h] = l0; hl1[h] = l1; hl2[h] = l2; hd0[h] = d0; hd1[h] = d1; hd2[h] = d2;
      }
      if (nemit_local) atomicAdd(&s_nemit, nemit_local);
      __syncthreads();
      tick(4);
      // ---- F: output slots for the pass's kept hits: the rest of the
      // workgroup's reserved chunk, then a new chunk (one global atomic per
      // kOutChunk hits)
      const uint32_t nemit = s_nemit;
      if (nemit) {
        if (tid == 0) {
          if (nemit <= s_out_left) {
            s_pass_a = s_out_next;
            s_pass_an = nemit;
            s_pass_b = 0;
            s_out_next += nemit;
            s_out_left -= nemit;
          } else {
            const uint32_t rest = nemit - s_out_left;
            const uint32_t chunk = rest > kOutChunk ? rest : kOutChunk;
            const unsigned long long nb = atomicAdd(out_count, (unsigned long long)chunk);
            s_pass_a = s_out_next;
            s_pass_an = s_out_left;
            s_pass_b = nb;
            s_out_next = nb + rest;
            s_out_left = chunk - rest;
          }
        }
        __syncthreads();
        tick(1);
        // emit kept hits in the record frame
        for (uint32_t i = tid; i < take; i += kP1Threads) {
          const uint4 e = sd[i];
          if (!((e.w >> 16) & F_EMIT)) continue;
          if constexpr (SEEDS) {
            const uint64_t ord = wseen + wk + i;
            UngHit h{};
            h.item = (uint32_t)item;
            h.q_seed = e.x;
            h.s_seed = e.y;
            h.q_start = (uint32_t)ord;
            h.q_end = (uint32_t)(ord >> 32);
            h.strand = ((e.w >> 16) & F_MINUS) ? 1u : 0u;
            const uint32_t k = atomicAdd(&s_emit_i, 1u);
            const unsigned long long pos = k < s_pass_an ? s_pass_a + k : s_pass_b + (k - s_pass_an);
            if (pos < P.out_cap) out[pos] = h;
            else atomicOr(err, (uint32_t)E_OUT_FULL);
            continue;
          }
          const bool minus = ((e.w >> 16) & F_MINUS) != 0;
          const int32_t es = minus ? lw - 1 + (int32_t)e.y : (int32_t)e.y;
          int32_t eq;
          uint32_t qb, qe;
          int32_t score;
          if constexpr (WORD) {
            // the frame from the seed record (diagonal = s_fwd - q0c, check
            // offset = s_fwd), the right end from the level a hit stores
            const uint32_t slot = e.w >> kHitSlotShift;
            const uint2 hv = slot < kHitLds ? s_hit[slot]
                                            : make_uint2(my_hitres[slot - kHitLds].x, my_hitres[slot - kHitLds].y);
            const int32_t q0c = (int32_t)(e.y - e.x);
            eq = minus ? q0c - c1 + lw - 1 : q0c;
            qb = hv.x;
            qe = e.z - (uint32_t)es + (uint32_t)eq;
            score = (int32_t)hv.y;
          } else {
            const uint4 hv = my_hitres[i];
            eq = (int32_t)hv.w;
            qb = hv.x;
            qe = hv.y;
            score = (int32_t)hv.z;
          }
          UngHit h;
          h.item = (uint32_t)item;
          h.score = score;
          if (!minus) {
            h.q_seed = (uint32_t)eq;
            h.s_seed = (uint32_t)es;
            h.q_start = qb;
            h.q_end = qe;
            h.s_start = (uint32_t)(es - (eq - (int32_t)qb));
            h.s_end = (uint32_t)es + (qe - (uint32_t)eq);
            h.strand = 0;
          } else {
            const uint32_t rs0 = (uint32_t)(es - (eq - (int32_t)qb));
            const uint32_t rs1 = (uint32_t)es + (qe - (uint32_t)eq);
            h.q_seed = (uint32_t)(q_len - 1 - eq);
            h.s_seed = (uint32_t)(s_len - 1 - es);
            h.q_start = (uint32_t)q_len - qe;
            h.q_end = (uint32_t)q_len - qb;
            h.s_start = (uint32_t)s_len - rs1;
            h.s_end = (uint32_t)s_len - rs0;
            h.strand = 1;
          }
          ++n_hits;
          const uint32_t k = atomicAdd(&s_emit_i, 1u);
          const unsigned long long pos = k < s_pass_an ? s_pass_a + k : s_pass_b + (k - s_pass_an);
          if (pos < P.out_cap) out[pos] = h;
          else atomicOr(err, (uint32_t)E_OUT_FULL);
        }
      }
      for (int b = tid; b < kBuckets; b += kP1Threads) bcnt[b] = 0;
      __syncthreads();
      tick(5);
      if (tid == 0) {  // read and updated only between this pass's E and F barriers
        s_nemit = 0;
        s_emit_i = 0;
        s_nhit = 0;    // written in B and C, read in F
      }
      wk += take;
      if (wk == wtotal) {
        w0 += kP1Threads;
        if constexpr (SEEDS) wseen += wtotal;
        scanned = false;
      }
    }
  }
  // mark the unused rest of the last reserved chunk
  for (uint32_t k = tid; k < s_out_left; k += kP1Threads) {
    const unsigned long long pos = s_out_next + k;
    if (pos < P.out_cap) out[pos].item = kNoItem;
  }
  atomicAdd(&counters[C_SEED_EVALS], n_evals);
  atomicAdd(&counters[C_UNG_STEPS], n_steps);
  atomicAdd(&counters[C_UNG_HITS], n_hits);
  if (tid == 0) {
    atomicAdd(&counters[C_SEEDS], s_nseeds);
    for (int k = 0; k < 6; ++k) atomicAdd(&counters[C_P1_STAGE + k], s_tick[k]);
  }
}

// Work queue shared by a launch: lane 0 of the wave takes `n` items from
// *next and every lane gets the first one's index.
__device__ __forceinline__ uint64_t wave_take(unsigned long long* next, uint64_t n) {
  unsigned long long base = 0;
  if ((threadIdx.x & 63u) == 0) base = atomicAdd(next, (unsigned long long)n);
  const uint32_t lo = __shfl((uint32_t)base, 0, 64), hi = __shfl((uint32_t)(base >> 32), 0, 64);
  return ((uint64_t)hi << 32) | lo;
}

// --------------------------------------------------------- X-drop DP ---------
// One lane runs one half alignment in the reference's row-major order, so every
// prune decision and band bound is the reference's.  The band [first_b, b_size)
// lives in a per-lane ring of 2^wlog2 cells indexed by the absolute column,
// cell k at ring[k * RS] (RS = 1: a lane's own stretch of global memory; RS =
// the workgroup size: an LDS ring interleaved across the lanes); a band wider
// than the ring sets E_BAND_WIDE and the host reruns the task with a wider
// ring.  With TB, one traceback byte per band cell is kept for rows up to
// stop_row (the best row from the score pass) and the walk runs at the end.
static constexpr uint8_t SCRIPT_GAP_IN_A = 0, SCRIPT_SUB = 3, SCRIPT_GAP_IN_B = 6, SCRIPT_OP_MASK = 7,
                         SCRIPT_EXTEND_GAP_A = 0x10, SCRIPT_EXTEND_GAP_B = 0x40;

// Bytes p[idx] (p[n - 1 - idx] when rev) for idx rising by about one per call,
// through 16-byte aligned blocks: the block after the current one (in the
// direction of travel) is requested as soon as the current one is entered, so
// a lane waits on memory only when it jumps.  Blocks are loaded only where they
// hold bytes of [p, p + n): an aligned block that holds a byte of an allocation
// lies within the allocation's 256-byte-aligned start and 64-byte tail.
struct SeqStream {
  uintptr_t lo, hi;  // readable bytes [lo, hi)
  int64_t cb;        // current block (address >> 4)
  uint4 cur, nxt;
  bool rev;
  __device__ __forceinline__ void init(const uint8_t* p, uint32_t n, bool r) {
    lo = (uintptr_t)p;
    hi = (uintptr_t)p + n;
    rev = r;
    cb = -4;
  }
  __device__ __forceinline__ uint4 load(int64_t b) const {
    if (b < (int64_t)(lo >> 4) || b > (int64_t)((hi - 1) >> 4)) return make_uint4(0u, 0u, 0u, 0u);
    return *reinterpret_cast<const uint4*>((uintptr_t)b << 4);
  }
  __device__ __forceinline__ uint32_t at(uint32_t idx) {  // idx < n
    const uintptr_t addr = rev ? hi - 1 - idx : lo + idx;
    const int64_t b = (int64_t)(addr >> 4), step = rev ? -1 : 1;
    if (b != cb) {
      cur = b == cb + step ? nxt : load(b);
      // cur is needed complete here, where the block changes: the wait for
      // it then sits in this branch, not at every read of a byte (a vector
      // memory wait also drains the lane's traceback stores queued since),
      // and the next block's load goes out after it
      __asm__ volatile("" ::"v"(cur.x), "v"(cur.y), "v"(cur.z), "v"(cur.w));
      nxt = load(b + step);
      cb = b;
    }
    // word (o >> 2) of cur by masks, which keeps cur in registers
    const uint32_t o = (uint32_t)addr & 15u;
    const uint32_t m8 = 0u - ((o >> 3) & 1u), m4 = 0u - ((o >> 2) & 1u);
    const uint32_t a0 = cur.x ^ ((cur.x ^ cur.z) & m8), a1 = cur.y ^ ((cur.y ^ cur.w) & m8);
    const uint32_t w = a0 ^ ((a0 ^ a1) & m4);
    return (w >> ((o & 3u) * 8u)) & 0xFFu;
  }
};

template <bool TB, uint32_t RS = 1>
__device__ __forceinline__ void dp_half(const DpTask& t, const int* mat, int32_t go, int32_t ge,
                                        int2* __restrict__ ring, uint8_t* __restrict__ bring, uint32_t wmask,
                                        DpResult& res,
                                        uint8_t* __restrict__ flat, uint32_t* __restrict__ rows, uint32_t stop_row,
                                        uint32_t flat_cap = 0xFFFFFFFFu, uint32_t rows_cap = 0xFFFFFFFFu) {
  res.score = 0; res.a_len = 0; res.b_len = 0; res.flags = 0; res.cells = 0; res.tb_bytes = 0; res.tb_rows = 0;
  const uint32_t m = t.m, n = t.n;
  if (m == 0 || n == 0) return;
  const bool rev = t.reverse != 0;
  const int32_t oe = go + ge;
  const int32_t xdrop = t.xdrop > oe ? t.xdrop : oe;
  // With rings in LDS (RS > 1) the b code of each band column rides in a byte
  // ring beside its cell, read from global memory once when the column joins
  // the band; the row loop then reads only LDS, one cell ahead.
  constexpr bool BR = RS > 1;
  SeqStream as, bs;
  as.init(t.a, m, rev);
  bs.init(t.b, n, rev);
  auto bcode = [&](uint32_t col) -> uint32_t { return col < n ? (bs.at(col) & 15u) : 0u; };
  uint32_t flat_len = 0;
  ring[0] = make_int2(0, -oe);
  if (BR) bring[0] = (uint8_t)bcode(0);
  if (TB) { rows[0] = 0; rows[1] = 0; flat[flat_len] = SCRIPT_GAP_IN_B; }
  ++flat_len;
  int32_t score = -oe;
  uint32_t b_size = 1;
  for (uint32_t j = 1; j <= n; ++j) {
    if (score < -xdrop) break;
    if (j > wmask) { res.flags |= E_BAND_WIDE; return; }
    ring[(j & wmask) * RS] = make_int2(score, score - oe);
    if (BR) bring[(j & wmask) * RS] = (uint8_t)bcode(j);
    if (TB) flat[flat_len] = SCRIPT_GAP_IN_A;
    ++flat_len;
    score -= ge;
    b_size = j + 1;
  }
  int32_t best = 0;
  uint32_t best_a = 0, best_b = 0, first_b = 0;
  uint64_t cells = 0;
  uint32_t bytes_at_best = flat_len, rows_at_best = 1;
  for (uint32_t ai = 0; ai < m; ++ai) {
    if (TB && ai >= stop_row) break;
    // a row adds one row-table entry and at most two rings' worth of bytes: its
    // band (at most wmask + 1 cells past the row's first column) plus the
    // columns its first cells drop
    if (TB && (ai + 2 > rows_cap || flat_len + 2 * wmask + 4 > flat_cap)) { res.flags |= E_TB_SPACE; return; }
    const int* row = mat + ((as.at(ai) & 15u) << 4);
    if (TB) { rows[2 * (ai + 1)] = flat_len; rows[2 * (ai + 1) + 1] = first_b; }
    int32_t sc = kMinInt, sgr = kMinInt;
    uint32_t last_b = first_b;
    bool best_here = false;
    cells += b_size - first_b;
    int2 cn = ring[(first_b & wmask) * RS];
    uint32_t bn = BR ? (uint32_t)bring[(first_b & wmask) * RS] : bcode(first_b);
    for (uint32_t bi = first_b; bi < b_size; ++bi) {
      int2 c = cn;
      const int32_t sub = row[bn];
      {  // the next cell's inputs, loaded while this one is computed (no write of this row touches its slot)
        const uint32_t ns = ((bi + 1) & wmask) * RS;
        cn = ring[ns];
        bn = BR ? (uint32_t)bring[ns] : bcode(bi + 1);
      }
      int32_t sgc = c.y;
      const int32_t next = c.x + sub;
      if constexpr (!TB) {
        // score only: both outcomes computed and selected, no branch per cell.
        // A leading pruned cell leaves the band, so rewriting its slot (the
        // column's next use writes it first) changes nothing.
        const int32_t cs = max(sc, max(sgc, sgr));
        const bool prune = best - cs > xdrop;
        const int32_t open = cs - oe;
        ring[(bi & wmask) * RS] = prune ? make_int2(kMinInt, sgc) : make_int2(cs, max(sgc - ge, open));
        first_b += (prune && bi == first_b) ? 1u : 0u;
        last_b = prune ? last_b : bi;
        const bool nb = cs > best;  // never a pruned cell
        best_a = nb ? ai + 1 : best_a;
        best_b = nb ? bi : best_b;
        best_here = best_here || nb;
        best = nb ? cs : best;
        sgr = prune ? sgr : max(sgr - ge, open);
        sc = next;
        ++flat_len;
        continue;
      }
      uint8_t op = SCRIPT_SUB;
      int32_t cs = sc;
      if (sgc > cs) { cs = sgc; op = SCRIPT_GAP_IN_B; }
      if (sgr > cs) { cs = sgr; op = SCRIPT_GAP_IN_A; }
      if (best - cs > xdrop) {
        if (bi == first_b) ++first_b;
        else ring[(bi & wmask) * RS].x = kMinInt;
      } else {
        last_b = bi;
        if (cs > best) { best = cs; best_a = ai + 1; best_b = bi; best_here = true; }
        sgr -= ge;
        sgc -= ge;
        const int32_t open = cs - oe;
        if (sgc < open) c.y = open;
        else { c.y = sgc; op |= SCRIPT_EXTEND_GAP_B; }
        c.x = cs;
        ring[(bi & wmask) * RS] = c;
        if (sgr < open) sgr = open;
        else op |= SCRIPT_EXTEND_GAP_A;
      }
      sc = next;
      if (TB) flat[flat_len] = op;
      ++flat_len;
    }
    if (first_b >= b_size) {
      if (best_here) { bytes_at_best = flat_len; rows_at_best = ai + 2; }
      break;
    }
    if (last_b + 1 < b_size) {
      b_size = last_b + 1;
    } else {
      while (sgr >= best - xdrop && b_size <= n) {
        if (b_size - first_b > wmask) { res.flags |= E_BAND_WIDE; return; }
        ring[(b_size & wmask) * RS] = make_int2(sgr, sgr - oe);
        if (BR) bring[(b_size & wmask) * RS] = (uint8_t)bcode(b_size);
        if (TB) flat[flat_len] = SCRIPT_GAP_IN_A;
        ++flat_len;
        sgr -= ge;
        ++b_size;
      }
    }
    if (b_size <= n) {
      if (b_size - first_b > wmask) { res.flags |= E_BAND_WIDE; return; }
      ring[(b_size & wmask) * RS] = make_int2(kMinInt, kMinInt);
      if (BR) bring[(b_size & wmask) * RS] = (uint8_t)bcode(b_size);
      ++b_size;
    }
    if (best_here) { bytes_at_best = flat_len; rows_at_best = ai + 2; }
  }
  res.score = best;
  res.a_len = best_a;
  res.b_len = best_b;
  res.cells = cells;
  res.tb_bytes = bytes_at_best;
  res.tb_rows = rows_at_best;
}

static constexpr uint32_t kDpLdsThreads = 64, kDpLdsLog2 = 6;
// wider one-pass rings (81 KB of LDS per workgroup, one per CU), for
// X-drops whose bands outgrow 64 cells
static constexpr uint32_t kDpLdsWideLog2 = 8;
// words of traceback bytes a lane gathers before writing them out at once
static constexpr uint32_t kTbChunkWords = 8;
// Row trailers of the traceback stream: a row holds at most twice the ring's
// cells, so one byte per field serves rings of up to 64 cells, two bytes wider
// ones.
__host__ __device__ constexpr uint32_t tb_trailer_bytes(uint32_t rlog2) { return rlog2 <= 6 ? 2u : 4u; }
__host__ __device__ constexpr uint32_t tb_trailer_max(uint32_t tr) { return tr == 2 ? 255u : 65535u; }
struct TbRowStart {  // where the best cell's row starts in the stream
  uint32_t off, fb, dfb;  // first byte, first column, first column's move from the row below
};

// The traceback DP of k_dp_onepass with 4-byte LDS ring cells, so twice as
// many lanes fit a CU: H and E as 16-bit (or, with C12, 12-bit) offsets from
// the best score at the start of the row that wrote the cell (every band cell
// is rewritten in each row it stays in the band, a pruned cell with its E
// re-based, so one base per row serves).  H of a pruned cell is the sentinel
// (the format's lowest value), and E is held one above it at the lowest: such
// an E lies more than 2,000 below the best score, beyond X-drop plus the open
// cost, so it loses every maximum it meets and never keeps a cell; the
// traceback code of a cell whose every candidate is that low is never read by
// the walk.  Every other value is exact, so every decision is dp_half's.  A
// task for which a kept value could leave the format's range comes back
// E_BAND_WIDE, for the retry with dp_half.  Without TB it is the score-only
// DP (the preliminary DPs' wide rings), rings interleaved across RS lanes.
template <uint32_t RLOG2, bool TB = true, uint32_t RS = kDpLdsThreads, bool C12 = false>
__device__ __forceinline__ void dp_half_tb16(const DpTask& t, const int* mat, int32_t go, int32_t ge,
                                             uint32_t* __restrict__ ring, uint8_t* __restrict__ bring,
                                             int32_t max_pos, DpResult& res, uint8_t* __restrict__ flat,
                                             uint32_t flat_cap, TbRowStart& at_best) {
  constexpr uint32_t wmask = (1u << RLOG2) - 1;
  constexpr uint32_t TR = tb_trailer_bytes(RLOG2);
  res.score = 0; res.a_len = 0; res.b_len = 0; res.flags = 0; res.cells = 0; res.tb_bytes = 0; res.tb_rows = 0;
  const uint32_t m = t.m, n = t.n;
  if (m == 0 || n == 0) return;
  const bool rev = t.reverse != 0;
  const int32_t oe = go + ge;
  const int32_t xdrop = t.xdrop > oe ? t.xdrop : oe;
  // C12: H and E as 12-bit offsets and the column's base code in bits 24-27
  // of the same cell, so no base-code ring is kept.  A cell of a row lies at
  // most the largest score above the row's base (its candidates come from
  // cells no higher than the best score before the row, plus one score), and
  // a kept cell at most X below it, so 12 bits hold every kept value while
  // X + open cost + the largest score stays below 2,000.
  constexpr uint32_t kBits = C12 ? 12u : 16u, kMask = (1u << kBits) - 1;
  if (C12 ? (int64_t)max_pos + xdrop + oe >= 2000 : (int64_t)max_pos * (wmask + 1) + xdrop + oe >= 30000) {
    res.flags |= E_BAND_WIDE;
    return;
  }
  constexpr int32_t kNegH = -(1 << (kBits - 1)), kLowE = kNegH + 1;
  auto enc = [](int32_t h, int32_t e) -> uint32_t {
    return ((uint32_t)h & kMask) | (((uint32_t)(e < kLowE ? kLowE : e) & kMask) << kBits);
  };
  auto dec_h = [](uint32_t c) -> int32_t { return (int32_t)(c << (32 - kBits)) >> (32 - kBits); };
  auto dec_e = [](uint32_t c) -> int32_t { return (int32_t)(c << (32 - 2 * kBits)) >> (32 - kBits); };
  // a new column's cell: its scores and, with C12, its base code
  auto put_col = [&](uint32_t col, uint32_t cell, uint32_t code) {
    if constexpr (C12) ring[(col & wmask) * RS] = cell | (code << 24);
    else {
      ring[(col & wmask) * RS] = cell;
      bring[(col & wmask) * RS] = (uint8_t)code;
    }
  };
  SeqStream as, bs;
  as.init(t.a, m, rev);
  bs.init(t.b, n, rev);
  auto bcode = [&](uint32_t col) -> uint32_t { return col < n ? (bs.at(col) & 15u) : 0u; };
  // The traceback stream: each row's bytes, then its trailer (the row's byte
  // count and how far its first column moved from the row before, TR/2 bytes
  // each), so the walk steps down a row from the trailer below the row it is
  // in, and no row table is written.  The stream leaves the lane in whole
  // aligned chunks of kTbChunkWords words, each written once: a lane's arena
  // lines are then never written back part-filled and written again.  Bytes
  // gather in `acc`, completed words shift through `q`; the last chunk is
  // padded, and its bytes past the end are never read.
  uint32_t flat_len = 0, acc = 0;
  uint32_t q[kTbChunkWords];
#pragma unroll
  for (uint32_t k = 0; k < kTbChunkWords; ++k) q[k] = 0;
  auto shift_in = [&](uint32_t w) {
#pragma unroll
    for (uint32_t k = 0; k + 1 < kTbChunkWords; ++k) q[k] = q[k + 1];
    q[kTbChunkWords - 1] = w;
  };
  auto store_chunk = [&](uint32_t at) {
#pragma unroll
    for (uint32_t k = 0; k < kTbChunkWords; k += 4)
      *reinterpret_cast<uint4*>(flat + at + 4 * k) = make_uint4(q[k], q[k + 1], q[k + 2], q[k + 3]);
  };
  auto put = [&](uint32_t v) {
    if constexpr (TB) {
      const uint32_t sh = (flat_len & 3u) * 8u;
      acc = (sh ? acc : 0u) | (v << sh);
      if (sh == 24u) {
        shift_in(acc);
        if (((flat_len + 1) & (4 * kTbChunkWords - 1)) == 0) store_chunk(flat_len + 1 - 4 * kTbChunkWords);
      }
      ++flat_len;
    }
  };
  uint32_t row_off = 0, row_fb = 0, row_dfb = 0;  // the current row's start, first column, its move
  auto put_trailer = [&]() -> bool {
    const uint32_t len = flat_len - row_off;
    if (len > tb_trailer_max(TR) || row_dfb > tb_trailer_max(TR)) return false;
    if constexpr (TR == 2) {
      put(len);
      put(row_dfb);
    } else {
      put(len & 255u);
      put(len >> 8);
      put(row_dfb & 255u);
      put(row_dfb >> 8);
    }
    return true;
  };
  put_col(0, enc(0, -oe), bcode(0));
  put(SCRIPT_GAP_IN_B);
  int32_t score = -oe;
  uint32_t b_size = 1;
  for (uint32_t j = 1; j <= n; ++j) {
    if (score < -xdrop) break;
    if (j > wmask) { res.flags |= E_BAND_WIDE; return; }
    put_col(j, enc(score, score - oe), bcode(j));
    put(SCRIPT_GAP_IN_A);
    score -= ge;
    b_size = j + 1;
  }
  int32_t best = 0, cbase = 0;  // cbase: base of the cells written in this row
  uint32_t best_a = 0, best_b = 0, first_b = 0;
  uint64_t cells = 0;
  uint32_t bytes_at_best = flat_len, rows_at_best = 1;
  if (TB) at_best = TbRowStart{0, 0, 0};
  for (uint32_t ai = 0; ai < m; ++ai) {
    if (TB) {
      // a row's bytes stay within twice the ring (its first column moves
      // right inside the row by at most what the ring held at the start)
      if (flat_len + 2 * (wmask + 1) + 2 * TR + 4 * kTbChunkWords > flat_cap || !put_trailer()) {
        res.flags |= E_TB_SPACE;
        return;
      }
      row_off = flat_len;
      row_dfb = first_b - row_fb;
      row_fb = first_b;
    }
    const int32_t pbase = cbase;  // the cells read in this row were written with it
    cbase = best;
    const int* row = mat + ((as.at(ai) & 15u) << 4);
    int32_t sc = kMinInt, sgr = kMinInt;
    uint32_t last_b = first_b;
    bool best_here = false;
    cells += b_size - first_b;
    uint32_t cn = ring[(first_b & wmask) * RS];
    uint32_t bn = C12 ? (cn >> 24) & 15u : bring[(first_b & wmask) * RS];
    for (uint32_t bi = first_b; bi < b_size; ++bi) {
      const uint32_t cv = cn;
      const int32_t sub = row[bn];
      {  // the next cell's inputs, loaded while this one is computed
        const uint32_t ns = ((bi + 1) & wmask) * RS;
        cn = ring[ns];
        bn = C12 ? (cn >> 24) & 15u : bring[ns];
      }
      const int32_t hr = dec_h(cv);
      const int32_t cx = hr == kNegH ? kMinInt : hr + pbase;
      int32_t sgc = dec_e(cv) + pbase;
      const uint32_t code_bits = C12 ? cv & 0x0F000000u : 0u;  // kept with the column's cell
      const int32_t next = cx + sub;
      if constexpr (!TB) {
        // score only: both outcomes computed and selected, as in dp_half
        const int32_t cs = max(sc, max(sgc, sgr));
        const bool prune = best - cs > xdrop;
        const int32_t open = cs - oe;
        ring[(bi & wmask) * RS] =
            (prune ? enc(kNegH, sgc - cbase) : enc(cs - cbase, max(sgc - ge, open) - cbase)) | code_bits;
        first_b += (prune && bi == first_b) ? 1u : 0u;
        last_b = prune ? last_b : bi;
        const bool nb = cs > best;  // never a pruned cell
        best_a = nb ? ai + 1 : best_a;
        best_b = nb ? bi : best_b;
        best_here = best_here || nb;
        best = nb ? cs : best;
        sgr = prune ? sgr : max(sgr - ge, open);
        sc = next;
        continue;
      }
      uint8_t op = SCRIPT_SUB;
      int32_t cs = sc;
      if (sgc > cs) { cs = sgc; op = SCRIPT_GAP_IN_B; }
      if (sgr > cs) { cs = sgr; op = SCRIPT_GAP_IN_A; }
      if (best - cs > xdrop) {
        if (bi == first_b) ++first_b;
        else ring[(bi & wmask) * RS] = enc(kNegH, sgc - cbase) | code_bits;
      } else {
        last_b = bi;
        if (cs > best) { best = cs; best_a = ai + 1; best_b = bi; best_here = true; }
        sgr -= ge;
        sgc -= ge;
        const int32_t open = cs - oe;
        int32_t cy;
        if (sgc < open) cy = open;
        else { cy = sgc; op |= SCRIPT_EXTEND_GAP_B; }
        ring[(bi & wmask) * RS] = enc(cs - cbase, cy - cbase) | code_bits;
        if (sgr < open) sgr = open;
        else op |= SCRIPT_EXTEND_GAP_A;
      }
      sc = next;
      put(op);
    }
    if (first_b >= b_size) {
      if (best_here) {
        bytes_at_best = flat_len;
        rows_at_best = ai + 2;
        if (TB) at_best = TbRowStart{row_off, row_fb, row_dfb};
      }
      break;
    }
    if (last_b + 1 < b_size) {
      b_size = last_b + 1;
    } else {
      while (sgr >= best - xdrop && b_size <= n) {
        if (b_size - first_b > wmask) { res.flags |= E_BAND_WIDE; return; }
        put_col(b_size, enc(sgr - cbase, sgr - oe - cbase), bcode(b_size));
        put(SCRIPT_GAP_IN_A);
        sgr -= ge;
        ++b_size;
      }
    }
    if (b_size <= n) {
      if (b_size - first_b > wmask) { res.flags |= E_BAND_WIDE; return; }
      put_col(b_size, enc(kNegH, kLowE), bcode(b_size));
      ++b_size;
    }
    if (best_here) {
      bytes_at_best = flat_len;
      rows_at_best = ai + 2;
      if (TB) at_best = TbRowStart{row_off, row_fb, row_dfb};
    }
  }
  if (TB) {
    if (flat_len & (4 * kTbChunkWords - 1)) {
      uint32_t words = (flat_len >> 2) & (kTbChunkWords - 1);
      if (flat_len & 3u) { shift_in(acc); ++words; }
      for (; words < kTbChunkWords; ++words) shift_in(0);
      store_chunk(flat_len & ~(4 * kTbChunkWords - 1));
    }
  }
  res.score = best;
  res.a_len = best_a;
  res.b_len = best_b;
  res.cells = cells;
  res.tb_bytes = bytes_at_best;
  res.tb_rows = rows_at_best;
}

// ------------------------------------------ wave-cooperative X-drop DP ------
// One wave runs one half alignment, a row at a time: lane l holds band column
// first_b + l (H and E of the row before, the column's b code), so a row of up
// to 64 cells is one vector step instead of a lane's serial walk along it.
// The row's serial dependencies are resolved exactly:
//  * G_j = max(diagonal, E_j) needs only the row before;
//  * the horizontal gap F obeys F_{j+1} = max(F_j - ge, G_j - oe) after a kept
//    cell and F_{j+1} = F_j after a pruned one (dp_half leaves sgr alone
//    there), an affine map in (max, +) composed by a prefix scan;
//  * the running best before cell j is max(best0, cs_k for k < j) over every
//    earlier cell, kept or not (a pruned cell lies below best - X), so it is
//    a prefix max;
//  * whether a cell is kept depends on the two scans, which depend on which
//    earlier cells are kept.  The row is iterated from a guess: one round's
//    decisions are exact up to the first cell whose guess was wrong, so the
//    rounds stop, at the latest after one per cell, with every decision
//    dp_half's.
// The band bookkeeping (first_b over the leading pruned run, b_size cut to the
// last kept cell or extended while F stays within X of best, the sentinel
// cell) and the traceback bytes and row table are dp_half's, so the result,
// the bytes and the walk are identical.  Bands wider than 64 cells come back
// E_BAND_WIDE, as from the 64-cell LDS ring.

// lane i takes lane i - 1's value of v; lane 0 takes `fill`
__device__ __forceinline__ int32_t wave_shr1(int32_t v, int32_t fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}
// v from the lane `s` rows of 16 lanes back within its row (row_shr:s), or id
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ int32_t dpp_or(int32_t v, int32_t id) {
  return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWMASK, 0xF, false);
}
// inclusive prefix max over the wave's 64 lanes
__device__ __forceinline__ int32_t wave_prefix_max(int32_t v, int32_t id) {
  v = max(v, dpp_or<0x111>(v, id));  // row_shr:1
  v = max(v, dpp_or<0x112>(v, id));  // row_shr:2
  v = max(v, dpp_or<0x114>(v, id));  // row_shr:4
  v = max(v, dpp_or<0x118>(v, id));  // row_shr:8
  v = max(v, dpp_or<0x142, 0xA>(v, id));  // row_bcast:15 into rows 1 and 3
  v = max(v, dpp_or<0x143, 0xC>(v, id));  // row_bcast:31 into rows 2 and 3
  return v;
}
// inclusive prefix composition of the maps x -> max(x + A, B), earlier lanes
// first: (A1, B1) then (A2, B2) is (A1 + A2, max(B1 + A2, B2)); (0, neg) is
// the identity
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ void affine_step(int32_t& A, int32_t& B, int32_t neg) {
  const int32_t pa = dpp_or<CTRL, ROWMASK>(A, 0), pb = dpp_or<CTRL, ROWMASK>(B, neg);
  B = max(pb + A, B);
  A = pa + A;
}
__device__ __forceinline__ void wave_prefix_affine(int32_t& A, int32_t& B, int32_t neg) {
  affine_step<0x111>(A, B, neg);
  affine_step<0x112>(A, B, neg);
  affine_step<0x114>(A, B, neg);
  affine_step<0x118>(A, B, neg);
  affine_step<0x142, 0xA>(A, B, neg);
  affine_step<0x143, 0xC>(A, B, neg);
}
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Byte `idx` of a sequence read as dp_half reads it (reversed when rev).
__device__ __forceinline__ uint32_t seq_code(const uint8_t* p, uint32_t len, bool rev, uint32_t idx) {
  return idx < len ? (uint32_t)(p[rev ? len - 1 - idx : idx] & 15u) : 0u;
}

struct WaveDpOut {
  DpResult res;
  uint32_t rounds;  // scan rounds over all rows (diagnostics)
};

// Requires a full wave in lockstep (all 64 lanes call it with the same task).
// Row table and traceback bytes as dp_half<true>; results as dp_half.
__device__ __forceinline__ void dp_half_wave(const DpTask& t, const int* __restrict__ mat, int32_t go, int32_t ge,
                                          DpResult& res, uint8_t* __restrict__ flat, uint32_t* __restrict__ rows,
                                          uint32_t flat_cap, uint32_t rows_cap, uint32_t& rounds) {
  constexpr uint32_t wmask = 63;
  constexpr int32_t kNeg = INT32_MIN + (1 << 27);  // below every reachable value, and safe to add to
  const uint32_t lane = threadIdx.x & 63u;
  res.score = 0; res.a_len = 0; res.b_len = 0; res.flags = 0; res.cells = 0; res.tb_bytes = 0; res.tb_rows = 0;
  const uint32_t m = t.m, n = t.n;
  if (m == 0 || n == 0) return;
  const bool rev = t.reverse != 0;
  const int32_t oe = go + ge;
  const int32_t xdrop = t.xdrop > oe ? t.xdrop : oe;
  // row 0: columns 0..J with H = -oe - (j - 1) ge while it stays within X
  uint32_t J;
  if (ge <= 0) J = n;
  else J = min(n, (uint32_t)((xdrop - oe) / ge) + 1u);
  if (J > wmask) { res.flags |= E_BAND_WIDE; return; }
  // b codes of columns [wb, wb + 192) in three registers, a codes of rows
  // [ab, ab + 64) in one: lanes hold one byte of each
  uint32_t wb = 0;
  uint32_t bw0 = seq_code(t.b, n, rev, lane), bw1 = seq_code(t.b, n, rev, 64 + lane),
           bw2 = seq_code(t.b, n, rev, 128 + lane);
  uint32_t ab = 0;
  uint32_t aw = seq_code(t.a, m, rev, lane), awn = seq_code(t.a, m, rev, 64 + lane);
  // col in [wb, wb + 128) (first_b - wb < 64 holds); bw2 is the next slide's
  // prefetch.  Every lane calls it.
  auto bcode_of = [&](uint32_t col) -> uint32_t {
    const uint32_t o = col - wb;
    const int src = (int)((o & 63u) << 2);
    const uint32_t v0 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)bw0);
    const uint32_t v1 = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)bw1);
    return o < 64 ? v0 : v1;
  };
  // the band: lane l holds column first_b + l while first_b + l < b_size
  uint32_t first_b = 0, b_size = J + 1;
  int32_t H = lane == 0 ? 0 : -oe - (int32_t)(lane - 1) * ge;
  int32_t E = H - oe;
  uint32_t bc = bcode_of(lane);
  if (lane == 0) { rows[0] = 0; rows[1] = 0; }
  if (lane <= J) flat[lane] = lane == 0 ? SCRIPT_GAP_IN_B : SCRIPT_GAP_IN_A;
  uint32_t flat_len = J + 1;
  int32_t best = 0;
  uint32_t best_a = 0, best_b = 0;
  uint64_t cells = 0;
  uint32_t bytes_at_best = flat_len, rows_at_best = 1;
  for (uint32_t ai = 0; ai < m; ++ai) {
    if (ai + 2 > rows_cap || flat_len + 2 * wmask + 4 > flat_cap) { res.flags |= E_TB_SPACE; return; }
    if (ai - ab == 64) {
      ab += 64;
      aw = awn;
      awn = seq_code(t.a, m, rev, ab + 64 + lane);
    }
    const uint32_t acode = (uint32_t)__builtin_amdgcn_readlane((int)aw, (int)(ai - ab));
    const uint32_t w = b_size - first_b;
    const bool act = lane < w;
    if (lane == 0) { rows[2 * (ai + 1)] = flat_len; rows[2 * (ai + 1) + 1] = first_b; }
    cells += w;
    const int32_t sub = mat[(acode << 4) | bc];
    const int32_t D = wave_shr1(H + sub, kMinInt);  // sc: the cell before's H + its substitution
    const int32_t V = E;
    const int32_t G = max(D, V);
    // iterate the keep decisions to the fixed point
    // (kept iff cs >= best - X, i.e. not best - cs > X)
    bool kept = act && G >= best - xdrop;
    int32_t F, cs, f_after;
    uint64_t kmask = wave_ballot(kept);
    for (uint32_t round = 0;; ++round) {
      // each round settles at least the first unsettled cell, so 65 rounds
      // settle a row of 64; more would mean a broken scan: hand the half back
      if (round > 64) { res.flags |= E_BAND_WIDE; return; }
      ++rounds;
      int32_t A = kept ? -ge : 0, B = kept ? G - oe : kNeg;
      wave_prefix_affine(A, B, kNeg);
      f_after = max(kMinInt + A, B);  // F after this lane's cell
      F = wave_shr1(f_after, kMinInt);
      cs = act ? max(G, F) : kNeg;
      const int32_t bb = max(best, wave_shr1(wave_prefix_max(cs, kNeg), kNeg));
      const bool k2 = act && cs >= bb - xdrop;
      const uint64_t m2 = wave_ballot(k2);
      if (m2 == kmask) break;
      kept = k2;
      kmask = m2;
    }
    // traceback byte of each cell (dp_half's choice and extension flags)
    uint8_t op = SCRIPT_SUB;
    {
      int32_t c0 = D;
      if (V > c0) { c0 = V; op = SCRIPT_GAP_IN_B; }
      if (F > c0) op = SCRIPT_GAP_IN_A;
      if (kept) {
        if (V - ge >= cs - oe) op |= SCRIPT_EXTEND_GAP_B;
        if (F - ge >= cs - oe) op |= SCRIPT_EXTEND_GAP_A;
      }
    }
    if (act) flat[flat_len + lane] = op;
    if (kmask == 0) break;  // every cell pruned: first_b reaches b_size
    const uint32_t lead = (uint32_t)__builtin_ctzll(kmask);
    const uint32_t last = 63u - (uint32_t)__builtin_clzll(kmask);
    // the row's best: the first cell holding the row maximum, if it beats best
    const int32_t rmax = __builtin_amdgcn_readlane(wave_prefix_max(cs, kNeg), (int)(w - 1));
    bool best_here = false;
    if (rmax > best) {
      const uint64_t at = wave_ballot(act && cs == rmax);
      best = rmax;
      best_a = ai + 1;
      best_b = first_b + (uint32_t)__builtin_ctzll(at);
      best_here = true;
    }
    // the row's cells for the next row: kept ones take (cs, E'), pruned ones
    // (not leading) (minimum, E unchanged)
    int32_t Hn = kept ? cs : kMinInt;
    int32_t En = kept ? max(V - ge, cs - oe) : V;
    flat_len += w;
    const uint32_t nfirst = first_b + lead;
    uint32_t nsize = b_size;
    uint32_t ext = 0;
    int32_t f_end = 0;
    if (last + 1 < w) {
      nsize = first_b + last + 1;
    } else {
      // F after the last cell, then the run of GAP_IN_A cells while F - k ge
      // stays within X of best
      f_end = __builtin_amdgcn_readlane(f_after, (int)(w - 1));
      const int32_t floor_v = best - xdrop;
      if (f_end >= floor_v) {
        const uint32_t room = n + 1 - b_size;  // columns b_size..n
        uint32_t k = ge <= 0 ? room : (uint32_t)((f_end - floor_v) / ge) + 1u;
        ext = min(k, room);
      }
      if (ext && (b_size + ext - 1) - nfirst > wmask) { res.flags |= E_BAND_WIDE; return; }
      if (lane < ext) flat[flat_len + lane] = SCRIPT_GAP_IN_A;
      flat_len += ext;
      nsize = b_size + ext;
    }
    uint32_t sentinel = 0;
    if (nsize <= n) {
      if (nsize - nfirst > wmask) { res.flags |= E_BAND_WIDE; return; }
      sentinel = 1;
    }
    // shift the band to start at nfirst: lane l takes column nfirst + l
    {
      const uint32_t col = nfirst + lane;
      const int src = (int)(((lane + lead) & 63u) << 2);
      const int32_t hs = __builtin_amdgcn_ds_bpermute(src, Hn);
      const int32_t es = __builtin_amdgcn_ds_bpermute(src, En);
      if (col < min(b_size, nsize)) { H = hs; E = es; }  // the row's cells still in the band
      else if (col < nsize) {  // extension cells
        const int32_t f = f_end - (int32_t)(col - b_size) * ge;
        H = f;
        E = f - oe;
      } else {
        H = kMinInt;
        E = kMinInt;
      }
    }
    first_b = nfirst;
    b_size = nsize + sentinel;
    if (first_b - wb >= 64) {  // slide the b window by 64 columns
      wb += 64;
      bw0 = bw1;
      bw1 = bw2;
      bw2 = seq_code(t.b, n, rev, wb + 128 + lane);
    }
    bc = bcode_of(first_b + lane);
    if (best_here) { bytes_at_best = flat_len; rows_at_best = ai + 2; }
  }
  res.score = best;
  res.a_len = best_a;
  res.b_len = best_b;
  res.cells = cells;
  res.tb_bytes = bytes_at_best;
  res.tb_rows = rows_at_best;
}

// dp_half_wave for bands of up to 256 cells: lane l holds band positions
// 4l..4l+3 (columns first_b + 4l + j), so each scan runs over the lane's four
// cells and then across the lanes, and the band moves by permutes whose
// source lane and register are the same for the whole wave.  The base codes
// of columns [wb, wb + 512) sit in the wave's LDS window `win`, reloaded when
// the band's start passes wb + 256.  Row table, traceback bytes and results as
// dp_half_wave; without TB, results only (flat and rows unused).
template <bool TB = true>
__device__ __forceinline__ void dp_half_wave4(const DpTask& t, const int* __restrict__ mat, int32_t go, int32_t ge,
                                              DpResult& res, uint8_t* __restrict__ flat, uint32_t* __restrict__ rows,
                                              uint32_t flat_cap, uint32_t rows_cap, uint32_t& rounds,
                                              uint8_t* __restrict__ win) {
  constexpr uint32_t K = 4, wmask = 64 * K - 1;
  constexpr int32_t kNeg = INT32_MIN + (1 << 27);  // below every reachable value, and safe to add to
  const uint32_t lane = threadIdx.x & 63u;
  res.score = 0; res.a_len = 0; res.b_len = 0; res.flags = 0; res.cells = 0; res.tb_bytes = 0; res.tb_rows = 0;
  const uint32_t m = t.m, n = t.n;
  if (m == 0 || n == 0) return;
  const bool rev = t.reverse != 0;
  const int32_t oe = go + ge;
  const int32_t xdrop = t.xdrop > oe ? t.xdrop : oe;
  uint32_t J;
  if (ge <= 0) J = n;
  else J = min(n, (uint32_t)((xdrop - oe) / ge) + 1u);
  if (J > wmask) { res.flags |= E_BAND_WIDE; return; }
  uint32_t wb = 0;
  auto load_win = [&]() {
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) win[lane + 64 * k] = (uint8_t)seq_code(t.b, n, rev, wb + lane + 64 * k);
  };
  load_win();
  uint32_t ab = 0;
  uint32_t aw = seq_code(t.a, m, rev, lane), awn = seq_code(t.a, m, rev, 64 + lane);
  uint32_t first_b = 0, b_size = J + 1;
  int32_t H[K], E[K];
  uint32_t bc[K];
#pragma unroll
  for (uint32_t j = 0; j < K; ++j) {
    const uint32_t p = K * lane + j;
    H[j] = p == 0 ? 0 : -oe - (int32_t)(p - 1) * ge;
    E[j] = H[j] - oe;
    bc[j] = win[p];
    if (TB && p <= J) flat[p] = p == 0 ? SCRIPT_GAP_IN_B : SCRIPT_GAP_IN_A;
  }
  if (TB && lane == 0) { rows[0] = 0; rows[1] = 0; }
  uint32_t flat_len = J + 1;
  int32_t best = 0;
  uint32_t best_a = 0, best_b = 0;
  uint64_t cells = 0;
  uint32_t bytes_at_best = flat_len, rows_at_best = 1;
  for (uint32_t ai = 0; ai < m; ++ai) {
    if (TB && (ai + 2 > rows_cap || flat_len + 2 * wmask + 4 > flat_cap)) { res.flags |= E_TB_SPACE; return; }
    if (ai - ab == 64) {
      ab += 64;
      aw = awn;
      awn = seq_code(t.a, m, rev, ab + 64 + lane);
    }
    const uint32_t acode = (uint32_t)__builtin_amdgcn_readlane((int)aw, (int)(ai - ab));
    const uint32_t w = b_size - first_b;
    if (TB && lane == 0) { rows[2 * (ai + 1)] = flat_len; rows[2 * (ai + 1) + 1] = first_b; }
    cells += w;
    bool act[K];
    int32_t D[K], V[K], G[K];
    {
      int32_t hs[K];
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        act[j] = K * lane + j < w;
        hs[j] = H[j] + mat[(acode << 4) | bc[j]];
      }
      D[0] = wave_shr1(hs[K - 1], kMinInt);  // the cell before's H + its substitution
#pragma unroll
      for (uint32_t j = 1; j < K; ++j) D[j] = hs[j - 1];
    }
    bool kept[K];
    uint64_t km[K];
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
      V[j] = E[j];
      G[j] = max(D[j], V[j]);
      kept[j] = act[j] && G[j] >= best - xdrop;
      km[j] = wave_ballot(kept[j]);
    }
    // iterate the keep decisions to the fixed point, as dp_half_wave
    int32_t F[K], cs[K], fa[K], rowmax;
    for (uint32_t round = 0;; ++round) {
      if (round > K * 64) { res.flags |= E_BAND_WIDE; return; }
      ++rounds;
      int32_t A[K], B[K];
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        A[j] = kept[j] ? -ge : 0;
        B[j] = kept[j] ? G[j] - oe : kNeg;
      }
#pragma unroll
      for (uint32_t j = 1; j < K; ++j) {  // the lane's own cells, in order
        B[j] = max(B[j - 1] + A[j], B[j]);
        A[j] = A[j - 1] + A[j];
      }
      int32_t LA = A[K - 1], LB = B[K - 1];
      wave_prefix_affine(LA, LB, kNeg);
      const int32_t EA = wave_shr1(LA, 0), EB = wave_shr1(LB, kNeg);  // the lanes before
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) fa[j] = max(kMinInt + (EA + A[j]), max(EB + A[j], B[j]));
      F[0] = wave_shr1(fa[K - 1], kMinInt);
#pragma unroll
      for (uint32_t j = 1; j < K; ++j) F[j] = fa[j - 1];
      int32_t pm[K];
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        cs[j] = act[j] ? max(G[j], F[j]) : kNeg;
        pm[j] = j ? max(pm[j - 1], cs[j]) : cs[j];
      }
      const int32_t incl = wave_prefix_max(pm[K - 1], kNeg);
      rowmax = __builtin_amdgcn_readlane(incl, 63);
      const int32_t ex = wave_shr1(incl, kNeg);
      bool same = true;
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        const int32_t bb = max(best, j ? max(ex, pm[j - 1]) : ex);
        const bool k2 = act[j] && cs[j] >= bb - xdrop;
        const uint64_t m2 = wave_ballot(k2);
        same = same && m2 == km[j];
        kept[j] = k2;
        km[j] = m2;
      }
      if (same) break;
    }
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {  // traceback byte of each cell
      uint8_t op = SCRIPT_SUB;
      int32_t c0 = D[j];
      if (V[j] > c0) { c0 = V[j]; op = SCRIPT_GAP_IN_B; }
      if (F[j] > c0) op = SCRIPT_GAP_IN_A;
      if (kept[j]) {
        if (V[j] - ge >= cs[j] - oe) op |= SCRIPT_EXTEND_GAP_B;
        if (F[j] - ge >= cs[j] - oe) op |= SCRIPT_EXTEND_GAP_A;
      }
      if (TB && act[j]) flat[flat_len + K * lane + j] = op;
    }
    uint32_t lead = 0xFFFFFFFFu, last = 0;
#pragma unroll
    for (uint32_t j = 0; j < K; ++j)
      if (km[j]) {
        lead = min(lead, K * (uint32_t)__builtin_ctzll(km[j]) + j);
        last = max(last, K * (63u - (uint32_t)__builtin_clzll(km[j])) + j);
      }
    if (lead == 0xFFFFFFFFu) break;  // every cell pruned: first_b reaches b_size
    bool best_here = false;
    if (rowmax > best) {  // the first cell holding the row maximum
      uint32_t at = 0xFFFFFFFFu;
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        const uint64_t bm = wave_ballot(act[j] && cs[j] == rowmax);
        if (bm) at = min(at, K * (uint32_t)__builtin_ctzll(bm) + j);
      }
      best = rowmax;
      best_a = ai + 1;
      best_b = first_b + at;
      best_here = true;
    }
    int32_t Hn[K], En[K];
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
      Hn[j] = kept[j] ? cs[j] : kMinInt;
      En[j] = kept[j] ? max(V[j] - ge, cs[j] - oe) : V[j];
    }
    flat_len += w;
    const uint32_t nfirst = first_b + lead;
    uint32_t nsize = b_size;
    uint32_t ext = 0;
    int32_t f_end = 0;
    if (last + 1 < w) {
      nsize = first_b + last + 1;
    } else {
      const uint32_t pj = (w - 1) & 3u;
      const int32_t fv = pj == 0 ? fa[0] : pj == 1 ? fa[1] : pj == 2 ? fa[2] : fa[3];
      f_end = __builtin_amdgcn_readlane(fv, (int)((w - 1) >> 2));
      const int32_t floor_v = best - xdrop;
      if (f_end >= floor_v) {
        const uint32_t room = n + 1 - b_size;  // columns b_size..n
        uint32_t k = ge <= 0 ? room : (uint32_t)((f_end - floor_v) / ge) + 1u;
        ext = min(k, room);
      }
      if (ext && (b_size + ext - 1) - nfirst > wmask) { res.flags |= E_BAND_WIDE; return; }
#pragma unroll
      for (uint32_t j = 0; j < K; ++j)
        if (TB && K * lane + j < ext) flat[flat_len + K * lane + j] = SCRIPT_GAP_IN_A;
      flat_len += ext;
      nsize = b_size + ext;
    }
    uint32_t sentinel = 0;
    if (nsize <= n) {
      if (nsize - nfirst > wmask) { res.flags |= E_BAND_WIDE; return; }
      sentinel = 1;
    }
    // shift the band to start at nfirst: position p takes old position
    // p + lead, i.e. lane + (j + lead) / 4, register (j + lead) % 4
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
      const uint32_t sj = (j + lead) & 3u, qj = (j + lead) >> 2;
      const int32_t hsrc = sj == 0 ? Hn[0] : sj == 1 ? Hn[1] : sj == 2 ? Hn[2] : Hn[3];
      const int32_t esrc = sj == 0 ? En[0] : sj == 1 ? En[1] : sj == 2 ? En[2] : En[3];
      const int src = (int)(((lane + qj) & 63u) << 2);
      const int32_t hs = __builtin_amdgcn_ds_bpermute(src, hsrc);
      const int32_t es = __builtin_amdgcn_ds_bpermute(src, esrc);
      const uint32_t col = nfirst + K * lane + j;
      if (col < min(b_size, nsize)) { H[j] = hs; E[j] = es; }  // the row's cells still in the band
      else if (col < nsize) {  // extension cells
        const int32_t f = f_end - (int32_t)(col - b_size) * ge;
        H[j] = f;
        E[j] = f - oe;
      } else {
        H[j] = kMinInt;
        E[j] = kMinInt;
      }
    }
    first_b = nfirst;
    b_size = nsize + sentinel;
    while (first_b - wb >= 256) {  // the band's columns stay inside [wb, wb + 512)
      wb += 256;
      load_win();
    }
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) bc[j] = win[first_b - wb + K * lane + j];
    if (best_here) { bytes_at_best = flat_len; rows_at_best = ai + 2; }
  }
  res.score = best;
  res.a_len = best_a;
  res.b_len = best_b;
  res.cells = cells;
  res.tb_bytes = bytes_at_best;
  res.tb_rows = rows_at_best;
}

// Where a row's traceback bytes lie, for the walk: from a row table (8 B per
// row: first byte, first column; a row ends where the next begins) ...
struct TbRowTable {
  const uint32_t* rows;
  uint32_t off, fb, end;
  __device__ TbRowTable(const uint32_t* r, uint32_t a, uint32_t flat_total)
      : rows(r), off(r[2 * a]), fb(r[2 * a + 1]), end(flat_total) {}
  __device__ void down(const uint8_t*, uint32_t a) {  // to row a from row a + 1
    end = off;
    off = rows[2 * a];
    fb = rows[2 * a + 1];
  }
};

// ... or from the stream's row trailers (dp_half_tb16): starting from the best
// row's start, each step down reads the trailer just below the current row.
template <uint32_t TR>
struct TbRowStream {
  uint32_t off, fb, end, dfb;
  __device__ TbRowStream(const TbRowStart& s, uint32_t flat_total) : off(s.off), fb(s.fb), end(flat_total), dfb(s.dfb) {}
  __device__ void down(const uint8_t* flat, uint32_t) {
    // every row below the best one was closed by a trailer, so off >= TR,
    // len <= end and dfb <= fb hold; the clamps only keep every read inside
    // [0, flat_total) whatever the bytes say
    fb = fb >= dfb ? fb - dfb : 0;
    end = off >= TR ? off - TR : 0;
    uint32_t len = 0, d = 0;
    if (off >= TR) {
      if constexpr (TR == 2) {
        len = flat[end];
        d = flat[end + 1];
      } else {
        len = flat[end] | ((uint32_t)flat[end + 1] << 8);
        d = flat[end + 2] | ((uint32_t)flat[end + 3] << 8);
      }
    }
    off = end - (len <= end ? len : end);
    dfb = d;
  }
};

// Traceback walk (gapped.rs:692-757) from (a, b), the best row and column,
// over the rows' traceback bytes: the ops, run-length encoded, collected
// backwards into ops[0, cap); the count is returned, E_TB_SPACE set in flags
// when they do not fit or the walk does not reach the origin.
template <class Rows>
__device__ __forceinline__ uint32_t tb_walk_rows(const uint8_t* __restrict__ flat, Rows rs, uint32_t a, uint32_t b,
                                                 uint32_t* __restrict__ ops, uint32_t cap, uint32_t& flags) {
  uint32_t nops = 0;
  uint8_t prev = SCRIPT_SUB;
  uint32_t last_kind = 3, run = 0;
  const uint32_t budget = a + b;
  for (uint32_t stepi = 0; stepi < budget && (a | b); ++stepi) {
    uint8_t raw = SCRIPT_GAP_IN_A;
    if (b >= rs.fb && rs.off + (b - rs.fb) < rs.end) raw = flat[rs.off + (b - rs.fb)];
    uint8_t op;
    if ((prev & SCRIPT_OP_MASK) == SCRIPT_GAP_IN_A)
      op = (raw & SCRIPT_EXTEND_GAP_A) ? SCRIPT_GAP_IN_A : (raw & SCRIPT_OP_MASK);
    else if ((prev & SCRIPT_OP_MASK) == SCRIPT_GAP_IN_B)
      op = (raw & SCRIPT_EXTEND_GAP_B) ? SCRIPT_GAP_IN_B : (raw & SCRIPT_OP_MASK);
    else
      op = raw & SCRIPT_OP_MASK;
    prev = op;
    uint32_t kind;
    if (op == SCRIPT_GAP_IN_A) {
      kind = 1;
      if (b) --b;
    } else {
      kind = op == SCRIPT_GAP_IN_B ? 2 : 0;
      if (a) rs.down(flat, --a);
      if (kind == 0 && b) --b;
    }
    if (kind == last_kind) {
      ++run;
    } else {
      if (run) {
        if (nops >= cap) { flags |= E_TB_SPACE; return nops; }
        ops[nops++] = last_kind | (run << 2);
      }
      last_kind = kind;
      run = 1;
    }
  }
  if (run) {
    if (nops >= cap) { flags |= E_TB_SPACE; return nops; }
    ops[nops++] = last_kind | (run << 2);
  }
  if (a | b) flags |= E_TB_SPACE;
  return nops;
}

__device__ __forceinline__ uint32_t tb_walk(const uint8_t* __restrict__ flat, const uint32_t* __restrict__ rows,
                                            uint32_t a, uint32_t b, uint32_t flat_total, uint32_t* __restrict__ ops,
                                            uint32_t cap, uint32_t& flags) {
  return tb_walk_rows(flat, TbRowTable(rows, a, flat_total), a, b, ops, cap, flags);
}

// Traceback-phase bands (X-drop 100 by default) stay within 64 cells: the
// first attempt gives each lane a 64-cell ring in LDS (64-lane workgroups);
// wider bands rerun with per-lane rings in global memory.

template <bool LDS_RING>
__global__ __launch_bounds__(LDS_RING ? kDpLdsThreads : 256) void k_dp_score(
    const DpTask* __restrict__ tasks, uint32_t n_tasks, const int32_t* __restrict__ gmat, int32_t go, int32_t ge,
    int2* __restrict__ ring, uint32_t wlog2, DpResult* __restrict__ out, unsigned long long* counters,
    uint32_t only_wide) {
  __shared__ int mat[256];
  __shared__ int2 lring[LDS_RING ? (kDpLdsThreads << kDpLdsLog2) : 1];
  __shared__ uint8_t lbring[LDS_RING ? (kDpLdsThreads << kDpLdsLog2) : 1];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) mat[i] = gmat[i];
  __syncthreads();
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  constexpr uint32_t RS = LDS_RING ? kDpLdsThreads : 1;
  int2* my = LDS_RING ? lring + threadIdx.x : ring + (size_t)gt * ring_stride(wlog2);
  uint8_t* myb = LDS_RING ? lbring + threadIdx.x : nullptr;
  const uint32_t wmask = LDS_RING ? (1u << kDpLdsLog2) - 1 : (1u << wlog2) - 1;
  unsigned long long cells = 0;
  for (uint32_t i = gt; i < n_tasks; i += nt) {
    // only_wide: rerun just the tasks an earlier pass flagged for their band
    if (only_wide && !(out[i].flags & E_BAND_WIDE)) continue;
    DpResult r;
    dp_half<false, RS>(tasks[i], mat, go, ge, my, myb, wmask, r, nullptr, nullptr, 0);
    out[i] = r;
    cells += r.cells;
  }
  atomicAdd(&counters[C_DP_CELLS], cells);
}

template <bool LDS_RING>
__global__ __launch_bounds__(LDS_RING ? kDpLdsThreads : 256) void k_dp_traceback(
    const TbTask* __restrict__ tasks, uint32_t n_tasks, const int32_t* __restrict__ gmat, int32_t go, int32_t ge,
    int2* __restrict__ ring, uint32_t wlog2, uint8_t* __restrict__ flat_pool, uint32_t* __restrict__ row_pool,
    uint32_t* __restrict__ ops_pool, TbResult* __restrict__ out) {
  __shared__ int mat[256];
  __shared__ int2 lring[LDS_RING ? (kDpLdsThreads << kDpLdsLog2) : 1];
  __shared__ uint8_t lbring[LDS_RING ? (kDpLdsThreads << kDpLdsLog2) : 1];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) mat[i] = gmat[i];
  __syncthreads();
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  constexpr uint32_t RS = LDS_RING ? kDpLdsThreads : 1;
  int2* my = LDS_RING ? lring + threadIdx.x : ring + (size_t)gt * ring_stride(wlog2);
  uint8_t* myb = LDS_RING ? lbring + threadIdx.x : nullptr;
  const uint32_t wmask = LDS_RING ? (1u << kDpLdsLog2) - 1 : (1u << wlog2) - 1;
  for (uint32_t i = gt; i < n_tasks; i += nt) {
    const TbTask T = tasks[i];
    uint8_t* flat = flat_pool + T.flat_off;
    uint32_t* rows = row_pool + 2 * T.row_off;
    uint32_t* ops = ops_pool + T.ops_off;
    DpResult r;
    dp_half<true, RS>(T.t, mat, go, ge, my, myb, wmask, r, flat, rows, T.stop_row);
    TbResult o{0, r.flags};
    if (!r.flags) {
      const uint32_t nops = tb_walk(flat, rows, T.stop_row, T.best_b, r.tb_bytes, ops, T.ops_cap, o.flags);
      for (uint32_t x = 0, y = nops ? nops - 1 : 0; x < y; ++x, --y) {
        const uint32_t tmp = ops[x];
        ops[x] = ops[y];
        ops[y] = tmp;
      }
      o.n_ops = nops;
    }
    out[i] = o;
  }
}

// Score and traceback in one pass: each lane keeps its half alignment's
// traceback stream (every row's bytes and trailer: the best row is known only
// at the end) and ops in its own arena (the launch passes rows_cap 0: no row
// table), walks back from the best cell, and appends the ops to one packed
// pool.  A half whose band, bytes or ops outgrow the LDS ring or the arena is
// rerun in a retry slot (dp_half, with a row table), else comes back flagged
// for the two-pass path.
template <bool LDS_RING, uint32_t RLOG2>
__global__ __launch_bounds__(kDpLdsThreads) void k_dp_onepass(
    const DpTask* __restrict__ tasks, uint32_t n_tasks, const int32_t* __restrict__ gmat, int32_t go, int32_t ge,
    int2* __restrict__ ring, uint32_t wlog2, uint8_t* __restrict__ arena, uint32_t flat_cap, uint32_t rows_cap,
    uint32_t ops_cap, uint32_t* __restrict__ ops_pool, unsigned long long* __restrict__ ops_count,
    unsigned long long ops_pool_cap, OnePassResult* __restrict__ out, unsigned long long* next, OnePassBig big,
    unsigned long long* counters) {
  __shared__ int mat[256];
  __shared__ uint32_t lring[kDpLdsThreads << RLOG2];  // 12-bit H and E and the column's base code
  __shared__ int s_max_pos;
  if (threadIdx.x == 0) s_max_pos = 0;
  __syncthreads();
  int my_max = 0;  // largest score among codes 0..14 (0 if none is positive)
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    const int v = gmat[i];
    mat[i] = v;
    if ((i >> 4) < 15 && (i & 15) < 15) my_max = max(my_max, v);
  }
  atomicMax(&s_max_pos, my_max);
  __syncthreads();
  const int32_t max_pos = s_max_pos;
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  static_assert(LDS_RING, "the one-pass kernel runs with LDS rings");
  (void)ring;
  (void)wlog2;
  uint32_t* my = lring + threadIdx.x;
  uint8_t* base = arena + (size_t)gt * onepass_stride(flat_cap, rows_cap, ops_cap);
  uint32_t* rows = reinterpret_cast<uint32_t*>(base);
  uint8_t* flat = base + (size_t)rows_cap * 8;
  uint32_t* ops = reinterpret_cast<uint32_t*>(flat + ((flat_cap + 15u) & ~15u));
  unsigned long long cells = 0;
  // tasks come longest first; each wave takes the next 64 from the queue, so
  // workgroups that start late (behind another stream's kernel) take less
  (void)nt;
  for (;;) {
    const uint64_t base = wave_take(next, 64);
    if (base >= n_tasks) break;
    const uint32_t i = (uint32_t)base + (threadIdx.x & 63u);
    if (i >= n_tasks) continue;
    const unsigned long long l0 = wall_clock64();
    DpResult r;
    TbRowStart best_row;
    dp_half_tb16<RLOG2, true, kDpLdsThreads, true>(tasks[i], mat, go, ge, my, nullptr, max_pos, r, flat, flat_cap,
                                                   best_row);
    uint8_t* tflat = flat;
    uint32_t *trows = rows, *tops = ops, tops_cap = ops_cap;
    bool stream = true;  // the traceback bytes with row trailers, else dp_half's row table
    if (r.flags) {  // band or traceback beyond the lane's storage: once more with a retry slot
      const uint32_t slot = atomicAdd(big.next, 1u);
      if (slot < big.slots) {
        stream = false;
        uint8_t* bb = big.arena + (size_t)slot * onepass_stride(big.flat_cap, big.rows_cap, big.ops_cap);
        trows = reinterpret_cast<uint32_t*>(bb);
        tflat = bb + (size_t)big.rows_cap * 8;
        tops = reinterpret_cast<uint32_t*>(tflat + ((big.flat_cap + 15u) & ~15u));
        tops_cap = big.ops_cap;
        dp_half<true, 1>(tasks[i], mat, go, ge, big.ring + (size_t)slot * ring_stride(big.ring_log2), nullptr,
                         (1u << big.ring_log2) - 1, r, tflat, trows, 0xFFFFFFFFu, big.flat_cap, big.rows_cap);
      }
    }
    const unsigned long long l1 = wall_clock64();
    OnePassResult o{r.score, r.a_len, r.b_len, 0, r.flags, 0};
    if (!r.flags) {
      cells += r.cells;
      if (r.a_len | r.b_len) {
        const uint32_t n =
            stream ? tb_walk_rows(tflat, TbRowStream<tb_trailer_bytes(RLOG2)>(best_row, r.tb_bytes), r.a_len,
                                  r.b_len, tops, tops_cap, o.flags)
                   : tb_walk(tflat, trows, r.a_len, r.b_len, r.tb_bytes, tops, tops_cap, o.flags);
        if (!o.flags) {
          const unsigned long long at = atomicAdd(ops_count, (unsigned long long)n);
          if (at + n <= ops_pool_cap) {
            for (uint32_t k = 0; k < n; ++k) ops_pool[at + k] = tops[n - 1 - k];
            o.n_ops = n;
            o.ops_at = at;
          } else {
            o.flags |= E_TB_SPACE;
          }
        }
      }
    }
    out[i] = o;
    {  // the wave's DP and walk times (its lanes finish together)
      const unsigned long long l2 = wall_clock64();
      if ((threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(__ballot(true))) {
        atomicMax(&counters[C_TB_TICKS + 2], l1 - l0);
        atomicMax(&counters[C_TB_TICKS + 3], l2 - l1);
      }
    }
  }
  atomicAdd(&counters[C_DP_CELLS], cells);
}

// The traceback halves that run one per wave (dp_half_wave, or with WIDE
// dp_half_wave4 and its base-code windows), four waves per workgroup with
// little LDS, so a SIMD holds several waves and each wave's row-to-row latency
// hides behind the others.  Each wave has its own arena (row table, traceback
// bytes, ops); a half beyond it, or with a band wider than 64 (WIDE: 256)
// cells, comes back flagged for the two-pass path.
static constexpr uint32_t kDpWaveThreads = 256;
template <bool WIDE>
__global__ __launch_bounds__(kDpWaveThreads) void k_dp_wave(
    const DpTask* __restrict__ tasks, uint32_t n_tasks, const int32_t* __restrict__ gmat, int32_t go, int32_t ge,
    uint8_t* __restrict__ arena, uint32_t rows_cap, uint32_t flat_cap, uint32_t ops_cap,
    uint32_t* __restrict__ ops_pool, unsigned long long* __restrict__ ops_count, unsigned long long ops_pool_cap,
    OnePassResult* __restrict__ out, unsigned long long* next, unsigned long long* counters) {
  __shared__ int mat[256];
  __shared__ uint8_t bwin[WIDE ? (kDpWaveThreads / 64) * 512 : 1];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) mat[i] = gmat[i];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  uint8_t* win = bwin + (WIDE ? (threadIdx.x >> 6) * 512 : 0);
  (void)win;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint8_t* wbase = arena + wave * onepass_stride(flat_cap, rows_cap, ops_cap);
  uint32_t* wrows = reinterpret_cast<uint32_t*>(wbase);
  uint8_t* wflat = wbase + (size_t)rows_cap * 8;
  uint32_t* wops = reinterpret_cast<uint32_t*>(wflat + ((flat_cap + 15u) & ~15u));
  unsigned long long cells = 0, wrows_n = 0;
  uint32_t rounds = 0;
  for (;;) {
    const uint64_t k = wave_take(next, 1);
    if (k >= n_tasks) break;
    const uint32_t i = (uint32_t)k;
    DpResult r;
    const unsigned long long t0 = wall_clock64();
    if constexpr (WIDE) dp_half_wave4(tasks[i], mat, go, ge, r, wflat, wrows, flat_cap, rows_cap, rounds, win);
    else dp_half_wave(tasks[i], mat, go, ge, r, wflat, wrows, flat_cap, rows_cap, rounds);
    // the walk reads bytes other lanes of the wave stored
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const unsigned long long t1 = wall_clock64();
    if (lane == 0) {
      atomicMax(&counters[C_TB_TICKS], t1 - t0);
      wrows_n += r.tb_rows;
      OnePassResult o{r.score, r.a_len, r.b_len, 0, r.flags, 0};
      if (!r.flags) {
        cells += r.cells;
        if (r.a_len | r.b_len) {
          const uint32_t nops = tb_walk(wflat, wrows, r.a_len, r.b_len, r.tb_bytes, wops, ops_cap, o.flags);
          if (!o.flags) {
            const unsigned long long at = atomicAdd(ops_count, (unsigned long long)nops);
            if (at + nops <= ops_pool_cap) {
              for (uint32_t q = 0; q < nops; ++q) ops_pool[at + q] = wops[nops - 1 - q];
              o.n_ops = nops;
              o.ops_at = at;
            } else {
              o.flags |= E_TB_SPACE;
            }
          }
        }
      }
      out[i] = o;
      atomicMax(&counters[C_TB_TICKS + 1], wall_clock64() - t1);
    }
  }
  if (lane == 0) {
    atomicAdd(&counters[C_DP_CELLS], cells);
    atomicAdd(&counters[C_WAVE_ROWS], wrows_n);
    atomicAdd(&counters[C_WAVE_ROUNDS], (unsigned long long)rounds);
  }
}

// Gathers each traceback's ops from its worst-case slot of the op pool into one
// packed list, so only the ops themselves cross PCIe.
__global__ __launch_bounds__(256) void k_pack_ops(const PackTask* __restrict__ tasks, uint32_t n_tasks,
                                                  const uint32_t* __restrict__ ops_pool, uint32_t* __restrict__ packed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_tasks) return;
  const PackTask t = tasks[i];
  for (uint32_t j = 0; j < t.n; ++j) packed[t.dst + j] = ops_pool[t.src + j];
}

// ------------------------------------------------- fused Phase 2a DPs -------
// Phase-2a pivot of an ungapped hit (engine.rs:1850-1895), record frame in,
// DP frame out: (pq, ps) splits the score-only alignment, (tq, ts) is the
// traceback seed kept for Phase 2b.
__device__ __forceinline__ void prelim_pivot(const UngHit& u, uint32_t qa_len, uint32_t sa_len, uint32_t lw,
                                             uint32_t& pq, uint32_t& ps, uint32_t& tq, uint32_t& ts) {
  // both strands computed into locals and selected, so no output is written
  // through a strand-dependent reference (which would go through scratch)
  const bool p_shift = u.s_end >= u.s_seed + 8;
  const uint32_t p_tq = u.q_seed + (p_shift ? 3u : 0u), p_ts = u.s_seed + (p_shift ? 3u : 0u);
  const uint32_t p_up = 4 - (p_ts & 3);
  const uint32_t p_ps0 = p_ts + p_up, p_pq0 = p_tq + p_up;
  const bool p_back = p_ps0 >= sa_len || p_pq0 >= qa_len;
  const uint32_t saved_fwd = sa_len - u.s_seed - lw;
  const bool m_shift = (sa_len - u.s_start) >= saved_fwd + 8;
  const uint32_t m_tq = m_shift ? u.q_seed + lw - 4 : u.q_seed, m_ts = m_shift ? u.s_seed + lw - 4 : u.s_seed;
  const uint32_t qn = (qa_len - u.q_seed - lw) + (m_shift ? 3u : 0u);
  const uint32_t sn = (sa_len - u.s_seed - lw) + (m_shift ? 3u : 0u);
  const uint32_t m_up = 4 - (sn & 3);
  const uint32_t m_pq0 = qn + m_up, m_ps0 = sn + m_up;
  const bool m_back = m_pq0 > qa_len || m_ps0 > sa_len;
  const bool minus = u.strand != 0;
  tq = minus ? m_tq : p_tq;
  ts = minus ? m_ts : p_ts;
  pq = minus ? m_pq0 - (m_back ? 4u : 0u) : p_pq0 - (p_back ? 4u : 0u);
  ps = minus ? m_ps0 - (m_back ? 4u : 0u) : p_ps0 - (p_back ? 4u : 0u);
}

// One lane per ungapped hit: pivot, both score-only halves (subject outer,
// gapped.rs:360-418), and the cutoff test.  The DP is a pure function of the
// hit, and a hit whose DP falls below the cutoff never changes the containment
// state of the ordered Phase-2a loop (engine.rs:1317-1411), so only the hits
// that reach the cutoff are kept for the ordered replay.  Preliminary bands
// (X-drop 30 by default) almost always fit 16 cells: the first pass gives
// each lane a ring of 2^kPrelimRingLog2 cells of LDS, interleaved across the
// workgroup's lanes, and lists the hits whose band outgrows it (by index) for
// a second pass over per-lane rings of 2^wlog2 cells in global memory, which
// hands the hits wider still back as records.  Output item ids name search
// records (P.rec_ids); the task count may come from device memory (P.n_dev),
// so the passes queue behind the seeding kernel without a host round trip.
static constexpr uint32_t kPrelimThreads = 128, kPrelimRingLog2 = 4, kPrelimRun = 1;
// wider rings for X-drops whose bands outgrow 16 cells: 64 cells of 4 bytes
// (H and E as 16-bit offsets, dp_half_tb16's score-only form) and a code
// byte per lane, in 64-lane workgroups of 21 KB of LDS, seven per CU
static constexpr uint32_t kPrelimWideLog2 = 6, kPrelimPackThreads = 64;

// A preliminary hit's two halves (gapped.rs's preliminary extension from the
// hit's pivot) and its item's record and subject.
struct PrelimHalves {
  uint32_t r, f, ql, sl, pq, ps, tq, ts;
  DpTask left, right;
};
__device__ __forceinline__ PrelimHalves prelim_halves(const PrelimParams& P, const UngHit& u,
                                                      const RecDesc* __restrict__ recs,
                                                      const SubjDesc* __restrict__ subjs,
                                                      const uint8_t* __restrict__ qfwd,
                                                      const uint8_t* __restrict__ qrc,
                                                      const uint8_t* __restrict__ splus) {
  PrelimHalves h;
  h.r = u.item / P.n_subj;
  h.f = u.item % P.n_subj;
  const RecDesc R = recs[h.r];
  const SubjDesc S = subjs[h.f];
  h.ql = R.q_len;
  h.sl = S.s_len;
  prelim_pivot(u, h.ql, h.sl, (uint32_t)P.lut_width, h.pq, h.ps, h.tq, h.ts);
  const uint8_t* qa = (u.strand ? qrc : qfwd) + R.q_base + 1;
  const uint8_t* sa = splus + S.s_base + 1;
  const int32_t xd = P.xdrop_gap < u.score ? P.xdrop_gap : u.score;
  h.left = DpTask{sa, qa, h.ps, h.pq, xd, 1};
  h.right = DpTask{sa + h.ps, qa + h.pq, h.sl > h.ps ? h.sl - h.ps : 0, h.ql > h.pq ? h.ql - h.pq : 0, xd, 0};
  return h;
}
// The hit's preliminary HSP, kept when its score reaches the cutoff.
__device__ __forceinline__ void prelim_emit(const PrelimParams& P, const UngHit& u, const PrelimHalves& hv,
                                            const DpResult& L, const DpResult& Rr, PreHit* __restrict__ out,
                                            unsigned long long* __restrict__ out_count) {
  const int32_t sc = L.score + Rr.score;
  if (sc < 0 || sc < P.cutoff) return;
  const uint32_t q0 = hv.pq - L.b_len, q1 = hv.pq + Rr.b_len, s0 = hv.ps - L.a_len, s1 = hv.ps + Rr.a_len;
  const uint32_t ql = hv.ql, sl = hv.sl;
  PreHit h;
  h.u = u;
  h.u.item = P.rec_ids[hv.r] * P.n_subj + hv.f;
  h.score = sc;
  if (u.strand) { h.q_start = ql - q1; h.q_end = ql - q0; h.s_start = sl - s1; h.s_end = sl - s0; }
  else { h.q_start = q0; h.q_end = q1; h.s_start = s0; h.s_end = s1; }
  h.tq = hv.tq;
  h.ts = hv.ts;
  const unsigned long long kk = atomicAdd(out_count, 1ull);
  if (kk < P.out_cap) out[kk] = h;
}

template <bool LDS_RING, uint32_t RLOG2, bool PACK = false>
__global__ __launch_bounds__(kPrelimThreads) void k_prelim_fused(
    PrelimParams P, const UngHit* __restrict__ hits, const uint32_t* __restrict__ idx, uint64_t n_tasks,
    const RecDesc* __restrict__ recs, const SubjDesc* __restrict__ subjs, const uint8_t* __restrict__ qfwd,
    const uint8_t* __restrict__ qrc, const uint8_t* __restrict__ splus, const int32_t* __restrict__ gmat,
    int2* __restrict__ gring, uint32_t wlog2, PreHit* __restrict__ out, unsigned long long* __restrict__ out_count,
    uint32_t* __restrict__ redo, UngHit* __restrict__ redo_hits, unsigned long long* __restrict__ redo_count,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ err) {
  constexpr uint32_t T = PACK ? kPrelimPackThreads : kPrelimThreads;  // the workgroup's lanes
  __shared__ int mat[256];
  __shared__ int2 lring[LDS_RING && !PACK ? (T << RLOG2) : 1];
  __shared__ uint32_t lring16[PACK ? (T << RLOG2) : 1];
  __shared__ uint8_t lbring[LDS_RING && !PACK ? (T << RLOG2) : 1];
  __shared__ int s_max_pos;
  static_assert(!PACK || LDS_RING, "packed rings live in LDS");
  if (threadIdx.x == 0) s_max_pos = 0;
  __syncthreads();
  int my_max = 0;  // largest score among codes 0..14 (0 if none is positive)
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    const int v = gmat[i];
    mat[i] = v;
    if ((i >> 4) < 15 && (i & 15) < 15) my_max = max(my_max, v);
  }
  if (PACK) atomicMax(&s_max_pos, my_max);
  __syncthreads();
  const int32_t max_pos = s_max_pos;
  uint32_t* my16 = lring16 + threadIdx.x;
  const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
  constexpr uint32_t RS = LDS_RING ? T : 1;
  int2* my = LDS_RING ? lring + threadIdx.x : gring + gt * ring_stride(wlog2);
  uint8_t* myb = LDS_RING ? lbring + threadIdx.x : nullptr;
  const uint32_t wmask = LDS_RING ? (1u << RLOG2) - 1 : (1u << wlog2) - 1;
  unsigned long long cells = 0, dps = 0;
  if (P.n_dev) n_tasks = *P.n_dev < n_tasks ? *P.n_dev : n_tasks;
  // waves take runs of kPrelimRun x 64 hits from the launch's queue, so
  // workgroups that start late (behind another stream's kernels) take less.
  // One hit per lane at a time measured fastest (10% below runs of four): a
  // run lasts as long as its longest DP, and shorter runs even out the waves.
  (void)nt;
  for (;;) {
   const uint64_t base = wave_take(P.next, 64 * kPrelimRun);
   if (base >= n_tasks) break;
   // the lane's next hit is requested before its current one is aligned
   const uint64_t k0 = base + (threadIdx.x & 63u);
   uint32_t i_next = k0 < n_tasks ? (idx ? idx[k0] : (uint32_t)k0) : 0u;
   UngHit u_next = hits[i_next];
   for (uint32_t run = 0; run < kPrelimRun; ++run) {
    const uint64_t k = base + 64 * run + (threadIdx.x & 63u);
    if (k >= n_tasks) break;
    const uint32_t i = i_next;
    const UngHit u = u_next;
    if (run + 1 < kPrelimRun && k + 64 < n_tasks) {
      i_next = idx ? idx[k + 64] : (uint32_t)(k + 64);
      u_next = hits[i_next];
    }
    if (u.item == kNoItem) continue;  // reserved slot the seeding kernel left unused
    const PrelimHalves hv = prelim_halves(P, u, recs, subjs, qfwd, qrc, splus);
    const uint32_t r = hv.r, f = hv.f;
    const DpTask &left = hv.left, &right = hv.right;
    DpResult L, Rr;
    if constexpr (PACK) {
      TbRowStart unused;
      dp_half_tb16<RLOG2, false, RS, true>(left, mat, P.gap_open, P.gap_extend, my16, nullptr, max_pos, L, nullptr,
                                           0u, unused);
      dp_half_tb16<RLOG2, false, RS, true>(right, mat, P.gap_open, P.gap_extend, my16, nullptr, max_pos, Rr, nullptr,
                                           0u, unused);
    } else {
      dp_half<false, RS>(left, mat, P.gap_open, P.gap_extend, my, myb, wmask, L, nullptr, nullptr, 0);
      dp_half<false, RS>(right, mat, P.gap_open, P.gap_extend, my, myb, wmask, Rr, nullptr, nullptr, 0);
    }
    if ((L.flags | Rr.flags) & E_BAND_WIDE) {
      const unsigned long long kk = atomicAdd(redo_count, 1ull);
      if (kk < P.redo_cap) {
        if (LDS_RING) {
          redo[kk] = i;
        } else {
          UngHit v = u;
          v.item = P.rec_ids[r] * P.n_subj + f;
          redo_hits[kk] = v;
        }
      }
      continue;
    }
    ++dps;
    cells += L.cells + Rr.cells;
    prelim_emit(P, u, hv, L, Rr, out, out_count);
   }
  }
  atomicAdd(&counters[C_PRELIM_CELLS], cells);
  atomicAdd(&counters[C_PRELIM_DPS], dps);
}

// Preliminary DPs of the hits pass 1 handed on (bands past its LDS ring), one
// hit per wave with dp_half_wave4: when pass 1 handed on at most wave_max
// hits, so that their longest DPs, not the lanes' throughput, set the time.
// A hit whose band outgrows 256 cells goes on to the global-ring pass (its
// index appended to `redo`); with more than wave_max hits the whole list goes
// on to it (copied to `redo`).
__global__ __launch_bounds__(kDpWaveThreads) void k_prelim_wave(
    PrelimParams P, const UngHit* __restrict__ hits, const uint32_t* __restrict__ idx, uint64_t wave_max,
    const RecDesc* __restrict__ recs, const SubjDesc* __restrict__ subjs, const uint8_t* __restrict__ qfwd,
    const uint8_t* __restrict__ qrc, const uint8_t* __restrict__ splus, const int32_t* __restrict__ gmat,
    PreHit* __restrict__ out, unsigned long long* __restrict__ out_count, uint32_t* __restrict__ redo,
    unsigned long long* __restrict__ redo_count, unsigned long long* __restrict__ counters) {
  __shared__ int mat[256];
  __shared__ uint8_t bwin[(kDpWaveThreads / 64) * 512];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) mat[i] = gmat[i];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  uint8_t* win = bwin + (threadIdx.x >> 6) * 512;
  uint64_t n = P.n_hits;
  if (P.n_dev) n = *P.n_dev < n ? *P.n_dev : n;
  if (n > wave_max) {  // too many for one per wave: all of them to the lane pass
    const uint64_t gt = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = gt; k < n && k < P.redo_cap; k += nt) redo[k] = idx[k];
    if (gt == 0) *redo_count = n;
    return;
  }
  unsigned long long cells = 0, dps = 0;
  uint32_t rounds = 0;
  for (;;) {
    const uint64_t k = wave_take(P.next, 1);
    if (k >= n) break;
    const uint32_t i = idx[k];
    const UngHit u = hits[i];
    const PrelimHalves hv = prelim_halves(P, u, recs, subjs, qfwd, qrc, splus);
    DpResult L, Rr;
    dp_half_wave4<false>(hv.left, mat, P.gap_open, P.gap_extend, L, nullptr, nullptr, 0u, 0u, rounds, win);
    dp_half_wave4<false>(hv.right, mat, P.gap_open, P.gap_extend, Rr, nullptr, nullptr, 0u, 0u, rounds, win);
    if (lane == 0) {
      if ((L.flags | Rr.flags) & E_BAND_WIDE) {
        const unsigned long long kk = atomicAdd(redo_count, 1ull);
        if (kk < P.redo_cap) redo[kk] = i;
      } else {
        ++dps;
        cells += L.cells + Rr.cells;
        prelim_emit(P, u, hv, L, Rr, out, out_count);
      }
    }
  }
  if (lane == 0) {
    atomicAdd(&counters[C_PRELIM_CELLS], cells);
    atomicAdd(&counters[C_PRELIM_DPS], dps);
  }
}

// ---------------------------------------------------------- launchers --------
hipError_t launch_prelim_fused(int blocks, bool lds_ring, PrelimParams P, const UngHit* hits, const uint32_t* idx,
                               uint64_t n_tasks, const RecDesc* recs, const SubjDesc* subjs, const uint8_t* qfwd,
                               const uint8_t* qrc, const uint8_t* splus, const int32_t* gmat, int2* gring,
                               uint32_t wlog2, PreHit* out, unsigned long long* out_count, uint32_t* redo,
                               UngHit* redo_hits, unsigned long long* redo_count, unsigned long long* counters,
                               uint32_t* err, bool wide, hipStream_t st) {
  if (lds_ring && wide)
    hipLaunchKernelGGL((k_prelim_fused<true, kPrelimWideLog2, true>), dim3(blocks), dim3(kPrelimPackThreads), 0, st, P, hits,
                       idx, n_tasks, recs, subjs, qfwd, qrc, splus, gmat, gring, wlog2, out, out_count, redo,
                       redo_hits, redo_count, counters, err);
  else if (lds_ring)
    hipLaunchKernelGGL((k_prelim_fused<true, kPrelimRingLog2>), dim3(blocks), dim3(kPrelimThreads), 0, st, P, hits,
                       idx, n_tasks, recs, subjs, qfwd, qrc, splus, gmat, gring, wlog2, out, out_count, redo,
                       redo_hits, redo_count, counters, err);
  else
    hipLaunchKernelGGL((k_prelim_fused<false, kPrelimRingLog2>), dim3(blocks), dim3(kPrelimThreads), 0, st, P, hits,
                       idx, n_tasks, recs, subjs, qfwd, qrc, splus, gmat, gring, wlog2, out, out_count, redo,
                       redo_hits, redo_count, counters, err);
  return hipGetLastError();
}

hipError_t launch_prelim_wave(int blocks, PrelimParams P, const UngHit* hits, const uint32_t* idx, uint64_t wave_max,
                              const RecDesc* recs, const SubjDesc* subjs, const uint8_t* qfwd, const uint8_t* qrc,
                              const uint8_t* splus, const int32_t* gmat, PreHit* out, unsigned long long* out_count,
                              uint32_t* redo, unsigned long long* redo_count, unsigned long long* counters,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_prelim_wave, dim3(blocks), dim3(kDpWaveThreads), 0, st, P, hits, idx, wave_max, recs, subjs,
                     qfwd, qrc, splus, gmat, out, out_count, redo, redo_count, counters);
  return hipGetLastError();
}

hipError_t launch_lut_keys(const RecDesc* recs, uint32_t n_recs, const uint64_t* elem_base, const uint8_t* qext,
                           int lw, int ws, int desc, uint64_t* keys, uint32_t* vals, uint32_t* counts,
                           uint64_t invalid_key, uint64_t total, hipStream_t st) {
  const uint32_t blocks = (uint32_t)((total + 255) / 256);
  hipLaunchKernelGGL(k_lut_keys, dim3(blocks), dim3(256), 0, st, recs, n_recs, elem_base, qext, lw, ws, desc, keys,
                     vals, counts, invalid_key, total);
  return hipGetLastError();
}

hipError_t sort_pairs_u64(void* temp, size_t& temp_bytes, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                          uint32_t* vout, uint64_t n, int end_bit, hipStream_t st) {
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0, end_bit, st);
}

hipError_t exclusive_sum_u32(void* temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, uint64_t n,
                             hipStream_t st) {
  return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, st);
}

hipError_t launch_phase1(int blocks, P1Params P, const RecDesc* recs, const SubjDesc* subjs, const uint8_t* qfwd,
                         const uint8_t* qrc, const uint8_t* qext, const uint8_t* splus, const uint8_t* src,
                         const uint32_t* cells, const uint32_t* entries, const int32_t* gmat, UngHit* out,
                         unsigned long long* out_count, uint2* spill, uint4* hitres, unsigned long long* queue,
                         unsigned long long* counters, uint32_t* err, bool seeds, hipStream_t st) {
  auto kern = seeds ? (P.known_word ? k_phase1<true, true> : k_phase1<true, false>)
                    : P.part_log2 ? (P.known_word ? k_phase1<false, true, true> : k_phase1<false, false, true>)
                                  : (P.known_word ? k_phase1<false, true> : k_phase1<false, false>);
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kP1Threads), 0, st, P, recs, subjs, qfwd, qrc, qext, splus, src, cells,
                     entries, gmat, out, out_count, spill, hitres, queue, counters, err);
  return hipGetLastError();
}

hipError_t launch_dp_score(int blocks, const DpTask* tasks, uint32_t n, const int32_t* gmat, int32_t go, int32_t ge,
                           int2* ring, uint32_t wlog2, DpResult* out, unsigned long long* counters, hipStream_t st) {
  // ring == nullptr: LDS rings, `blocks` workgroups of kDpLdsThreads lanes
  if (!ring)
    hipLaunchKernelGGL(k_dp_score<true>, dim3(blocks), dim3(kDpLdsThreads), 0, st, tasks, n, gmat, go, ge, ring,
                       wlog2, out, counters, 0u);
  else
    hipLaunchKernelGGL(k_dp_score<false>, dim3(blocks), dim3(256), 0, st, tasks, n, gmat, go, ge, ring, wlog2, out,
                       counters, 0u);
  return hipGetLastError();
}

// ---- score-only X-drop DPs without device allocation ------------------------
//
// One fixed scratch array per GPU (a __device__ array: each device holds its
// own copy until the process ends) holds everything these DPs need beyond the
// caller's inputs and outputs: per-half results, the counters, an error word
// and the per-lane band rings of the rerun of halves whose band outgrows the
// 64-cell LDS ring.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
constexpr uint32_t kNoAllocMaxHalves = 65536;
constexpr uint32_t kNoAllocRerunBlocks = 8;  // x 256 lanes, each with a global ring
constexpr uint32_t kNoAllocRingLog2 = 10;
constexpr size_t kNoAllocCountersOff = (size_t)kNoAllocMaxHalves * sizeof(DpResult);
constexpr size_t kNoAllocErrOff = kNoAllocCountersOff + C_NCOUNTERS * sizeof(unsigned long long);
constexpr size_t kNoAllocRingOff = (kNoAllocErrOff + sizeof(uint32_t) + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes =
    kNoAllocRingOff + (size_t)kNoAllocRerunBlocks * 256 * ring_stride(kNoAllocRingLog2) * sizeof(int2);
static_assert(kNoAllocBytes < (24u << 20), "no-alloc scratch stays modest");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

__global__ void k_noalloc_clear(unsigned long long* __restrict__ counters, uint32_t* __restrict__ err) {
  if (threadIdx.x < C_NCOUNTERS) counters[threadIdx.x] = 0;
  if (threadIdx.x == 0) *err = 0;
}

__global__ void k_noalloc_finish(const DpResult* __restrict__ res, uint32_t n, int32_t* __restrict__ out,
                                 uint32_t* __restrict__ err) {
  uint32_t flags = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const DpResult r = res[i];
    out[3 * i] = r.score;
    out[3 * i + 1] = (int32_t)r.a_len;
    out[3 * i + 2] = (int32_t)r.b_len;
    flags |= r.flags;
  }
  if (flags) atomicOr(err, flags);
}

size_t noalloc_max_halves() { return kNoAllocMaxHalves; }

// The scratch array's address on the current device.
hipError_t noalloc_scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

// Score-only DPs of n halves (n <= kNoAllocMaxHalves) with `base`, the
// scratch address of the current device: LDS-ring pass, rerun of the halves
// flagged for their band with 2^10-cell rings, results to out (3 int32 per
// half), then the OR of every half's flags into *err_host (a synchronous
// read on st).
hipError_t launch_noalloc_score(void* base, int cu_count, const DpTask* tasks, uint32_t n, const int32_t* gmat,
                                int32_t go, int32_t ge, int32_t* out, uint32_t* err_host, hipStream_t st) {
  uint8_t* b = static_cast<uint8_t*>(base);
  DpResult* res = reinterpret_cast<DpResult*>(b);
  unsigned long long* counters = reinterpret_cast<unsigned long long*>(b + kNoAllocCountersOff);
  uint32_t* err = reinterpret_cast<uint32_t*>(b + kNoAllocErrOff);
  int2* ring = reinterpret_cast<int2*>(b + kNoAllocRingOff);
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, st, counters, err);
  const uint32_t blocks = std::min<uint32_t>((uint32_t)cu_count * 4, (n + kDpLdsThreads - 1) / kDpLdsThreads);
  hipLaunchKernelGGL(k_dp_score<true>, dim3(blocks), dim3(kDpLdsThreads), 0, st, tasks, n, gmat, go, ge, nullptr,
                     0u, res, counters, 0u);
  hipLaunchKernelGGL(k_dp_score<false>, dim3(kNoAllocRerunBlocks), dim3(256), 0, st, tasks, n, gmat, go, ge, ring,
                     kNoAllocRingLog2, res, counters, 1u);
  hipLaunchKernelGGL(k_noalloc_finish, dim3(std::min<uint32_t>(1024, (n + 255) / 256)), dim3(256), 0, st, res, n, out,
                     err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(err_host, err, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

hipError_t launch_dp_onepass(int blocks, const DpTask* tasks, uint32_t n, const int32_t* gmat, int32_t go,
                             int32_t ge, uint8_t* arena, uint32_t flat_cap, uint32_t rows_cap, uint32_t ops_cap,
                             uint32_t* ops_pool, unsigned long long* ops_count, unsigned long long ops_pool_cap,
                             OnePassResult* out, unsigned long long* next, OnePassBig big,
                             unsigned long long* counters, bool wide, hipStream_t st) {
  if (wide)
    hipLaunchKernelGGL((k_dp_onepass<true, kDpLdsWideLog2>), dim3(blocks), dim3(kDpLdsThreads), 0, st, tasks, n, gmat,
                       go, ge, nullptr, 0, arena, flat_cap, rows_cap, ops_cap, ops_pool, ops_count, ops_pool_cap, out,
                       next, big, counters);
  else
    hipLaunchKernelGGL((k_dp_onepass<true, kDpLdsLog2>), dim3(blocks), dim3(kDpLdsThreads), 0, st, tasks, n, gmat, go,
                       ge, nullptr, 0, arena, flat_cap, rows_cap, ops_cap, ops_pool, ops_count, ops_pool_cap, out,
                       next, big, counters);
  return hipGetLastError();
}

// `blocks` workgroups of four waves; `arena` holds blocks x 4 wave arenas.
hipError_t launch_dp_wave(int blocks, const DpTask* tasks, uint32_t n, const int32_t* gmat, int32_t go, int32_t ge,
                          uint8_t* arena, uint32_t rows_cap, uint32_t flat_cap, uint32_t ops_cap, uint32_t* ops_pool,
                          unsigned long long* ops_count, unsigned long long ops_pool_cap, OnePassResult* out,
                          unsigned long long* next, unsigned long long* counters, bool wide, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (wide)
    hipLaunchKernelGGL(k_dp_wave<true>, dim3(blocks), dim3(kDpWaveThreads), 0, st, tasks, n, gmat, go, ge, arena,
                       rows_cap, flat_cap, ops_cap, ops_pool, ops_count, ops_pool_cap, out, next, counters);
  else
    hipLaunchKernelGGL(k_dp_wave<false>, dim3(blocks), dim3(kDpWaveThreads), 0, st, tasks, n, gmat, go, ge, arena,
                       rows_cap, flat_cap, ops_cap, ops_pool, ops_count, ops_pool_cap, out, next, counters);
  return hipGetLastError();
}

hipError_t launch_pack_ops(const PackTask* tasks, uint32_t n, const uint32_t* ops_pool, uint32_t* packed,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_ops, dim3((n + 255) / 256), dim3(256), 0, st, tasks, n, ops_pool, packed);
  return hipGetLastError();
}

hipError_t launch_dp_traceback(int blocks, const TbTask* tasks, uint32_t n, const int32_t* gmat, int32_t go,
                               int32_t ge, int2* ring, uint32_t wlog2, uint8_t* flat, uint32_t* rows, uint32_t* ops,
                               TbResult* out, hipStream_t st) {
  // ring == nullptr: LDS rings, `blocks` workgroups of kDpLdsThreads lanes
  if (!ring)
    hipLaunchKernelGGL(k_dp_traceback<true>, dim3(blocks), dim3(kDpLdsThreads), 0, st, tasks, n, gmat, go, ge, ring,
                       wlog2, flat, rows, ops, out);
  else
    hipLaunchKernelGGL(k_dp_traceback<false>, dim3(blocks), dim3(256), 0, st, tasks, n, gmat, go, ge, ring, wlog2,
                       flat, rows, ops, out);
  return hipGetLastError();
}

}  // namespace rmb
