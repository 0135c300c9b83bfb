// Ordered host replay of the reference's per-(query, subject) control flow.
// The GPU computes every extension and DP; these routines make the
// order-dependent decisions on the compact results: containment
// (search/itree.rs), pivots (engine.rs:1850-1895), endpoint purges
// (engine.rs:1969-2083, 2576-2769), seed improvement (engine.rs:383-572),
// complexity adjustment (complexity_adjust.rs) and the cut-remainder
// re-evaluation (engine.rs:2268-2567).
#pragma once
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace rmb {

struct Scoring {
  int32_t s[16][16];
  double freqs[16];
  double lambda;
};

enum : uint8_t { K_SUB = 0, K_GAP_Q = 1, K_GAP_S = 2 };  // Sub, GapInQuery, GapInSubject
struct Run { uint8_t kind; uint32_t len; };
using Script = std::vector<Run>;
void script_push(Script& s, uint8_t kind, uint32_t len);

extern const uint8_t kBlastnaComplement[16];

// Record-frame ungapped hit (engine.rs:100-118).
struct UngRec { uint32_t q_seed, s_seed, q_start, q_end, s_start, s_end; int32_t score; uint8_t minus; };
// Preliminary HSP (engine.rs:158-169).
struct PreRec { uint32_t q_seed, s_seed, q_start, q_end, s_start, s_end; int32_t score; uint8_t minus; };
// Final HSP of one subject; minus-strand subject coordinates on the plus strand.
struct FinalRec {
  int32_t score;
  uint32_t q_start, q_end, s_start, s_end;
  uint8_t minus;
  Script ops;
};

// NCBI ungapped-hit order (engine.rs:1280-1309).
bool ung_before(const UngRec& a, const UngRec& b, uint32_t q_len, uint32_t s_len);
void order_ungapped(std::vector<UngRec>& v, uint32_t q_len, uint32_t s_len);
// Phase-2a pivot and traceback seed (engine.rs:1850-1895).
void pivot_of(const UngRec& u, uint32_t qa_len, uint32_t sa_len, uint32_t lw, uint32_t& pq, uint32_t& ps,
              uint32_t& tq, uint32_t& ts);
// Blast_HSPListPurgeHSPsWithCommonEndpoints on prelims (engine.rs:1969-2083).
void purge_prelims(std::vector<PreRec>& v, uint32_t sa_len);
// NCBI's query split (rmblastn/src/main.rs:1150-1193): chunks of `size` bases
// starting every `step`; one chunk below 1,999,800 bases.
struct QuerySplit { uint32_t size, count, step; };
QuerySplit query_split(uint32_t n);
// Folds a later chunk's prelims (query coordinates) into the accumulated ones
// at chunk start `split` (rmblastn/src/main.rs:1016-1085).
void merge_chunk_prelims(std::vector<PreRec>& acc, std::vector<PreRec>&& later, uint32_t split);
// Traceback seed inside a prelim (engine.rs:383-572).
void improve_seed(const uint8_t* q, size_t qlen, const uint8_t* s, size_t slen, uint32_t q_seed, uint32_t s_seed,
                  uint32_t pqs, uint32_t pqe, uint32_t pss, uint32_t pse, bool minus, uint32_t& oq, uint32_t& os);
// complexity_adjust.rs:57-132; false = drop.
bool complexity_adjust(int32_t raw, const uint8_t* query_wrapped, size_t qw_len, uint32_t q_start, const Script& ops,
                       const Scoring& m, int32_t cutoff, bool minus, int32_t& out);

// 2-D containment tree (search/itree.rs).
class HspTree {
 public:
  HspTree(uint32_t q_hi, uint32_t s_hi, bool rc_query = false);
  bool add(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus);  // with endpoint dedup
  void add_plain(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc);
  bool covers(uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus) const;

 private:
  struct Box { uint32_t qs, qe, ss, se; int32_t sc; };
  struct Node { uint32_t lo, hi, left, mid, right, box; };  // box == kInner for inner nodes
  static constexpr uint32_t kInner = 0xFFFFFFFFu;
  std::vector<Node> nodes_;
  std::vector<Box> boxes_;
  std::vector<uint8_t> dead_;
  std::map<uint64_t, std::pair<int32_t, uint32_t>> lend_, rend_;
  uint32_t s_hi_;
  bool rcq_;
  uint32_t new_inner(uint32_t lo, uint32_t hi);
  void place(uint32_t leaf);
  void retire(uint32_t idx, bool plus);
  bool test(uint32_t idx, uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc, bool plus) const;
  bool subject_tree_covers(uint32_t root, uint32_t qs, uint32_t qe, uint32_t ss, uint32_t se, int32_t sc,
                           bool plus) const;
};

// Phase 2b tail for one subject: common-endpoint cut, re-evaluation of the
// cut remainders, delete pass and the final containment pass
// (engine.rs:1637-1759).  `subj_fwd`/`subj_rc` are the ambiguity-restored
// subject strands (sentinel-wrapped), n_mask/n_mask_rc as in seq/mod.rs:38-100.
void finish_subject(std::vector<FinalRec>& hsps, const uint8_t* query_wrapped, uint32_t qa_len,
                    const uint8_t* subj_fwd, const uint8_t* subj_rc, uint32_t sa_len,
                    const std::vector<uint8_t>& n_mask, const std::vector<uint8_t>& n_mask_rc, const Scoring& m,
                    int32_t gap_open, int32_t gap_extend, int32_t cutoff);

}  // namespace rmb
