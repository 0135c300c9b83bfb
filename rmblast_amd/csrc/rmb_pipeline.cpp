// librmblast_hip: context, device memory, stage orchestration and the C ABI.
//
// Stage chain of one rmb_search call (all records x all subjects, both strands):
//   1. lookup tables for every record on the GPU (k_lut_keys + radix sort + scan)
//   2. k_phase1 over (record, subject) items from a device queue -> ungapped hits
//   3. per item, NCBI order of the hits; Phase 2a as ordered rounds: the host
//      walks each item's hits against its containment trees, the GPU runs the
//      score-only X-drop DPs of the hits that need one (k_dp_score)
//   4. Phase 2b: every prelim's traceback DP on the GPU (score pass, then the
//      traceback pass with exact-size storage), ordered replay on the host
//      (containment, complexity adjustment, common-endpoint purge)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <sstream>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <sched.h>

#include <thread>
#include <vector>

#include "../../include/rmblast_hip.h"
#include "rmb_device.h"
#include "rmb_dust.h"
#include "rmb_ka.h"

#include <map>
#include <unordered_map>
#include "rmb_replay.h"
#include "rmb_report.h"

namespace rmb {
hipError_t launch_lut_keys(const RecDesc*, uint32_t, const uint64_t*, const uint8_t*, int, int, int, uint64_t*,
                           uint32_t*, uint32_t*, uint64_t, uint64_t, hipStream_t);
hipError_t sort_pairs_u64(void*, size_t&, const uint64_t*, uint64_t*, const uint32_t*, uint32_t*, uint64_t, int,
                          hipStream_t);
hipError_t exclusive_sum_u32(void*, size_t&, const uint32_t*, uint32_t*, uint64_t, hipStream_t);
hipError_t launch_phase1(int, P1Params, const RecDesc*, const SubjDesc*, const uint8_t*, const uint8_t*,
                         const uint8_t*, const uint8_t*, const uint8_t*, const uint32_t*, const uint32_t*,
                         const int32_t*, UngHit*, unsigned long long*, uint2*, uint4*, unsigned long long*,
                         unsigned long long*, uint32_t*, bool, hipStream_t);
hipError_t noalloc_scratch_address(void**);
hipError_t launch_resolve_ambiguity(uint8_t*, const SubjDesc*, uint32_t, hipStream_t);
hipError_t launch_revcomp_segments(const uint8_t*, uint8_t*, const uint64_t*, const uint32_t*, const uint2*, uint32_t,
                                   hipStream_t);
uint32_t revcomp_chunk();
size_t noalloc_max_halves();
hipError_t launch_noalloc_score(void*, int, const DpTask*, uint32_t, const int32_t*, int32_t, int32_t, int32_t*,
                                uint32_t*, hipStream_t);
hipError_t launch_pack_ops(const PackTask*, uint32_t, const uint32_t*, uint32_t*, hipStream_t);
hipError_t launch_dp_score(int, const DpTask*, uint32_t, const int32_t*, int32_t, int32_t, int2*, uint32_t,
                           DpResult*, unsigned long long*, hipStream_t);
hipError_t launch_dp_traceback(int, const TbTask*, uint32_t, const int32_t*, int32_t, int32_t, int2*, uint32_t,
                               uint8_t*, uint32_t*, uint32_t*, TbResult*, hipStream_t);
hipError_t launch_dp_onepass(int, const DpTask*, uint32_t, const int32_t*, int32_t, int32_t, uint8_t*, uint32_t,
                             uint32_t, uint32_t, uint32_t*, unsigned long long*, unsigned long long, OnePassResult*,
                             unsigned long long*, OnePassBig, unsigned long long*, bool, hipStream_t);
hipError_t launch_prelim_wave(int, PrelimParams, const UngHit*, const uint32_t*, uint64_t, const RecDesc*,
                              const SubjDesc*, const uint8_t*, const uint8_t*, const uint8_t*, const int32_t*, PreHit*,
                              unsigned long long*, uint32_t*, unsigned long long*, unsigned long long*, hipStream_t);
hipError_t launch_dp_wave(int, const DpTask*, uint32_t, const int32_t*, int32_t, int32_t, uint8_t*, uint32_t, uint32_t,
                          uint32_t, uint32_t*, unsigned long long*, unsigned long long, OnePassResult*,
                          unsigned long long*, unsigned long long*, bool, hipStream_t);
hipError_t launch_prelim_fused(int, bool, PrelimParams, const UngHit*, const uint32_t*, uint64_t, const RecDesc*,
                               const SubjDesc*, const uint8_t*, const uint8_t*, const uint8_t*, const int32_t*, int2*,
                               uint32_t, PreHit*, unsigned long long*, uint32_t*, UngHit*, unsigned long long*,
                               unsigned long long*, uint32_t*, bool, hipStream_t);

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Failure(RMB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  template <class T>
  T* get(size_t n_elems) {
    const size_t bytes = std::max<size_t>(n_elems * sizeof(T), 16);
    if (bytes > cap) {
      if (p) hip_check(hipFree(p), "hipFree");
      p = nullptr;
      // 64 bytes of tail: the kernels' aligned 16-byte sequence windows may
      // read up to 15 bytes past the last sentinel of a pool.
      hip_check(hipMalloc(&p, bytes + 64), "hipMalloc");
      cap = bytes;
    }
    return (T*)p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Pinned host staging buffer, grown on demand (the query pools: uploads run as
// DMA on a stream instead of through the runtime's pageable-copy path).
struct HostBuf {
  uint8_t* p = nullptr;
  size_t cap = 0;
  uint8_t* get(size_t bytes) {
    if (bytes > cap) {
      if (p) hip_check(hipHostFree(p), "hipHostFree");
      p = nullptr;
      hip_check(hipHostMalloc((void**)&p, std::max<size_t>(bytes, 16), hipHostMallocDefault), "hipHostMalloc");
      cap = std::max<size_t>(bytes, 16);
    }
    return p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// A sentinel-wrapped query (code 15, the bases, code 15) inside a pool.
struct SeqView {
  const uint8_t* p = nullptr;
  size_t n = 0;
  const uint8_t* data() const { return p; }
  size_t size() const { return n; }
  const uint8_t* begin() const { return p; }
  const uint8_t* end() const { return p + n; }
};

// Host worker threads: RMB_THREADS if set, else every CPU the process may run
// on (its affinity mask: a job's share of the host on a shared node).
static unsigned host_threads() {
  static const unsigned n = [] {
    if (const char* e = std::getenv("RMB_THREADS")) {
      const long v = std::strtol(e, nullptr, 10);
      if (v > 0) return (unsigned)std::min<long>(v, 1024);
    }
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0 && CPU_COUNT(&set) > 0) return (unsigned)CPU_COUNT(&set);
    return std::max(1u, std::thread::hardware_concurrency());
  }();
  return n;
}

static void parallel_for(size_t n, const std::function<void(size_t)>& fn) {
  if (n == 0) return;
  unsigned nt = host_threads();
  nt = (unsigned)std::min<size_t>(nt, n);
  if (nt == 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  std::exception_ptr failure;
  std::atomic<bool> failed{false};
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      try {
        for (size_t i; !failed && (i = next.fetch_add(1)) < n;) fn(i);
      } catch (...) {
        if (!failed.exchange(true)) failure = std::current_exception();
      }
    });
  for (auto& th : pool) th.join();
  if (failure) std::rethrow_exception(failure);
}

// NCBI lookup width thresholds for a two-strand query of `e` entries.
static int lut_width_for(int ws, size_t e) {
  switch (ws) {
    case 4: case 5: case 6: return ws;
    case 7: return e < 250 ? 6 : 7;
    case 8: return e < 8500 ? 7 : 8;
    case 9: return e < 1250 ? 7 : (e < 21000 ? 8 : 9);
    case 10: return e < 1250 ? 7 : (e < 8500 ? 8 : (e < 18000 ? 9 : 10));
    case 11: return e < 12000 ? 8 : (e < 180000 ? 10 : 11);
    case 12: return e < 8500 ? 8 : (e < 18000 ? 9 : (e < 60000 ? 10 : (e < 900000 ? 11 : 12)));
    default: return e < 8500 ? 8 : (e < 300000 ? 11 : 12);
  }
}

struct Subject {  // host views; the resolved strands live on the device only
  std::vector<uint8_t> align, align_rc;  // sentinel-wrapped original codes and their reverse complement
  std::vector<uint8_t> n_mask, n_mask_rc;
};

// Device outputs a queued Phase-1 batch keeps until the host has read them;
// the context holds two, so one batch is read while the next one runs.
struct P1Slot {
  DevBuf pre, redo_hits, cnt, hitcnt, ctr, err;
  std::vector<hipEvent_t> ev;  // per group: k_phase1, prelim pass 1, prelim pass 2 (start, stop)
  hipEvent_t done = nullptr;
};

static std::vector<uint8_t> wrap(const uint8_t* s, uint32_t n) {
  std::vector<uint8_t> v(n + 2);
  v[0] = 15;
  for (uint32_t i = 0; i < n; ++i) v[1 + i] = s[i] & 15;
  v[n + 1] = 15;
  return v;
}
static std::vector<uint8_t> revcomp_wrapped(const std::vector<uint8_t>& v) {
  std::vector<uint8_t> r(v.size());
  for (size_t i = 0; i < v.size(); ++i) r[i] = kBlastnaComplement[v[v.size() - 1 - i] & 15];
  return r;
}

}  // namespace rmb

using namespace rmb;

struct rmb_ctx {
  int device = 0;
  hipStream_t st = nullptr;   // Phase 2b and the taps
  hipStream_t st1 = nullptr;  // seeding and preliminary DPs
  P1Slot slot[2];
  double hits_per_item = 0;   // most ungapped hits per searched item seen, sizes the hit buffers
  rmb_params P{};
  Scoring M{};
  std::string err;
  int32_t* d_mat = nullptr;
  int cu_count = 256;
  void* noalloc_scratch = nullptr;  // this device's copy of the no-alloc DPs' scratch array
  rmb_tuning tune{};                 // schedule choices (0 fields: the defaults)
  // library
  std::vector<Subject> subj;
  std::vector<SubjDesc> sdesc;
  uint64_t avg_subj_len = 0;  // total residues / sequences, truncated (BlastSeqSrcGetAvgSeqLen)
  uint64_t db_letters = 0;    // total residues of the loaded subjects
  DevBuf d_splus, d_src, d_salign, d_sdesc;
  // results of the last search
  std::vector<rmb_hsp> hsps;
  std::vector<rmb_op> ops;
  uint64_t stats[RMB_NSTATS] = {0};
  uint64_t host_prelim_dps = 0;  // preliminary DPs of the last search run through run_dp_score
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // scratch
  HostBuf h_qfwd;  // host forward query pool of the last search
  std::vector<const uint8_t*> last_q;  // the last search's whole queries (wrapped) in h_qfwd, for the report
  std::vector<uint32_t> last_qlen;
  ReportText report;                   // text of the last rmb_report_tabular
  bool tap_pre_on = false;             // the prelims tap collects Phase 2a's output into tap_pre
  std::vector<PreRec> tap_pre;
  DevBuf b_qfwd, b_qrc, b_recs, b_elem, b_keys, b_keys2, b_vals, b_vals2, b_counts, b_cells, b_sort_tmp;
  DevBuf b_hits, b_hitcount, b_spill, b_hitres, b_queue, b_counters, b_err;
  DevBuf b_tasks, b_dpres, b_ring, b_ring_big, b_tbtasks, b_tbres, b_flat, b_rows, b_ops;
  DevBuf b_pre, b_redo, b_redo2, b_precount, b_packtasks, b_packed, b_items, b_qext, b_recids;
  DevBuf b_arena, b_arena_big, b_arena_wave, b_oppool, b_opres, b_opcount;  // one-pass traceback
  DevBuf b_seg_base, b_seg_len, b_seg_work;  // reverse-complement segments of a pool
};

namespace rmb {

static const int kMaxBatches = 8;                   // Phase-1 batches a search is cut into (RMB_BATCHES overrides)
static const uint64_t kBatchMinBases = 4000000;     // bases per batch at least
static const uint32_t kRingLog2 = 10;   // per-lane band ring for the common case
static const uint32_t kRingBigLog2 = 17;

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  uint64_t us() const {
    return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Runs one kernel launch between two events on the context's stream and adds
// its duration in nanoseconds to stats[slot].
template <class F>
static void timed_launch(rmb_ctx* c, int slot, F&& launch) {
  hip_check(hipEventRecord(c->ev0, c->st), "event");
  launch();
  hip_check(hipEventRecord(c->ev1, c->st), "event");
  hip_check(hipEventSynchronize(c->ev1), "event sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "event time");
  c->stats[slot] += (uint64_t)((double)ms * 1e6);
}

// Lanes of a wave run their DPs in lockstep, so a wave lasts as long as its
// longest DP: tasks are launched longest first (by `work`, an estimate),
// which puts DPs of like size in one wave.  The order is a counting sort on
// quarter-octave classes of the work (within 19% of each other), stable.
static std::vector<uint32_t> longest_first(const std::vector<uint64_t>& work) {
  auto cls = [](uint64_t w) -> uint32_t {
    if (w < 4) return (uint32_t)w;
    const uint32_t e = 63u - (uint32_t)__builtin_clzll(w);  // >= 2
    return 4 * (e - 1) + (uint32_t)((w >> (e - 2)) & 3u);   // 4 .. 255
  };
  std::vector<uint32_t> start(257, 0), ord(work.size());
  for (uint64_t w : work) ++start[255 - cls(w) + 1];
  for (int k = 0; k < 256; ++k) start[k + 1] += start[k];
  for (uint32_t i = 0; i < work.size(); ++i) ord[start[255 - cls(work[i])]++] = i;
  return ord;
}

// Run a list of lane-serial score-only DPs; tasks whose band outgrows the
// common ring are rerun once with a ring of 2^kRingBigLog2 cells.  With
// `work`, tasks run longest first.
static void run_dp_score(rmb_ctx* c, const std::vector<DpTask>& tasks_in, std::vector<DpResult>& res,
                         const std::vector<uint64_t>* work = nullptr) {
  Timer tt;
  res.assign(tasks_in.size(), DpResult{});
  if (tasks_in.empty()) return;
  std::vector<uint32_t> ord;
  std::vector<DpTask> sorted;
  if (work) {
    ord = longest_first(*work);
    sorted.resize(tasks_in.size());
    for (size_t k = 0; k < ord.size(); ++k) sorted[k] = tasks_in[ord[k]];
  }
  const std::vector<DpTask>& tasks = work ? sorted : tasks_in;
  std::vector<DpResult> sres;
  std::vector<DpResult>& res_o = work ? sres : res;
  if (work) sres.assign(tasks.size(), DpResult{});
  const uint64_t t_prep = tt.us();
  unsigned long long* ctr = c->b_counters.get<unsigned long long>(C_NCOUNTERS);
  const size_t chunk = 1u << 20;
  for (size_t base = 0; base < tasks.size(); base += chunk) {
    const uint32_t n = (uint32_t)std::min(chunk, tasks.size() - base);
    DpTask* dt = c->b_tasks.get<DpTask>(n);
    DpResult* dr = c->b_dpres.get<DpResult>(n);
    hip_check(hipMemcpyAsync(dt, tasks.data() + base, n * sizeof(DpTask), hipMemcpyHostToDevice, c->st), "copy tasks");
    // 64-lane workgroups with a 32 KB LDS band ring each, four per CU
    const uint32_t blocks = std::min<uint32_t>((uint32_t)c->cu_count * 4, (n + 63) / 64);
    timed_launch(c, 14, [&] {
      hip_check(launch_dp_score((int)blocks, dt, n, c->d_mat, c->P.gap_open, c->P.gap_extend, nullptr, 0, dr, ctr,
                                c->st),
                "k_dp_score");
    });
    hip_check(hipMemcpyAsync(res_o.data() + base, dr, n * sizeof(DpResult), hipMemcpyDeviceToHost, c->st), "copy dp");
  }
  hip_check(hipStreamSynchronize(c->st), "dp sync");
  const uint64_t t_gpu = tt.us();
  std::vector<uint32_t> wide;
  for (size_t i = 0; i < res_o.size(); ++i)
    if (res_o[i].flags & E_BAND_WIDE) wide.push_back((uint32_t)i);
  if (!wide.empty()) {
    std::vector<DpTask> wt(wide.size());
    for (size_t i = 0; i < wide.size(); ++i) wt[i] = tasks[wide[i]];
    const uint32_t n = (uint32_t)wt.size();
    DpTask* dt = c->b_tasks.get<DpTask>(n);
    DpResult* dr = c->b_dpres.get<DpResult>(n);
    hip_check(hipMemcpyAsync(dt, wt.data(), n * sizeof(DpTask), hipMemcpyHostToDevice, c->st), "copy wide");
    const uint32_t lanes = std::min<uint32_t>(4096, ((n + 255) / 256) * 256);
    int2* ring = c->b_ring_big.get<int2>((size_t)lanes * ring_stride(kRingBigLog2));
    timed_launch(c, 14, [&] {
      hip_check(launch_dp_score((int)(lanes / 256), dt, n, c->d_mat, c->P.gap_open, c->P.gap_extend, ring,
                                kRingBigLog2, dr, ctr, c->st),
                "k_dp_score wide");
    });
    std::vector<DpResult> wr(n);
    hip_check(hipMemcpyAsync(wr.data(), dr, n * sizeof(DpResult), hipMemcpyDeviceToHost, c->st), "copy wide res");
    hip_check(hipStreamSynchronize(c->st), "dp wide sync");
    for (size_t i = 0; i < wide.size(); ++i) {
      if (wr[i].flags & E_BAND_WIDE)
        throw Failure(RMB_ERR_CAPACITY, "X-drop band wider than 2^17 cells");
      res_o[wide[i]] = wr[i];
    }
  }
  if (work)
    for (size_t k = 0; k < ord.size(); ++k) res[ord[k]] = sres[k];
  if (work && std::getenv("RMB_STAGE_TIMES"))
    std::fprintf(stderr, "[rmb] score DPs ms: order %.1f  launch+copy %.1f  wide+reorder %.1f  (%zu tasks, %zu wide)\n",
                 t_prep / 1e3, (t_gpu - t_prep) / 1e3, (tt.us() - t_gpu) / 1e3, tasks.size(), wide.size());
}

// One traceback half: score, lengths, and its run-length ops (kind | len << 2,
// in alignment order from the seed outwards) at ops_at in the op pool.
struct HalfTb { int32_t score; uint32_t a_len, b_len, n_ops; uint64_t ops_at; };

// Traceback DPs in two passes: score pass for sizes, then the traceback pass
// in batches bounded by the device storage they need.
static void run_dp_traceback_2pass(rmb_ctx* c, const std::vector<DpTask>& tasks, std::vector<HalfTb>& out,
                                   std::vector<uint32_t>& op_pool, const std::vector<uint64_t>* work = nullptr) {
  out.assign(tasks.size(), HalfTb{0, 0, 0, 0, 0});
  op_pool.clear();
  Timer tt;
  uint64_t t_tb = 0, t_unpack = 0;
  std::vector<DpResult> sres;
  run_dp_score(c, tasks, sres, work);
  const uint64_t t_score = tt.us();
  std::vector<uint32_t> need;
  for (size_t i = 0; i < tasks.size(); ++i) {
    out[i].score = sres[i].score;
    out[i].a_len = sres[i].a_len;
    out[i].b_len = sres[i].b_len;
    if (sres[i].a_len > 0 || sres[i].b_len > 0) need.push_back((uint32_t)i);
  }
  // the score pass measured each band: the traceback pass runs longest first
  {
    std::vector<uint64_t> tb(need.size());
    for (size_t k = 0; k < need.size(); ++k) tb[k] = sres[need[k]].tb_bytes;
    const std::vector<uint32_t> ord = longest_first(tb);
    std::vector<uint32_t> sorted(need.size());
    for (size_t k = 0; k < ord.size(); ++k) sorted[k] = need[ord[k]];
    need.swap(sorted);
  }
  const uint64_t kBudget = 1ull << 31;
  size_t pos = 0;
  while (pos < need.size()) {
    std::vector<TbTask> batch;
    uint64_t flat = 0, rows = 0, opsn = 0;
    while (pos < need.size()) {
      const uint32_t i = need[pos];
      const DpResult& r = sres[i];
      const uint64_t ocap = (uint64_t)r.a_len + r.b_len + 1;
      if (!batch.empty() && flat + r.tb_bytes > kBudget) break;
      TbTask t{};
      t.t = tasks[i];
      t.flat_off = flat;
      t.row_off = rows;
      t.ops_off = opsn;
      t.stop_row = r.a_len;
      t.best_b = r.b_len;
      t.ops_cap = (uint32_t)ocap;
      flat += r.tb_bytes;
      rows += r.tb_rows;
      opsn += ocap;
      batch.push_back(t);
      ++pos;
    }
    const uint32_t n = (uint32_t)batch.size();
    const uint64_t t_b0 = tt.us();
    TbTask* dt = c->b_tbtasks.get<TbTask>(n);
    TbResult* dr = c->b_tbres.get<TbResult>(n);
    uint8_t* dflat = c->b_flat.get<uint8_t>(flat + 16);
    uint32_t* drows = c->b_rows.get<uint32_t>(2 * rows + 16);
    uint32_t* dops = c->b_ops.get<uint32_t>(opsn + 16);
    hip_check(hipMemcpyAsync(dt, batch.data(), n * sizeof(TbTask), hipMemcpyHostToDevice, c->st), "copy tb tasks");
    std::vector<TbResult> tr(n);
    const uint32_t blocks = std::min<uint32_t>((uint32_t)c->cu_count * 4, (n + 63) / 64);
    timed_launch(c, 14, [&] {
      hip_check(launch_dp_traceback((int)blocks, dt, n, c->d_mat, c->P.gap_open, c->P.gap_extend, nullptr, 0, dflat,
                                    drows, dops, dr, c->st),
                "k_dp_traceback");
    });
    hip_check(hipMemcpyAsync(tr.data(), dr, n * sizeof(TbResult), hipMemcpyDeviceToHost, c->st), "copy tb res");
    hip_check(hipStreamSynchronize(c->st), "tb sync");
    // rerun band-wide tasks with the big ring
    std::vector<uint32_t> redo;
    for (uint32_t k = 0; k < n; ++k)
      if (tr[k].flags & E_BAND_WIDE) redo.push_back(k);
    if (!redo.empty()) {
      std::vector<TbTask> rb(redo.size());
      for (size_t k = 0; k < redo.size(); ++k) rb[k] = batch[redo[k]];
      const uint32_t rn = (uint32_t)rb.size();
      TbTask* rdt = c->b_tbtasks.get<TbTask>(std::max<size_t>(n, rn));
      hip_check(hipMemcpyAsync(rdt, rb.data(), rn * sizeof(TbTask), hipMemcpyHostToDevice, c->st), "copy tb redo");
      const uint32_t rl = std::min<uint32_t>(4096, ((rn + 255) / 256) * 256);
      int2* big = c->b_ring_big.get<int2>((size_t)rl * ring_stride(kRingBigLog2));
      timed_launch(c, 14, [&] {
        hip_check(launch_dp_traceback((int)(rl / 256), rdt, rn, c->d_mat, c->P.gap_open, c->P.gap_extend, big,
                                      kRingBigLog2, dflat, drows, dops, dr, c->st),
                  "k_dp_traceback wide");
      });
      std::vector<TbResult> rr(rn);
      hip_check(hipMemcpyAsync(rr.data(), dr, rn * sizeof(TbResult), hipMemcpyDeviceToHost, c->st), "copy redo");
      hip_check(hipStreamSynchronize(c->st), "tb redo sync");
      for (size_t k = 0; k < redo.size(); ++k) tr[redo[k]] = rr[k];
    }
    // pack the ops on the GPU and copy back only those
    std::vector<PackTask> pk(n);
    uint64_t packed_n = 0;
    for (uint32_t k = 0; k < n; ++k) {
      if (tr[k].flags) throw Failure(RMB_ERR_CAPACITY, "traceback storage or band limit exceeded");
      pk[k] = PackTask{batch[k].ops_off, packed_n, tr[k].n_ops, 0};
      packed_n += tr[k].n_ops;
    }
    PackTask* dpk = c->b_packtasks.get<PackTask>(n);
    uint32_t* dpacked = c->b_packed.get<uint32_t>(packed_n + 1);
    hip_check(hipMemcpyAsync(dpk, pk.data(), n * sizeof(PackTask), hipMemcpyHostToDevice, c->st), "copy pack");
    hip_check(launch_pack_ops(dpk, n, dops, dpacked, c->st), "k_pack_ops");
    const uint64_t pool_at = op_pool.size();
    op_pool.resize(pool_at + packed_n);
    hip_check(hipMemcpyAsync(op_pool.data() + pool_at, dpacked, packed_n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->st),
              "copy ops");
    hip_check(hipStreamSynchronize(c->st), "ops sync");
    const uint64_t t_b1 = tt.us();
    for (uint32_t k = 0; k < n; ++k) {
      HalfTb& h = out[need[pos - n + k]];
      h.ops_at = pool_at + pk[k].dst;
      h.n_ops = tr[k].n_ops;
    }
    t_tb += t_b1 - t_b0;
    t_unpack += tt.us() - t_b1;
  }
  if (std::getenv("RMB_STAGE_TIMES"))
    std::fprintf(stderr, "[rmb] traceback DPs ms: score pass %.1f  traceback pass %.1f  unpack %.1f  total %.1f\n",
                 t_score / 1e3, t_tb / 1e3, t_unpack / 1e3, tt.us() / 1e3);
}

// Per-lane arena of the one-pass traceback: the traceback stream (each row's
// bytes and its trailer) and the ops of one half alignment.  Halves beyond it
// (and bands wider than the LDS ring) take a retry slot or the two-pass path.
constexpr uint32_t kOnePassFlat = 256u << 10, kOnePassOps = 4096;
// lane arenas of the wide-ring launches (two workgroups per CU, so 32,768 lanes)
constexpr uint32_t kOnePassWideFlat = 1u << 20, kOnePassWideOps = 16384;
constexpr unsigned long long kOnePassPool = 64ull << 20;  // packed ops per launch
constexpr uint32_t kOnePassPerCu = 9;  // 64-lane workgroups with ~17 KB of LDS: nine fit a CU
constexpr uint32_t kOnePassWidePerCu = 2;  // with the 256-cell rings, ~65 KB: two fit a CU
// A gap of g bases costs open + g x extend, so a band reaches about
// (X - open) / extend cells to either side of its best cell: past 48 cells
// in all the one-pass kernel takes its wider rings (RMB_TB_WIDE=0 keeps the
// 64-cell rings).
static bool onepass_wide(const rmb_ctx* c) {
  if (c->tune.tb_ring) return c->tune.tb_ring == 256;
  if (const char* e = std::getenv("RMB_TB_WIDE"))
    if (std::atoi(e) == 0) return false;
  const int ge = std::max(1, (int)c->P.gap_extend);
  return 2 * std::max(0, (int)c->P.xdrop_gap_final - (int)c->P.gap_open) / ge + 1 > 48;
}
// retry slots for halves beyond the LDS ring or the lane arena: global rings of
// 2^kRingBigLog2 cells and 70 MB arenas (RMB_TB_RETRY_SLOTS lowers the count,
// down to 0, which sends every such half to the two-pass path)
constexpr uint32_t kBigSlots = 32, kBigRows = 1u << 18, kBigFlat = 64u << 20, kBigOps = 1u << 20;
static uint32_t retry_slots() {
  if (const char* e = std::getenv("RMB_TB_RETRY_SLOTS")) return (uint32_t)std::min<long>(std::max(0L, std::atol(e)), kBigSlots);
  return kBigSlots;
}

// A launch of at most kTbWaveHalvesPerCu x CUs halves runs every half one
// per wave, more halves one per lane.  A lane's half costs about 35
// instructions per band cell, and a lane-run wave lasts as long as the most
// cells among its 64 halves; a wave-run half costs about 150 per band row
// (2 scan rounds) whatever the band (up to 64 cells; four cells a lane up
// to 256, about 300 per row).  With few halves the
// lane-run kernel is one row of waves whose slowest lanes set the time
// (1 Mbp x 50 families at 4 kbp records, 14 k halves: 27 ms lane-run, 10 ms
// wave-run); with many, the lanes' throughput wins (the 100 Mbp x 1,000
// family workload, 117 k halves a batch: 0.99 s lane-run, 1.13 s wave-run).
constexpr int32_t kTbWaveHalvesPerCu = 112;
// wave arenas: row table, traceback bytes (at most two rings' worth per row)
// and ops of one half
constexpr uint32_t kWaveRows = 8192, kWaveFlat = kWaveRows * 128 + 256, kWaveOps = 4096;
constexpr uint32_t kWavePerCu = 8;  // workgroups of four waves per CU
// with bands of up to 256 cells (four per lane): rows of up to 512 bytes, the
// arena holding 2 MB of them, four workgroups per CU (8.6 GB of arenas)
constexpr uint32_t kWaveWideRows = 8192, kWaveWideFlat = 2u << 20, kWaveWideOps = 4096;
constexpr uint32_t kWaveWidePerCu = 4;
// Preliminary hits past the first pass's LDS ring: up to this many per CU run
// one per wave (RMB_PRELIM_WAVE_HITS sets the total; 0 sends every one to the
// global-ring lane pass), in workgroups of four waves, four per CU.
constexpr uint64_t kPrelimWaveHitsPerCu = 128;
constexpr uint32_t kPrelimWaveWgPerCu = 4;
static uint64_t prelim_wave_hits(const rmb_ctx* c) {
  if (const char* e = std::getenv("RMB_PRELIM_WAVE_HITS")) return (uint64_t)std::max(0LL, std::atoll(e));
  return kPrelimWaveHitsPerCu * (uint64_t)c->cu_count;
}

// Traceback DPs: one pass (score, traceback bytes, walk) per half in a
// per-lane arena, longest first (the longest one per wave, the rest one per
// lane); the halves it flags run in two passes.
static void run_dp_traceback(rmb_ctx* c, const std::vector<DpTask>& tasks, std::vector<HalfTb>& out,
                             std::vector<uint32_t>& op_pool, const std::vector<uint64_t>* work = nullptr) {
  out.assign(tasks.size(), HalfTb{0, 0, 0, 0, 0});
  op_pool.clear();
  if (tasks.empty()) return;
  Timer tt;
  std::vector<uint32_t> ord;
  if (work) ord = longest_first(*work);
  else {
    ord.resize(tasks.size());
    for (uint32_t i = 0; i < ord.size(); ++i) ord[i] = i;
  }
  std::vector<DpTask> sorted(tasks.size());
  for (size_t k = 0; k < ord.size(); ++k) sorted[k] = tasks[ord[k]];
  // few halves: every one runs one per wave (RMB_TB_WAVE_HALVES sets the
  // limit; 0 runs every half one per lane)
  int64_t wave_halves = c->tune.tb_wave_halves ? c->tune.tb_wave_halves : (int64_t)kTbWaveHalvesPerCu * c->cu_count;
  if (const char* e = std::getenv("RMB_TB_WAVE_HALVES")) wave_halves = std::atoll(e);
  const size_t n_wave_all = wave_halves > 0 && ord.size() <= (size_t)wave_halves ? ord.size() : 0;
  unsigned long long* ctr = c->b_counters.get<unsigned long long>(C_NCOUNTERS);
  std::vector<uint32_t> fallback;
  const size_t chunk = 1u << 20;
  for (size_t base = 0; base < sorted.size(); base += chunk) {
    const uint32_t n = (uint32_t)std::min(chunk, sorted.size() - base);
    DpTask* dt = c->b_tasks.get<DpTask>(n);
    OnePassResult* dr = c->b_opres.get<OnePassResult>(n);
    const bool wide = onepass_wide(c);
    // lane arenas are per workgroup: at most 9 (2 with the wide rings) per CU
    const uint32_t per_cu = c->tune.tb_wg_per_cu ? std::min<uint32_t>((uint32_t)c->tune.tb_wg_per_cu, wide ? 2u : 9u)
                                                 : (wide ? kOnePassWidePerCu : kOnePassPerCu);
    // the first n_wave halves run one per wave, the rest one per lane
    const uint32_t n_wave = base < n_wave_all ? (uint32_t)std::min<size_t>(n, n_wave_all - base) : 0u;
    const uint32_t n_lane = n - n_wave;
    const uint32_t blocks = std::min<uint32_t>((uint32_t)c->cu_count * per_cu, (n_lane + 63) / 64);
    // the lane arenas hold no row table: the traceback stream carries the
    // rows' trailers (the retry slots keep theirs)
    const uint32_t flat_cap = wide ? kOnePassWideFlat : kOnePassFlat, rows_cap = 0,
                   ops_cap = wide ? kOnePassWideOps : kOnePassOps;
    const uint64_t stride = onepass_stride(flat_cap, rows_cap, ops_cap);
    uint8_t* arena = n_lane ? c->b_arena.get<uint8_t>((size_t)blocks * 64 * stride) : nullptr;
    const uint32_t wrows = wide ? kWaveWideRows : kWaveRows, wflat = wide ? kWaveWideFlat : kWaveFlat,
                   wops = wide ? kWaveWideOps : kWaveOps;
    const uint32_t wblocks =
        std::min<uint32_t>((uint32_t)c->cu_count * (wide ? kWaveWidePerCu : kWavePerCu), (n_wave + 3) / 4);
    uint8_t* warena =
        n_wave ? c->b_arena_wave.get<uint8_t>((size_t)wblocks * 4 * onepass_stride(wflat, wrows, wops)) : nullptr;
    uint32_t* pool = c->b_oppool.get<uint32_t>(kOnePassPool);
    // [0] ops used, [1] work queue, [2] retry slots taken (low 32 bits), [3] wave queue
    unsigned long long* pcount = c->b_opcount.get<unsigned long long>(4);
    hip_check(hipMemsetAsync(pcount, 0, 4 * sizeof(unsigned long long), c->st), "clear op count");
    OnePassBig big{};
    big.ring = c->b_ring_big.get<int2>((size_t)kBigSlots * ring_stride(kRingBigLog2));
    big.arena = c->b_arena_big.get<uint8_t>((size_t)kBigSlots * onepass_stride(kBigFlat, kBigRows, kBigOps));
    big.next = reinterpret_cast<unsigned int*>(pcount + 2);
    big.ring_log2 = kRingBigLog2;
    big.slots = retry_slots();
    big.flat_cap = kBigFlat;
    big.rows_cap = kBigRows;
    big.ops_cap = kBigOps;
    hip_check(hipMemcpyAsync(dt, sorted.data() + base, n * sizeof(DpTask), hipMemcpyHostToDevice, c->st),
              "copy tasks");
    timed_launch(c, 14, [&] {
      hip_check(launch_dp_wave((int)wblocks, dt, n_wave, c->d_mat, c->P.gap_open, c->P.gap_extend, warena, wrows,
                               wflat, wops, pool, pcount, kOnePassPool, dr, pcount + 3, ctr, wide, c->st),
                "k_dp_wave");
      if (n_lane)
        hip_check(launch_dp_onepass((int)blocks, dt + n_wave, n_lane, c->d_mat, c->P.gap_open, c->P.gap_extend, arena,
                                    flat_cap, rows_cap, ops_cap, pool, pcount, kOnePassPool, dr + n_wave, pcount + 1,
                                    big, ctr, wide, c->st),
                  "k_dp_onepass");
    });
    std::vector<OnePassResult> r(n);
    unsigned long long used = 0;
    hip_check(hipMemcpyAsync(r.data(), dr, n * sizeof(OnePassResult), hipMemcpyDeviceToHost, c->st), "copy onepass");
    hip_check(hipMemcpyAsync(&used, pcount, sizeof(used), hipMemcpyDeviceToHost, c->st), "copy op count");
    hip_check(hipStreamSynchronize(c->st), "onepass sync");
    used = std::min(used, kOnePassPool);
    const uint64_t pool_at = op_pool.size();
    op_pool.resize(pool_at + used);
    if (used) {
      hip_check(hipMemcpyAsync(op_pool.data() + pool_at, pool, used * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st),
                "copy ops");
      hip_check(hipStreamSynchronize(c->st), "ops sync");
    }
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = ord[base + k];
      if (r[k].flags) {
        fallback.push_back(i);
        continue;
      }
      out[i] = HalfTb{r[k].score, r[k].a_len, r[k].b_len, r[k].n_ops, pool_at + r[k].ops_at};
    }
  }
  const uint64_t t_one = tt.us();
  if (!fallback.empty()) {
    std::vector<DpTask> ft(fallback.size());
    std::vector<uint64_t> fw;
    for (size_t k = 0; k < fallback.size(); ++k) ft[k] = tasks[fallback[k]];
    if (work) {
      fw.resize(fallback.size());
      for (size_t k = 0; k < fallback.size(); ++k) fw[k] = (*work)[fallback[k]];
    }
    std::vector<HalfTb> fo;
    std::vector<uint32_t> fpool;
    run_dp_traceback_2pass(c, ft, fo, fpool, work ? &fw : nullptr);
    const uint64_t at = op_pool.size();
    op_pool.insert(op_pool.end(), fpool.begin(), fpool.end());
    for (size_t k = 0; k < fallback.size(); ++k) {
      fo[k].ops_at += at;
      out[fallback[k]] = fo[k];
    }
  }
  if (std::getenv("RMB_STAGE_TIMES"))
    std::fprintf(stderr,
                 "[rmb] traceback DPs ms: one pass %.1f (%zu halves, %zu wave-run)  two-pass fallback %.1f (%zu "
                 "halves)\n",
                 t_one / 1e3, tasks.size(), n_wave_all, (tt.us() - t_one) / 1e3, fallback.size());
}

struct ItemState {
  uint32_t item = 0;
  std::vector<uint32_t> surv;    // indices of the item's surviving hits
  std::vector<PreRec> pre;
  std::vector<FinalRec> fin;
};

// ---- Phase 1 and the preliminary DPs, queued per batch of search records ----
//
// A batch runs on the context's first stream as one queue of kernels per
// lookup-width group: table build, k_phase1, and the two preliminary-DP
// passes, which read their task counts from device memory, so nothing waits
// on the host.  The host reads a batch back while the next batch runs, and
// meanwhile replays and aligns the batch before (Phase 2a/2b on the second
// stream and the host threads).

struct P1Group {
  int lw = 0;
  std::vector<uint32_t> recs_idx;  // search records
  std::vector<uint32_t> items;     // group-local items to search (listed pairs); unused when !listed
  bool listed = false;
  uint64_t hits_cap = 0;
  // host copies that queued transfers read
  std::vector<RecDesc> recs;
  std::vector<uint64_t> elem_base;
};

struct P1Batch {
  int slot = 0;
  std::vector<P1Group> groups;
  uint64_t redo_cap = 0;
  uint32_t q0 = 0, q1 = 0;  // the whole queries of its records
};

// Lookup tables of a group's records (blast/nalookup.rs, blast/mblookup.rs):
// one (record, word, offset) key per window of the two-strand query buffer,
// radix-sorted, then per (record, word) cell the start of its offsets:
// cells [R x 4^lw + 1] and entries (query offsets in the two-strand buffer,
// the cell's order), both device arrays of the context.
struct GroupLut {
  const uint32_t* cells;
  const uint32_t* entries;
  uint64_t n_entries;  // windows, valid or not (the valid ones come first)
};
static GroupLut build_group_lut(rmb_ctx* c, hipStream_t st, P1Group& g, const uint8_t* d_qext) {
  const int ws = c->P.word_size, lw = g.lw;
  const uint32_t R = (uint32_t)g.recs_idx.size();
  const uint64_t cells = 1ull << (2 * lw);
  const uint64_t E = g.elem_base[R];
  RecDesc* drecs = c->b_recs.get<RecDesc>(R);
  uint64_t* delem = c->b_elem.get<uint64_t>(R + 1);
  hip_check(hipMemcpyAsync(drecs, g.recs.data(), R * sizeof(RecDesc), hipMemcpyHostToDevice, st), "copy recs");
  hip_check(hipMemcpyAsync(delem, g.elem_base.data(), (R + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st),
            "copy elem");
  uint64_t* keys = c->b_keys.get<uint64_t>(E);
  uint64_t* keys2 = c->b_keys2.get<uint64_t>(E);
  uint32_t* vals = c->b_vals.get<uint32_t>(E);
  uint32_t* vals2 = c->b_vals2.get<uint32_t>(E);
  const uint64_t ncell = (uint64_t)R * cells + 1;
  uint32_t* counts = c->b_counts.get<uint32_t>(ncell);
  uint32_t* cstart = c->b_cells.get<uint32_t>(ncell);
  hip_check(hipMemsetAsync(counts, 0, ncell * sizeof(uint32_t), st), "memset counts");
  const uint64_t invalid = (uint64_t)R << (2 * lw);
  int end_bit = 2 * lw;
  while ((1ull << (end_bit - 2 * lw)) <= R) ++end_bit;
  const int desc = lw > 8 ? 1 : 0;
  hip_check(launch_lut_keys(drecs, R, delem, d_qext, lw, ws, desc, keys, vals, counts, invalid, E, st), "lut keys");
  size_t tmp1 = 0, tmp2 = 0;
  hip_check(sort_pairs_u64(nullptr, tmp1, keys, keys2, vals, vals2, E, end_bit, st), "sort size");
  hip_check(exclusive_sum_u32(nullptr, tmp2, counts, cstart, ncell, st), "scan size");
  void* tmp = c->b_sort_tmp.get<uint8_t>(std::max(tmp1, tmp2));
  hip_check(sort_pairs_u64(tmp, tmp1, keys, keys2, vals, vals2, E, end_bit, st), "radix sort");
  hip_check(exclusive_sum_u32(tmp, tmp2, counts, cstart, ncell, st), "scan");
  return GroupLut{cstart, vals2, E};
}

// The reverse complement of every sentinel-wrapped segment (base, bases) of
// the device pool `in` into `out`, on stream st (k_revcomp_segments).
static void device_revcomp(rmb_ctx* c, hipStream_t st, const uint8_t* in, uint8_t* out,
                           const std::vector<uint64_t>& base, const std::vector<uint32_t>& len) {
  const uint32_t chunk = revcomp_chunk();
  std::vector<uint2> work;
  for (uint32_t s = 0; s < (uint32_t)base.size(); ++s)
    for (uint64_t q = 0; q < (uint64_t)len[s] + 2; q += chunk) work.push_back(make_uint2(s, (uint32_t)q));
  if (work.empty()) return;
  uint64_t* db = c->b_seg_base.get<uint64_t>(base.size());
  uint32_t* dl = c->b_seg_len.get<uint32_t>(len.size());
  uint2* dw = c->b_seg_work.get<uint2>(work.size());
  hip_check(hipMemcpyAsync(db, base.data(), base.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st), "copy segs");
  hip_check(hipMemcpyAsync(dl, len.data(), len.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st), "copy segs");
  hip_check(hipMemcpyAsync(dw, work.data(), work.size() * sizeof(uint2), hipMemcpyHostToDevice, st), "copy work");
  hip_check(launch_revcomp_segments(in, out, db, dl, dw, (uint32_t)work.size(), st), "k_revcomp_segments");
  // the host vectors are read by the queued copies: wait for them
  hip_check(hipStreamSynchronize(st), "revcomp sync");
}

// Per-group seeding launch: lookup table for the group's records, then
// k_phase1 writing up to `cap` hits to dh (count at *hitcount); with `seeds`,
// the seeds the diagonal hash lets through instead of the hits.
static void launch_group_phase1(rmb_ctx* c, hipStream_t st, P1Group& g, const uint8_t* d_qfwd, const uint8_t* d_qrc,
                                const uint8_t* d_qext, UngHit* dh, uint64_t cap, unsigned long long* hitcount,
                                unsigned long long* ctr, uint32_t* err, hipEvent_t ev_a, hipEvent_t ev_b,
                                bool seeds = false) {
  const int ws = c->P.word_size, lw = g.lw;
  const uint32_t R = (uint32_t)g.recs_idx.size();
  const GroupLut lut = build_group_lut(c, st, g, d_qext);
  RecDesc* drecs = c->b_recs.get<RecDesc>(R);
  const uint32_t nsub = (uint32_t)c->subj.size();
  P1Params P{};
  P.word_size = ws;
  P.lut_width = lw;
  P.scan_step = ws - lw + 1;
  P.xdrop_ungap = c->P.xdrop_ungap;
  P.n_subj = nsub;
  P.n_items = (uint64_t)R * nsub;
  P.item_list = nullptr;
  P.out_cap = cap;
  bool pos_match = true;
  for (int b = 0; b < 4; ++b) pos_match = pos_match && c->M.s[b][b] > 0;
  P.known_word = ws >= 7 && lw == ws && pos_match ? 1 : 0;
  if (g.listed) {  // group-local item ids (record index in the group * nsub + subject)
    uint32_t* dl = c->b_items.get<uint32_t>(g.items.size());
    hip_check(hipMemcpyAsync(dl, g.items.data(), g.items.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
              "copy items");
    P.item_list = dl;
    P.n_items = g.items.size();
  }
  // A search of few items would leave most of the GPU idle, each item's seeds
  // replayed through the diagonal hash by one workgroup: its items are then
  // split over up to 16 workgroups, each owning an equal range of the 512
  // hash buckets (the buckets are independent, so the split is exact), as
  // far as the resident workgroups allow; items fewer than four times the
  // resident workgroups are split in two, so the last round of items is not
  // a few workgroups' (one 10 Mbp record as ten chunks x 300 families, 3,000
  // items: 3% off the step; four or eight parts cost more than they balance).
  const uint64_t slots = (uint64_t)c->cu_count * kP1PerCu;
  while (!seeds && P.part_log2 < 4 && (P.n_items << (P.part_log2 + 1)) <= slots) ++P.part_log2;
  if (!seeds && P.part_log2 == 0 && P.n_items < 4 * slots) P.part_log2 = 1;
  if (const char* e = std::getenv("RMB_P1_PARTS_LOG2")) P.part_log2 = seeds ? 0u : (uint32_t)std::min(4, std::max(0, std::atoi(e)));
  // ~30 KB of LDS per workgroup: kP1PerCu resident per CU.  Workgroups are
  // dealt round-robin over the XCDs, so blockIdx % kXcds names the XCD.
  const int blocks = std::max(
      1, std::min<int>(c->cu_count * kP1PerCu, (int)std::min<uint64_t>(P.n_items << P.part_log2, 1u << 20)));
  uint2* spill = c->b_spill.get<uint2>((size_t)blocks * kBuckets * kHashGlobal);
  uint4* hitres = c->b_hitres.get<uint4>((size_t)blocks * kSeedCap);
  unsigned long long* queue = c->b_queue.get<unsigned long long>(kXcds);
  hip_check(hipMemsetAsync(queue, 0, kXcds * sizeof(unsigned long long), st), "memset queue");
  if (ev_a) hip_check(hipEventRecord(ev_a, st), "event");
  hip_check(launch_phase1(blocks, P, drecs, (const SubjDesc*)c->d_sdesc.p, d_qfwd, d_qrc, d_qext,
                          (const uint8_t*)c->d_splus.p, (const uint8_t*)c->d_src.p, lut.cells, lut.entries, c->d_mat,
                          dh, hitcount, spill, hitres, queue, ctr, err, seeds, st),
            "k_phase1");
  if (ev_b) hip_check(hipEventRecord(ev_b, st), "event");
}

// Estimated hit capacity for a group: hits per searched item seen so far.
static uint64_t hits_estimate(const rmb_ctx* c, const P1Group& g) {
  const uint64_t n_items = g.listed ? g.items.size() : (uint64_t)g.recs_idx.size() * c->subj.size();
  const double per = c->hits_per_item > 0 ? c->hits_per_item : 0.0;
  return std::max<uint64_t>(1u << 20, (uint64_t)(per * 1.25 * (double)n_items) + (1u << 16));
}

static void ensure_events(std::vector<hipEvent_t>& ev, size_t n) {
  while (ev.size() < n) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    ev.push_back(e);
  }
}

// Queue a batch on the seeding stream.
static void enqueue_batch(rmb_ctx* c, P1Batch& b, const uint8_t* d_qfwd, const uint8_t* d_qrc,
                          const uint8_t* d_qext) {
  hipStream_t st = c->st1;
  P1Slot& S = c->slot[b.slot];
  const size_t ng = b.groups.size();
  uint64_t max_cap = 0, pre_cap = 0;
  for (auto& g : b.groups) {
    if (!g.hits_cap) g.hits_cap = hits_estimate(c, g);
    max_cap = std::max(max_cap, g.hits_cap);
    pre_cap += g.hits_cap;
  }
  if (!b.redo_cap) b.redo_cap = std::max<uint64_t>(4096, pre_cap / 64);
  if (max_cap > 0xFFFFFFFFull) throw Failure(RMB_ERR_CAPACITY, "more than 2^32 ungapped hits in one group");
  PreHit* dpre = S.pre.get<PreHit>(pre_cap);
  UngHit* dredoh = S.redo_hits.get<UngHit>(b.redo_cap);
  // [0] prelims, [2] wide hits, [4 + 2g] and [5 + 2g] group g's first and
  // global-ring preliminary passes' work queues, [4 + 2ng + 2g] its wave-run
  // pass's queue and [5 + 2ng + 2g] the hits that pass hands on
  unsigned long long* cnt = S.cnt.get<unsigned long long>(4 + 4 * ng);
  unsigned long long* hitcnt = S.hitcnt.get<unsigned long long>(2 * ng + 2);  // per group: hits, wide (pass 1)
  unsigned long long* ctr = S.ctr.get<unsigned long long>(C_NCOUNTERS);
  uint32_t* err = S.err.get<uint32_t>(1);
  hip_check(hipMemsetAsync(cnt, 0, (4 + 4 * ng) * sizeof(unsigned long long), st), "memset counts");
  hip_check(hipMemsetAsync(hitcnt, 0, (2 * ng + 2) * sizeof(unsigned long long), st), "memset hit counts");
  hip_check(hipMemsetAsync(ctr, 0, C_NCOUNTERS * sizeof(unsigned long long), st), "memset counters");
  hip_check(hipMemsetAsync(err, 0, sizeof(uint32_t), st), "memset err");
  ensure_events(S.ev, 6 * ng);
  UngHit* dh = c->b_hits.get<UngHit>(max_cap);
  uint32_t* dredo = c->b_redo.get<uint32_t>(max_cap);
  uint32_t* dredo2 = c->b_redo2.get<uint32_t>(max_cap);
  const uint32_t nsub = (uint32_t)c->subj.size();
  for (size_t gi = 0; gi < ng; ++gi) {
    P1Group& g = b.groups[gi];
    if (g.listed && g.items.empty()) continue;
    launch_group_phase1(c, st, g, d_qfwd, d_qrc, d_qext, dh, g.hits_cap, hitcnt + 2 * gi, ctr, err, S.ev[6 * gi],
                        S.ev[6 * gi + 1]);
    const RecDesc* drecs = (const RecDesc*)c->b_recs.p;
    uint32_t* dids = c->b_recids.get<uint32_t>(g.recs_idx.size());
    hip_check(hipMemcpyAsync(dids, g.recs_idx.data(), g.recs_idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                             st),
              "copy record ids");
    PrelimParams PP{};
    PP.xdrop_gap = c->P.xdrop_gap;
    PP.gap_open = c->P.gap_open;
    PP.gap_extend = c->P.gap_extend;
    PP.cutoff = std::max(c->P.min_raw_gapped_score, 1);
    PP.lut_width = g.lw;
    PP.n_subj = nsub;
    PP.out_cap = pre_cap;
    PP.rec_ids = dids;
    // pass 1: 128-lane workgroups with a 16 KB LDS band ring each, eight per CU
    PP.n_hits = g.hits_cap;
    PP.n_dev = hitcnt + 2 * gi;
    PP.redo_cap = g.hits_cap;
    PP.next = cnt + 4 + 2 * gi;
    hip_check(hipEventRecord(S.ev[6 * gi + 2], st), "event");
    // past 12 band cells ((X - open) / extend to either side), 64-cell rings
    // of 4-byte cells (12-bit H and E, the column's base code) in 64-lane
    // workgroups, nine per CU (RMB_PRELIM_WIDE=0 keeps the 16-cell rings)
    const char* pw_env = std::getenv("RMB_PRELIM_WIDE");
    const bool pwide = c->tune.prelim_ring ? c->tune.prelim_ring == 64
                                           : !(pw_env && std::atoi(pw_env) == 0) &&
                                                 2 * std::max(0, (int)c->P.xdrop_gap - (int)c->P.gap_open) /
                                                         std::max(1, (int)c->P.gap_extend) + 1 > 12;
    const int pre_per_cu = c->tune.prelim_wg_per_cu ? c->tune.prelim_wg_per_cu : (pwide ? 9 : 8);
    hip_check(launch_prelim_fused(c->cu_count * pre_per_cu, true, PP, dh, nullptr, g.hits_cap, drecs,
                                  (const SubjDesc*)c->d_sdesc.p, d_qfwd, d_qrc, (const uint8_t*)c->d_splus.p,
                                  c->d_mat, nullptr, 0, dpre, cnt, dredo, nullptr, hitcnt + 2 * gi + 1, ctr, err, pwide, st),
              "k_prelim_fused");
    hip_check(hipEventRecord(S.ev[6 * gi + 3], st), "event");
    // pass 2: the hits whose band outgrew the LDS ring.  Few of them run one
    // per wave (bands of up to 256 cells); the rest, and wider bands, with
    // per-lane rings of 2^kRingLog2 cells in global memory; wider still go
    // back as records
    hip_check(hipEventRecord(S.ev[6 * gi + 4], st), "event");
    {
      PrelimParams PW = PP;
      PW.n_dev = hitcnt + 2 * gi + 1;
      PW.redo_cap = g.hits_cap;
      PW.next = cnt + 4 + 2 * ng + 2 * gi;
      hip_check(launch_prelim_wave((int)c->cu_count * kPrelimWaveWgPerCu, PW, dh, dredo, prelim_wave_hits(c), drecs,
                                   (const SubjDesc*)c->d_sdesc.p, d_qfwd, d_qrc, (const uint8_t*)c->d_splus.p,
                                   c->d_mat, dpre, cnt, dredo2, cnt + 5 + 2 * ng + 2 * gi, ctr, st),
                "k_prelim_wave");
    }
    const uint64_t lanes = (uint64_t)c->cu_count * 4 * 128;
    int2* ring = c->b_ring.get<int2>((size_t)lanes * ring_stride(kRingLog2));
    PP.n_dev = cnt + 5 + 2 * ng + 2 * gi;
    PP.redo_cap = b.redo_cap;
    PP.next = cnt + 5 + 2 * gi;
    hip_check(launch_prelim_fused((int)(lanes / 128), false, PP, dh, dredo2, g.hits_cap, drecs,
                                  (const SubjDesc*)c->d_sdesc.p, d_qfwd, d_qrc, (const uint8_t*)c->d_splus.p,
                                  c->d_mat, ring, kRingLog2, dpre, cnt, nullptr, dredoh, cnt + 2, ctr, err, false, st),
              "k_prelim_fused wide");
    hip_check(hipEventRecord(S.ev[6 * gi + 5], st), "event");
  }
  if (!S.done) hip_check(hipEventCreate(&S.done), "hipEventCreate");
  hip_check(hipEventRecord(S.done, st), "event");
}

// Wait for a queued batch and read it back: the prelims (item ids naming
// search records) are appended to `out`.  A group whose hits outgrew their
// buffer, or wide hits beyond theirs, rerun the batch with larger buffers.
static void complete_batch(rmb_ctx* c, P1Batch& b, const std::vector<RecDesc>& recs, const uint8_t* d_qfwd,
                           const uint8_t* d_qrc, const uint8_t* d_qext, std::vector<PreHit>& out,
                           uint64_t p1_ctr[C_NCOUNTERS]) {
  P1Slot& S = c->slot[b.slot];
  const size_t ng = b.groups.size();
  for (int attempt = 0;; ++attempt) {
    // Reads go through the Phase-2b stream after the batch's own event: a
    // copy on the null stream would also wait for the batch queued behind.
    hip_check(hipEventSynchronize(S.done), "batch sync");
    std::vector<unsigned long long> hc(2 * ng + 2), cnt(4), ctr(C_NCOUNTERS);
    uint32_t e = 0;
    hip_check(hipMemcpyAsync(hc.data(), S.hitcnt.p, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->st),
              "copy hit counts");
    hip_check(hipMemcpyAsync(cnt.data(), S.cnt.p, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->st),
              "copy counts");
    hip_check(hipMemcpyAsync(ctr.data(), S.ctr.p, ctr.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->st),
              "copy counters");
    hip_check(hipMemcpyAsync(&e, S.err.p, sizeof e, hipMemcpyDeviceToHost, c->st), "copy err");
    hip_check(hipStreamSynchronize(c->st), "copy sync");
    if (e & E_HASH_FULL) throw Failure(RMB_ERR_CAPACITY, "diagonal hash bucket exceeded 256 cells");
    bool rerun = false;
    for (size_t gi = 0; gi < ng; ++gi) {
      P1Group& g = b.groups[gi];
      const uint64_t n_items = g.listed ? g.items.size() : (uint64_t)g.recs_idx.size() * c->subj.size();
      if (n_items) c->hits_per_item = std::max(c->hits_per_item, (double)hc[2 * gi] / (double)n_items);
      if (hc[2 * gi] > g.hits_cap) {
        g.hits_cap = hc[2 * gi] + hc[2 * gi] / 4;
        rerun = true;
      }
    }
    if (cnt[2] > b.redo_cap) {
      b.redo_cap = cnt[2] + cnt[2] / 4;
      rerun = true;
    }
    if (rerun) {
      if (attempt >= 2) throw Failure(RMB_ERR_CAPACITY, "ungapped hit buffer could not be sized");
      enqueue_batch(c, b, d_qfwd, d_qrc, d_qext);
      continue;
    }
    for (int k = 0; k < C_NCOUNTERS; ++k) p1_ctr[k] += ctr[k];
    for (size_t gi = 0; gi < ng; ++gi) {
      const P1Group& g = b.groups[gi];
      if (g.listed && g.items.empty()) continue;
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, S.ev[6 * gi], S.ev[6 * gi + 1]), "event time");
      c->stats[13] += (uint64_t)((double)ms * 1e6);
      c->stats[15] += 1;
      float pms[2] = {0, 0};
      for (int p = 0; p < 2; ++p) {
        hip_check(hipEventElapsedTime(&ms, S.ev[6 * gi + 2 + 2 * p], S.ev[6 * gi + 3 + 2 * p]), "event time");
        c->stats[14] += (uint64_t)((double)ms * 1e6);
        c->stats[23] += (uint64_t)((double)ms * 1e6);
        pms[p] = ms;
      }
      if (std::getenv("RMB_STAGE_TIMES"))
        std::fprintf(stderr, "[rmb] preliminary DPs ms: %.1f (%llu hits), wider bands %.1f (%llu hits)\n", pms[0],
                     hc[2 * gi], pms[1], hc[2 * gi + 1]);
    }
    const size_t base = out.size();
    out.resize(base + cnt[0]);
    hip_check(hipMemcpyAsync(out.data() + base, S.pre.p, cnt[0] * sizeof(PreHit), hipMemcpyDeviceToHost, c->st),
              "copy prelims");
    std::vector<UngHit> redo(cnt[2]);
    if (cnt[2])
      hip_check(hipMemcpyAsync(redo.data(), S.redo_hits.p, cnt[2] * sizeof(UngHit), hipMemcpyDeviceToHost, c->st),
                "copy redo");
    hip_check(hipStreamSynchronize(c->st), "copy sync");
    if (!cnt[2]) return;
    // bands wider than 2^kRingLog2 cells: score-only DPs with wider rings
    c->host_prelim_dps += redo.size();
    const uint32_t nsub = (uint32_t)c->subj.size();
    const int32_t cutoff = std::max(c->P.min_raw_gapped_score, 1);
    std::vector<DpTask> tasks;
    std::vector<std::array<uint32_t, 4>> piv;
    for (const UngHit& h : redo) {
      const uint32_t r = h.item / nsub, f = h.item % nsub;
      const uint32_t ql = recs[r].q_len, sl = c->sdesc[f].s_len;
      UngRec u{h.q_seed, h.s_seed, h.q_start, h.q_end, h.s_start, h.s_end, h.score, (uint8_t)h.strand};
      uint32_t pq, ps, tq, ts;
      pivot_of(u, ql, sl, (uint32_t)lut_width_for(c->P.word_size, 2ull * ql), pq, ps, tq, ts);
      piv.push_back({pq, ps, tq, ts});
      const uint8_t* qa = (h.strand ? d_qrc : d_qfwd) + recs[r].q_base + 1;
      const uint8_t* sa = (const uint8_t*)c->d_splus.p + c->sdesc[f].s_base + 1;
      const int32_t xd = std::min(c->P.xdrop_gap, h.score);
      tasks.push_back(DpTask{sa, qa, ps, pq, xd, 1});
      tasks.push_back(DpTask{sa + ps, qa + pq, sl > ps ? sl - ps : 0, ql > pq ? ql - pq : 0, xd, 0});
    }
    std::vector<DpResult> res;
    run_dp_score(c, tasks, res);
    for (size_t k = 0; k < redo.size(); ++k) {
      const UngHit& h = redo[k];
      const DpResult &L = res[2 * k], &R = res[2 * k + 1];
      const int32_t sc = L.score + R.score;
      if (sc < 0 || sc < cutoff) continue;
      const uint32_t r = h.item / nsub, f = h.item % nsub;
      const uint32_t ql = recs[r].q_len, sl = c->sdesc[f].s_len;
      const auto& pv = piv[k];
      const uint32_t q0 = pv[0] - L.b_len, q1 = pv[0] + R.b_len, s0 = pv[1] - L.a_len, s1 = pv[1] + R.a_len;
      PreHit p;
      p.u = h;
      p.score = sc;
      if (h.strand) { p.q_start = ql - q1; p.q_end = ql - q0; p.s_start = sl - s1; p.s_end = sl - s0; }
      else { p.q_start = q0; p.q_end = q1; p.s_start = s0; p.s_end = s1; }
      p.tq = pv[2];
      p.ts = pv[3];
      out.push_back(p);
    }
    return;
  }
}

// Lookup-width groups of a batch's search records [k0, k1), bounded by the
// cell table size and the entry count; with `pairs`, only the listed items.
static std::vector<P1Group> group_records(const rmb_ctx* c, const std::vector<RecDesc>& recs,
                                          const std::vector<uint32_t>& first, uint32_t k0, uint32_t k1,
                                          const std::vector<uint64_t>* pairs) {
  const int ws = c->P.word_size;
  const uint32_t nsub = (uint32_t)c->subj.size();
  std::vector<P1Group> out;
  auto flush = [&](std::vector<uint32_t>& grp, int lw) {
    if (grp.empty()) return;
    P1Group g;
    g.lw = lw;
    g.recs_idx = grp;
    const uint32_t R = (uint32_t)grp.size();
    const uint64_t cells = 1ull << (2 * lw);
    g.recs.resize(R);
    g.elem_base.assign(R + 1, 0);
    for (uint32_t i = 0; i < R; ++i) {
      g.recs[i] = recs[grp[i]];
      g.recs[i].cell_base = (uint64_t)i * cells;
      g.elem_base[i + 1] = g.elem_base[i] + 2ull * g.recs[i].q_len + 1;
    }
    if (pairs) {  // pairs name queries: search every record of the query
      g.listed = true;
      std::unordered_map<uint32_t, uint32_t> pos;
      for (uint32_t k = 0; k < R; ++k) pos[grp[k]] = k;
      for (uint64_t it : *pairs) {
        const uint32_t q = (uint32_t)(it / nsub), f = (uint32_t)(it % nsub);
        for (uint32_t k = first[q]; k < first[q + 1]; ++k) {
          auto p = pos.find(k);
          if (p != pos.end()) g.items.push_back(p->second * nsub + f);
        }
      }
      std::sort(g.items.begin(), g.items.end());
    }
    out.push_back(std::move(g));
    grp.clear();
  };
  for (int lw = 4; lw <= 12; ++lw) {
    std::vector<uint32_t> grp;
    const uint64_t cells = 1ull << (2 * lw);
    uint64_t elems = 0;
    for (uint32_t i = k0; i < k1; ++i) {
      if (lut_width_for(ws, 2ull * recs[i].q_len) != lw) continue;
      if (!grp.empty() && ((grp.size() + 1) * cells > (1ull << 28) ||
                           elems + 2ull * recs[i].q_len + 1 > (1ull << 31))) {
        flush(grp, lw);
        elems = 0;
      }
      grp.push_back(i);
      elems += 2ull * recs[i].q_len + 1;
    }
    flush(grp, lw);
  }
  return out;
}

// The records a query is searched as through Phase 2a: the whole query, or
// the chunks of NCBI's split (rmblastn/src/main.rs:1238-1300).
struct SearchRecs {
  std::vector<RecDesc> recs;        // search records, grouped by query in order
  std::vector<uint32_t> query;      // query of each search record
  std::vector<uint32_t> offset;     // its first base in the query
  std::vector<uint32_t> first;      // search records of query i: [first[i], first[i + 1])
  std::vector<uint64_t> whole_base; // pool offset of each whole query (Phase 2b)
  uint64_t pool = 0;                // pool bytes
};

// DUST-masked copy of the query pool (filter/dust.rs:462-484: masked bases set
// to 14), uploaded to the context's masked pool.  A split query is masked
// whole, with one more base after each interval (NCBI's chunked-interval
// off-by-one, dust.rs:603-631), and its chunks are cut from that mask.
static const uint8_t* upload_dusted(rmb_ctx* c, const std::vector<SeqView>& qf,
                                    const SearchRecs& sr) {
  std::vector<uint8_t> pm(sr.pool, 15);
  parallel_for(qf.size(), [&](size_t i) {
    const uint32_t n = (uint32_t)qf[i].size() - 2;
    const uint32_t k0 = sr.first[i], k1 = sr.first[i + 1];
    if (k1 - k0 == 1) {
      uint8_t* dst = pm.data() + sr.recs[k0].q_base;
      std::memcpy(dst, qf[i].data(), qf[i].size());
      if (n) dust_mask(dst + 1, n);
      return;
    }
    std::vector<uint8_t> m(qf[i].begin() + 1, qf[i].end() - 1);
    for (const auto& iv : dust_mask(m.data(), n))
      if (iv.second + 1 < n) m[iv.second + 1] = 14;
    for (uint32_t k = k0; k < k1; ++k)
      std::memcpy(pm.data() + sr.recs[k].q_base + 1, m.data() + sr.offset[k], sr.recs[k].q_len);
  });
  uint8_t* d = c->b_qext.get<uint8_t>(sr.pool);
  hip_check(hipMemcpy(d, pm.data(), sr.pool, hipMemcpyHostToDevice), "copy masked query");
  return d;
}

// Phase 2a of one batch: ordered replay of the hits whose preliminary DP
// reached the cutoff, in NCBI's ungapped-hit order per item
// (engine.rs:1280-1411), then split queries' chunks folded together.
static std::vector<ItemState> phase2a(rmb_ctx* c, const SearchRecs& sr, std::vector<PreHit>& survivors) {
  const uint32_t nsub = (uint32_t)c->subj.size();
  const std::vector<RecDesc>& recs = sr.recs;
  const uint32_t nsr = (uint32_t)recs.size(), nq = (uint32_t)sr.whole_base.size();
  std::vector<ItemState> items;
  {
    std::vector<uint64_t> keyed(survivors.size());
    for (size_t i = 0; i < survivors.size(); ++i) keyed[i] = ((uint64_t)survivors[i].u.item << 32) | i;
    std::sort(keyed.begin(), keyed.end());
    for (size_t i = 0; i < keyed.size();) {
      const uint32_t it = (uint32_t)(keyed[i] >> 32);
      ItemState s;
      s.item = it;
      size_t j = i;
      for (; j < keyed.size() && (uint32_t)(keyed[j] >> 32) == it; ++j) s.surv.push_back((uint32_t)keyed[j]);
      items.push_back(std::move(s));
      i = j;
    }
  }
  parallel_for(items.size(), [&](size_t k) {
    ItemState& s = items[k];
    const uint32_t ql = recs[s.item / nsub].q_len, sl = c->sdesc[s.item % nsub].s_len;
    std::vector<UngRec> u(s.surv.size());
    for (size_t i = 0; i < s.surv.size(); ++i) {
      const UngHit& h = survivors[s.surv[i]].u;
      u[i] = UngRec{h.q_seed, h.s_seed, h.q_start, h.q_end, h.s_start, h.s_end, h.score, (uint8_t)h.strand};
    }
    std::vector<uint32_t> ord(u.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (uint32_t)i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return ung_before(u[a], u[b], ql, sl); });
    HspTree tp(ql + 1, sl + 1), tm(ql + 1, sl + 1);
    for (uint32_t i : ord) {
      const UngRec& x = u[i];
      const uint32_t css = x.minus ? sl - x.s_end : x.s_start, cse = x.minus ? sl - x.s_start : x.s_end;
      HspTree& t = x.minus ? tm : tp;
      if (t.covers(x.q_start, x.q_end, css, cse, x.score, !x.minus)) continue;
      const PreHit& p = survivors[s.surv[i]];
      const uint32_t nss = x.minus ? sl - p.s_end : p.s_start, nse = x.minus ? sl - p.s_start : p.s_end;
      t.add(p.q_start, p.q_end, nss, nse, p.score, !x.minus);
      s.pre.push_back(PreRec{p.tq, p.ts, p.q_start, p.q_end, p.s_start, p.s_end, p.score, x.minus});
    }
    purge_prelims(s.pre, sl);
  });
  survivors.clear();
  survivors.shrink_to_fit();
  // ---- items from search records to queries: a split query's prelims move
  // to query coordinates and fold chunk by chunk into those of the chunks
  // before (main.rs:1270-1300; items are in search-record order, so the
  // chunks of a query arrive in order)
  if (nsr != nq) {
    std::vector<ItemState> merged;
    merged.reserve(items.size());
    for (size_t k = 0; k < items.size();) {
      const uint32_t q = sr.query[items[k].item / nsub];
      if (sr.first[q + 1] - sr.first[q] == 1) {
        items[k].item = q * nsub + items[k].item % nsub;
        merged.push_back(std::move(items[k++]));
        continue;
      }
      std::map<uint32_t, ItemState> acc;  // by subject
      for (; k < items.size() && sr.query[items[k].item / nsub] == q; ++k) {
        ItemState& s = items[k];
        const uint32_t chunk = s.item / nsub, f = s.item % nsub, off = sr.offset[chunk];
        for (PreRec& h : s.pre) {
          h.q_seed += off;
          h.q_start += off;
          h.q_end += off;
        }
        ItemState& a = acc[f];
        a.item = q * nsub + f;
        if (chunk == sr.first[q]) a.pre = std::move(s.pre);
        else merge_chunk_prelims(a.pre, std::move(s.pre), off);
      }
      for (auto& kv : acc) merged.push_back(std::move(kv.second));
    }
    items = std::move(merged);
  }
  return items;
}

// Phase 2b of one batch: traceback DPs on the GPU, then the per-subject
// replay on the host (engine.rs:1438-1762).
static void phase2b(rmb_ctx* c, const SearchRecs& sr, const std::vector<SeqView>& qf,
                    const uint32_t* lens, std::vector<ItemState>& items, const uint8_t* d_qfwd, const uint8_t* d_qrc,
                    uint64_t tb_ms[4]) {
  const uint32_t nsub = (uint32_t)c->subj.size();
  const int32_t cutoff = std::max(c->P.min_raw_gapped_score, 1);
  Timer t3;
  uint64_t tb_marks[4] = {0, 0, 0, 0};  // sort, seeds + tasks, DP on the GPU, replay (us since t3)
  if (c->tap_pre_on)  // the prelims tap: Phase 2a's output, before the traceback order
    for (const ItemState& s : items) c->tap_pre.insert(c->tap_pre.end(), s.pre.begin(), s.pre.end());
  {
    struct TbRef { uint32_t item; uint32_t pre; uint32_t tq, ts; };
    std::vector<size_t> first(items.size() + 1, 0);
    parallel_for(items.size(), [&](size_t k) {
      ItemState& s = items[k];
      const uint32_t ql = lens[s.item / nsub];
      const uint32_t sl = (uint32_t)c->subj[s.item % nsub].align.size() - 2;
      std::stable_sort(s.pre.begin(), s.pre.end(), [&](const PreRec& a, const PreRec& b) {
        const uint32_t fsa = a.minus ? sl - a.s_end : a.s_start, fsb = b.minus ? sl - b.s_end : b.s_start;
        const uint32_t fea = a.minus ? sl - a.s_start : a.s_end, feb = b.minus ? sl - b.s_start : b.s_end;
        const uint32_t qoa = a.minus ? ql - a.q_end : a.q_start, qob = b.minus ? ql - b.q_end : b.q_start;
        const uint32_t qea = a.minus ? ql - a.q_start : a.q_end, qeb = b.minus ? ql - b.q_start : b.q_end;
        if (a.score != b.score) return a.score > b.score;
        if (fsa != fsb) return fsa < fsb;
        if (fea != feb) return fea > feb;
        if (qoa != qob) return qoa < qob;
        return qea > qeb;
      });
    });
    tb_marks[0] = t3.us();
    for (size_t k = 0; k < items.size(); ++k) first[k + 1] = first[k] + items[k].pre.size();
    std::vector<TbRef> refs(first.back());
    std::vector<DpTask> tasks(2 * first.back());
    std::vector<uint64_t> work(tasks.size());  // rows each half is expected to cover: the prelim's extent
    const uint8_t* d_salign = (const uint8_t*)c->d_salign.p;
    parallel_for(items.size(), [&](size_t k) {
      ItemState& s = items[k];
      const uint32_t r = s.item / nsub, f = s.item % nsub;
      const uint32_t ql = lens[r];
      const Subject& S = c->subj[f];
      const uint32_t sl = (uint32_t)S.align.size() - 2;
      for (size_t p = 0; p < s.pre.size(); ++p) {
        const PreRec& h = s.pre[p];
        const std::vector<uint8_t>& sv = h.minus ? S.align_rc : S.align;
        uint32_t nq, ns;
        improve_seed(qf[r].data() + 1, ql, sv.data() + 1, sl, h.q_seed, h.s_seed, h.q_start, h.q_end, h.s_start,
                     h.s_end, h.minus, nq, ns);
        const uint32_t tq = h.minus ? ql - 1 - nq : nq, ts = h.minus ? sl - 1 - ns : ns;
        const size_t g = first[k] + p;
        refs[g] = TbRef{(uint32_t)k, (uint32_t)p, tq, ts};
        const uint8_t* qa = (h.minus ? d_qrc : d_qfwd) + sr.whole_base[r] + 1;
        const uint8_t* sa = d_salign + c->sdesc[f].s_base + 1;
        tasks[2 * g] = DpTask{qa, sa, tq + 1, ts + 1, c->P.xdrop_gap_final, 1};
        const uint32_t rm = ql > tq + 1 ? ql - tq - 1 : 0, rn = sl > ts + 1 ? sl - ts - 1 : 0;
        tasks[2 * g + 1] = DpTask{qa + tq + 1, sa + ts + 1, rn ? rm : 0, rm ? rn : 0, c->P.xdrop_gap_final, 0};
        const uint32_t fqs = h.minus ? ql - h.q_end : h.q_start, fqe = h.minus ? ql - h.q_start : h.q_end;
        work[2 * g] = tq + 1 > fqs ? tq + 1 - fqs : 1;
        work[2 * g + 1] = fqe > tq + 1 ? fqe - tq - 1 : 1;
      }
    });
    tb_marks[1] = t3.us();
    std::vector<HalfTb> halves;
    std::vector<uint32_t> op_pool;
    run_dp_traceback(c, tasks, halves, op_pool, &work);
    tb_marks[2] = t3.us();
    c->stats[3] += refs.size();
    const int32_t go = c->P.gap_open, ge = c->P.gap_extend;
    parallel_for(items.size(), [&](size_t k) {
      ItemState& s = items[k];
      const uint32_t r = s.item / nsub, f = s.item % nsub;
      const uint32_t ql = lens[r];
      const Subject& S = c->subj[f];
      const uint32_t sl = (uint32_t)S.align.size() - 2;
      HspTree tbp(ql + 1, sl + 1), tbm(ql + 1, sl + 1, true);
      for (size_t p = 0; p < s.pre.size(); ++p) {
        const PreRec& h = s.pre[p];
        const bool plus = !h.minus;
        const uint32_t pss = plus ? h.s_start : sl - h.s_end, pse = plus ? h.s_end : sl - h.s_start;
        const uint32_t cqs = plus ? h.q_start : ql - h.q_end, cqe = plus ? h.q_end : ql - h.q_start;
        if ((plus ? tbp : tbm).covers(cqs, cqe, pss, pse, h.score, plus)) continue;
        const size_t g = first[k] + p;
        const TbRef& ref = refs[g];
        const HalfTb& L = halves[2 * g];
        const HalfTb& R = halves[2 * g + 1];
        int32_t sc = L.score + R.score;
        uint32_t q0 = ref.tq + 1 - L.a_len, s0 = ref.ts + 1 - L.b_len;
        uint32_t q1 = ref.tq + 1 + R.a_len, s1 = ref.ts + 1 + R.b_len;
        Script ops;
        for (uint32_t j = L.n_ops; j-- > 0;) {
          const uint32_t o = op_pool[L.ops_at + j];
          script_push(ops, (uint8_t)(o & 3), o >> 2);
        }
        for (uint32_t j = 0; j < R.n_ops; ++j) {
          const uint32_t o = op_pool[R.ops_at + j];
          script_push(ops, (uint8_t)(o & 3), o >> 2);
        }
        while (!ops.empty() && ops.front().kind != K_SUB) {  // engine.rs:3016-3033
          sc += go + (int32_t)ops.front().len * ge;
          if (ops.front().kind == K_GAP_Q) s0 += ops.front().len; else q0 += ops.front().len;
          ops.erase(ops.begin());
        }
        while (!ops.empty() && ops.back().kind != K_SUB) {
          sc += go + (int32_t)ops.back().len * ge;
          if (ops.back().kind == K_GAP_Q) s1 -= ops.back().len; else q1 -= ops.back().len;
          ops.pop_back();
        }
        if (sc < 0) continue;
        FinalRec fr;
        fr.minus = h.minus;
        if (h.minus) {
          fr.q_start = ql - q1;
          fr.q_end = ql - q0;
          fr.s_start = s0;  // forward subject
          fr.s_end = s1;
          std::reverse(ops.begin(), ops.end());
        } else {
          fr.q_start = q0;
          fr.q_end = q1;
          fr.s_start = s0;
          fr.s_end = s1;
        }
        if (c->P.complexity_adjust) {
          int32_t adj;
          if (!complexity_adjust(sc, qf[r].data(), qf[r].size(), fr.q_start, ops, c->M, cutoff, h.minus, adj))
            continue;
          sc = adj;
        }
        fr.score = sc;
        fr.ops = std::move(ops);
        const uint32_t tqs = plus ? fr.q_start : ql - fr.q_end, tqe = plus ? fr.q_end : ql - fr.q_start;
        (plus ? tbp : tbm).add(tqs, tqe, fr.s_start, fr.s_end, sc, true);
        s.fin.push_back(std::move(fr));
      }
      finish_subject(s.fin, qf[r].data(), ql, S.align.data(), S.align_rc.data(), sl, S.n_mask, S.n_mask_rc, c->M,
                     go, ge, cutoff);
    });
  }
  tb_marks[3] = t3.us();
  tb_ms[0] += tb_marks[0];
  tb_ms[1] += tb_marks[1] - tb_marks[0];
  tb_ms[2] += tb_marks[2] - tb_marks[1];
  tb_ms[3] += tb_marks[3] - tb_marks[2];
}

// Final HSPs of a batch's items, in item order, into the context's results.
static void emit_results(rmb_ctx* c, const std::vector<ItemState>& items) {
  const uint32_t nsub = (uint32_t)c->subj.size();
  for (const auto& s : items) {
    for (const auto& h : s.fin) {
      rmb_hsp o{};
      o.query = s.item / nsub;
      o.subject = s.item % nsub;
      o.score = h.score;
      o.q_start = h.q_start;
      o.q_end = h.q_end;
      o.s_start = h.s_start;
      o.s_end = h.s_end;
      o.strand = h.minus;
      o.ops_offset = (uint32_t)c->ops.size();
      o.ops_count = (uint32_t)h.ops.size();
      for (const Run& x : h.ops) c->ops.push_back((uint32_t)x.kind | (x.len << 2));
      c->hsps.push_back(o);
    }
  }
}

// pairs: null for every (query, subject) pair, else the item ids q * nsub + s to search
// Failure injection for tests of the error path (RMB_FORCE_ERROR=<stage>):
// the host fails the search at that stage, after the stage's kernels ended
// normally.
static void fail_point(const char* stage) {
  const char* e = std::getenv("RMB_FORCE_ERROR");
  if (e && std::strcmp(e, stage) == 0) throw Failure(RMB_ERR_DEVICE, std::string("forced failure after ") + stage);
}

// Ungapped keep threshold of a query record of full_len bases: the caller's
// fixed value, else the KA-derived cutoff rmblastn resolves per query before
// its search (main.rs:1154-1170, ka_cutoff.rs:126-161), else cutoff_score_max.
static int32_t record_cutoff(const rmb_ctx* c, uint64_t full_len) {
  if (c->P.ungapped_cutoff >= 0) return c->P.ungapped_cutoff;
  const int32_t ka = ka_ungapped_cutoff(c->P.ka_lambda, c->P.ka_k, c->P.min_raw_gapped_score, full_len,
                                        c->avg_subj_len, nullptr, nullptr);
  return ka >= 0 ? ka : cutoff_score_max(c->P.min_raw_gapped_score);
}

// Worker threads of the report written during a search: RMB_REPORT_THREADS
// if set, else a quarter of the host threads (it runs beside the search's
// own host work, which has the rest).
static unsigned report_threads() {
  static const unsigned n = [] {
    if (const char* e = std::getenv("RMB_REPORT_THREADS")) {
      const long v = std::strtol(e, nullptr, 10);
      if (v > 0) return (unsigned)std::min<long>(v, 1024);
    }
    return std::max(1u, host_threads() / 4);
  }();
  return n;
}

// The report written while a search runs (rmb_search_report): after a
// batch's HSPs are final, a worker thread writes its queries' rows while the
// search goes on with the next batches; one batch's worker at a time, so the
// text stays in query order.
struct ReportPipe {
  const rmb_report_spec* spec = nullptr;
  ReportStream* stream = nullptr;
  std::thread worker;
  std::exception_ptr failure;
  // the running batch's snapshot (its HSPs with ops rebased to its own ops)
  std::vector<rmb_hsp> hsps;
  std::vector<rmb_op> ops;
  std::vector<const uint8_t*> sp, sr;
  std::vector<uint32_t> sl;
  void join() {
    if (worker.joinable()) worker.join();
    if (failure) {
      std::exception_ptr f = failure;
      failure = nullptr;
      std::rethrow_exception(f);
    }
  }
  ~ReportPipe() {
    if (worker.joinable()) worker.join();
  }
};

static void search_impl(rmb_ctx* c, uint32_t nq, const uint8_t* const* qs, const uint32_t* lens,
                        const std::vector<uint64_t>* pairs = nullptr, ReportPipe* rp = nullptr) {
  Timer total;
  if (c->subj.empty()) throw Failure(RMB_ERR_STATE, "no library loaded");
  std::memset(c->stats, 0, sizeof c->stats);
  c->host_prelim_dps = 0;
  c->hsps.clear();
  c->ops.clear();
  c->last_q.clear();
  c->last_qlen.clear();
  c->report = ReportText{};
  const int ws = c->P.word_size;
  const uint32_t nsub = (uint32_t)c->subj.size();
  // ---- query pools.  A query of 1,999,800 bases or more is searched
  // through Phase 2a as NCBI's overlapping chunks, each its own sentinel-
  // wrapped record; Phase 2b aligns on the whole query, stored after them.
  std::vector<SeqView> qf(nq);  // the whole queries, wrapped, in the host forward pool
  SearchRecs sr;
  sr.first.assign(nq + 1, 0);
  sr.whole_base.assign(nq, 0);
  std::vector<QuerySplit> split(nq);
  for (uint32_t i = 0; i < nq; ++i) {
    if (lens[i] > 0x7FFFFFFFu) throw Failure(RMB_ERR_ARG, "query longer than 2^31-1 bases");
    split[i] = query_split(lens[i]);
    sr.first[i] = (uint32_t)sr.recs.size();
    if (split[i].count <= 1) {
      sr.whole_base[i] = sr.pool;
      sr.recs.push_back(RecDesc{sr.pool, 0, lens[i], record_cutoff(c, lens[i])});
      sr.query.push_back(i);
      sr.offset.push_back(0);
      sr.pool += lens[i] + 2;
      continue;
    }
    for (uint32_t k = 0; k < split[i].count; ++k) {
      const uint32_t cs = k * split[i].step;
      if (cs >= lens[i]) break;
      const uint32_t ce = std::min<uint64_t>((uint64_t)cs + split[i].size, lens[i]);
      if (ce - cs < (uint32_t)ws) break;  // main.rs:1256-1258
      sr.recs.push_back(RecDesc{sr.pool, 0, ce - cs, record_cutoff(c, lens[i])});
      sr.query.push_back(i);
      sr.offset.push_back(cs);
      sr.pool += ce - cs + 2;
    }
    sr.whole_base[i] = sr.pool;
    sr.pool += lens[i] + 2;
  }
  sr.first[nq] = (uint32_t)sr.recs.size();
  const uint32_t nsr = (uint32_t)sr.recs.size();
  std::vector<RecDesc>& recs = sr.recs;
  uint8_t* d_qfwd = c->b_qfwd.get<uint8_t>(sr.pool);
  uint8_t* d_qrc = c->b_qrc.get<uint8_t>(sr.pool);
  {
    // the forward pool filled in parallel from the caller's buffers (pinned),
    // uploaded on the seeding stream; the reverse-complement pool is built
    // from it on the device, and the first kernels queue behind both
    uint8_t* pf = c->h_qfwd.get(sr.pool);
    auto put = [&](uint64_t base, const uint8_t* bases, size_t n) {  // n bases between two sentinels
      uint8_t* f = pf + base;
      f[0] = f[n + 1] = 15;
      for (size_t j = 0; j < n; ++j) f[1 + j] = bases[j] & 15;
    };
    parallel_for(nq, [&](size_t i) {
      put(sr.whole_base[i], qs[i], lens[i]);
      if (sr.first[i + 1] - sr.first[i] > 1)
        for (uint32_t k = sr.first[i]; k < sr.first[i + 1]; ++k)
          put(recs[k].q_base, qs[i] + sr.offset[k], recs[k].q_len);
    });
    for (uint32_t i = 0; i < nq; ++i) qf[i] = SeqView{pf + sr.whole_base[i], (size_t)lens[i] + 2};
    c->last_q.resize(nq);
    c->last_qlen.assign(lens, lens + nq);
    for (uint32_t i = 0; i < nq; ++i) c->last_q[i] = qf[i].data();
    hip_check(hipMemcpyAsync(d_qfwd, pf, sr.pool, hipMemcpyHostToDevice, c->st1), "copy qfwd");
    std::vector<uint64_t> sb;
    std::vector<uint32_t> sl;
    for (uint32_t i = 0; i < nq; ++i) {
      sb.push_back(sr.whole_base[i]);
      sl.push_back(lens[i]);
      if (sr.first[i + 1] - sr.first[i] > 1)
        for (uint32_t k = sr.first[i]; k < sr.first[i + 1]; ++k) {
          sb.push_back(recs[k].q_base);
          sl.push_back(recs[k].q_len);
        }
    }
    device_revcomp(c, c->st1, d_qfwd, d_qrc, sb, sl);
  }
  // With DUST on, the lookup tables and the exact word extension read the
  // masked query; ungapped and gapped extension read the query itself
  // (engine.rs:606-619).
  const uint8_t* d_qext = d_qfwd;
  if (c->P.dust) d_qext = upload_dusted(c, qf, sr);
  unsigned long long* ctr = c->b_counters.get<unsigned long long>(C_NCOUNTERS);
  hip_check(hipMemset(ctr, 0, C_NCOUNTERS * sizeof(unsigned long long)), "memset counters");
  // ---- batches of whole queries, pipelined: the GPU seeds batch b + 1 while
  // the host replays and aligns batch b.  Whatever of the last batch's work
  // the host does after the seeding stream drains is not overlapped, so the
  // last batch gets a quarter of the others' size: cuts fall at S, 2S, ...,
  // (B - 1)S with S = total / (B - 3/4).
  uint64_t total_bases = 0;
  for (uint32_t i = 0; i < nsr; ++i) total_bases += recs[i].q_len;
  int max_batches = kMaxBatches;
  if (const char* e = std::getenv("RMB_BATCHES")) max_batches = std::max(1, std::atoi(e));
  if (c->tune.batches) max_batches = c->tune.batches;
  const double span = std::max<double>((double)kBatchMinBases,
                                       max_batches > 1 ? (double)total_bases / (max_batches - 0.75)
                                                       : (double)total_bases);
  std::vector<P1Batch> batches;
  {
    uint32_t k0 = 0, qb = 0;
    uint64_t acc = 0;
    for (uint32_t q = 0; q < nq; ++q) {
      for (uint32_t k = sr.first[q]; k < sr.first[q + 1]; ++k) acc += recs[k].q_len;
      const double cut = span * (double)(batches.size() + 1);
      if ((double)acc >= cut || q + 1 == nq) {
        const uint32_t k1 = sr.first[q + 1];
        if (k1 > k0) {
          P1Batch b;
          b.slot = (int)(batches.size() % 2);
          b.groups = group_records(c, recs, sr.first, k0, k1, pairs);
          b.q0 = qb;
          b.q1 = q + 1;
          batches.push_back(std::move(b));
          qb = q + 1;
        }
        k0 = k1;
      }
    }
  }
  if (rp) {  // the subjects' views for the report, once per search
    rp->sp.clear();
    rp->sr.clear();
    rp->sl.clear();
    for (const Subject& S : c->subj) {
      rp->sp.push_back(S.align.data());
      rp->sr.push_back(S.align_rc.data());
      rp->sl.push_back((uint32_t)S.align.size() - 2);
    }
  }
  uint64_t p1_ctr[C_NCOUNTERS] = {0};
  uint64_t t_seed = 0, t_2a = 0, t_2b = 0, tb_ms[4] = {0, 0, 0, 0};
  const uint64_t t_setup = total.us();
  // Two batches in flight on the seeding stream: a batch is queued as soon as
  // the batch before it in its slot has been read back, so the stream never
  // waits for the host's Phase 2a/2b work.
  {
    Timer te;
    for (size_t b = 0; b < std::min<size_t>(2, batches.size()); ++b)
      enqueue_batch(c, batches[b], d_qfwd, d_qrc, d_qext);
    t_seed += te.us();
  }
  for (size_t b = 0; b < batches.size(); ++b) {
    Timer t1;
    std::vector<PreHit> survivors;
    complete_batch(c, batches[b], recs, d_qfwd, d_qrc, d_qext, survivors, p1_ctr);
    if (b + 2 < batches.size()) enqueue_batch(c, batches[b + 2], d_qfwd, d_qrc, d_qext);
    t_seed += t1.us();
    Timer t2;
    fail_point("phase1");
    std::vector<ItemState> items = phase2a(c, sr, survivors);
    fail_point("phase2a");
    t_2a += t2.us();
    Timer t3;
    phase2b(c, sr, qf, lens, items, d_qfwd, d_qrc, tb_ms);
    fail_point("phase2b");
    const size_t h0 = c->hsps.size(), o0 = c->ops.size();
    emit_results(c, items);
    t_2b += t3.us();
    if (rp) {
      // the batch's queries are complete: their rows are written on a worker
      // thread over a copy of the batch's HSPs while the next batches run
      rp->join();
      rp->hsps.assign(c->hsps.begin() + (ptrdiff_t)h0, c->hsps.end());
      rp->ops.assign(c->ops.begin() + (ptrdiff_t)o0, c->ops.end());
      for (rmb_hsp& h : rp->hsps) h.ops_offset -= (uint32_t)o0;
      const uint32_t q0 = batches[b].q0, q1 = batches[b].q1;
      rp->worker = std::thread([c, rp, q0, q1]() {
        try {
          ReportData d;
          d.hsps = rp->hsps.data();
          d.n_hsps = rp->hsps.size();
          d.ops = rp->ops.data();
          d.n_queries = (uint32_t)c->last_q.size();
          d.queries = c->last_q.data();
          d.query_lens = c->last_qlen.data();
          d.n_subjects = (uint32_t)c->subj.size();
          d.subjects = rp->sp.data();
          d.subjects_rc = rp->sr.data();
          d.subject_lens = rp->sl.data();
          d.ka = rp->spec->ka;
          d.evalue_max = rp->spec->evalue_max;
          d.db_length = (int64_t)c->db_letters;
          d.db_num_seqs = (int32_t)c->subj.size();
          rp->stream->add(d, q0, q1, report_threads());
        } catch (...) {
          rp->failure = std::current_exception();
        }
      });
    }
  }
  if (rp) {
    Timer tr;
    rp->join();
    c->report = rp->stream->finish();
    c->stats[24] = tr.us();
  }
  c->stats[9] = t_seed;
  c->stats[10] = t_2a;
  c->stats[11] = t_2b;
  if (std::getenv("RMB_STAGE_TIMES"))
    std::fprintf(stderr,
                 "[rmb] %zu batches; setup ms %.1f; phase2b ms: sort %.1f  seeds+tasks %.1f  dp %.1f  replay %.1f  "
                 "(prelims %zu)\n",
                 batches.size(), t_setup / 1e3, tb_ms[0] / 1e3, tb_ms[1] / 1e3, tb_ms[2] / 1e3, tb_ms[3] / 1e3,
                 (size_t)c->stats[3]);
  unsigned long long hc[C_NCOUNTERS];
  hip_check(hipMemcpy(hc, ctr, sizeof hc, hipMemcpyDeviceToHost), "copy counters");
  c->stats[0] = p1_ctr[C_SEED_EVALS];
  c->stats[1] = p1_ctr[C_UNG_HITS];
  c->stats[7] = p1_ctr[C_UNG_STEPS];
  c->stats[2] = p1_ctr[C_PRELIM_DPS] + c->host_prelim_dps;
  c->stats[4] = p1_ctr[C_PRELIM_CELLS] + hc[C_DP_CELLS];  // DP cells of both phases (score passes)
  c->stats[5] = hc[C_DP_CELLS];
  c->stats[6] = c->hsps.size();
  c->stats[12] = total.us();
  c->stats[16] = p1_ctr[C_SEEDS];
  for (int k = 0; k < 6; ++k) c->stats[17 + k] = p1_ctr[C_P1_STAGE + k];
  if (std::getenv("RMB_STAGE_TIMES"))
    std::fprintf(stderr,
                 "[rmb] wave-run traceback rows %llu, scan rounds %llu; longest (us): wave-run DP %.1f walk %.1f, "
                 "lane-run DP %.1f walk %.1f\n",
                 hc[C_WAVE_ROWS], hc[C_WAVE_ROUNDS], hc[C_TB_TICKS] / 100.0, hc[C_TB_TICKS + 1] / 100.0,
                 hc[C_TB_TICKS + 2] / 100.0, hc[C_TB_TICKS + 3] / 100.0);
}
}  // namespace rmb

namespace rmb {
// RMBlast matrix file (matrix.rs:76-213): FREQS comment, header row, rows.
static bool parse_matrix(const std::string& path, rmb_matrix& m, std::string& err) {
  std::ifstream f(path);
  if (!f) { err = "cannot open matrix " + path; return false; }
  auto slot = [](char c) -> int {
    const char* s = "ACGTRYMKWSBDHVN";
    const char* p = std::strchr(s, c);
    return (p && c) ? (int)(p - s) : 15;
  };
  for (auto& r : m.scores)
    for (auto& v : r) v = -32768;
  std::memset(m.freqs, 0, sizeof m.freqs);
  std::vector<char> cols;
  int row = 0;
  bool freqs = false;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tok;
    if (!(is >> tok)) continue;
    if (tok[0] == '#') {
      size_t at = line.find("FREQS");
      if (at == std::string::npos) continue;
      std::istringstream fs(line.substr(at + 5));
      std::string b;
      double v;
      int got = 0;
      while (fs >> b >> v) {
        const int k = slot((char)std::toupper((unsigned char)b[0]));
        if (k > 3 || b.size() != 1) { err = "bad FREQS"; return false; }
        m.freqs[k] = v;
        ++got;
      }
      if (got != 4) { err = "FREQS needs A, C, G and T"; return false; }
      freqs = true;
      continue;
    }
    if (cols.empty()) {
      do cols.push_back((char)std::toupper((unsigned char)tok[0])); while (is >> tok);
      continue;
    }
    if (row >= (int)cols.size()) { err = "more rows than columns"; return false; }
    std::vector<int> vals;
    bool first = true;
    do {
      char* e = nullptr;
      long v = std::strtol(tok.c_str(), &e, 10);
      if (*e) {
        if (first && tok.size() == 1) { first = false; continue; }  // row label
        err = "score '" + tok + "' is not an integer";
        return false;
      }
      first = false;
      vals.push_back((int)v);
    } while (is >> tok);
    if (vals.size() != cols.size()) { err = "row length differs from the header"; return false; }
    for (size_t c = 0; c < cols.size(); ++c) m.scores[slot(cols[row])][slot(cols[c])] = vals[c];
    ++row;
  }
  if (!freqs) { err = "no FREQS line"; return false; }
  if (row != (int)cols.size()) { err = "matrix must be square"; return false; }
  for (int i = 0; i < 16; ++i) { m.scores[15][i] = INT32_MIN / 2; m.scores[i][15] = INT32_MIN / 2; }
  double expected = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) expected += m.freqs[i] * m.freqs[j] * m.scores[i][j];
  if (expected >= 0) { err = "expected pair score must be negative"; return false; }
  // lambda: sum_ij f_i f_j exp(lambda s_ij) = 1, doubling then bisection to 1e-5 (matrix.rs:306-354)
  auto g = [&](double lam) {
    double acc = 0;
    for (int i = 0; i < 16; ++i)
      for (int j = 0; j < 16; ++j)
        if (m.freqs[i] > 0.0 && m.freqs[j] > 0.0) acc += m.freqs[i] * m.freqs[j] * std::exp(lam * (double)m.scores[i][j]);
    return acc;
  };
  double lo = 0, x = 0.5;
  m.lambda = 0;
  while (g(x) < 1.0) {
    lo = x;
    x *= 2;
    if (x > 1000.0) return true;
  }
  double hi = x;
  while (hi - lo > 0.00001) {
    x = (lo + hi) / 2;
    if (g(x) >= 1.0) hi = x; else lo = x;
  }
  m.lambda = x;
  return true;
}

}  // namespace rmb

// ============================================================== C ABI =====
// A search that fails part way drains the work it queued before the error
// propagates, and leaves no partial results.
template <class F>
static void search_or_drain(rmb_ctx* c, F&& fn) {
  try {
    fn();
  } catch (...) {
    (void)hipStreamSynchronize(c->st1);
    (void)hipStreamSynchronize(c->st);
    c->hsps.clear();
    c->ops.clear();
    throw;
  }
}

template <class F>
static int guarded(rmb_ctx* c, F&& fn) {
  try {
    fn();
    return RMB_OK;
  } catch (const Failure& f) {
    if (c) c->err = f.what();
    return f.code;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return RMB_ERR_DEVICE;
  } catch (...) {
    if (c) c->err = "unknown failure";
    return RMB_ERR_DEVICE;
  }
}

static std::vector<std::string> spec_fields(const rmb_report_spec* spec);

extern "C" {

int rmb_matrix_load(const char* path, rmb_matrix* out, char* err, size_t err_len) {
  if (!path || !out) return RMB_ERR_ARG;
  std::string e;
  double karlin[5];
  bool has_karlin = false;
  // a "# KARLIN" comment line must parse too (matrix.rs:100-104,248-266)
  if (!rmb::parse_matrix(path, *out, e) || !rmb::read_karlin_comment(path, karlin, has_karlin, e)) {
    if (err && err_len) { std::strncpy(err, e.c_str(), err_len - 1); err[err_len - 1] = 0; }
    return RMB_ERR_ARG;
  }
  return RMB_OK;
}

int rmb_ctx_create(int device, const rmb_params* params, const rmb_matrix* matrix, rmb_ctx** out) {
  if (!params || !matrix || !out) return RMB_ERR_ARG;
  *out = nullptr;
  rmb_ctx* c = new rmb_ctx();
  int rc = guarded(c, [&]() {
    if (params->dust != 0 && params->dust != 1) throw Failure(RMB_ERR_ARG, "dust must be 0 or 1");
    if (params->word_size < 4) throw Failure(RMB_ERR_ARG, "word_size below 4");
    if (params->gap_extend < 1 || params->gap_open < 0) throw Failure(RMB_ERR_ARG, "bad gap costs");
    if (!(params->ka_lambda >= 0.0) || !(params->ka_k >= 0.0))
      throw Failure(RMB_ERR_ARG, "ka_lambda and ka_k must be 0 or positive");
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw Failure(RMB_ERR_ARG, "no such HIP device");
    c->device = device;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    c->cu_count = prop.multiProcessorCount;
    hip_check(hipStreamCreate(&c->st), "hipStreamCreate");
    hip_check(hipStreamCreate(&c->st1), "hipStreamCreate");
    hip_check(hipEventCreate(&c->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&c->ev1), "hipEventCreate");
    c->P = *params;
    std::memcpy(c->M.s, matrix->scores, sizeof c->M.s);
    std::memcpy(c->M.freqs, matrix->freqs, sizeof c->M.freqs);
    c->M.lambda = matrix->lambda;
    hip_check(noalloc_scratch_address(&c->noalloc_scratch), "hipGetSymbolAddress (no-alloc scratch)");
    hip_check(hipMalloc(&c->d_mat, 256 * sizeof(int32_t)), "hipMalloc matrix");
    hip_check(hipMemcpy(c->d_mat, matrix->scores, 256 * sizeof(int32_t), hipMemcpyHostToDevice), "copy matrix");
  });
  if (rc != RMB_OK) {
    // keep the context so the caller can read the message
    *out = c;
    return rc;
  }
  *out = c;
  return RMB_OK;
}

int rmb_ctx_destroy(rmb_ctx* c) {
  if (!c) return RMB_OK;
  (void)hipSetDevice(c->device);
  for (DevBuf* b : {&c->d_splus, &c->d_src, &c->d_salign, &c->d_sdesc, &c->b_qfwd, &c->b_qrc, &c->b_recs, &c->b_elem,
                    &c->b_keys, &c->b_keys2, &c->b_vals, &c->b_vals2, &c->b_counts, &c->b_cells, &c->b_sort_tmp,
                    &c->b_hits, &c->b_hitcount, &c->b_spill, &c->b_hitres, &c->b_queue, &c->b_counters, &c->b_err, &c->b_tasks,
                    &c->b_dpres, &c->b_ring, &c->b_ring_big, &c->b_tbtasks, &c->b_tbres, &c->b_flat, &c->b_rows,
                    &c->b_ops, &c->b_pre, &c->b_redo, &c->b_redo2, &c->b_precount, &c->b_packtasks, &c->b_packed, &c->b_items, &c->b_qext,
                    &c->b_recids, &c->b_arena, &c->b_arena_big, &c->b_arena_wave, &c->b_oppool, &c->b_opres, &c->b_opcount})
    b->release();
  c->h_qfwd.release();
  for (P1Slot& S : c->slot) {
    for (DevBuf* b : {&S.pre, &S.redo_hits, &S.cnt, &S.hitcnt, &S.ctr, &S.err}) b->release();
    for (hipEvent_t e : S.ev) (void)hipEventDestroy(e);
    if (S.done) (void)hipEventDestroy(S.done);
  }
  if (c->d_mat) (void)hipFree(c->d_mat);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->st1) (void)hipStreamDestroy(c->st1);
  delete c;
  return RMB_OK;
}

int rmb_ctx_set_tuning(rmb_ctx* c, const rmb_tuning* t) {
  if (!c || !t) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    auto bad = [](const char* what) { throw Failure(RMB_ERR_ARG, std::string("tuning: ") + what); };
    if (t->batches < 0 || t->batches > 64) bad("batches must be 0 (default) or 1..64");
    if (t->prelim_ring != 0 && t->prelim_ring != 16 && t->prelim_ring != 64) bad("prelim_ring must be 0, 16 or 64");
    if (t->prelim_wg_per_cu < 0 || t->prelim_wg_per_cu > 16) bad("prelim_wg_per_cu must be 0..16");
    if (t->tb_ring != 0 && t->tb_ring != 64 && t->tb_ring != 256) bad("tb_ring must be 0, 64 or 256");
    if (t->tb_wg_per_cu < 0 || t->tb_wg_per_cu > 16) bad("tb_wg_per_cu must be 0..16");
    if (t->tb_wave_halves < -1 || t->tb_wave_halves > (1 << 30)) bad("tb_wave_halves must be -1, 0 or 1..2^30");
    c->tune = *t;
  });
}

int rmb_last_error(const rmb_ctx* c, char* buf, size_t len) {
  if (!c || !buf || len == 0) return RMB_ERR_ARG;
  std::strncpy(buf, c->err.c_str(), len - 1);
  buf[len - 1] = 0;
  return RMB_OK;
}

int rmb_library_load(rmb_ctx* c, uint32_t n, const uint8_t* const* seqs, const uint32_t* lens) {
  if (!c) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (n == 0 || !seqs || !lens) throw Failure(RMB_ERR_ARG, "empty library");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    c->hsps.clear();  // results of an earlier search name the old library's subjects
    c->ops.clear();
    c->last_q.clear();
    c->last_qlen.clear();
    c->subj.assign(n, Subject{});
    c->sdesc.assign(n, SubjDesc{});
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (lens[i] > 0x7FFFFFFFu) throw Failure(RMB_ERR_ARG, "subject longer than 2^31-1 bases");
      c->sdesc[i].s_base = off;
      c->sdesc[i].s_len = lens[i];
      off += lens[i] + 2;
    }
    // host views: the original codes, their reverse complement and the
    // ambiguity masks (extract_aligned restores them)
    std::vector<uint8_t> pa(off);
    parallel_for(n, [&](size_t i) {
      Subject& S = c->subj[i];
      S.align = wrap(seqs[i], lens[i]);
      S.align_rc = revcomp_wrapped(S.align);
      for (uint32_t k = 0; k < lens[i]; ++k) {
        const uint8_t code = S.align[1 + k];
        if (code >= 4 && code <= 14) {
          if (S.n_mask.empty()) S.n_mask.assign(lens[i], 0);
          S.n_mask[k] = code;
        }
      }
      if (!S.n_mask.empty()) {
        S.n_mask_rc.resize(lens[i]);
        for (uint32_t k = 0; k < lens[i]; ++k) {
          const uint8_t code = S.n_mask[lens[i] - 1 - k];
          S.n_mask_rc[k] = code ? kBlastnaComplement[code] : 0;
        }
      }
      std::memcpy(pa.data() + c->sdesc[i].s_base, S.align.data(), S.align.size());
    });
    c->db_letters = off - 2ull * n;
    c->avg_subj_len = c->db_letters / n;
    // the device builds the scanned strands from the one upload: the plus
    // strand with its ambiguity codes resolved, and its reverse complement
    uint8_t* dal = c->d_salign.get<uint8_t>(off);
    uint8_t* dpl = c->d_splus.get<uint8_t>(off);
    uint8_t* drc = c->d_src.get<uint8_t>(off);
    hip_check(hipMemcpy(dal, pa.data(), off, hipMemcpyHostToDevice), "copy subjects");
    hip_check(hipMemcpy(c->d_sdesc.get<SubjDesc>(n), c->sdesc.data(), n * sizeof(SubjDesc), hipMemcpyHostToDevice),
              "copy subject table");
    hip_check(hipMemcpyAsync(dpl, dal, off, hipMemcpyDeviceToDevice, c->st), "copy plus");
    hip_check(launch_resolve_ambiguity(dpl, (const SubjDesc*)c->d_sdesc.p, n, c->st), "k_resolve_ambiguity");
    std::vector<uint64_t> sb(n);
    std::vector<uint32_t> sl(n);
    for (uint32_t i = 0; i < n; ++i) {
      sb[i] = c->sdesc[i].s_base;
      sl[i] = lens[i];
    }
    device_revcomp(c, c->st, dpl, drc, sb, sl);
  });
}

int rmb_search(rmb_ctx* c, uint32_t nq, const uint8_t* const* qs, const uint32_t* lens) {
  if (!c) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (nq == 0 || !qs || !lens) throw Failure(RMB_ERR_ARG, "empty query batch");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    search_or_drain(c, [&] { search_impl(c, nq, qs, lens); });
  });
}

int rmb_search_report(rmb_ctx* c, uint32_t nq, const uint8_t* const* qs, const uint32_t* lens,
                      const rmb_report_spec* spec, uint64_t* n_bytes) {
  if (!c || !spec || !n_bytes) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (nq == 0 || !qs || !lens) throw Failure(RMB_ERR_ARG, "empty query batch");
    const std::vector<std::string> fields = spec_fields(spec);
    if (!spec->query_ids || (!spec->subject_ids && !c->subj.empty()))
      throw Failure(RMB_ERR_ARG, "query and subject ids are required");
    ReportStream stream(fields, spec->query_ids, spec->subject_ids, spec->mask_level);
    ReportPipe rp;
    rp.spec = spec;
    rp.stream = &stream;
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    search_or_drain(c, [&] { search_impl(c, nq, qs, lens, nullptr, &rp); });
    *n_bytes = c->report.size;
  });
}

int rmb_search_pairs(rmb_ctx* c, uint32_t nq, const uint8_t* const* qs, const uint32_t* lens, uint64_t n_pairs,
                     const uint32_t* pair_query, const uint32_t* pair_subject) {
  if (!c) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (nq == 0 || !qs || !lens) throw Failure(RMB_ERR_ARG, "empty query batch");
    if (n_pairs && (!pair_query || !pair_subject)) throw Failure(RMB_ERR_ARG, "null pair list");
    if (c->subj.empty()) throw Failure(RMB_ERR_STATE, "no library loaded");
    const uint64_t nsub = c->subj.size();
    if ((uint64_t)nq * nsub > UINT32_MAX) throw Failure(RMB_ERR_ARG, "more than 2^32 query x subject pairs");
    std::vector<uint64_t> items(n_pairs);
    for (uint64_t k = 0; k < n_pairs; ++k) {
      if (pair_query[k] >= nq || pair_subject[k] >= nsub) throw Failure(RMB_ERR_ARG, "pair index out of range");
      items[k] = (uint64_t)pair_query[k] * nsub + pair_subject[k];
    }
    std::sort(items.begin(), items.end());
    items.erase(std::unique(items.begin(), items.end()), items.end());
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    search_or_drain(c, [&] { search_impl(c, nq, qs, lens, &items); });
  });
}

// The last search's HSPs and sequences for the host tail.
static ReportData report_data(const rmb_ctx* c, const rmb_report_spec* spec, std::vector<const uint8_t*>& sp,
                              std::vector<const uint8_t*>& sr, std::vector<uint32_t>& sl) {
  for (const Subject& S : c->subj) {
    sp.push_back(S.align.data());
    sr.push_back(S.align_rc.data());
    sl.push_back((uint32_t)S.align.size() - 2);
  }
  ReportData d;
  d.hsps = c->hsps.data();
  d.n_hsps = c->hsps.size();
  d.ops = c->ops.data();
  d.n_queries = (uint32_t)c->last_q.size();
  d.queries = c->last_q.data();
  d.query_lens = c->last_qlen.data();
  d.n_subjects = (uint32_t)c->subj.size();
  d.subjects = sp.data();
  d.subjects_rc = sr.data();
  d.subject_lens = sl.data();
  if (spec) {
    d.ka = spec->ka;
    d.evalue_max = spec->evalue_max;
  }
  d.db_length = (int64_t)c->db_letters;
  d.db_num_seqs = (int32_t)c->subj.size();
  return d;
}

int rmb_align_stats_of(const char* q, const char* s, uint64_t n, rmb_align_stats* out) {
  if ((!q || !s) && n) return RMB_ERR_ARG;
  if (!out) return RMB_ERR_ARG;
  *out = align_stats(q, s, n);
  return RMB_OK;
}

int rmb_mask_level(const rmb_hsp* hsps, uint64_t n, uint32_t q_len, int32_t mask_level, uint8_t* keep) {
  if ((!hsps || !keep) && n) return RMB_ERR_ARG;
  try {
    ReportData d;
    d.hsps = hsps;
    d.n_hsps = n;
    std::vector<uint64_t> rows(n);
    for (uint64_t i = 0; i < n; ++i) rows[i] = i;
    if (mask_level >= 0 && mask_level < 100) mask_level_filter(d, rows, mask_level, q_len);
    std::fill(keep, keep + n, (uint8_t)0);
    for (uint64_t r : rows) keep[r] = 1;
    return RMB_OK;
  } catch (...) {
    return RMB_ERR_DEVICE;
  }
}

int rmb_report_order(rmb_ctx* c, const rmb_report_spec* spec, uint64_t* rows, uint64_t* n) {
  if (!c || !spec || !n) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    std::vector<const uint8_t*> sp, sr;
    std::vector<uint32_t> sl;
    const ReportData d = report_data(c, spec, sp, sr, sl);
    const auto order = report_order(d, spec->mask_level, host_threads());
    uint64_t k = 0;
    for (const auto& q : order) {
      if (rows)
        for (uint64_t r : q) rows[k++] = r;
      else
        k += q.size();
    }
    *n = k;
  });
}

}  // extern "C"

// The outfmt 6 field names of a report spec, checked.
static std::vector<std::string> spec_fields(const rmb_report_spec* spec) {
  if (!spec->fields) throw Failure(RMB_ERR_ARG, "no outfmt fields");
  std::vector<std::string> fields;
  std::string w;
  for (const char* p = spec->fields;; ++p) {
    if (*p == ' ' || *p == '\t' || *p == 0) {
      if (!w.empty()) fields.push_back(w);
      w.clear();
      if (!*p) break;
    } else {
      w += *p;
    }
  }
  if (fields.empty()) throw Failure(RMB_ERR_ARG, "no outfmt fields");
  try {
    (void)parse_report_fields(fields);
  } catch (const std::invalid_argument& e) {
    throw Failure(RMB_ERR_ARG, e.what());
  }
  return fields;
}

extern "C" {

int rmb_report_tabular(rmb_ctx* c, const rmb_report_spec* spec, uint64_t* n_bytes) {
  if (!c || !spec || !n_bytes) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    const std::vector<std::string> fields = spec_fields(spec);
    if ((!spec->query_ids && !c->last_q.empty()) || (!spec->subject_ids && !c->subj.empty()))
      throw Failure(RMB_ERR_ARG, "query and subject ids are required");
    std::vector<const uint8_t*> sp, sr;
    std::vector<uint32_t> sl;
    const ReportData d = report_data(c, spec, sp, sr, sl);
    const unsigned th = host_threads();
    c->report = ReportText{};
    try {
      c->report = report_tabular(d, report_order(d, spec->mask_level, th), fields, spec->query_ids,
                                 spec->subject_ids, th);
    } catch (const std::invalid_argument& e) {
      throw Failure(RMB_ERR_ARG, e.what());
    }
    *n_bytes = c->report.size;
  });
}

int rmb_report_copy(const rmb_ctx* c, char* out) {
  if (!c || (!out && c->report.size)) return RMB_ERR_ARG;
  if (c->report.size) std::memcpy(out, c->report.data.get(), c->report.size);
  return RMB_OK;
}

int rmb_report_text(const rmb_ctx* c, const char** text, uint64_t* n_bytes) {
  if (!c || !text || !n_bytes) return RMB_ERR_ARG;
  *text = c->report.size ? c->report.data.get() : "";
  *n_bytes = c->report.size;
  return RMB_OK;
}

int rmb_ka_stats_resolve(const char* matrix_path, int32_t gap_open, int32_t gap_extend,
                         const rmb_ka_stats_options* opt, rmb_ka_stats* out, char* err, size_t err_len) {
  if (!matrix_path || !opt || !out) return RMB_ERR_ARG;
  std::string msg;
  if (!ka_stats_resolve(matrix_path, gap_open, gap_extend, *opt, *out, msg)) {
    if (err && err_len) {
      std::strncpy(err, msg.c_str(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return RMB_ERR_ARG;
  }
  return RMB_OK;
}

int rmb_ka_search_space(const rmb_ka_stats* ka, int32_t query_length, int64_t db_length, int32_t db_num_seqs,
                        int64_t* eff_searchsp, int32_t* la) {
  if (!ka || !eff_searchsp || !la) return RMB_ERR_ARG;
  ka_search_space(*ka, query_length, db_length, db_num_seqs, *eff_searchsp, *la);
  return RMB_OK;
}

int rmb_ka_evalue(const rmb_ka_stats* ka, int64_t eff_searchsp, int32_t score, double* evalue, double* bit_score) {
  if (!ka) return RMB_ERR_ARG;
  if (evalue) *evalue = ka_evalue(*ka, eff_searchsp, score);
  if (bit_score) *bit_score = ka_bit_score(*ka, score);
  return RMB_OK;
}

int rmb_ka_lookup(const char* matrix_name, int32_t gap_open, int32_t gap_extend, double out[2], int32_t* found) {
  if (!matrix_name || !out || !found) return RMB_ERR_ARG;
  double lambda = 0, k = 0;
  *found = ka_lookup(matrix_name, gap_open, gap_extend, lambda, k) ? 1 : 0;
  out[0] = lambda;
  out[1] = k;
  return RMB_OK;
}

int rmb_ungapped_cutoff(const char* matrix_name, int32_t gap_open, int32_t gap_extend, int32_t min_raw_gapped_score,
                        uint64_t query_length, uint64_t avg_subj_length, int32_t* cutoff, rmb_ka_cutoff_info* info) {
  if (!matrix_name || !cutoff) return RMB_ERR_ARG;
  double lambda = 0, k = 0;
  const bool hit = ka_lookup(matrix_name, gap_open, gap_extend, lambda, k);
  int32_t es = 0;
  uint64_t searchsp = 0;
  const int32_t ka = hit ? ka_ungapped_cutoff(lambda, k, min_raw_gapped_score, query_length, avg_subj_length, &es,
                                              &searchsp)
                         : -1;
  *cutoff = ka >= 0 ? ka : cutoff_score_max(min_raw_gapped_score);
  if (info) *info = rmb_ka_cutoff_info{ka >= 0 ? 1 : 0, es, lambda, k, searchsp};
  return RMB_OK;
}

int rmb_dust_intervals(const uint8_t* bases, uint64_t n, uint64_t* out, uint64_t* n_out) {
  if ((!bases && n) || !n_out) return RMB_ERR_ARG;
  try {
    const std::vector<std::pair<size_t, size_t>> iv = dust_intervals(bases, n);
    if (out) {
      const uint64_t cap = *n_out;
      for (size_t k = 0; k < iv.size() && k < cap; ++k) {
        out[2 * k] = iv[k].first;
        out[2 * k + 1] = iv[k].second;
      }
    }
    *n_out = iv.size();
    return RMB_OK;
  } catch (...) {
    return RMB_ERR_DEVICE;
  }
}

int rmb_results_size(const rmb_ctx* c, uint64_t* n_hsps, uint64_t* n_ops) {
  if (!c || !n_hsps || !n_ops) return RMB_ERR_ARG;
  *n_hsps = c->hsps.size();
  *n_ops = c->ops.size();
  return RMB_OK;
}

int rmb_results_copy(const rmb_ctx* c, rmb_hsp* hsps, rmb_op* ops) {
  if (!c) return RMB_ERR_ARG;
  if (hsps && !c->hsps.empty()) std::memcpy(hsps, c->hsps.data(), c->hsps.size() * sizeof(rmb_hsp));
  if (ops && !c->ops.empty()) std::memcpy(ops, c->ops.data(), c->ops.size() * sizeof(rmb_op));
  return RMB_OK;
}

int rmb_stats(const rmb_ctx* c, uint64_t* out, int n) {
  if (!c || !out) return RMB_ERR_ARG;
  for (int i = 0; i < n && i < RMB_NSTATS; ++i) out[i] = c->stats[i];
  return RMB_OK;
}

// The query as one record of a one-record group, uploaded (both strands and,
// with DUST, the masked copy) for the stage taps.
struct TapQuery {
  P1Group g;
  const uint8_t *d_qfwd = nullptr, *d_qrc = nullptr, *d_qext = nullptr;
};
static TapQuery tap_query(rmb_ctx* c, const uint8_t* query, uint32_t q_len) {
  TapQuery t;
  std::vector<uint8_t> qf = wrap(query, q_len), qr = revcomp_wrapped(qf);
  uint8_t* d_qfwd = c->b_qfwd.get<uint8_t>(qf.size());
  uint8_t* d_qrc = c->b_qrc.get<uint8_t>(qr.size());
  hip_check(hipMemcpy(d_qfwd, qf.data(), qf.size(), hipMemcpyHostToDevice), "copy q");
  hip_check(hipMemcpy(d_qrc, qr.data(), qr.size(), hipMemcpyHostToDevice), "copy qrc");
  std::vector<RecDesc> recs(1);
  recs[0].q_base = 0;
  recs[0].q_len = q_len;
  recs[0].cutoff = record_cutoff(c, q_len);
  SearchRecs one;  // the query as one record, whatever its length
  one.recs = recs;
  one.query = {0};
  one.offset = {0};
  one.first = {0, 1};
  one.whole_base = {0};
  one.pool = qf.size();
  t.d_qfwd = d_qfwd;
  t.d_qrc = d_qrc;
  t.d_qext = c->P.dust ? upload_dusted(c, {SeqView{qf.data(), qf.size()}}, one) : d_qfwd;
  t.g.lw = lut_width_for(c->P.word_size, 2ull * q_len);
  t.g.recs_idx = {0};
  t.g.recs = recs;
  t.g.elem_base = {0, 2ull * q_len + 1};
  return t;
}

// k_phase1 over the query against every subject: its hits, or with `seeds`
// its seed records, in the kernel's output order.
static std::vector<UngHit> tap_phase1(rmb_ctx* c, TapQuery& t, uint32_t q_len, bool seeds) {
  std::vector<UngHit> hits;
  if (q_len < (uint32_t)c->P.word_size) return hits;
  unsigned long long* ctr = c->b_counters.get<unsigned long long>(C_NCOUNTERS);
  hip_check(hipMemset(ctr, 0, C_NCOUNTERS * sizeof(unsigned long long)), "memset counters");
  unsigned long long* hitcount = c->b_hitcount.get<unsigned long long>(1);
  uint32_t* err = c->b_err.get<uint32_t>(1);
  for (uint64_t cap = 1u << 16;;) {
    UngHit* dh = c->b_hits.get<UngHit>(cap);
    hip_check(hipMemsetAsync(hitcount, 0, sizeof(unsigned long long), c->st1), "memset hitcount");
    hip_check(hipMemsetAsync(err, 0, sizeof(uint32_t), c->st1), "memset err");
    launch_group_phase1(c, c->st1, t.g, t.d_qfwd, t.d_qrc, t.d_qext, dh, cap, hitcount, ctr, err, nullptr, nullptr,
                        seeds);
    unsigned long long nh = 0;
    uint32_t e = 0;
    hip_check(hipMemcpyAsync(&nh, hitcount, sizeof nh, hipMemcpyDeviceToHost, c->st1), "copy hitcount");
    hip_check(hipMemcpyAsync(&e, err, sizeof e, hipMemcpyDeviceToHost, c->st1), "copy err");
    hip_check(hipStreamSynchronize(c->st1), "tap sync");
    if (e & E_HASH_FULL) throw Failure(RMB_ERR_CAPACITY, "diagonal hash bucket exceeded 256 cells");
    if (nh > cap) {
      cap = nh + nh / 4;
      continue;
    }
    hits.resize(nh);
    hip_check(hipMemcpy(hits.data(), dh, nh * sizeof(UngHit), hipMemcpyDeviceToHost), "copy hits");
    return hits;
  }
}

int rmb_tap_ungapped(rmb_ctx* c, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out, uint64_t* n) {
  if (!c || !query || !n) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (subject >= c->subj.size()) throw Failure(RMB_ERR_ARG, "subject index out of range");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    TapQuery t = tap_query(c, query, q_len);
    const std::vector<UngHit> hits = tap_phase1(c, t, q_len, false);
    std::vector<UngRec> v;
    for (auto& h : hits)
      if (h.item != kNoItem && h.item % c->subj.size() == subject)  // skip reserved slots left unused
        v.push_back(UngRec{h.q_seed, h.s_seed, h.q_start, h.q_end, h.s_start, h.s_end, h.score, (uint8_t)h.strand});
    order_ungapped(v, q_len, c->sdesc[subject].s_len);
    if (out) {
      const uint64_t cap = *n;
      for (size_t i = 0; i < v.size() && i < cap; ++i) {
        uint32_t* o = out + 8 * i;
        o[0] = v[i].q_seed; o[1] = v[i].s_seed; o[2] = v[i].q_start; o[3] = v[i].q_end;
        o[4] = v[i].s_start; o[5] = v[i].s_end; o[6] = (uint32_t)v[i].score; o[7] = v[i].minus;
      }
    }
    *n = v.size();
  });
}

int rmb_tap_seeds(rmb_ctx* c, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out, uint64_t* n) {
  if (!c || !query || !n) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (subject >= c->subj.size()) throw Failure(RMB_ERR_ARG, "subject index out of range");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    TapQuery t = tap_query(c, query, q_len);
    const std::vector<UngHit> recs = tap_phase1(c, t, q_len, true);
    // {ordinal, diagonal, check offset, strand} of the subject's seeds, in scan order
    struct S { uint64_t ord; int32_t diag; uint32_t chk, minus; };
    std::vector<S> v;
    for (auto& h : recs)
      if (h.item != kNoItem && h.item % c->subj.size() == subject)
        v.push_back(S{(uint64_t)h.q_start | ((uint64_t)h.q_end << 32), (int32_t)h.q_seed, h.s_seed, h.strand});
    std::sort(v.begin(), v.end(), [](const S& a, const S& b) { return a.ord < b.ord; });
    if (out) {
      const uint64_t cap = *n;
      const int64_t ws = c->P.word_size;
      for (size_t i = 0; i < v.size() && i < cap; ++i) {
        uint32_t* o = out + 3 * i;
        // plus: the extended word's (q, s) on diagonal s - q; minus (engine.rs:1223-1236): the
        // forward-query offset less the extension still needed, and the forward-subject offset
        const int64_t q = v[i].minus ? 2 * (int64_t)q_len + 1 - ws + v[i].diag - (int64_t)v[i].chk
                                     : (int64_t)(int32_t)(v[i].chk - (uint32_t)v[i].diag);
        o[0] = (uint32_t)std::max<int64_t>(q, 0);
        o[1] = v[i].chk;
        o[2] = v[i].minus;
      }
    }
    *n = v.size();
  });
}

int rmb_tap_subject(const rmb_ctx* c, uint32_t subject, uint8_t* plus, uint8_t* rc, uint32_t* n) {
  if (!c || !n) return RMB_ERR_ARG;
  return guarded(const_cast<rmb_ctx*>(c), [&]() {
    if (subject >= c->subj.size()) throw Failure(RMB_ERR_ARG, "subject index out of range");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const SubjDesc& S = c->sdesc[subject];
    if (plus && *n >= S.s_len && S.s_len)
      hip_check(hipMemcpy(plus, (const uint8_t*)c->d_splus.p + S.s_base + 1, S.s_len, hipMemcpyDeviceToHost), "copy plus");
    if (rc && *n >= S.s_len && S.s_len)
      hip_check(hipMemcpy(rc, (const uint8_t*)c->d_src.p + S.s_base + 1, S.s_len, hipMemcpyDeviceToHost), "copy rc");
    *n = S.s_len;
  });
}

int rmb_tap_lookup(rmb_ctx* c, const uint8_t* query, uint32_t q_len, int32_t* lut_width, uint32_t* cells,
                   uint64_t* n_cells, uint32_t* entries, uint64_t* n_entries) {
  if (!c || !query || !lut_width || !n_cells || !n_entries) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    TapQuery t = tap_query(c, query, q_len);
    *lut_width = t.g.lw;
    const uint64_t nc = (1ull << (2 * t.g.lw)) + 1;
    if (q_len < (uint32_t)c->P.word_size) {  // no table (engine.rs:751-753): every cell empty
      if (cells && *n_cells >= nc) std::fill(cells, cells + nc, 0u);
      *n_cells = nc;
      *n_entries = 0;
      return;
    }
    const GroupLut lut = build_group_lut(c, c->st1, t.g, t.d_qext);
    std::vector<uint32_t> hc(nc);
    hip_check(hipMemcpyAsync(hc.data(), lut.cells, nc * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st1),
              "copy cells");
    hip_check(hipStreamSynchronize(c->st1), "lookup sync");
    const uint64_t ne = hc[nc - 1];  // valid windows
    if (cells && *n_cells >= nc) std::copy(hc.begin(), hc.end(), cells);
    if (entries && *n_entries >= ne && ne)
      hip_check(hipMemcpy(entries, lut.entries, ne * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy entries");
    *n_cells = nc;
    *n_entries = ne;
  });
}

int rmb_tap_prelims(rmb_ctx* c, const uint8_t* query, uint32_t q_len, uint32_t subject, uint32_t* out, uint64_t* n) {
  if (!c || !query || !n) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (subject >= c->subj.size()) throw Failure(RMB_ERR_ARG, "subject index out of range");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const uint8_t* qs[1] = {query};
    const uint32_t ls[1] = {q_len};
    const std::vector<uint64_t> items{(uint64_t)subject};
    c->tap_pre.clear();
    c->tap_pre_on = true;
    try {
      search_or_drain(c, [&] { search_impl(c, 1, qs, ls, &items); });
    } catch (...) {
      c->tap_pre_on = false;
      throw;
    }
    c->tap_pre_on = false;
    if (out) {
      const uint64_t cap = *n;
      for (size_t i = 0; i < c->tap_pre.size() && i < cap; ++i) {
        const PreRec& h = c->tap_pre[i];
        uint32_t* o = out + 8 * i;
        o[0] = h.q_seed; o[1] = h.s_seed; o[2] = h.q_start; o[3] = h.q_end;
        o[4] = h.s_start; o[5] = h.s_end; o[6] = (uint32_t)h.score; o[7] = h.minus;
      }
    }
    *n = c->tap_pre.size();
  });
}

static void tap_upload(rmb_ctx* c, const uint8_t* a, uint32_t m, const uint8_t* b, uint32_t n, const uint8_t*& da,
                       const uint8_t*& db) {
  uint8_t* d = c->b_qfwd.get<uint8_t>((size_t)m + n + 2);
  hip_check(hipMemcpy(d, a, m, hipMemcpyHostToDevice), "copy a");
  hip_check(hipMemcpy(d + m + 1, b, n, hipMemcpyHostToDevice), "copy b");
  da = d;
  db = d + m + 1;
}

int rmb_tap_align_score_only(rmb_ctx* c, const uint8_t* a, uint32_t m, const uint8_t* b, uint32_t n, int32_t xdrop,
                             int32_t reverse, int64_t out[3]) {
  if (!c || !out || (m && !a) || (n && !b)) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const uint8_t *da, *db;
    tap_upload(c, a, m, b, n, da, db);
    std::vector<DpTask> t{DpTask{da, db, m, n, xdrop, (uint32_t)(reverse != 0)}};
    std::vector<DpResult> r;
    run_dp_score(c, t, r);
    out[0] = r[0].score;
    out[1] = r[0].a_len;
    out[2] = r[0].b_len;
  });
}

int rmb_tap_align_ex(rmb_ctx* c, const uint8_t* a, uint32_t m, const uint8_t* b, uint32_t n, int32_t xdrop,
                     int32_t reverse, int64_t out[4], rmb_op* ops, uint32_t max_ops) {
  if (!c || !out || (m && !a) || (n && !b)) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const uint8_t *da, *db;
    tap_upload(c, a, m, b, n, da, db);
    std::vector<DpTask> t{DpTask{da, db, m, n, xdrop, (uint32_t)(reverse != 0)}};
    std::vector<HalfTb> r;
    std::vector<uint32_t> pool;
    run_dp_traceback(c, t, r, pool);
    Script sc;
    for (uint32_t j = 0; j < r[0].n_ops; ++j) {
      const uint32_t o = pool[r[0].ops_at + j];
      script_push(sc, (uint8_t)(o & 3), o >> 2);
    }
    out[0] = r[0].score;
    out[1] = r[0].a_len;
    out[2] = r[0].b_len;
    out[3] = (int64_t)sc.size();
    for (size_t i = 0; i < sc.size() && i < max_ops; ++i) ops[i] = sc[i].kind | (sc[i].len << 2);
  });
}

static_assert(sizeof(rmb_dp_half) == sizeof(DpTask) && offsetof(rmb_dp_half, m) == offsetof(DpTask, m) &&
                  offsetof(rmb_dp_half, xdrop) == offsetof(DpTask, xdrop) &&
                  offsetof(rmb_dp_half, reverse) == offsetof(DpTask, reverse),
              "rmb_dp_half is the DP kernels' task record");

int rmb_align_score_only_noalloc(rmb_ctx* c, const rmb_dp_half* halves, uint32_t n_halves, int32_t* out) {
  if (!c) return RMB_ERR_ARG;
  return guarded(c, [&]() {
    if (n_halves == 0) return;
    if (!halves || !out) throw Failure(RMB_ERR_ARG, "halves and out must be device pointers");
    if (n_halves > noalloc_max_halves())
      throw Failure(RMB_ERR_ARG, "n_halves exceeds RMB_NOALLOC_MAX_HALVES (65,536)");
    if (!c->noalloc_scratch) throw Failure(RMB_ERR_STATE, "context has no scratch address");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    uint32_t err = 0;
    hip_check(launch_noalloc_score(c->noalloc_scratch, c->cu_count, reinterpret_cast<const DpTask*>(halves),
                                   n_halves, c->d_mat, c->P.gap_open, c->P.gap_extend, out, &err, c->st),
              "no-alloc score DPs");
    if (err & E_BAND_WIDE) throw Failure(RMB_ERR_CAPACITY, "an X-drop band outgrew 1,024 cells");
  });
}

}  // extern "C"
