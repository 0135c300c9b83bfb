// Input preparation on the device (gfx950): ambiguity resolution of the
// subject library and the reverse-complement pools of subjects and queries.
//
//  k_resolve_ambiguity  one lane per subject: makeblastdb's lagged-Fibonacci
//                       generator (seq/twobit.rs:50-134), seeded with the
//                       subject's length, one draw per ambiguous base 5' to 3',
//                       patches the resolved plus strand in place
//  k_revcomp_segments   the reverse complement of every sentinel-wrapped
//                       segment of a pool (sentinels map to sentinels)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rmb_device.h"

namespace rmb {

// BLASTNA complement (A<->T, C<->G, and the IUPAC codes with them; N, sentinel
// and gap stay)
__constant__ uint8_t c_prep_complement[16] = {3, 2, 1, 0, 5, 4, 7, 6, 8, 9, 13, 12, 11, 10, 14, 15};
// candidate bases per ambiguity code: count, then the bases (N: any of 4)
__constant__ uint8_t c_amb_sets[16][4] = {{0}, {0}, {0}, {0}, {2, 0, 2}, {2, 1, 3}, {2, 0, 1}, {2, 2, 3},
                                          {2, 0, 3}, {2, 1, 2}, {3, 1, 2, 3}, {3, 0, 2, 3}, {3, 0, 1, 3},
                                          {3, 0, 1, 2}, {4}, {0}};

constexpr uint32_t kResolveThreads = 64;

// plus holds the wrapped raw codes on entry; ambiguous bases (codes 4..14) of
// subject i are replaced by the generator's choices.  The generator's 33
// words live in LDS, one slice per lane.
__global__ __launch_bounds__(kResolveThreads) void k_resolve_ambiguity(uint8_t* __restrict__ plus,
                                                                      const SubjDesc* __restrict__ sd, uint32_t n) {
  __shared__ uint32_t st[33][kResolveThreads];
  const uint32_t t = threadIdx.x, i = blockIdx.x * kResolveThreads + t;
  if (i >= n) return;
  const SubjDesc S = sd[i];
  uint8_t* p = plus + S.s_base + 1;
  bool seeded = false;
  int j = 12, k = 32;
  auto next = [&]() -> uint32_t {
    const uint32_t r = st[k][t] + st[j][t];
    st[k][t] = r;
    --k;
    --j;
    if (k < 0) k = 32;
    else if (j < 0) j = 32;
    return r;
  };
  for (uint32_t x = 0; x < S.s_len; ++x) {
    const uint8_t code = p[x];
    if (code < 4 || code > 14) continue;
    if (!seeded) {  // the generator starts at the subject's first ambiguous base
      st[0][t] = S.s_len;
      for (int q = 1; q < 33; ++q) st[q][t] = 1103515245u * st[q - 1][t] + 12345u;
      for (int q = 0; q < 330; ++q) next();
      seeded = true;
    }
    const uint32_t r = next() >> 1;
    p[x] = code == 14 ? (uint8_t)(r & 3u) : c_amb_sets[code][1 + r % c_amb_sets[code][0]];
  }
}

// out[b + q] = complement(in[b + len + 1 - q]) for q in [0, len + 2) of each
// segment (b, len); work items are (segment, first byte) pairs of at most
// kRevcompChunk bytes.
constexpr uint32_t kRevcompChunk = 16384;
__global__ __launch_bounds__(256) void k_revcomp_segments(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                         const uint64_t* __restrict__ seg_base,
                                                         const uint32_t* __restrict__ seg_len,
                                                         const uint2* __restrict__ work, uint32_t n_work) {
  const uint32_t w = blockIdx.x;
  if (w >= n_work) return;
  const uint2 it = work[w];
  const uint64_t b = seg_base[it.x];
  const uint32_t span = seg_len[it.x] + 2;
  const uint32_t q1 = it.y + kRevcompChunk < span ? it.y + kRevcompChunk : span;
  for (uint32_t q = it.y + threadIdx.x; q < q1; q += blockDim.x) out[b + q] = c_prep_complement[in[b + span - 1 - q] & 15u];
}

hipError_t launch_resolve_ambiguity(uint8_t* plus, const SubjDesc* sd, uint32_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_resolve_ambiguity, dim3((n + kResolveThreads - 1) / kResolveThreads), dim3(kResolveThreads), 0,
                     st, plus, sd, n);
  return hipGetLastError();
}

// The (segment, first byte) work list of k_revcomp_segments for segments of
// the given lengths (bases, sentinels excluded).
uint32_t revcomp_chunk() { return kRevcompChunk; }

hipError_t launch_revcomp_segments(const uint8_t* in, uint8_t* out, const uint64_t* seg_base, const uint32_t* seg_len,
                                   const uint2* work, uint32_t n_work, hipStream_t st) {
  if (!n_work) return hipSuccess;
  hipLaunchKernelGGL(k_revcomp_segments, dim3(n_work), dim3(256), 0, st, in, out, seg_base, seg_len, work, n_work);
  return hipGetLastError();
}

}  // namespace rmb
