// HIP kernels of the learned triage (gfx950): canonical 6-mer profiles, a BF16
// MFMA GEMM with FP32 accumulation for the towers and the head, the head's
// forward/backward with the logistic loss, and Adam.
//
// Model (two towers, dot-plus-two-layer head), per (window w, family f):
//   e_w = W2w relu(W1w x_w + b1w) + b2w          x: log1p canonical 6-mer counts
//   e_f = W2f relu(W1f x_f + b1f) + b2f          (2080 -> 512 -> 256 per tower)
//   logit = a * <e_w, e_f> + c * cos(x_w, x_f) + wh2 . relu(Wh1 (e_w * e_f) + bh1) + bh2
// (head 256 -> 64 -> 1; cos is the cosine of the raw profiles)
// Matrices are held as FP32 masters and BF16 copies; GEMM operands are BF16,
// accumulation and everything else FP32.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rmb_triage_dev.h"

namespace rmb {
namespace tri {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept a NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------ profiles -----
// One workgroup per sequence window: LDS histogram of canonical 6-mers (words
// with a code >= 4 are skipped), then log1p of each count as BF16.
__global__ __launch_bounds__(256) void k_kmer_profile(const uint8_t* __restrict__ seq,
                                                      const uint64_t* __restrict__ starts,
                                                      const uint32_t* __restrict__ lens, uint32_t n,
                                                      const uint16_t* __restrict__ canon, uint16_t* __restrict__ out) {
  __shared__ uint32_t cnt[kKmerDim];
  const uint32_t w = blockIdx.x;
  for (int i = threadIdx.x; i < kKmerDim; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  const uint8_t* s = seq + starts[w];
  const uint32_t len = lens[w];
  for (uint32_t p = threadIdx.x; p + 6 <= len; p += blockDim.x) {
    uint32_t code = 0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const uint32_t b = s[p + k];
      ok = ok && b < 4;
      code = (code << 2) | (b & 3u);
    }
    if (ok) atomicAdd(&cnt[canon[code]], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kKmerDim; i += blockDim.x)
    out[(size_t)w * kKmerDim + i] = f2bf(log1pf((float)cnt[i]));
}

// --------------------------------------------------------------- GEMM ------
// C[M x N] (+)= op(A)[M x K] * op(B)[K x N] with gfx950's v_mfma_f32_32x32x16_bf16.
//   A: TA ? A[k * lda + m] : A[m * lda + k]
//   B: TB ? B[n * ldb + k] : B[k * ldb + n]      (TB: weights stored [out][in])
// 64 x 64 tile per workgroup, one 32 x 32 MFMA tile per wave, K in steps of 32
// staged through LDS (zero-filled past the edges).  Optional BF16 copy of C.
// Operand layout of the 32x32x16 MFMA: lane l holds A[l % 32][8 (l / 32) + j]
// and B[8 (l / 32) + j][l % 32], j = 0..7 (one 16-byte LDS read each); result
// register v of lane l is C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, int accumulate, uint16_t* __restrict__ Cbf) {
  constexpr int BM = 64, BN = 64, BK = 32, PAD = 8;  // 80-byte rows keep the 16-byte reads aligned
  __shared__ __align__(16) uint16_t As[BM][BK + PAD];
  __shared__ __align__(16) uint16_t Bs[BN][BK + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  for (int k0 = 0; k0 < K; k0 += BK) {
    for (int e = tid; e < BM * BK; e += 256) {
      int r, c;
      if (TA) { c = e / BM; r = e % BM; } else { r = e / BK; c = e % BK; }
      const int m = m0 + r, k = k0 + c;
      uint16_t v = 0;
      if (m < M && k < K) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      As[r][c] = v;
    }
    for (int e = tid; e < BN * BK; e += 256) {
      int r, c;
      if (TB) { r = e / BK; c = e % BK; } else { c = e / BN; r = e % BN; }
      const int nn = n0 + r, k = k0 + c;
      uint16_t v = 0;
      if (nn < N && k < K) v = TB ? B[(size_t)nn * ldb + k] : B[(size_t)k * ldb + nn];
      Bs[r][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kq = kk + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + (lane & 31)][kq]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int m = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    const int nn = n0 + wn + (lane & 31);
    if (m < M && nn < N) {
      float x = acc[v];
      if (accumulate) x += C[(size_t)m * ldc + nn];
      C[(size_t)m * ldc + nn] = x;
      if (Cbf) Cbf[(size_t)m * ldc + nn] = f2bf(x);
    }
  }
}

// ------------------------------------------------------------ elementwise ---
// X[m][n] += bias[n], optional ReLU, BF16 copy.
__global__ void k_bias_act(float* __restrict__ X, int M, int N, const float* __restrict__ bias, int relu,
                           uint16_t* __restrict__ Xbf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)M * N) return;
  float x = X[i] + bias[i % N];
  if (relu && x < 0.f) x = 0.f;
  X[i] = x;
  if (Xbf) Xbf[i] = f2bf(x);
}

// dX = post > 0 ? dX : 0 (ReLU backward from the post-activation), BF16 copy.
__global__ void k_relu_bwd(float* __restrict__ dX, const float* __restrict__ post, size_t n,
                           uint16_t* __restrict__ dXbf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = post[i] > 0.f ? dX[i] : 0.f;
  dX[i] = d;
  dXbf[i] = f2bf(d);
}

__global__ void k_to_bf16(const float* __restrict__ X, size_t n, uint16_t* __restrict__ Y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Y[i] = f2bf(X[i]);
}

// out[n] += sum over rows of X[m][n]; blockIdx.y takes a chunk of rows.
__global__ void k_colsum(const float* __restrict__ X, int M, int N, float* __restrict__ out, int rows_per_block) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int m0 = blockIdx.y * rows_per_block, m1 = min(M, m0 + rows_per_block);
  float s = 0.f;
  for (int m = m0; m < m1; ++m) s += X[(size_t)m * N + n];
  atomicAdd(&out[n], s);
}

// P[i * nf + j][d] = e_w[i][d] * e_f[j][d]
__global__ void k_pair_product(const float* __restrict__ Ew, const float* __restrict__ Ef, int nw, int nf,
                               uint16_t* __restrict__ P) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nw * nf * kEmb) return;
  const int d = (int)(t % kEmb);
  const size_t p = t / kEmb;
  const int i = (int)(p / nf), j = (int)(p % nf);
  P[t] = f2bf(Ew[(size_t)i * kEmb + d] * Ef[(size_t)j * kEmb + d]);
}

// Head, one thread per pair.  Hp holds Wh1 (e_w * e_f) (no bias), cosr the
// raw-profile cosine.  Writes the logit; with labels also the loss terms and
// the gradients: g = dL/dlogit, dH (pre-activation gradient of the hidden
// layer, FP32 and BF16) and the workgroup's partial sums of dwh2, dbh2, da and
// dc (dbh1 is left to a column sum of dH).  Loss: weighted logistic loss,
// weight pos_weight for label 1, 1 for label 0, 0 for label 2 (ignored),
// normalised by wsum.
__global__ __launch_bounds__(256) void k_head(const float* __restrict__ Hp, const float* __restrict__ Ew,
                                              const float* __restrict__ Ef, const float* __restrict__ cosr,
                                              int nw, int nf,
                                              const float* __restrict__ hp,  // bh1[64], wh2[64], bh2, a, c
                                              const uint8_t* __restrict__ labels, float pos_weight, float wsum,
                                              float* __restrict__ logits, float* __restrict__ g_out,
                                              float* __restrict__ dH, uint16_t* __restrict__ dHbf,
                                              float* __restrict__ grad_hp,  // dwh2[64], dbh2, da, dc (accumulated)
                                              float* __restrict__ loss) {
  __shared__ float s_dwh2[kHead];
  __shared__ float s_misc[3];
  __shared__ float s_loss;
  for (int k = threadIdx.x; k < kHead; k += blockDim.x) s_dwh2[k] = 0.f;
  if (threadIdx.x < 3) s_misc[threadIdx.x] = 0.f;
  if (threadIdx.x == 0) s_loss = 0.f;
  __syncthreads();
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t npairs = (size_t)nw * nf;
  if (p < npairs) {
    const int i = (int)(p / nf), j = (int)(p % nf);
    const float* bh1 = hp;
    const float* wh2 = hp + kHead;
    const float bh2 = hp[2 * kHead], a = hp[2 * kHead + 1], c = hp[2 * kHead + 2];
    float dot = 0.f;
    for (int d = 0; d < kEmb; ++d) dot += Ew[(size_t)i * kEmb + d] * Ef[(size_t)j * kEmb + d];
    float z = bh2 + a * dot + c * cosr[p];
    const float* h = Hp + p * kHead;
    for (int k = 0; k < kHead; ++k) {
      const float hk = h[k] + bh1[k];
      if (hk > 0.f) z += wh2[k] * hk;
    }
    logits[p] = z;
    if (labels) {
      const float y = labels[p] == 1 ? 1.f : 0.f;
      const float wt = labels[p] == 1 ? pos_weight : (labels[p] == 0 ? 1.f : 0.f);
      // softplus(z) - y z, computed stably
      const float l = fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)));
      const float sg = 1.f / (1.f + expf(-z));
      const float g = wt * (sg - y) / wsum;
      g_out[p] = g;
      atomicAdd(&s_loss, wt * l / wsum);
      atomicAdd(&s_misc[0], g);        // dbh2
      atomicAdd(&s_misc[1], g * dot);  // da
      atomicAdd(&s_misc[2], g * cosr[p]);  // dc
      for (int k = 0; k < kHead; ++k) {
        const float hk = h[k] + bh1[k];
        const float act = hk > 0.f ? hk : 0.f;
        atomicAdd(&s_dwh2[k], g * act);
        const float dh = hk > 0.f ? g * wh2[k] : 0.f;
        dH[p * kHead + k] = dh;
        dHbf[p * kHead + k] = f2bf(dh);
      }
    }
  }
  __syncthreads();
  if (labels) {
    for (int k = threadIdx.x; k < kHead; k += blockDim.x) atomicAdd(&grad_hp[k], s_dwh2[k]);
    if (threadIdx.x == 0) {
      atomicAdd(&grad_hp[kHead], s_misc[0]);
      atomicAdd(&grad_hp[kHead + 1], s_misc[1]);
      atomicAdd(&grad_hp[kHead + 2], s_misc[2]);
      atomicAdd(loss, s_loss);
    }
  }
}

// Row norms of BF16 profile rows.
__global__ void k_row_norm(const uint16_t* __restrict__ X, int n, float* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float s = 0.f;
  for (int k = 0; k < kKmerDim; ++k) {
    const float x = __uint_as_float((uint32_t)X[(size_t)r * kKmerDim + k] << 16);
    s += x * x;
  }
  out[r] = sqrtf(s);
}

// cos[i nf + j] = D[i][j] / (|x_w[i]| |x_f[j]|), in place on D
__global__ void k_cosine(float* __restrict__ D, const float* __restrict__ nw_norm, const float* __restrict__ nf_norm,
                         int nw, int nf) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (size_t)nw * nf) return;
  const float den = nw_norm[p / nf] * nf_norm[p % nf];
  D[p] = den > 0.f ? D[p] / den : 0.f;
}

// dE_w[i][d] = sum_j (dP[i nf + j][d] + a g[i nf + j]) e_f[j][d], and the
// same for the family side over i.
__global__ void k_embed_grad(const float* __restrict__ dP, const float* __restrict__ g, const float* __restrict__ Ew,
                             const float* __restrict__ Ef, int nw, int nf, const float* __restrict__ hp,
                             float* __restrict__ dEw, float* __restrict__ dEf) {
  const float a = hp[2 * kHead + 1];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nwd = (size_t)nw * kEmb, nfd = (size_t)nf * kEmb;
  if (t < nwd) {
    const int i = (int)(t / kEmb), d = (int)(t % kEmb);
    float s = 0.f;
    for (int j = 0; j < nf; ++j) {
      const size_t p = (size_t)i * nf + j;
      s += (dP[p * kEmb + d] + a * g[p]) * Ef[(size_t)j * kEmb + d];
    }
    dEw[t] = s;
  } else if (t < nwd + nfd) {
    const size_t u = t - nwd;
    const int j = (int)(u / kEmb), d = (int)(u % kEmb);
    float s = 0.f;
    for (int i = 0; i < nw; ++i) {
      const size_t p = (size_t)i * nf + j;
      s += (dP[p * kEmb + d] + a * g[p]) * Ew[(size_t)i * kEmb + d];
    }
    dEf[u] = s;
  }
}

// Adam on the flat FP32 parameter vector; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wbf, const float* __restrict__ grad,
                       float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float x = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = x;
  wbf[i] = f2bf(x);
}

// ------------------------------------------------------------ launchers ----
static inline unsigned nblk(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_kmer_profile(const uint8_t* seq, const uint64_t* starts, const uint32_t* lens, uint32_t n,
                               const uint16_t* canon, uint16_t* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kmer_profile, dim3(n), dim3(256), 0, st, seq, starts, lens, n, canon, out);
  return hipGetLastError();
}

hipError_t launch_gemm(bool ta, bool tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                       float* C, int ldc, bool accumulate, uint16_t* Cbf, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const dim3 grid((N + 63) / 64, (M + 63) / 64), blk(256);
  const int acc = accumulate ? 1 : 0;
  if (!ta && tb) hipLaunchKernelGGL((k_gemm_bf16<false, true>), grid, blk, 0, st, M, N, K, A, lda, B, ldb, C, ldc, acc, Cbf);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm_bf16<true, false>), grid, blk, 0, st, M, N, K, A, lda, B, ldb, C, ldc, acc, Cbf);
  else if (!ta && !tb) hipLaunchKernelGGL((k_gemm_bf16<false, false>), grid, blk, 0, st, M, N, K, A, lda, B, ldb, C, ldc, acc, Cbf);
  else hipLaunchKernelGGL((k_gemm_bf16<true, true>), grid, blk, 0, st, M, N, K, A, lda, B, ldb, C, ldc, acc, Cbf);
  return hipGetLastError();
}

hipError_t launch_bias_act(float* X, int M, int N, const float* bias, bool relu, uint16_t* Xbf, hipStream_t st) {
  const size_t n = (size_t)M * N;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_bias_act, dim3(nblk(n)), dim3(256), 0, st, X, M, N, bias, relu ? 1 : 0, Xbf);
  return hipGetLastError();
}

hipError_t launch_relu_bwd(float* dX, const float* post, size_t n, uint16_t* dXbf, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, st, dX, post, n, dXbf);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* X, size_t n, uint16_t* Y, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(256), 0, st, X, n, Y);
  return hipGetLastError();
}

hipError_t launch_colsum(const float* X, int M, int N, float* out, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int rows = 256;
  hipLaunchKernelGGL(k_colsum, dim3(nblk((size_t)N), (unsigned)((M + rows - 1) / rows)), dim3(256), 0, st, X, M, N,
                     out, rows);
  return hipGetLastError();
}

hipError_t launch_pair_product(const float* Ew, const float* Ef, int nw, int nf, uint16_t* P, hipStream_t st) {
  const size_t n = (size_t)nw * nf * kEmb;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_pair_product, dim3(nblk(n)), dim3(256), 0, st, Ew, Ef, nw, nf, P);
  return hipGetLastError();
}

hipError_t launch_head(const float* Hp, const float* Ew, const float* Ef, const float* cosr, int nw, int nf,
                       const float* hp, const uint8_t* labels, float pos_weight, float wsum, float* logits, float* g,
                       float* dH, uint16_t* dHbf, float* grad_hp, float* loss, hipStream_t st) {
  const size_t n = (size_t)nw * nf;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_head, dim3(nblk(n)), dim3(256), 0, st, Hp, Ew, Ef, cosr, nw, nf, hp, labels, pos_weight,
                     wsum, logits, g, dH, dHbf, grad_hp, loss);
  return hipGetLastError();
}

hipError_t launch_cosine(const uint16_t* Xw, int nw, const uint16_t* Xf, int nf, float* norms, float* D,
                         hipStream_t st) {
  if (!nw || !nf) return hipSuccess;
  hipError_t e = launch_gemm(false, true, nw, nf, kKmerDim, Xw, kKmerDim, Xf, kKmerDim, D, nf, false, nullptr, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_row_norm, dim3(nblk((size_t)nw)), dim3(256), 0, st, Xw, nw, norms);
  hipLaunchKernelGGL(k_row_norm, dim3(nblk((size_t)nf)), dim3(256), 0, st, Xf, nf, norms + nw);
  hipLaunchKernelGGL(k_cosine, dim3(nblk((size_t)nw * nf)), dim3(256), 0, st, D, norms, norms + nw, nw, nf);
  return hipGetLastError();
}

hipError_t launch_embed_grad(const float* dP, const float* g, const float* Ew, const float* Ef, int nw, int nf,
                             const float* hp, float* dEw, float* dEf, hipStream_t st) {
  const size_t n = (size_t)(nw + nf) * kEmb;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_embed_grad, dim3(nblk(n)), dim3(256), 0, st, dP, g, Ew, Ef, nw, nf, hp, dEw, dEf);
  return hipGetLastError();
}

hipError_t launch_adam(float* w, uint16_t* wbf, const float* grad, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, st, w, wbf, grad, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace tri
}  // namespace rmb
