// Host side of the learned triage: parameters and optimiser state on the
// device, forward/backward orchestration over the kernels of
// triage_kernels.hip, and the C ABI of include/rmblast_triage.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rmblast_triage.h"
#include "rmb_triage_dev.h"

namespace rmb {
namespace tri {

hipError_t launch_kmer_profile(const uint8_t*, const uint64_t*, const uint32_t*, uint32_t, const uint16_t*, uint16_t*,
                               hipStream_t);
hipError_t launch_gemm(bool, bool, int, int, int, const uint16_t*, int, const uint16_t*, int, float*, int, bool,
                       uint16_t*, hipStream_t);
hipError_t launch_bias_act(float*, int, int, const float*, bool, uint16_t*, hipStream_t);
hipError_t launch_relu_bwd(float*, const float*, size_t, uint16_t*, hipStream_t);
hipError_t launch_to_bf16(const float*, size_t, uint16_t*, hipStream_t);
hipError_t launch_colsum(const float*, int, int, float*, hipStream_t);
hipError_t launch_pair_product(const float*, const float*, int, int, uint16_t*, hipStream_t);
hipError_t launch_head(const float*, const float*, const float*, const float*, int, int, const float*, const uint8_t*,
                       float, float, float*, float*, float*, uint16_t*, float*, float*, hipStream_t);
hipError_t launch_cosine(const uint16_t*, int, const uint16_t*, int, float*, float*, hipStream_t);
hipError_t launch_embed_grad(const float*, const float*, const float*, const float*, int, int, const float*, float*,
                             float*, hipStream_t);
hipError_t launch_adam(float*, uint16_t*, const float*, float*, float*, size_t, float, float, float, float, float,
                       float, hipStream_t);

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static void ck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(2, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct Buf {
  T* p = nullptr;
  size_t cap = 0;
  T* get(size_t n) {
    if (n > cap) {
      if (p) ck(hipFree(p), "hipFree");
      p = nullptr;
      ck(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
      cap = n;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Canonical 6-mer index: min(word, reverse complement) ranked among all
// canonical words (4096 words -> 2080 classes).
static std::vector<uint16_t> canonical_map() {
  std::vector<uint32_t> canon(4096);
  std::vector<uint32_t> classes;
  for (uint32_t x = 0; x < 4096; ++x) {
    uint32_t rc = 0;
    for (int i = 0; i < 6; ++i) rc = (rc << 2) | (3u - ((x >> (2 * i)) & 3u));
    canon[x] = std::min(x, rc);
    if (canon[x] == x) classes.push_back(x);
  }
  if (classes.size() != (size_t)kKmerDim) throw Error(2, "canonical 6-mer table size");
  std::vector<uint16_t> out(4096);
  for (uint32_t x = 0; x < 4096; ++x)
    out[x] = (uint16_t)(std::lower_bound(classes.begin(), classes.end(), canon[x]) - classes.begin());
  return out;
}

static inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace tri
}  // namespace rmb

using namespace rmb::tri;

struct rmb_triage {
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  float *w = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wbf = nullptr;
  uint16_t* canon = nullptr;
  uint64_t step = 0;
  bool have_grad = false;
  // scratch
  Buf<uint16_t> xw, xf, P, h1bf_w, h1bf_f, dHbf, dEbf, dH1bf, prof;
  Buf<float> h1_w, h1_f, e_w, e_f, Hp, logits, g, dH, dP, dEw, dEf, dH1, loss, cosr, norms;
  Buf<uint8_t> labels, seq;
  Buf<uint64_t> starts;
  Buf<uint32_t> lens;
};

namespace {

template <class F>
int guarded(rmb_triage* t, F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    if (t) t->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (t) t->err = e.what();
    return 2;
  } catch (...) {
    if (t) t->err = "unknown error";
    return 2;
  }
}

// tower t over n rows of X: H1 (post-ReLU, FP32 + BF16) and E (FP32)
void tower_forward(rmb_triage* t, int tw, const uint16_t* X, int n, float* H1, uint16_t* H1bf, float* E) {
  ck(launch_gemm(false, true, n, kH1, kKmerDim, X, kKmerDim, t->wbf + Layout::w1(tw), kKmerDim, H1, kH1, false,
                 nullptr, t->st), "gemm w1");
  ck(launch_bias_act(H1, n, kH1, t->w + Layout::b1(tw), true, H1bf, t->st), "bias relu");
  ck(launch_gemm(false, true, n, kEmb, kH1, H1bf, kH1, t->wbf + Layout::w2(tw), kH1, E, kEmb, false, nullptr, t->st),
     "gemm w2");
  ck(launch_bias_act(E, n, kEmb, t->w + Layout::b2(tw), false, nullptr, t->st), "bias");
}

void tower_backward(rmb_triage* t, int tw, const uint16_t* X, int n, const float* H1, const uint16_t* H1bf,
                    const float* dE) {
  uint16_t* dEbf = t->dEbf.get((size_t)n * kEmb);
  float* dH1 = t->dH1.get((size_t)n * kH1);
  uint16_t* dH1bf = t->dH1bf.get((size_t)n * kH1);
  ck(launch_to_bf16(dE, (size_t)n * kEmb, dEbf, t->st), "to bf16");
  ck(launch_colsum(dE, n, kEmb, t->grad + Layout::b2(tw), t->st), "db2");
  ck(launch_gemm(true, false, kEmb, kH1, n, dEbf, kEmb, H1bf, kH1, t->grad + Layout::w2(tw), kH1, false, nullptr,
                 t->st), "dW2");
  ck(launch_gemm(false, false, n, kH1, kEmb, dEbf, kEmb, t->wbf + Layout::w2(tw), kH1, dH1, kH1, false, nullptr,
                 t->st), "dH1");
  ck(launch_relu_bwd(dH1, H1, (size_t)n * kH1, dH1bf, t->st), "relu bwd");
  ck(launch_colsum(dH1, n, kH1, t->grad + Layout::b1(tw), t->st), "db1");
  ck(launch_gemm(true, false, kH1, kKmerDim, n, dH1bf, kH1, X, kKmerDim, t->grad + Layout::w1(tw), kKmerDim, false,
                 nullptr, t->st), "dW1");
}

// forward over all pairs; with labels also the backward pass into t->grad
void run(rmb_triage* t, const uint16_t* hxw, uint32_t nw, const uint16_t* hxf, uint32_t nf, const uint8_t* hlab,
         float pos_weight, float* loss_out, float* logits_out) {
  if (nw == 0 || nf == 0) throw Error(1, "empty batch");
  const size_t np = (size_t)nw * nf;
  if (np > (1ull << 31) / kEmb) throw Error(1, "batch too large");
  ck(hipSetDevice(t->device), "hipSetDevice");
  uint16_t* xw = t->xw.get((size_t)nw * kKmerDim);
  uint16_t* xf = t->xf.get((size_t)nf * kKmerDim);
  ck(hipMemcpyAsync(xw, hxw, (size_t)nw * kKmerDim * 2, hipMemcpyHostToDevice, t->st), "copy xw");
  ck(hipMemcpyAsync(xf, hxf, (size_t)nf * kKmerDim * 2, hipMemcpyHostToDevice, t->st), "copy xf");
  float* h1w = t->h1_w.get((size_t)nw * kH1);
  float* h1f = t->h1_f.get((size_t)nf * kH1);
  uint16_t* h1bw = t->h1bf_w.get((size_t)nw * kH1);
  uint16_t* h1bf = t->h1bf_f.get((size_t)nf * kH1);
  float* ew = t->e_w.get((size_t)nw * kEmb);
  float* ef = t->e_f.get((size_t)nf * kEmb);
  tower_forward(t, 0, xw, (int)nw, h1w, h1bw, ew);
  tower_forward(t, 1, xf, (int)nf, h1f, h1bf, ef);
  uint16_t* P = t->P.get(np * kEmb);
  float* Hp = t->Hp.get(np * kHead);
  float* logits = t->logits.get(np);
  float* cosr = t->cosr.get(np);
  ck(launch_cosine(xw, (int)nw, xf, (int)nf, t->norms.get((size_t)nw + nf), cosr, t->st), "cosine");
  ck(launch_pair_product(ew, ef, (int)nw, (int)nf, P, t->st), "pair product");
  ck(launch_gemm(false, true, (int)np, kHead, kEmb, P, kEmb, t->wbf + Layout::wh1(), kEmb, Hp, kHead, false, nullptr,
                 t->st), "gemm wh1");
  uint8_t* lab = nullptr;
  float wsum = 1.f;
  float *g = nullptr, *dH = nullptr, *loss = nullptr;
  uint16_t* dHbf = nullptr;
  if (hlab) {
    lab = t->labels.get(np);
    ck(hipMemcpyAsync(lab, hlab, np, hipMemcpyHostToDevice, t->st), "copy labels");
    size_t npos = 0, nneg = 0;
    for (size_t p = 0; p < np; ++p) {
      npos += hlab[p] == 1 ? 1 : 0;
      nneg += hlab[p] == 0 ? 1 : 0;
    }
    wsum = std::max(1.f, (float)npos * pos_weight + (float)nneg);
    g = t->g.get(np);
    dH = t->dH.get(np * kHead);
    dHbf = t->dHbf.get(np * kHead);
    loss = t->loss.get(1);
    ck(hipMemsetAsync(t->grad, 0, Layout::TOTAL * sizeof(float), t->st), "zero grad");
    ck(hipMemsetAsync(loss, 0, sizeof(float), t->st), "zero loss");
  }
  ck(launch_head(Hp, ew, ef, cosr, (int)nw, (int)nf, t->w + Layout::hp(), lab, pos_weight, wsum, logits, g, dH, dHbf,
                 t->grad + Layout::hp() + kHead, loss, t->st), "head");
  if (hlab) {
    // head backward
    ck(launch_colsum(dH, (int)np, kHead, t->grad + Layout::hp(), t->st), "dbh1");
    ck(launch_gemm(true, false, kHead, kEmb, (int)np, dHbf, kHead, P, kEmb, t->grad + Layout::wh1(), kEmb, false,
                   nullptr, t->st), "dWh1");
    float* dP = t->dP.get(np * kEmb);
    ck(launch_gemm(false, false, (int)np, kEmb, kHead, dHbf, kHead, t->wbf + Layout::wh1(), kEmb, dP, kEmb, false,
                   nullptr, t->st), "dP");
    float* dEw = t->dEw.get((size_t)nw * kEmb);
    float* dEf = t->dEf.get((size_t)nf * kEmb);
    ck(launch_embed_grad(dP, g, ew, ef, (int)nw, (int)nf, t->w + Layout::hp(), dEw, dEf, t->st), "embed grad");
    tower_backward(t, 0, xw, (int)nw, h1w, h1bw, dEw);
    tower_backward(t, 1, xf, (int)nf, h1f, h1bf, dEf);
    ck(hipMemcpyAsync(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost, t->st), "copy loss");
    t->have_grad = true;
  }
  if (logits_out) ck(hipMemcpyAsync(logits_out, logits, np * sizeof(float), hipMemcpyDeviceToHost, t->st), "logits");
  ck(hipStreamSynchronize(t->st), "sync");
}

void init_params(rmb_triage* t, uint64_t seed) {
  std::vector<float> h(Layout::TOTAL, 0.f);
  uint64_t s = seed;
  auto uni = [&](float a) { return (float)(((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
  for (int tw = 0; tw < 2; ++tw) {
    const float a1 = std::sqrt(6.f / kKmerDim), a2 = std::sqrt(6.f / (kH1 + kEmb));
    for (size_t i = 0; i < Layout::W1; ++i) h[Layout::w1(tw) + i] = uni(a1);
    for (size_t i = 0; i < Layout::W2; ++i) h[Layout::w2(tw) + i] = uni(a2);
  }
  const float ah = std::sqrt(6.f / kEmb), ao = std::sqrt(6.f / (kHead + 1));
  for (size_t i = 0; i < (size_t)kHead * kEmb; ++i) h[Layout::wh1() + i] = uni(ah);
  for (int k = 0; k < kHead; ++k) h[Layout::hp() + kHead + k] = uni(ao);
  h[Layout::hp() + 2 * kHead] = -2.f;  // bh2: start near a 12% prior
  h[Layout::hp() + 2 * kHead + 1] = 0.f;  // a
  h[Layout::hp() + 2 * kHead + 2] = 0.f;  // c
  rmb_triage_set_params(t, h.data());
}

}  // namespace

extern "C" {

int rmb_triage_create(int device, uint64_t seed, rmb_triage** out) {
  if (!out) return 1;
  *out = nullptr;
  rmb_triage* t = new rmb_triage();
  t->device = device;
  const int rc = guarded(t, [&]() {
    int n = 0;
    ck(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw Error(2, "no such HIP device");
    ck(hipSetDevice(device), "hipSetDevice");
    ck(hipStreamCreate(&t->st), "hipStreamCreate");
    ck(hipMalloc(&t->w, Layout::TOTAL * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&t->grad, Layout::TOTAL * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&t->m, Layout::TOTAL * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&t->v, Layout::TOTAL * sizeof(float)), "hipMalloc");
    ck(hipMalloc(&t->wbf, Layout::TOTAL * sizeof(uint16_t)), "hipMalloc");
    const std::vector<uint16_t> cm = canonical_map();
    ck(hipMalloc(&t->canon, cm.size() * sizeof(uint16_t)), "hipMalloc");
    ck(hipMemcpy(t->canon, cm.data(), cm.size() * sizeof(uint16_t), hipMemcpyHostToDevice), "copy canon");
    init_params(t, seed);
  });
  if (rc) {
    rmb_triage_destroy(t);
    return rc;
  }
  *out = t;
  return 0;
}

int rmb_triage_destroy(rmb_triage* t) {
  if (!t) return 0;
  (void)hipSetDevice(t->device);
  for (void* p : {(void*)t->w, (void*)t->grad, (void*)t->m, (void*)t->v, (void*)t->wbf, (void*)t->canon})
    if (p) (void)hipFree(p);
  t->xw.release(); t->xf.release(); t->P.release(); t->h1bf_w.release(); t->h1bf_f.release(); t->dHbf.release();
  t->dEbf.release(); t->dH1bf.release(); t->prof.release(); t->h1_w.release(); t->h1_f.release(); t->e_w.release();
  t->e_f.release(); t->cosr.release(); t->norms.release(); t->Hp.release(); t->logits.release(); t->g.release(); t->dH.release(); t->dP.release();
  t->dEw.release(); t->dEf.release(); t->dH1.release(); t->loss.release(); t->labels.release(); t->seq.release();
  t->starts.release(); t->lens.release();
  if (t->st) (void)hipStreamDestroy(t->st);
  delete t;
  return 0;
}

int rmb_triage_last_error(const rmb_triage* t, char* buf, size_t len) {
  if (!t || !buf || !len) return 1;
  std::snprintf(buf, len, "%s", t->err.c_str());
  return 0;
}

int rmb_triage_n_params(const rmb_triage* t, uint64_t* n) {
  if (!t || !n) return 1;
  *n = Layout::TOTAL;
  return 0;
}

int rmb_triage_get_params(const rmb_triage* t, float* out) {
  if (!t || !out) return 1;
  return guarded(const_cast<rmb_triage*>(t), [&]() {
    ck(hipSetDevice(t->device), "hipSetDevice");
    ck(hipMemcpy(out, t->w, Layout::TOTAL * sizeof(float), hipMemcpyDeviceToHost), "copy params");
  });
}

int rmb_triage_set_params(rmb_triage* t, const float* in) {
  if (!t || !in) return 1;
  return guarded(t, [&]() {
    ck(hipSetDevice(t->device), "hipSetDevice");
    ck(hipMemcpy(t->w, in, Layout::TOTAL * sizeof(float), hipMemcpyHostToDevice), "copy params");
    ck(launch_to_bf16(t->w, Layout::TOTAL, t->wbf, t->st), "bf16 copy");
    ck(hipMemsetAsync(t->m, 0, Layout::TOTAL * sizeof(float), t->st), "zero m");
    ck(hipMemsetAsync(t->v, 0, Layout::TOTAL * sizeof(float), t->st), "zero v");
    ck(hipStreamSynchronize(t->st), "sync");
    t->step = 0;
    t->have_grad = false;
  });
}

int rmb_kmer_profiles(rmb_triage* t, const uint8_t* seq, uint64_t seq_len, const uint64_t* starts,
                      const uint32_t* lens, uint32_t n, uint16_t* out) {
  if (!t || (n && (!seq || !starts || !lens || !out))) return 1;
  return guarded(t, [&]() {
    for (uint32_t k = 0; k < n; ++k)
      if (starts[k] > seq_len || lens[k] > seq_len - starts[k]) throw Error(1, "window outside the sequence");
    if (!n) return;
    ck(hipSetDevice(t->device), "hipSetDevice");
    uint8_t* ds = t->seq.get(seq_len + 8);
    uint64_t* dst = t->starts.get(n);
    uint32_t* dl = t->lens.get(n);
    uint16_t* dout = t->prof.get((size_t)n * kKmerDim);
    ck(hipMemcpyAsync(ds, seq, seq_len, hipMemcpyHostToDevice, t->st), "copy seq");
    ck(hipMemcpyAsync(dst, starts, n * sizeof(uint64_t), hipMemcpyHostToDevice, t->st), "copy starts");
    ck(hipMemcpyAsync(dl, lens, n * sizeof(uint32_t), hipMemcpyHostToDevice, t->st), "copy lens");
    ck(launch_kmer_profile(ds, dst, dl, n, t->canon, dout, t->st), "k_kmer_profile");
    ck(hipMemcpyAsync(out, dout, (size_t)n * kKmerDim * 2, hipMemcpyDeviceToHost, t->st), "copy profiles");
    ck(hipStreamSynchronize(t->st), "sync");
  });
}

int rmb_triage_forward_backward(rmb_triage* t, const uint16_t* xw, uint32_t nw, const uint16_t* xf, uint32_t nf,
                                const uint8_t* labels, float pos_weight, float* loss, float* grads) {
  if (!t || !xw || !xf || !labels || !loss) return 1;
  return guarded(t, [&]() {
    run(t, xw, nw, xf, nf, labels, pos_weight, loss, nullptr);
    if (grads) ck(hipMemcpy(grads, t->grad, Layout::TOTAL * sizeof(float), hipMemcpyDeviceToHost), "copy grads");
  });
}

int rmb_triage_adam_step(rmb_triage* t, float lr) {
  if (!t) return 1;
  if (!t->have_grad) {
    t->err = "adam step without a gradient";
    return 3;
  }
  return guarded(t, [&]() {
    ck(hipSetDevice(t->device), "hipSetDevice");
    ++t->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)t->step), c2 = 1.f - std::pow(b2, (float)t->step);
    ck(launch_adam(t->w, t->wbf, t->grad, t->m, t->v, Layout::TOTAL, lr, b1, b2, 1e-8f, c1, c2, t->st), "adam");
    ck(hipStreamSynchronize(t->st), "sync");
  });
}

int rmb_triage_logits(rmb_triage* t, const uint16_t* xw, uint32_t nw, const uint16_t* xf, uint32_t nf,
                      float* logits) {
  if (!t || !xw || !xf || !logits) return 1;
  return guarded(t, [&]() { run(t, xw, nw, xf, nf, nullptr, 1.f, nullptr, logits); });
}

int rmb_triage_tap_gemm(rmb_triage* t, int ta, int tb, int M, int N, int K, const uint16_t* A, int lda,
                        const uint16_t* B, int ldb, float* C) {
  if (!t || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return 1;
  if (lda < (ta ? M : K) || ldb < (tb ? K : N)) return 1;
  return guarded(t, [&]() {
    ck(hipSetDevice(t->device), "hipSetDevice");
    const size_t na = (size_t)(ta ? K : M) * lda, nb = (size_t)(tb ? N : K) * ldb, nc = (size_t)M * N;
    Buf<uint16_t> a, b;
    Buf<float> c;
    uint16_t* da = a.get(na);
    uint16_t* db = b.get(nb);
    float* dc = c.get(nc);
    ck(hipMemcpy(da, A, na * 2, hipMemcpyHostToDevice), "copy A");
    ck(hipMemcpy(db, B, nb * 2, hipMemcpyHostToDevice), "copy B");
    ck(launch_gemm(ta != 0, tb != 0, M, N, K, da, lda, db, ldb, dc, N, false, nullptr, t->st), "gemm");
    ck(hipStreamSynchronize(t->st), "sync");
    ck(hipMemcpy(C, dc, nc * sizeof(float), hipMemcpyDeviceToHost), "copy C");
    a.release();
    b.release();
    c.release();
  });
}

}  // extern "C"
