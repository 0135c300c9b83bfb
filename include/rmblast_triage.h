/*
 * rmblast_triage.h — C ABI of the learned triage in librmblast_hip.so.
 *
 * The triage has no counterpart in the reference: it is the pruning step the
 * reference does not have, placed before seeding.  It scores (query window,
 * family) pairs from canonical 6-mer profiles with a two-tower MLP
 * (2080 -> 512 -> 256 per tower, dot-plus-two-layer head: 256 -> 64 -> 1, plus
 * a learned weight on the cosine of the raw profiles),
 * BF16 weights and GEMM operands with FP32 accumulation, trained on the GPU
 * by the library's own forward and backward kernels.  The caller prunes
 * whole (query record, family) pairs with its scores (a pair passes when one
 * of the record's windows does) and searches the passed pairs with
 * rmb_search_pairs (rmblast_hip.h), whose search of a pair is the plain
 * search's for that pair: every reported HSP is one the plain search of those
 * records reports, and pruning can only drop HSPs.
 *
 * Conventions as in rmblast_hip.h: int status (0 ok, 1 argument, 2 device,
 * 3 state), message from rmb_triage_last_error, nothing thrown across the
 * boundary, caller-owned host buffers.  BF16 values are passed as their
 * 16-bit patterns (the upper half of the IEEE-754 single).
 */
#ifndef RMBLAST_TRIAGE_H
#define RMBLAST_TRIAGE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMB_TRIAGE_KMER_DIM 2080

typedef struct rmb_triage rmb_triage;

/* Create a model on HIP device `device` with parameters drawn from `seed`
 * (Glorot/He uniform, zero biases). */
int rmb_triage_create(int device, uint64_t seed, rmb_triage** out);
int rmb_triage_destroy(rmb_triage* t);
int rmb_triage_last_error(const rmb_triage* t, char* buf, size_t len);

/* Flat FP32 parameter vector (window tower W1 b1 W2 b2, family tower, head
 * Wh1 bh1 wh2 bh2 a c; matrices row-major [out][in]).  set_params resets the
 * optimiser state. */
int rmb_triage_n_params(const rmb_triage* t, uint64_t* n);
int rmb_triage_get_params(const rmb_triage* t, float* out);
int rmb_triage_set_params(rmb_triage* t, const float* in);

/* Profiles of n windows of one BLASTNA sequence: window k is
 * seq[starts[k] .. starts[k] + lens[k]); out receives n x 2080 BF16 values,
 * log1p of the canonical 6-mer counts (words with an ambiguity code skipped). */
int rmb_kmer_profiles(rmb_triage* t, const uint8_t* seq, uint64_t seq_len, const uint64_t* starts,
                      const uint32_t* lens, uint32_t n, uint16_t* out);

/* Forward and backward over all nw x nf pairs of one batch (pair p = i*nf + j,
 * labels[p] 0 negative, 1 positive, 2 ignored): weighted logistic loss
 * (weight pos_weight on positives, 1 on negatives, normalised by the total
 * weight).  The gradient is kept for
 * rmb_triage_adam_step and copied to grads (n_params floats) when non-NULL. */
int rmb_triage_forward_backward(rmb_triage* t, const uint16_t* xw, uint32_t nw, const uint16_t* xf, uint32_t nf,
                                const uint8_t* labels, float pos_weight, float* loss, float* grads);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradient. */
int rmb_triage_adam_step(rmb_triage* t, float lr);

/* Logits of all nw x nf pairs (pair p = i*nf + j). */
int rmb_triage_logits(rmb_triage* t, const uint16_t* xw, uint32_t nw, const uint16_t* xf, uint32_t nf,
                      float* logits);

/* Parity tap of the BF16 MFMA GEMM the model runs on:
 * C[M x N] = op(A)[M x K] * op(B)[K x N], FP32 result;
 * A is TA ? K rows of lda : M rows of lda, B is TB ? N rows of ldb : K rows of ldb. */
int rmb_triage_tap_gemm(rmb_triage* t, int ta, int tb, int M, int N, int K, const uint16_t* A, int lda,
                        const uint16_t* B, int ldb, float* C);

#ifdef __cplusplus
}
#endif
#endif
