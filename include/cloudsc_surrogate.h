/*
 * cloudsc_surrogate.h -- learned surrogate of CLOUDSC and the hybrid dispatcher (gfx950).
 *
 * No counterpart in the reference: OpenIFS has no learned microphysics.  The surrogate is a per-column
 * multilayer perceptron (BF16 weights and activations, FP32 accumulation on the matrix cores, FP32 master
 * weights and Adam) trained on input/output pairs of the plain CLOUDSC kernel (cloudsc_blocked_sp_c).
 * The hybrid entry runs the network on every column and reruns a seeded random fraction of columns, plus
 * every column its gate flags, through the plain kernel, whose results replace the network's.
 *
 * All pointers to arrays are device pointers unless a comment says otherwise.  Status codes are those of
 * cloudsc_hip.h; cloudsc_last_error() gives the text.
 */
#ifndef CLOUDSC_SURROGATE_H
#define CLOUDSC_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Generic BF16 MLP: x[n_in] -> (Linear, SiLU) x n_hidden_layers -> Linear -> y[n_out].
 * Widths and batches are multiples of 128.  BF16 buffers are uint16_t (bfloat16 bit patterns).
 * ------------------------------------------------------------------------- */
typedef struct cloudsc_mlp cloudsc_mlp_t;

int cloudsc_mlp_create(cloudsc_mlp_t **out, int n_in, int n_hidden, int n_hidden_layers, int n_out, int max_batch,
                       uint64_t seed);
int cloudsc_mlp_destroy(cloudsc_mlp_t *m);
/* n_layers = n_hidden_layers + 1; din/dout (may be NULL) receive each linear layer's widths. */
int cloudsc_mlp_dims(const cloudsc_mlp_t *m, int *n_layers, int *din, int *dout);
/* y[M x n_out] (fp32) = network(x[M x n_in] bf16). */
int cloudsc_mlp_forward(cloudsc_mlp_t *m, const uint16_t *x, int M, float *y, void *stream);
/* Forward + backward for L = sum_{m,n} wgt[n] (y[m,n] - t[m,n])^2 / M; x and its transpose xT[n_in x M]. */
int cloudsc_mlp_gradients(cloudsc_mlp_t *m, const uint16_t *x, const uint16_t *xT, const float *t, const float *wgt,
                          int M, void *stream);
/* One Adam step on the gradients of the last cloudsc_mlp_gradients call. */
int cloudsc_mlp_adam(cloudsc_mlp_t *m, float lr, float beta1, float beta2, float eps, void *stream);
/* cloudsc_mlp_gradients + cloudsc_mlp_adam(lr, 0.9, 0.999, 1e-8). */
int cloudsc_mlp_train_step(cloudsc_mlp_t *m, const uint16_t *x, const uint16_t *xT, const float *t, const float *wgt,
                           int M, float lr, void *stream);
/* Host copy of a parameter: which 0 = W [dout x din], 1 = b, 2 = dW, 3 = db, 4 = Y of the last training forward,
 * 5 = loss of the last cloudsc_mlp_gradients; set = 1 writes W or b from host (synchronous). */
int cloudsc_mlp_param(cloudsc_mlp_t *m, int layer, int which, float *host, size_t n, int set);
/* Test product: C[M x N] = A[M x K] * B[N x K]^T (bf16 in, fp32 out, A and B K-contiguous). */
int cloudsc_mlp_gemm_tn(const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C, int ldc, int M, int N,
                        int K, void *stream);

/* ---------------------------------------------------------------------------
 * CLOUDSC surrogate.
 *
 * Features per column: 14 per level, feature-major (index f*KLEV + k), zero-padded to a multiple of 128:
 *   0 PT, 1 PQ, 2 PA, 3-6 log10(max(PCLV(l,i,r,s),0) + 1e-10), 7 PTENDENCY_CML_T, 8 PTENDENCY_CML_Q,
 *   9 PTENDENCY_CML_A, 10-11 PTENDENCY_CML_CLD(l,i), 12 PQ / qsat_mix(PT, PAP) (relative humidity over the
 *   mixed phase, FOEEWM), 13 log10(max(PLUDE,0) + 1e-10).
 * Targets per column: 6 per level, target-major (index t*KLEV + k), zero-padded to a multiple of 128:
 *   0-2 PTENDENCY_LOC_T/Q/A, 3-4 PTENDENCY_LOC_CLD(l,i), 5 PFPLSL+PFPLSN at half level k+1.
 * Both are standardised per (feature or target, level) with the statistics held by the object.
 * ------------------------------------------------------------------------- */
#define CLOUDSC_SURR_NFEAT 14
#define CLOUDSC_SURR_NTARG 6

typedef struct cloudsc_surrogate cloudsc_surrogate_t;

int cloudsc_surrogate_create(cloudsc_surrogate_t **out, int klev, int n_hidden, int n_hidden_layers, int max_batch,
                             uint64_t seed);
int cloudsc_surrogate_destroy(cloudsc_surrogate_t *s);
cloudsc_mlp_t *cloudsc_surrogate_mlp(cloudsc_surrogate_t *s);
/* Padded widths: n_in = ceil128(14*KLEV), n_out = ceil128(6*KLEV). */
int cloudsc_surrogate_dims(const cloudsc_surrogate_t *s, int *klev, int *n_in, int *n_out, int *max_batch);
/* Normalisation statistics (host arrays): mean and scale of the 14*KLEV raw features and the 6*KLEV raw targets,
 * and the training weight of each target entry (wgt_y may be NULL).  get = 1 copies the object's values out. */
int cloudsc_surrogate_norm(cloudsc_surrogate_t *s, float *mu_x, float *sd_x, float *mu_y, float *sd_y, float *wgt_y,
                           int get);
/* Computes the statistics from a blocked field set holding plain-kernel inputs and outputs (all NGPTOT columns):
 * means and standard deviations per entry; constant features are left unscaled and other features get at least
 * 1e-3 of their largest level spread; target entries whose spread is below 1e-3 of their target's largest level
 * spread get training weight 0 and that floor as scale.  Standardised values are clipped to +-10. */
int cloudsc_surrogate_fit_norm(cloudsc_surrogate_t *s, CLOUDSC_BLOCKED_ARGS(float));
/* Trains on all columns of a blocked field set holding plain-kernel inputs and outputs, one Adam step per
 * consecutive batch of `batch` columns (a multiple of 128, <= max_batch; a trailing partial batch is skipped).
 * Returns the mean training loss over the batches in *loss (host). */
int cloudsc_surrogate_train(cloudsc_surrogate_t *s, float lr, int batch, float *loss, CLOUDSC_BLOCKED_ARGS(float));
/* Network-only CLOUDSC: writes PTENDENCY_LOC_T/Q/A, PTENDENCY_LOC_CLD(l,i) and PFPLSL/PFPLSN/PFHPSL/PFHPSN
 * from the network; PFPLSL/PFPLSN split the predicted total by the phase of the level above (liquid where
 * T > RTT); PTENDENCY_LOC_CLD(r,s), PRAINFRAC_TOPRFZ, the ten DDH flux profiles and PCOVPTOT at levels
 * JK >= NCLDTOP (the levels CLOUDSC writes) are set to zero; PLUDE is left unchanged. */
int cloudsc_surrogate_predict_sp(cloudsc_surrogate_t *s, CLOUDSC_BLOCKED_ARGS(float));
/* Hybrid: cloudsc_surrogate_predict_sp on every column, then the plain kernel (exact outputs, every array
 * cloudsc_blocked_sp_c writes) on the columns whose Philox(rerun_seed, column) uniform is below rerun_fraction
 * or whose gate score exceeds gate_threshold.  Gate score of a column: the largest q / qsat_mix(T, p) over the
 * levels with T < 238 K (FOEEWM, fcttre.func.h).  *n_plain (host, may be NULL) receives the number of
 * columns run through the plain kernel. */
int cloudsc_hybrid_sp_c(cloudsc_surrogate_t *s, float rerun_fraction, uint64_t rerun_seed, float gate_threshold,
                        int *n_plain, CLOUDSC_BLOCKED_ARGS(float));
/* Gate scores of all columns (fp32 [NGPTOT]), for calibrating gate_threshold. */
int cloudsc_surrogate_gate_scores(cloudsc_surrogate_t *s, float *score, CLOUDSC_BLOCKED_ARGS(float));

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_SURROGATE_H */
