/*
 * cloudsc_synth.h -- deterministic synthetic atmospheres for CLOUDSC tests
 * and benchmarks.  No counterpart in the reference (which drives CLOUDSC
 * from a full model state); provided for callers that need reproducible
 * CLOUDSC inputs of a given size: parity tests, bench.py, training-data
 * generation.  Fields are written in the blocked layout of
 * cloudsc_blocked_*_c (see cloudsc_hip.h); every value depends only on
 * (seed, field, global column, level), so host and device generation agree
 * to the last bit of the transcendental functions they call.
 */
#ifndef CLOUDSC_SYNTH_H
#define CLOUDSC_SYNTH_H

#include <stdint.h>

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cloudsc_synth_desc_t {
  uint64_t seed;
  int32_t ngptot;   /* columns (the padded tail of the last block is filled too) */
  int32_t nproma;
  int32_t klev;
  int32_t ngpblks;  /* must equal ceil(ngptot/nproma) */
  const double *ahalf, *bhalf; /* hybrid coefficients of the KLEV+1 half levels */
} cloudsc_synth_desc_t;

#define CLOUDSC_SYNTH_FIELDS(R)                                                                     \
  R *pt, R *pq, R *ptendency_cml_t, R *ptendency_cml_q, R *ptendency_cml_a, R *ptendency_cml_cld,   \
      R *pvfl, R *pvfi, R *pap, R *paph, R *plsm, int32_t *ldcum, int32_t *ktype, R *peis, R *plu,  \
      R *plude, R *pludeli, R *psnde, R *pmfu, R *pmfd, R *pa, R *pclv

/* Device generation (all pointers device memory; ahalf/bhalf device too). Asynchronous on stream. */
int cloudsc_synth_sp(const cloudsc_params_t *p, const cloudsc_synth_desc_t *d, CLOUDSC_SYNTH_FIELDS(float),
                     void *stream);
int cloudsc_synth_dp(const cloudsc_params_t *p, const cloudsc_synth_desc_t *d, CLOUDSC_SYNTH_FIELDS(double),
                     void *stream);
/* Host generation of global columns [col0, col0+ncol) (all pointers host memory, blocked layout). */
int cloudsc_synth_host_sp(const cloudsc_params_t *p, const cloudsc_synth_desc_t *d, int64_t col0, int64_t ncol,
                          CLOUDSC_SYNTH_FIELDS(float));
int cloudsc_synth_host_dp(const cloudsc_params_t *p, const cloudsc_synth_desc_t *d, int64_t col0, int64_t ncol,
                          CLOUDSC_SYNTH_FIELDS(double));
/* Philox4x32-10 known-answer check: out[4] = philox(ctr[4], key[2]). */
void cloudsc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_SYNTH_H */
