/*
 * cloudsc_satadj.h -- C-ABI of the gfx950 CLOUD_SATADJ (cloud saturation adjustment).
 *
 * Reference interface replaced: SUBROUTINE CLOUD_SATADJ
 *   ifs-source/arpifs/phys_ec/cloud_satadj.F90:13-26 (dummy arguments), declarations :90-130,
 *   called from CALLPAR (ifs-source/arpifs/phys_ec/callpar.F90:1253-1275) just before CLOUD_LAYER/CLOUDSC.
 *
 * Conventions as in cloudsc_hip.h: Fortran-ordered arrays (JL fastest), the derived-type arguments
 * (YDECLDP, YDEPHLI, YDECUMF, YDEPHY, YDSPP_CONFIG) flattened into cloudsc_params_t, LOGICALs as int32,
 * return value 0 or a negative CLOUDSC_E* code.  Default configuration only: LSPP and the budget
 * diagnostics (LCLDBUD*, LBUD23) are rejected with CLOUDSC_EUNSUPPORTED.  Arguments the default path does
 * not read (KTYPE, PMFU, PMFD, PGP2DSPP, PEXTRA) may be NULL.  Outputs are written for JL = KIDIA..KFDIA
 * on every level (zero above NCLDTOP, cloud_satadj.F90:338-348).
 *
 * Variants: _dp (FP64), _sp (FP32), _mixed (FP32 storage and bulk arithmetic; FP64 for the two-iteration
 * Newton saturation adjustment QSATMIXADJ, cloud_satadj.F90:948-1017, and for CLOUD_SUPERSATCHECK).
 */
#ifndef CLOUDSC_SATADJ_H
#define CLOUDSC_SATADJ_H

#include <stdint.h>

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-block drop-in (host pointers), the reference's positional argument list. */
#define CLOUDSC_SATADJ_ARGS(R)                                                                          \
  const cloudsc_params_t *ydparams, int kidia, int kfdia, int klon, int klev, const int32_t *ktype,     \
      int lddiag, R ptsphy, const R *pap, const R *paph, const R *pt, const R *pq, const R *pa,         \
      const R *pl, const R *pi, const R *ptendency_cml_t, const R *ptendency_cml_q,                     \
      const R *ptendency_cml_a, const R *ptendency_cml_l, const R *ptendency_cml_i, const R *phrsw,     \
      const R *phrlw, const R *ptendency_vdf_t, const R *ptendency_vdf_q, const R *pmfu,                 \
      const R *pmfd, const R *pludeli, const R *pvervel, const R *pgp2dspp, R *ptendency_loc_t,         \
      R *ptendency_loc_q, R *ptendency_loc_a, R *ptendency_loc_l, R *ptendency_loc_i, R *pextra, int kfldx

int cloudsc_satadj_dp_c(CLOUDSC_SATADJ_ARGS(double));
int cloudsc_satadj_sp_c(CLOUDSC_SATADJ_ARGS(float));
int cloudsc_satadj_mixed_c(CLOUDSC_SATADJ_ARGS(float));

/* Blocked, device-resident entry over all NGPBLKS blocks (device pointers, IFS blocked fields as in
 * cloudsc_hip.h).  The liquid/ice state, CML tendencies and local tendencies are the species 1 and 2
 * slices of the (NPROMA,KLEV,5,NGPBLKS) fields PCLV, PTENDENCY_CML_CLD and PTENDENCY_LOC_CLD, as CALLPAR
 * passes STATE%CLD(:,:,NCLDQL) etc.; the other species are neither read nor written.  Asynchronous on
 * `stream`. */
#define CLOUDSC_SATADJ_BLOCKED_ARGS(R)                                                                  \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pap, const R *paph, const R *pt, const R *pq, const R *pa, const R *pclv,                \
      const R *ptendency_cml_t, const R *ptendency_cml_q, const R *ptendency_cml_a,                     \
      const R *ptendency_cml_cld, const R *phrsw, const R *phrlw, const R *ptendency_vdf_t,             \
      const R *ptendency_vdf_q, const R *pludeli, const R *pvervel, R *ptendency_loc_t,                 \
      R *ptendency_loc_q, R *ptendency_loc_a, R *ptendency_loc_cld, void *stream

int cloudsc_satadj_blocked_dp_c(CLOUDSC_SATADJ_BLOCKED_ARGS(double));
int cloudsc_satadj_blocked_sp_c(CLOUDSC_SATADJ_BLOCKED_ARGS(float));
int cloudsc_satadj_blocked_mixed_c(CLOUDSC_SATADJ_BLOCKED_ARGS(float));

/* Fused CALLPAR sequence on device-resident blocked fields, one launch: CLOUD_SATADJ, the increment of the
 * cumulated tendencies by its result (STATE_INCREMENT, callpar.F90:1278-1279), the copy of its result into
 * TENDENCY_SATADJ (STATE_COPY, :1289) and CLOUDSC (CLOUD_LAYER, :1292-1296), with each column's adjustment done
 * level by level inside CLOUDSC's sweep.  Arguments: those of cloudsc_blocked_*_c (the CML tendencies are INOUT
 * here: they return incremented, as TENDENCY_CML after STATE_INCREMENT; PLUDELI carries all four species),
 * then the SATADJ-only inputs and the TENDENCY_SATADJ outputs (T, Q, A and species 1-2 of the five-species
 * field; may all be NULL).  Results equal cloudsc_satadj_blocked_*_c, the increment and cloudsc_blocked_*_c in
 * sequence, bit for bit. */
#define CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(R)                                                          \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pt, const R *pq, R *ptendency_cml_t, R *ptendency_cml_q, R *ptendency_cml_a,             \
      R *ptendency_cml_cld, R *ptendency_loc_t, R *ptendency_loc_q, R *ptendency_loc_a,                 \
      R *ptendency_loc_cld, const R *pvfl, const R *pvfi, const R *pap, const R *paph, const R *plsm,   \
      const int32_t *ldcum, const int32_t *ktype, R *peis, const R *plu, R *plude, const R *pludeli,    \
      const R *psnde, const R *pmfu, const R *pmfd, const R *pa, const R *pclv, R *pcovptot,            \
      R *prainfrac_toprfz, R *pfsqlf, R *pfsqif, R *pfcqnng, R *pfcqlng, R *pfsqrf, R *pfsqsf,          \
      R *pfcqrng, R *pfcqsng, R *pfsqltur, R *pfsqitur, R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn,     \
      const R *phrsw, const R *phrlw, const R *ptendency_vdf_t, const R *ptendency_vdf_q,               \
      const R *pvervel, R *ptendency_satadj_t, R *ptendency_satadj_q, R *ptendency_satadj_a,            \
      R *ptendency_satadj_cld, void *stream

int cloudsc_satadj_cloudsc_blocked_dp_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(double));
int cloudsc_satadj_cloudsc_blocked_sp_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(float));
int cloudsc_satadj_cloudsc_blocked_mixed_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(float));

/* Synthetic forcings for CLOUD_SATADJ on a blocked field set (device pointers), seeded per column as
 * cloudsc_synth_*: PHRSW ~ U(0, 2e-5) K/s, PHRLW ~ N(-2e-5, 1e-5) K/s, PTENDENCY_VDF_T ~ N(0, 1e-5) K/s,
 * PTENDENCY_VDF_Q ~ N(0, 1e-8) /s, PVERVEL ~ N(0, 0.3) Pa/s, PLUDELI(:,:,3) = PLUDE * 1e-2 (detrained vapour)
 * and PLUDELI(:,:,4) = PLUDE * 1e3 K (detrained heat), PLUDE read from `plude`. */
int cloudsc_satadj_synth_sp(uint64_t seed, int ngptot, int nproma, int klev, int ngpblks, const float *plude,
                            float *phrsw, float *phrlw, float *ptendency_vdf_t, float *ptendency_vdf_q,
                            float *pvervel, float *pludeli, void *stream);
int cloudsc_satadj_synth_dp(uint64_t seed, int ngptot, int nproma, int klev, int ngpblks, const double *plude,
                            double *phrsw, double *phrlw, double *ptendency_vdf_t, double *ptendency_vdf_q,
                            double *pvervel, double *pludeli, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_SATADJ_H */
