/*
 * cloudsc2.h -- C-ABI of the gfx950 CLOUDSC2, the simplified prognostic cloud scheme of the linearised physics.
 *
 * Reference interface replaced:
 *   SUBROUTINE CLOUDSC2 (ifs-source/arpifs/phys_ec/cloudsc2.F90:10-17), called from CLOUD_S_LAYER
 *   (cloud_s_layer.F90:126-139) when LEPCLD2, with its final saturation adjustment CUADJTQ (KCALL = 0,
 *   cuadjtq.F90:250-274) or, with LPHYLIN or LDRAIN1D, CUADJTQS (KCALL = 0, cuadjtqs.F90:217-250).
 *
 * Conventions as in cloudsc_hip.h: Fortran order (JL fastest), LOGICALs as int32, return 0 or a negative CLOUDSC_E*
 * code.  The thermodynamic and TECLDP constants come from cloudsc_params_t (RG, RD, RCPD, RETV, RTT, RLVTT, RLSTT,
 * RLMLT, R2ES..R5IES, R5ALVCP, R5ALSCP, RALVDCP, RALSDCP, RTWAT, RTICE, RTWAT_RTICE_R, RKCONV, RLMIN, RPECONS);
 * the few components CLOUDSC does not read are in cloudsc2_params_t.  CETA (YDECLD%CETA, the reference eta of each
 * level, sucld.F90:126) is an array argument.  Variants: _dp (FP64) and _sp (FP32).
 */
#ifndef CLOUDSC2_H
#define CLOUDSC2_H

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cloudsc2_params_t {
  double rclcrit;  /* TECLDP: critical autoconversion threshold (sucldp.F90:264) */
  double rlptrc;   /* TEPHLI: phase-transition temperature of the linearised physics (suphli.F90:125) */
  double rlpal1;   /* TEPHLI: smoothing of the phase function in CUADJTQ (suphli.F90:126) */
  double rlpal2;   /* TEPHLI: smoothing of the saturation clip in CUADJTQ (suphli.F90:127) */
  double rvtmp2;   /* TTHF: RCPV/RCPD-1, set to 0 (suphec.F90:229) */
  int32_t lphylin;  /* TEPHLI: linearised-physics thermodynamics (CUADJTQS, tanh phase function) */
  int32_t levapls2; /* TPHNC: evaporation of large-scale precipitation */
  int32_t lepcld2;  /* TPHNC: CLOUD_S_LAYER runs CLOUDSC2 (prognostic) if set, CLOUDST (diagnostic) if not */
  int32_t pad_;
} cloudsc2_params_t;

/* Defaults of the TECLDP/TEPHLI/TPHNC/TTHF components above for the given cloudsc_params_t (RTICE, RTWAT). */
int cloudsc2_params_default(cloudsc2_params_t *p2, const cloudsc_params_t *p);

/* One NPROMA block, host arrays: the reference's argument list after the derived types.  PAPHP1 is
 * (KLON,KLEV+1), PFPLSL/PFPLSN/PFHPSL/PFHPSN (KLON,KLEV+1), the rest (KLON,KLEV); CETA (KLEV).  Written as the
 * reference writes them: PCLC and PCOVPTOT on every level, the tendencies on KTDIA..KLEV, PFPLSL/PFPLSN at the
 * top and below KTDIA..KLEV, PFHPSL/PFHPSN on every half level from the precipitation fluxes. */
#define CLOUDSC2_ARGS(R)                                                                                \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int kidia,        \
      int kfdia, int klon, int ktdia, int klev, int ldrain1d, R ptsphy, const R *paphp1, const R *papp1, \
      const R *pqm1, const R *pqs, const R *ptm1, const R *pl, const R *pi, const R *plude,               \
      const R *plu, const R *pmfu, const R *pmfd, R *ptent, const R *pgtent, R *ptenq, const R *pgtenq,  \
      R *ptenl, const R *pgtenl, R *pteni, const R *pgteni, R *pclc, R *pfplsl, R *pfplsn, R *pfhpsl,    \
      R *pfhpsn, R *pcovptot

int cloudsc2_dp_c(CLOUDSC2_ARGS(double));
int cloudsc2_sp_c(CLOUDSC2_ARGS(float));

/* All blocks of device-resident blocked fields (NPROMA,KLEV[+1],NGPBLKS) in one launch, as CLOUD_S_LAYER would
 * be called from the EC_PHYS_DRV block loop; CETA is a device array (KLEV).  Asynchronous on `stream`. */
#define CLOUDSC2_BLOCKED_ARGS(R)                                                                        \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot,       \
      int nproma, int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R *paphp1,            \
      const R *papp1, const R *pqm1, const R *pqs, const R *ptm1, const R *pl, const R *pi,               \
      const R *plude, const R *plu, const R *pmfu, const R *pmfd, R *ptent, const R *pgtent, R *ptenq,   \
      const R *pgtenq, R *ptenl, const R *pgtenl, R *pteni, const R *pgteni, R *pclc, R *pfplsl,         \
      R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot, void *stream

int cloudsc2_blocked_dp_c(CLOUDSC2_BLOCKED_ARGS(double));
int cloudsc2_blocked_sp_c(CLOUDSC2_BLOCKED_ARGS(float));

/* CLOUDST (cloudst.F90:9-16), the diagnostic scheme, one NPROMA block of host arrays (Fortran order): the
 * reference's argument list after the derived types, as cloudsc2_*_c. */
#define CLOUDST_ARGS(R)                                                                                 \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int kidia,        \
      int kfdia, int klon, int ktdia, int klev, int ldrain1d, R ptsphy, const R *paphp1, const R *papp1, \
      const R *pqm1, const R *pqs, const R *ptm1, const R *plude, const R *plu, R *ptent,               \
      const R *pgtent, R *ptenq, const R *pgtenq, R *pclc, R *pqiwp, R *pqlwp, R *pfplsl, R *pfplsn,   \
      R *pfhpsl, R *pfhpsn, R *pcovptot

int cloudst_dp_c(CLOUDST_ARGS(double));
int cloudst_sp_c(CLOUDST_ARGS(float));

/* CLOUDST over all blocks of device-resident blocked fields:
 * CLOUDSC2's arrays without the condensate and mass-flux ones, plus the diagnostic ice and liquid water PQIWP and
 * PQLWP; every array is required. */
#define CLOUDST_BLOCKED_ARGS(R)                                                                         \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot,       \
      int nproma, int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R *paphp1,            \
      const R *papp1, const R *pqm1, const R *pqs, const R *ptm1, const R *plude, const R *plu,         \
      R *ptent, const R *pgtent, R *ptenq, const R *pgtenq, R *pclc, R *pqiwp, R *pqlwp, R *pfplsl,     \
      R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot, void *stream

int cloudst_blocked_dp_c(CLOUDST_BLOCKED_ARGS(double));
int cloudst_blocked_sp_c(CLOUDST_BLOCKED_ARGS(float));

/* CLOUD_S_LAYER (cloud_s_layer.F90:112-176), all blocks of device-resident blocked fields in one launch.
 * - SATUR (KFLAG 2) at PT + TSPHY*PTENDENCY_T gives PQSAT.
 * - The scheme runs on it: CLOUDSC2 if ydparams2->lepcld2, else CLOUDST.
 * - The wrap-up writes PTENDENCY_LOC_A = (PCLC - PA)/TSPHY on every level, zero rain and snow tendencies on every
 *   column and level, and zero PPRECTYPE and PFZRA. With CLOUDST it also writes PTENDENCY_LOC_L/I as
 *   (PQLWP - PL)/TSPHY and (PQIWP - PI)/TSPHY.
 * Arguments are as CLOUD_S_LAYER passes them to CLOUDSC2: PRS1 = PAPHP1, PRSF1 = PAPP1, PQ = PQM1, PT = PTM1,
 * PTENDENCY_* = PGTEN*, PTENDENCY_LOC_* = PTEN*. PQSAT is PDIAG%PQSAT; PCLC, PQIWP and PQLWP are AUXL%ZPATMP,
 * ZPLSMP and ZPLTMP.
 * These may be NULL where the caller's key mask has no such field: PTENDENCY_LOC_A (which needs PA),
 * PTENDENCY_LOC_R/S, PQSAT, PCLC, PPRECTYPE and PFZRA (both (NPROMA,NGPBLKS)), and PQIWP and PQLWP.
 * PMFU, PMFD and PTENDENCY_L/I are read by CLOUDSC2 only. */
#define CLOUD_S_LAYER_BLOCKED_ARGS(R)                                                                   \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot,       \
      int nproma, int ktdia, int klev, int ngpblks, int ldrain1d, R tsphy, const R *prs1,              \
      const R *prsf1, const R *pq, const R *pt, const R *pa, const R *pl, const R *pi, const R *plude,  \
      const R *plu, const R *pmfu, const R *pmfd, const R *ptendency_t, const R *ptendency_q,           \
      const R *ptendency_l, const R *ptendency_i, R *ptendency_loc_t, R *ptendency_loc_q,              \
      R *ptendency_loc_a, R *ptendency_loc_l, R *ptendency_loc_i, R *ptendency_loc_r,                  \
      R *ptendency_loc_s, R *pqsat, R *pclc, R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot,  \
      R *pprectype, R *pfzra, R *pqiwp, R *pqlwp, void *stream

int cloud_s_layer_blocked_dp_c(CLOUD_S_LAYER_BLOCKED_ARGS(double));
int cloud_s_layer_blocked_sp_c(CLOUD_S_LAYER_BLOCKED_ARGS(float));

/* CLOUD_S_LAYER on one NPROMA block of host arrays (Fortran order, (KLON,KLEV[+1]) and (KLON) for PPRECTYPE/PFZRA):
 * the arrays of CLOUD_S_LAYER_BLOCKED_ARGS with the same optional ones (a NULL required array: CLOUDSC_ENULL).
 * Outputs the layer does not write come back unchanged. */
#define CLOUD_S_LAYER_ARGS(R)                                                                           \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int kidia,        \
      int kfdia, int klon, int ktdia, int klev, int ldrain1d, R tsphy, const R *prs1, const R *prsf1,   \
      const R *pq, const R *pt, const R *pa, const R *pl, const R *pi, const R *plude, const R *plu,     \
      const R *pmfu, const R *pmfd, const R *ptendency_t, const R *ptendency_q, const R *ptendency_l,   \
      const R *ptendency_i, R *ptendency_loc_t, R *ptendency_loc_q, R *ptendency_loc_a,                 \
      R *ptendency_loc_l, R *ptendency_loc_i, R *ptendency_loc_r, R *ptendency_loc_s, R *pqsat, R *pclc, \
      R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot, R *pprectype, R *pfzra, R *pqiwp,       \
      R *pqlwp

int cloud_s_layer_dp_c(CLOUD_S_LAYER_ARGS(double));
int cloud_s_layer_sp_c(CLOUD_S_LAYER_ARGS(float));

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC2_H */
