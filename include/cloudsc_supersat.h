/*
 * cloudsc_supersat.h -- C-ABI of the gfx950 CLOUD_SUPERSATCHECK and of SLTEND's supersaturation adjustment.
 *
 * Reference interfaces replaced:
 *   SUBROUTINE CLOUD_SUPERSATCHECK (ifs-source/arpifs/phys_ec/cloud_supersatcheck.F90:9-56), called by
 *   CLOUDSC (cloudsc.F90:3774), CLOUD_SATADJ (cloud_satadj.F90:770) and SLTEND (sltend.F90:392);
 *   section 3 of SUBROUTINE SLTEND (sltend.F90:370-409), the adjustment of the semi-Lagrangian physics
 *   tendencies for any supersaturation of the updated state.
 *
 * Conventions as in cloudsc_hip.h (Fortran order, return 0 or a negative CLOUDSC_E* code).  Variants: _dp (FP64),
 * _sp (FP32), _mixed (FP32 storage, CLOUD_SUPERSATCHECK in FP64).
 */
#ifndef CLOUDSC_SUPERSAT_H
#define CLOUDSC_SUPERSAT_H

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CLOUD_SUPERSATCHECK for one level of KLON points (host pointers), the reference's argument list: KFTLIQICE 1
 * partitions the condensate with FOEALFA, 2 makes it liquid where PT+PT_ADJ >= RTHOMO and ice below.  All arrays
 * have KLON elements and are read and written at JL = KIDIA..KFDIA (element JL-1); KLEV and PSURF are accepted and,
 * as in the reference, do not change the result. */
#define CLOUDSC_SUPERSATCHECK_ARGS(R)                                                                   \
  const cloudsc_params_t *ydparams, int kidia, int kfdia, int klon, int klev, int kftliqice,            \
      const R *pt, const R *pq, const R *pa, const R *pp, const R *psurf, R *pt_adj, R *pq_adj,         \
      R *pa_adj, R *pl_adj, R *pi_adj

int cloudsc_supersatcheck_dp_c(CLOUDSC_SUPERSATCHECK_ARGS(double));
int cloudsc_supersatcheck_sp_c(CLOUDSC_SUPERSATCHECK_ARGS(float));
int cloudsc_supersatcheck_mixed_c(CLOUDSC_SUPERSATCHECK_ARGS(float));

/* SLTEND section 3 over all blocks of device-resident blocked fields (NPROMA,KLEV[+1],[5,]NGPBLKS): on levels
 * NCLDTOP..KLEV, PTENT/PTENQ/PTENA and species 1-2 of PTENCLD (INOUT) receive the CLOUD_SUPERSATCHECK
 * (KFTLIQICE = 1) adjustments of the state PT + TSPHY*(PCENT+PTENT) etc. divided by TSPHY; PADJA/PADJL/PADJI
 * (may be NULL) receive the adjustments alone, the budget terms SLTEND keeps (ZADJA/ZADJL/ZADJI).  Levels above
 * NCLDTOP are not touched.  Asynchronous on `stream`. */
#define CLOUDSC_SLTEND_SUPERSAT_ARGS(R)                                                                 \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R tsphy,             \
      const R *pt, const R *pq, const R *pa, const R *prsf1, const R *prs1, const R *pcent,             \
      const R *pcenq, const R *pcena, R *ptent, R *ptenq, R *ptena, R *ptencld, R *padja, R *padjl,     \
      R *padji, void *stream

int cloudsc_sltend_supersat_blocked_dp_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(double));
int cloudsc_sltend_supersat_blocked_sp_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(float));
int cloudsc_sltend_supersat_blocked_mixed_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(float));

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_SUPERSAT_H */
