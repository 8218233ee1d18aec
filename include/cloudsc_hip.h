/*
 * cloudsc_hip.h -- C-ABI of the MI355X (gfx950) CLOUDSC drop-in.
 *
 * Reference interface replaced: SUBROUTINE CLOUDSC
 *   ifs-source/arpifs/phys_ec/cloudsc.F90:14-44 (dummy arguments),
 *   declarations :194-277, called from CLOUD_LAYER
 *   ifs-source/arpifs/phys_ec/cloud_layer.F90:169-198.
 *
 * Conventions
 *   - Every array argument is a plain pointer to Fortran-ordered storage
 *     (JL, the column index, fastest).  Shapes follow cloudsc.F90:207-277:
 *     "KLEV array" = (KLON,KLEV), "half-level array" = (KLON,KLEV+1),
 *     PTENDENCY_*_CLD / PCLV = (KLON,KLEV,NCLV=5), PLUDELI = (KLON,KLEV,4),
 *     PSNDE = (KLON,KLEV,2), PEXTRA = (KLON,KLEV,KFLDX).
 *   - The seven derived-type arguments (YDECLDP, YDECUMF, YDEPHLI, YDERAD,
 *     YDEPHY, YDVDF, YDSPP_CONFIG) are flattened into one cloudsc_params_t.
 *   - LOGICAL arguments (LDCUM(KLON), LDDIAG, LDSLPHY) are passed as int32
 *     (0 = .FALSE., nonzero = .TRUE.).
 *   - Arguments that the default configuration never reads (budget-only,
 *     aerosol-only, SPP-only, inhomogeneity-only: PVFA, PDYN*, PHRSW, PHRLW,
 *     PVERVEL, PGEOM, PGAW, KCTOP, KPBLTYPE, PGP2DSPP, PLCRIT_AER, PICRIT_AER,
 *     PRE_ICE, PCCN, PNICE, PEXTRA) are kept in the list so the call is a
 *     positional drop-in; they may be NULL unless the budget switches
 *     (PVFA..PHRLW, KPBLTYPE, KCTOP, PEXTRA) or LRAD_CLOUD_INHOMOG (PGAW) read
 *     them.  PFSD is written only when LRAD_CLOUD_INHOMOG is set.
 *   - Write set is the reference's: PTENDENCY_LOC_CLD(:,:,5) is not written;
 *     PCOVPTOT is written only for JK >= NCLDTOP; PRAINFRAC_TOPRFZ is zeroed
 *     for all KLON columns; PLUDE is rescaled in place for NCLDTOP<=JK<KLEV
 *     (cloudsc.F90:971-977, 3815, 1010, 1367/1384).
 *   - Return value: 0 on success, a negative CLOUDSC_E* code otherwise.  The
 *     reference has no status argument; its only fatal path is ABOR1
 *     (budget diagnostics), so a Fortran shim maps nonzero to ABOR1.
 *
 * Three arithmetic variants share one argument list:
 *   _dp    : FP64 storage and arithmetic (the ifs_dp build).
 *   _sp    : FP32 storage and arithmetic (the ifs_sp build).
 *   _mixed : FP32 storage and bulk arithmetic; FP64 for the water budget of
 *            each level: the first-guess amounts ZQXFG, the explicit
 *            source/sink matrix ZSOLQA, the sinks limited by ZQXFG, the sink
 *            truncation (cloudsc.F90:3455-3536) and the right-hand side of the
 *            implicit solve (:3576-3587).  The elimination and
 *            CLOUD_SUPERSATCHECK stay in FP32.
 */
#ifndef CLOUDSC_HIP_H
#define CLOUDSC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLOUDSC_NCLV 5
#define CLOUDSC_NCLDQL 1
#define CLOUDSC_NCLDQI 2
#define CLOUDSC_NCLDQR 3
#define CLOUDSC_NCLDQS 4
#define CLOUDSC_NCLDQV 5

/* Status codes */
#define CLOUDSC_OK 0
#define CLOUDSC_EDIMS (-1)        /* inconsistent KIDIA/KFDIA/KLON/KLEV/NCLDTOP */
#define CLOUDSC_EUNSUPPORTED (-2) /* a configuration switch this build does not implement */
#define CLOUDSC_ENULL (-3)        /* a pointer the default path reads or writes is NULL */
#define CLOUDSC_EHIP (-4)         /* a HIP runtime call failed */
#define CLOUDSC_ENODEVICE (-5)    /* no gfx950 device visible */
#define CLOUDSC_ENOMEM (-6)       /* device allocation failed */
#define CLOUDSC_ESTATE (-7)       /* a field is not present where it is asked for (include/cloudsc_field.h) */
#define CLOUDSC_ENONFINITE (-8)   /* the per-block call wrote non-finite outputs (count in cloudsc_last_error()) */

/*
 * Flattened configuration: every TECLDP / TTHF / TCST / TECUMF / TERAD /
 * TEPHY / TVDF / TSPP_CONFIG component CLOUDSC and CLOUD_SUPERSATCHECK read.
 * Names are the reference's component names in lower case
 * (yoecldp.F90:112-296, yoethf.F90:102-128, yomcst.F90:508-629).
 * Values are FP64; the FP32 variants round them once at launch.
 */
typedef struct cloudsc_params_t {
  /* TCST (yomcst.F90) */
  double rg, rd, rv, rcpd, retv, rtt, rlvtt, rlstt, rlmlt, ra, rpi;
  /* TTHF (suphec.F90:231-254) */
  double r2es, r3les, r3ies, r4les, r4ies, r5les, r5ies, r5alvcp, r5alscp;
  double ralvdcp, ralsdcp, ralfdcp, rtwat, rtice, rtwat_rtice_r, rkoop1, rkoop2;
  /* TECLDP (sucldp.F90:243-655) */
  double ramid, ramin, rlmin, rcldiff, rcldiff_convi, rclcrit_sea, rclcrit_land;
  double rkconv, rprc1, rcl_inhomogaut, rcl_inhomogacc;
  double rsnowlin1, rsnowlin2, rlcritsnow, rvice, rvrain, rvsnow;
  double rxfallspeedrain, rxfallspeedice, rnice, rccn;
  double rcldtopcf, rdepliqrefrate, rdepliqrefdepth, rcl_overlapliqice, rcl_effrime;
  double riceinit, rthomo, rprecrhmax, rpecons, rvrfactor, rcovpmin;
  double rtaumel, rkooptau, rssicefactor, rdenswat, rdensref;
  double rcl_dr, rcl_x1r, rcl_x2r, rcl_x1s, rcl_lam1r, rcl_lam2r, rcl_lam1s, rcl_lam2s;
  double rcl_cdenom1, rcl_cdenom2, rcl_cdenom3, rcl_ka273;
  double rcl_const1r, rcl_const2r, rcl_const3r, rcl_const4r, rcl_const5r;
  double rcl_const6r, rcl_const7r, rcl_const9r, rcl_const10r, rcl_eff_racw;
  double rcl_kkaac, rcl_kkbac, rcl_kkaau, rcl_kkbauq, rcl_kkbaun;
  double rcl_kk_cloud_num_sea, rcl_kk_cloud_num_land, rcl_fzrab;
  double rcl_const1i, rcl_const2i, rcl_const3i, rcl_const4i, rcl_const5i, rcl_const6i, rcl_const7i;
  double rcl_apb1, rcl_apb2, rcl_apb3;
  double rcl_const1s, rcl_const2s, rcl_const3s, rcl_const4s, rcl_const5s;
  double rcl_const6s, rcl_const7s, rcl_const8s;
  /* TECUMF (sucumf.F90:203,228-229) */
  double rmfadvw, rmfadvwdd;
  /* TERAD (suecrad.F90:745) and TVDF (suvdf.F90:135) */
  double rcloud_frac_std, reisthsc;
  /* integer / logical switches */
  int32_t ncldtop;          /* first level the cloud scheme treats (sucldp.F90:631-634) */
  int32_t lmfdsnow;         /* TECUMF: convective snow/rain detrainment */
  int32_t lcloud_inhomog;   /* TECLDP: KK2000 enhancement from the condensate FSD (per-block and diag entries only) */
  int32_t lrad_cloud_inhomog; /* TEPHY: condensate FSD into PFSD (per-block and cloudsc_blocked_diag_*_c only) */
  int32_t laeraware;        /* any of LAERLIQAUTOLSP/LAERLIQCOLL/LAERICESED/LAERICEAUTO: must be 0 */
  int32_t lspp;             /* TSPP_CONFIG%LSPP: must be 0 */
  int32_t lcldbud;          /* budget diagnostics, CLOUDSC_BUD_* bits below (0 = all off, the default) */
  int32_t pad_;
} cloudsc_params_t;

/*
 * Fill *p with the reference defaults of SUPHEC/SUCLDP/SUCUMF/SU0PHY/SUECRAD/
 * SUVDF for physics time step ptsphy (RSSICEFACTOR = MIN(TSPHY/RKOOPTAU,1),
 * sucldp.F90:655) and NCLDTOP from a standard-atmosphere full-level pressure
 * profile stpre[0..nflevg-1] (Pa) with RCLDTOPP = 100 Pa (sucldp.F90:631-634).
 * Replaces the setup call chain SUPHEC -> SUCLDP (suphec.F90:406).
 */
int cloudsc_params_default(cloudsc_params_t* p, double ptsphy, int nflevg, const double* stpre);

/* ---------------------------------------------------------------------------
 * Per-block drop-in (host pointers).  One call = one NPROMA block, exactly
 * like the reference call from CLOUD_LAYER.  Copies the block to the device,
 * runs the kernel and copies the outputs back.  Reentrant: OpenMP threads of
 * the EC_PHYS_DRV block loop (ec_phys_drv.F90:430-439) may call concurrently.
 * ------------------------------------------------------------------------- */
#define CLOUDSC_ARGS(R)                                                                                 \
  const cloudsc_params_t *ydparams, int kidia, int kfdia, int klon, int klev, int lddiag, R ptsphy,     \
      const R *pt, const R *pq, const R *ptendency_cml_t, const R *ptendency_cml_q,                     \
      const R *ptendency_cml_a, const R *ptendency_cml_cld, R *ptendency_loc_t, R *ptendency_loc_q,     \
      R *ptendency_loc_a, R *ptendency_loc_cld, const R *pvfa, const R *pvfl, const R *pvfi,            \
      const R *pdyna, const R *pdynl, const R *pdyni, const R *pdynr, const R *pdyns, const R *phrsw,  \
      const R *phrlw, const R *pvervel, const R *pap, const R *paph, const R *pgeom, const R *plsm,     \
      const R *pgaw, const int32_t *ldcum, const int32_t *kctop, const int32_t *ktype,                  \
      const int32_t *kpbltype, R *peis, const R *plu, R *plude, const R *pludeli, const R *psnde,       \
      const R *pmfu, const R *pmfd, const R *pgp2dspp, int ldslphy, const R *pa, const R *pclv,         \
      const R *plcrit_aer, const R *picrit_aer, const R *pre_ice, const R *pccn, const R *pnice,        \
      R *pcovptot, R *pfsd, R *prainfrac_toprfz, R *pfsqlf, R *pfsqif, R *pfcqnng, R *pfcqlng,          \
      R *pfsqrf, R *pfsqsf, R *pfcqrng, R *pfcqsng, R *pfsqltur, R *pfsqitur, R *pfplsl, R *pfplsn,     \
      R *pfhpsl, R *pfhpsn, R *pextra, int kfldx

/* Replaces CLOUDSC of the ifs_dp build (JPRB = FP64). */
int cloudsc_dp_c(CLOUDSC_ARGS(double));
/* Replaces CLOUDSC of the ifs_sp build (JPRB = FP32). */
int cloudsc_sp_c(CLOUDSC_ARGS(float));
/* Replaces CLOUDSC of the ifs_sp build with FP64 for the water budget of each level (the first-guess amounts,
 * the source/sink matrix, the limited sinks, the sink truncation and the solve's right-hand side): within 1e-5
 * relative RMS of the ifs_dp build, where the ifs_sp build is ~1e-2 away. */
int cloudsc_mixed_c(CLOUDSC_ARGS(float));
/* cloudsc_mixed_c without device allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (hosts that forbid hipMalloc, tight serving loops).  The block is staged in a fixed
 * 32 MiB device array of the library (one per GPU, held until the process ends) instead of the calling thread's
 * workspace, and runs on `stream` (a hipStream_t; NULL = the null stream) of the current device.  A block needs
 * 4 B x KLON x (38 KLEV + 15 (KLEV+1) + 4) plus the budget inputs and KFLDX PEXTRA fields when budgets are on;
 * one that needs more than the array returns CLOUDSC_EDIMS (at KLEV 137 without budgets: KLON up to about 1,150).
 * Returns after `stream` has finished the block.  All calls on one GPU share its array: calls that overlap on one
 * GPU race (other threads must use cloudsc_mixed_c or wait for each other). */
int cloudsc_mixed_noalloc_c(CLOUDSC_ARGS(float), void *stream);

/* ---------------------------------------------------------------------------
 * Blocked, device-resident entry: one launch over all NGPBLKS blocks of the
 * EC_PHYS_DRV loop (ec_phys_drv.F90:430-441).  Every pointer is a device
 * pointer.  Arrays are the IFS blocked fields (ec_phys_drv.F90:146):
 * (NPROMA,KLEV,NGPBLKS), (NPROMA,KLEV+1,NGPBLKS), (NPROMA,KLEV,n,NGPBLKS),
 * per-column arrays (NPROMA,NGPBLKS).  Block IBL covers global columns
 * (IBL-1)*NPROMA+1 .. MIN(IBL*NPROMA, NGPTOT) with KIDIA = 1 and
 * KFDIA = MIN(NPROMA, NGPTOT-(IBL-1)*NPROMA) (ec_phys_drv.F90:441-465).
 * stream: a hipStream_t (NULL = the null stream).  The call is asynchronous.
 * ------------------------------------------------------------------------- */
/* Bits of cloudsc_params_t::lcldbud: the LCLDBUD* switches of YDECLDP (cloudsc.F90:706-713), LBUD23 of YDEPHY
 * (PEXTRA fields start after 26, :3832-3836) and LCLDBUDGET (its check, :1032-1082 and :4146-4212, has no output,
 * so the bit is accepted and changes nothing).  The C, L, I, T and VERTINT budgets write PEXTRA: they are honoured
 * by the per-block entries (PEXTRA and KFLDX are in the argument list) and by cloudsc_blocked_diag_*_c, and
 * rejected with CLOUDSC_EUNSUPPORTED by the other entries.  VERTINT needs KLEV >= 83 (the slots the reference
 * writes) and all of them need KFLDX >= the number of fields they fill. */
#define CLOUDSC_BUD_C 1         /* LCLDBUDC: cloud fraction budget, 12 fields */
#define CLOUDSC_BUD_L 2         /* LCLDBUDL: cloud liquid budget, 22 fields */
#define CLOUDSC_BUD_I 4         /* LCLDBUDI: cloud ice budget, 20 fields */
#define CLOUDSC_BUD_T 8         /* LCLDBUDT: cloud process temperature budget, 13 fields */
#define CLOUDSC_BUD_VERTINT 16  /* LCLDBUD_VERTINT: vertical integrals in one field */
#define CLOUDSC_BUD_TIMEINT 32  /* LCLDBUD_TIMEINT: accumulate rate * PTSPHY instead of the rate */
#define CLOUDSC_BUD_LBUD23 64   /* LBUD23: the first 26 PEXTRA fields belong to other budgets */
#define CLOUDSC_BUD_CHECK 128   /* LCLDBUDGET */

#define CLOUDSC_BLOCKED_ARGS(R)                                                                         \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pt, const R *pq, const R *ptendency_cml_t, const R *ptendency_cml_q,                     \
      const R *ptendency_cml_a, const R *ptendency_cml_cld, R *ptendency_loc_t, R *ptendency_loc_q,     \
      R *ptendency_loc_a, R *ptendency_loc_cld, const R *pvfl, const R *pvfi, const R *pap,             \
      const R *paph, const R *plsm, const int32_t *ldcum, const int32_t *ktype, R *peis,                \
      const R *plu, R *plude, const R *pludeli, const R *psnde, const R *pmfu, const R *pmfd,           \
      const R *pa, const R *pclv, R *pcovptot, R *prainfrac_toprfz, R *pfsqlf, R *pfsqif,              \
      R *pfcqnng, R *pfcqlng, R *pfsqrf, R *pfsqsf, R *pfcqrng, R *pfcqsng, R *pfsqltur, R *pfsqitur,   \
      R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, void *stream

int cloudsc_blocked_dp_c(CLOUDSC_BLOCKED_ARGS(double));
int cloudsc_blocked_sp_c(CLOUDSC_BLOCKED_ARGS(float));
int cloudsc_blocked_mixed_c(CLOUDSC_BLOCKED_ARGS(float));

/* Blocked entry with the diagnostics the plain entries refuse: the arguments of cloudsc_blocked_*_c, then the
 * inputs only the budgets read (PVFA, PDYNA, PDYNL, PDYNI, PDYNS, PHRSW, PHRLW as (NPROMA,KLEV,NGPBLKS); KPBLTYPE,
 * KCTOP as (NPROMA,NGPBLKS)), PEXTRA (NPROMA,KLEV,KFLDX,NGPBLKS) and KFLDX, then PGAW (NPROMA,NGPBLKS) and PFSD
 * (NPROMA,KLEV,NGPBLKS).  params->lcldbud selects the budgets, which accumulate into PEXTRA as the reference does
 * (cloudsc.F90:3826-4101); params->lrad_cloud_inhomog writes the condensate FSD for the radiation into PFSD on
 * every level (:1937-1971, :3818-3822, :4112-4126, :4307-4359).  Arrays of a diagnostic that is off may be NULL.
 * Replaces the CLOUDSC call of CLOUD_LAYER in runs with those switches. */
#define CLOUDSC_BLOCKED_DIAG_ARGS(R)                                                                    \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pt, const R *pq, const R *ptendency_cml_t, const R *ptendency_cml_q,                     \
      const R *ptendency_cml_a, const R *ptendency_cml_cld, R *ptendency_loc_t, R *ptendency_loc_q,     \
      R *ptendency_loc_a, R *ptendency_loc_cld, const R *pvfl, const R *pvfi, const R *pap,             \
      const R *paph, const R *plsm, const int32_t *ldcum, const int32_t *ktype, R *peis,                \
      const R *plu, R *plude, const R *pludeli, const R *psnde, const R *pmfu, const R *pmfd,           \
      const R *pa, const R *pclv, R *pcovptot, R *prainfrac_toprfz, R *pfsqlf, R *pfsqif,              \
      R *pfcqnng, R *pfcqlng, R *pfsqrf, R *pfsqsf, R *pfcqrng, R *pfcqsng, R *pfsqltur, R *pfsqitur,   \
      R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, const R *pvfa, const R *pdyna, const R *pdynl,        \
      const R *pdyni, const R *pdyns, const R *phrsw, const R *phrlw, const int32_t *kpbltype,          \
      const int32_t *kctop, R *pextra, int kfldx, const R *pgaw, R *pfsd, void *stream

int cloudsc_blocked_diag_dp_c(CLOUDSC_BLOCKED_DIAG_ARGS(double));
int cloudsc_blocked_diag_sp_c(CLOUDSC_BLOCKED_DIAG_ARGS(float));
int cloudsc_blocked_diag_mixed_c(CLOUDSC_BLOCKED_DIAG_ARGS(float));

/* Last HIP error text of the calling thread (empty string if none). */
const char *cloudsc_last_error(void);

/* Replaces the reference's negative-condensate warnings WRITE(NULOUT,*) 'CLOUDSC: WARNING Negative cloud
 * liquid!' / 'ice!' (cloudsc.F90:3794, :3802): the number of points of the calling thread's last per-block
 * CLOUDSC call (cloudsc_{dp,sp,mixed}_c) where the condensate fell below -RLMIN before the clean-up.  The
 * reference prints one line per point; a host that wants the lines prints them from these counts. */
int cloudsc_last_warnings(int64_t *n_negative_liquid, int64_t *n_negative_ice);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_HIP_H */
