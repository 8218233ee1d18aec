// C entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY (see
// cloudsc_oracle.hpp).  Built into oracle/liboracle_cloudsc.so by
// oracle/Makefile; loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py, never by the product library.
//
//   oracle_cloudsc_{dp,sp,mixed}        one NPROMA block, CLOUDSC's argument list
//   oracle_cloudsc_blocked_{dp,sp,mixed} all blocks, OpenMP over blocks with
//       SCHEDULE(DYNAMIC,1) as the reference driver (ec_phys_drv.F90:430-439)
//   oracle_cloudsc_blocked_diag_{dp,sp,mixed}  the same with the LCLDBUD* budgets into PEXTRA and the PFSD of
//       LRAD_CLOUD_INHOMOG
//   oracle_satadj_{dp,sp,mixed}         one block of CLOUD_SATADJ, the argument list of
//       include/cloudsc_satadj.h
//   oracle_satadj_blocked_{dp,sp,mixed}  all blocks, OpenMP over blocks
//   oracle_sltend_supersat_blocked_{dp,sp,mixed}  SLTEND's supersaturation adjustment, all blocks
//   oracle_cloudsc2_{dp,sp}, oracle_cloudsc2_blocked_{dp,sp}  CLOUDSC2 (cloudsc2_oracle.hpp), one block / all blocks
//   oracle_cloud_s_layer_blocked_{dp,sp}                     CLOUD_S_LAYER with LEPCLD2: SATUR, CLOUDSC2, wrap-up
//   oracle_supersatcheck_*, oracle_supersatcheck_k_dp (with KFTLIQICE), oracle_lu5_dp,
//   oracle_truncate_sinks_dp, oracle_thermo_dp, oracle_qsatmixadj_dp  unit-level entry points for the tests
#include <omp.h>

#include "../include/cloudsc_satadj.h"
#include "cloudsc_oracle.hpp"
#include "cloudsc2_oracle.hpp"
#include "satadj_oracle.hpp"
#include "sltend_oracle.hpp"

namespace {

constexpr int32_t kBudOut = CLOUDSC_BUD_C | CLOUDSC_BUD_L | CLOUDSC_BUD_I | CLOUDSC_BUD_T | CLOUDSC_BUD_VERTINT;

// The same rules as the product library: budgets that write PEXTRA, the PFSD of LRAD_CLOUD_INHOMOG and the PGAW
// that LCLOUD_INHOMOG reads only where PEXTRA, PGAW and PFSD are arguments.
int check_config(const cloudsc_params_t* p, bool diag = false) {
  if (!p) return CLOUDSC_ENULL;
  if (p->laeraware || p->lspp) return CLOUDSC_EUNSUPPORTED;
  if (((p->lcldbud & kBudOut) || p->lrad_cloud_inhomog || p->lcloud_inhomog) && !diag) return CLOUDSC_EUNSUPPORTED;
  if (p->lcldbud & ~255) return CLOUDSC_EUNSUPPORTED;
  return CLOUDSC_OK;
}

// PEXTRA fields the switches fill: the value of IS after the last budget block (cloudsc.F90:3832-4101).
int budget_fields(int32_t f) {
  int is = (f & CLOUDSC_BUD_LBUD23) ? 26 : 0;
  if (f & CLOUDSC_BUD_VERTINT) is += 1;
  if (f & CLOUDSC_BUD_C) is += 12;
  if (f & CLOUDSC_BUD_L) is += 22;
  if (f & CLOUDSC_BUD_I) is += 20;
  if (f & CLOUDSC_BUD_T) is += 13;
  return is;
}

template <typename B>
int check_budget(const cloudsc_params_t* p, int klev, const B& b) {
  if (p->lrad_cloud_inhomog && (!b.pgaw || !b.pfsd)) return CLOUDSC_ENULL;
  if (p->lcloud_inhomog && !b.pgaw) return CLOUDSC_ENULL;
  if (!(p->lcldbud & kBudOut)) return CLOUDSC_OK;
  const void* req[] = {b.pvfa, b.pdyna, b.pdynl, b.pdyni, b.pdyns, b.phrsw, b.phrlw, b.kpbltype, b.kctop, b.pextra};
  for (const void* x : req)
    if (!x) return CLOUDSC_ENULL;
  if (b.kfldx < budget_fields(p->lcldbud)) return CLOUDSC_EDIMS;
  if ((p->lcldbud & CLOUDSC_BUD_VERTINT) && klev < 83) return CLOUDSC_EDIMS;  // slots 4..83 of the field
  return CLOUDSC_OK;
}

// Negative-condensate warning counts of the calling thread's last per-block call (oracle_last_warnings).
thread_local int64_t g_last_warn[2] = {0, 0};

int check_dims(const cloudsc_params_t* p, int kidia, int kfdia, int klon, int klev) {
  if (klon < 1 || klev < 2 || kidia < 1 || kfdia < kidia || kfdia > klon) return CLOUDSC_EDIMS;
  // CLOUD_SUPERSATCHECK reads the pressure section from KIDIA (cloudsc.F90:3776): stay inside KLON.
  if (kidia + kfdia - 1 > klon) return CLOUDSC_EDIMS;
  if (p->ncldtop < 2 || p->ncldtop > klev) return CLOUDSC_EDIMS;
  return CLOUDSC_OK;
}

template <typename R, typename RD>
int run_block(const cloudsc_params_t* p, int kidia, int kfdia, int klon, int klev, R ptsphy, const R* pt,
              const R* pq, const R* tcml_t, const R* tcml_q, const R* tcml_a, const R* tcml_cld, R* tloc_t,
              R* tloc_q, R* tloc_a, R* tloc_cld, const R* pvfl, const R* pvfi, const R* pap, const R* paph,
              const R* plsm, const int32_t* ldcum, const int32_t* ktype, const R* peis, const R* plu, R* plude,
              const R* pludeli, const R* psnde, const R* pmfu, const R* pmfd, const R* pa, const R* pclv,
              R* pcovptot, R* prainfrac_toprfz, R* pfsqlf, R* pfsqif, R* pfcqnng, R* pfcqlng, R* pfsqrf,
              R* pfsqsf, R* pfcqrng, R* pfcqsng, R* pfsqltur, R* pfsqitur, R* pfplsl, R* pfplsn, R* pfhpsl,
              R* pfhpsn, const oracle::BudArgs<R>& bud) {
  int rc = check_config(p, true);
  if (rc) return rc;
  rc = check_dims(p, kidia, kfdia, klon, klev);
  if (rc) return rc;
  rc = check_budget(p, klev, bud);
  if (rc) return rc;
  oracle::Args<R> a{kidia,  kfdia,  klon,    klev,    ptsphy,  pt,       pq,      tcml_t,  tcml_q,  tcml_a,
                    tcml_cld, tloc_t, tloc_q, tloc_a, tloc_cld, pvfl, pvfi, pap, paph, plsm,
                    ldcum,  ktype,  peis,    plu,     plude,   pludeli,  psnde,   pmfu,    pmfd,    pa,
                    pclv,   pcovptot, prainfrac_toprfz, pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf,
                    pfcqrng, pfcqsng, pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn,
                    ((p->lcldbud & kBudOut) || p->lrad_cloud_inhomog || p->lcloud_inhomog) ? &bud : nullptr};
  g_last_warn[0] = g_last_warn[1] = 0;
  a.nwarn = g_last_warn;
  oracle::cloudsc<R, RD>(*p, a);
  return CLOUDSC_OK;
}

#define ORACLE_BLOCK_CALL(R, RD)                                                                              \
  (void)lddiag, (void)pdynr, (void)pvervel, (void)pgeom, (void)pgaw, (void)pgp2dspp, (void)ldslphy,          \
      (void)plcrit_aer, (void)picrit_aer, (void)pre_ice, (void)pccn, (void)pnice;                               \
  const oracle::BudArgs<R> bud{pvfa, pdyna, pdynl, pdyni, pdyns, phrsw, phrlw, kpbltype, kctop, pextra, kfldx,  \
                               pgaw, pfsd};                                                                     \
  return run_block<R, RD>(ydparams, kidia, kfdia, klon, klev, ptsphy, pt, pq, ptendency_cml_t, ptendency_cml_q, \
                          ptendency_cml_a, ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, ptendency_loc_a, \
                          ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis, plu, plude,        \
                          pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz, pfsqlf, pfsqif,      \
                          pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng, pfsqltur, pfsqitur, pfplsl,        \
                          pfplsn, pfhpsl, pfhpsn, bud)

template <typename R, typename RD>
int run_blocked(const cloudsc_params_t* p, int ngptot, int nproma, int klev, int ngpblks, R ptsphy, const R* pt,
                const R* pq, const R* tcml_t, const R* tcml_q, const R* tcml_a, const R* tcml_cld, R* tloc_t,
                R* tloc_q, R* tloc_a, R* tloc_cld, const R* pvfl, const R* pvfi, const R* pap, const R* paph,
                const R* plsm, const int32_t* ldcum, const int32_t* ktype, const R* peis, const R* plu, R* plude,
                const R* pludeli, const R* psnde, const R* pmfu, const R* pmfd, const R* pa, const R* pclv,
                R* pcovptot, R* prainfrac_toprfz, R* pfsqlf, R* pfsqif, R* pfcqnng, R* pfcqlng, R* pfsqrf,
                R* pfsqsf, R* pfcqrng, R* pfcqsng, R* pfsqltur, R* pfsqitur, R* pfplsl, R* pfplsn, R* pfhpsl,
                R* pfhpsn, int nthreads, const oracle::BudArgs<R>* bud = nullptr) {
  int rc = check_config(p, bud != nullptr);
  if (rc) return rc;
  if (nproma < 1 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  rc = check_dims(p, 1, nproma, nproma, klev);
  if (rc) return rc;
  if (bud) {
    rc = check_budget(p, klev, *bud);
    if (rc) return rc;
  }
  const bool budgets = bud && ((p->lcldbud & kBudOut) || p->lrad_cloud_inhomog || p->lcloud_inhomog);
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1), s1 = nproma;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh, o1 = ibl * s1, o5 = ibl * s2 * 5, o4 = ibl * s2 * 4,
                 o2b = ibl * s2 * 2;
    oracle::BudArgs<R> bb{};
    if (budgets) {  // PEXTRA block IBL is (NPROMA,KLEV,KFLDX); arrays of a diagnostic that is off may be NULL
      auto at = [](auto* x, size_t n) { return x ? x + n : x; };
      bb = {at(bud->pvfa, o2),     at(bud->pdyna, o2),    at(bud->pdynl, o2), at(bud->pdyni, o2),
            at(bud->pdyns, o2),    at(bud->phrsw, o2),    at(bud->phrlw, o2), at(bud->kpbltype, o1),
            at(bud->kctop, o1),    at(bud->pextra, o2 * bud->kfldx), bud->kfldx, at(bud->pgaw, o1),
            at(bud->pfsd, o2)};
    }
    oracle::Args<R> a{1,
                      kfdia,
                      nproma,
                      klev,
                      ptsphy,
                      pt + o2,
                      pq + o2,
                      tcml_t + o2,
                      tcml_q + o2,
                      tcml_a + o2,
                      tcml_cld + o5,
                      tloc_t + o2,
                      tloc_q + o2,
                      tloc_a + o2,
                      tloc_cld + o5,
                      pvfl + o2,
                      pvfi + o2,
                      pap + o2,
                      paph + oh,
                      plsm + o1,
                      ldcum + o1,
                      ktype + o1,
                      peis + o1,
                      plu + o2,
                      plude + o2,
                      pludeli + o4,
                      psnde + o2b,
                      pmfu + o2,
                      pmfd + o2,
                      pa + o2,
                      pclv + o5,
                      pcovptot + o2,
                      prainfrac_toprfz + o1,
                      pfsqlf + oh,
                      pfsqif + oh,
                      pfcqnng + oh,
                      pfcqlng + oh,
                      pfsqrf + oh,
                      pfsqsf + oh,
                      pfcqrng + oh,
                      pfcqsng + oh,
                      pfsqltur + oh,
                      pfsqitur + oh,
                      pfplsl + oh,
                      pfplsn + oh,
                      pfhpsl + oh,
                      pfhpsn + oh,
                      budgets ? &bb : nullptr};
    oracle::cloudsc<R, RD>(*p, a);
  }
  return CLOUDSC_OK;
}

#define ORACLE_BLOCKED_CALL(R, RD)                                                                              \
  return run_blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy, pt, pq, ptendency_cml_t,          \
                            ptendency_cml_q, ptendency_cml_a, ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, \
                            ptendency_loc_a, ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis,  \
                            plu, plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz, pfsqlf, \
                            pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng, pfsqltur, pfsqitur,       \
                            pfplsl, pfplsn, pfhpsl, pfhpsn, nthreads)

#define ORACLE_BLOCKED_ARGS(R)                                                                          \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pt, const R *pq, const R *ptendency_cml_t, const R *ptendency_cml_q,                     \
      const R *ptendency_cml_a, const R *ptendency_cml_cld, R *ptendency_loc_t, R *ptendency_loc_q,     \
      R *ptendency_loc_a, R *ptendency_loc_cld, const R *pvfl, const R *pvfi, const R *pap,             \
      const R *paph, const R *plsm, const int32_t *ldcum, const int32_t *ktype, R *peis,                \
      const R *plu, R *plude, const R *pludeli, const R *psnde, const R *pmfu, const R *pmfd,           \
      const R *pa, const R *pclv, R *pcovptot, R *prainfrac_toprfz, R *pfsqlf, R *pfsqif,              \
      R *pfcqnng, R *pfcqlng, R *pfsqrf, R *pfsqsf, R *pfcqrng, R *pfcqsng, R *pfsqltur, R *pfsqitur,   \
      R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, int nthreads

// The blocked oracle with the diagnostics: the arguments of cloudsc_blocked_diag_*_c in their order (host arrays),
// then the thread count.
#define ORACLE_BLOCKED_DIAG_ARGS(R)                                                                   \
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
      const int32_t *kctop, R *pextra, int kfldx, const R *pgaw, R *pfsd, int nthreads

#define ORACLE_BLOCKED_DIAG_CALL(R, RD)                                                                         \
  const oracle::BudArgs<R> bud{pvfa, pdyna, pdynl, pdyni, pdyns, phrsw, phrlw, kpbltype, kctop, pextra, kfldx,    \
                               pgaw, pfsd};                                                                       \
  return run_blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy, pt, pq, ptendency_cml_t,          \
                            ptendency_cml_q, ptendency_cml_a, ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, \
                            ptendency_loc_a, ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis,  \
                            plu, plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz, pfsqlf, \
                            pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng, pfsqltur, pfsqitur,       \
                            pfplsl, pfplsn, pfhpsl, pfhpsn, nthreads, &bud)

// ---- CLOUD_SATADJ ----
template <typename R>
int satadj_check(const cloudsc_params_t* p, int kidia, int kfdia, int klon, int klev) {
  if (!p) return CLOUDSC_ENULL;
  if (p->lspp || p->lcldbud) return CLOUDSC_EUNSUPPORTED;
  if (klon < 1 || klev < 2 || kidia < 1 || kfdia < kidia || kfdia > klon) return CLOUDSC_EDIMS;
  if (p->ncldtop < 2 || p->ncldtop > klev) return CLOUDSC_EDIMS;
  return CLOUDSC_OK;
}

#define ORACLE_SATADJ_CALL(R, RD)                                                                              \
  (void)ktype, (void)lddiag, (void)pmfu, (void)pmfd, (void)pgp2dspp, (void)pextra, (void)kfldx;                 \
  const int rc = satadj_check<R>(ydparams, kidia, kfdia, klon, klev);                                          \
  if (rc) return rc;                                                                                           \
  oracle::SatadjArgs<R> a{kidia, kfdia, klon, klev, ptsphy, pap, paph, pt, pq, pa, pl, pi, ptendency_cml_t,    \
                          ptendency_cml_q, ptendency_cml_a, ptendency_cml_l, ptendency_cml_i, phrsw, phrlw,     \
                          ptendency_vdf_t, ptendency_vdf_q, pludeli, pvervel, ptendency_loc_t, ptendency_loc_q, \
                          ptendency_loc_a, ptendency_loc_l, ptendency_loc_i};                                  \
  oracle::satadj<R, RD>(*ydparams, a);                                                                         \
  return CLOUDSC_OK

template <typename R, typename RD>
int satadj_blocked(const cloudsc_params_t* p, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,
                   const R* pap, const R* paph, const R* pt, const R* pq, const R* pa, const R* pclv,
                   const R* tcml_t, const R* tcml_q, const R* tcml_a, const R* tcml_cld, const R* phrsw,
                   const R* phrlw, const R* tvdf_t, const R* tvdf_q, const R* pludeli, const R* pvervel, R* tloc_t,
                   R* tloc_q, R* tloc_a, R* tloc_cld, int nthreads) {
  int rc = satadj_check<R>(p, 1, nproma, nproma, klev);
  if (rc) return rc;
  if (ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh, o5 = ibl * s2 * 5, o4 = ibl * s2 * 4;
    oracle::SatadjArgs<R> a{1, kfdia, nproma, klev, ptsphy, pap + o2, paph + oh, pt + o2, pq + o2, pa + o2,
                            pclv + o5, pclv + o5 + s2, tcml_t + o2, tcml_q + o2, tcml_a + o2, tcml_cld + o5,
                            tcml_cld + o5 + s2, phrsw + o2, phrlw + o2, tvdf_t + o2, tvdf_q + o2, pludeli + o4,
                            pvervel + o2, tloc_t + o2, tloc_q + o2, tloc_a + o2, tloc_cld + o5, tloc_cld + o5 + s2};
    oracle::satadj<R, RD>(*p, a);
  }
  return CLOUDSC_OK;
}

#define ORACLE_SATADJ_BLOCKED_ARGS(R)                                                                   \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R ptsphy,            \
      const R *pap, const R *paph, const R *pt, const R *pq, const R *pa, const R *pclv,                \
      const R *ptendency_cml_t, const R *ptendency_cml_q, const R *ptendency_cml_a,                     \
      const R *ptendency_cml_cld, const R *phrsw, const R *phrlw, const R *ptendency_vdf_t,             \
      const R *ptendency_vdf_q, const R *pludeli, const R *pvervel, R *ptendency_loc_t,                 \
      R *ptendency_loc_q, R *ptendency_loc_a, R *ptendency_loc_cld, int nthreads

#define ORACLE_SATADJ_BLOCKED_CALL(R, RD)                                                                      \
  return satadj_blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy, pap, paph, pt, pq, pa, pclv,     \
                               ptendency_cml_t, ptendency_cml_q, ptendency_cml_a, ptendency_cml_cld, phrsw,      \
                               phrlw, ptendency_vdf_t, ptendency_vdf_q, pludeli, pvervel, ptendency_loc_t,       \
                               ptendency_loc_q, ptendency_loc_a, ptendency_loc_cld, nthreads)

template <typename R, typename RD>
int sltend_blocked(const cloudsc_params_t* p, int ngptot, int nproma, int klev, int ngpblks, R tsphy, const R* pt,
                   const R* pq, const R* pa, const R* prsf1, const R* prs1, const R* pcent, const R* pcenq,
                   const R* pcena, R* ptent, R* ptenq, R* ptena, R* ptencld, R* adja, R* adjl, R* adji,
                   int nthreads) {
  if (!p) return CLOUDSC_ENULL;
  if (nproma < 1 || klev < 2 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  if (p->ncldtop < 1 || p->ncldtop > klev) return CLOUDSC_EDIMS;
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh, o5 = ibl * s2 * 5;
    oracle::SltendArgs<R> a{1, kfdia, nproma, klev, tsphy, pt + o2, pq + o2, pa + o2, prsf1 + o2, prs1 + oh,
                            pcent + o2, pcenq + o2, pcena + o2, ptent + o2, ptenq + o2, ptena + o2, ptencld + o5,
                            adja ? adja + o2 : nullptr, adjl ? adjl + o2 : nullptr, adji ? adji + o2 : nullptr};
    oracle::sltend_supersat<R, RD>(*p, a);
  }
  return CLOUDSC_OK;
}

#define ORACLE_SLTEND_ARGS(R)                                                                                  \
  const cloudsc_params_t *ydparams, int ngptot, int nproma, int klev, int ngpblks, R tsphy, const R *pt,       \
      const R *pq, const R *pa, const R *prsf1, const R *prs1, const R *pcent, const R *pcenq, const R *pcena, \
      R *ptent, R *ptenq, R *ptena, R *ptencld, R *padja, R *padjl, R *padji, int nthreads

#define ORACLE_SLTEND_CALL(R, RD)                                                                              \
  return sltend_blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, tsphy, pt, pq, pa, prsf1, prs1, pcent, \
                               pcenq, pcena, ptent, ptenq, ptena, ptencld, padja, padjl, padji, nthreads)

// ---- CLOUDSC2 ----
template <typename R>
int cloudsc2_check(const cloudsc_params_t* p, const cloudsc2_params_t* p2, int kidia, int kfdia, int klon, int ktdia,
                   int klev) {
  if (!p || !p2) return CLOUDSC_ENULL;
  if (klon < 1 || klev < 2 || kidia < 1 || kfdia < kidia || kfdia > klon || ktdia < 1 || ktdia > klev)
    return CLOUDSC_EDIMS;
  return CLOUDSC_OK;
}

#define ORACLE_CLOUDSC2_CALL(R)                                                                                   \
  int rc = cloudsc2_check<R>(ydparams, ydparams2, kidia, kfdia, klon, ktdia, klev);                               \
  if (rc) return rc;                                                                                              \
  oracle::Cloudsc2Args<R> a{kidia,  kfdia, klon,  ktdia, klev,   ldrain1d != 0, ptsphy, ceta,   paphp1, papp1,   \
                            pqm1,   pqs,   ptm1,  pl,    pi,     plude,         plu,    pmfu,   pmfd,   ptent,   \
                            pgtent, ptenq, pgtenq, ptenl, pgtenl, pteni,        pgteni, pclc,   pfplsl, pfplsn,  \
                            pfhpsl, pfhpsn, pcovptot};                                                             \
  oracle::cloudsc2<R>(*ydparams, *ydparams2, a);                                                                  \
  return CLOUDSC_OK;

// All blocks of blocked host arrays (NGPBLKS outermost), OpenMP over blocks; the CLOUDSC2_BLOCKED_ARGS order
// without the stream, then the thread count.
template <typename R>
int cloudsc2_blocked(const cloudsc_params_t* p, const cloudsc2_params_t* p2, const R* ceta, int ngptot, int nproma,
                     int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R* paphp1, const R* papp1,
                     const R* pqm1, const R* pqs, const R* ptm1, const R* pl, const R* pi, const R* plude,
                     const R* plu, const R* pmfu, const R* pmfd, R* ptent, const R* pgtent, R* ptenq,
                     const R* pgtenq, R* ptenl, const R* pgtenl, R* pteni, const R* pgteni, R* pclc, R* pfplsl,
                     R* pfplsn, R* pfhpsl, R* pfhpsn, R* pcovptot, int nthreads) {
  int rc = cloudsc2_check<R>(p, p2, 1, nproma, nproma, ktdia, klev);
  if (rc) return rc;
  if (ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh;
    oracle::Cloudsc2Args<R> a{1,           kfdia,       nproma,      ktdia,       klev,         ldrain1d != 0,
                              ptsphy,      ceta,        paphp1 + oh, papp1 + o2,  pqm1 + o2,    pqs + o2,
                              ptm1 + o2,   pl + o2,     pi + o2,     plude + o2,  plu + o2,     pmfu + o2,
                              pmfd + o2,   ptent + o2,  pgtent + o2, ptenq + o2,  pgtenq + o2,  ptenl + o2,
                              pgtenl + o2, pteni + o2,  pgteni + o2, pclc + o2,   pfplsl + oh,  pfplsn + oh,
                              pfhpsl + oh, pfhpsn + oh, pcovptot + o2};
    oracle::cloudsc2<R>(*p, *p2, a);
  }
  return CLOUDSC_OK;
}

#define ORACLE_CLOUDSC2_BLOCKED_ARGS(R)                                                                 \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot,       \
      int nproma, int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R *paphp1,            \
      const R *papp1, const R *pqm1, const R *pqs, const R *ptm1, const R *pl, const R *pi,               \
      const R *plude, const R *plu, const R *pmfu, const R *pmfd, R *ptent, const R *pgtent, R *ptenq,   \
      const R *pgtenq, R *ptenl, const R *pgtenl, R *pteni, const R *pgteni, R *pclc, R *pfplsl,         \
      R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot, int nthreads

#define ORACLE_CLOUDSC2_BLOCKED_CALL(R)                                                                        \
  return cloudsc2_blocked<R>(ydparams, ydparams2, ceta, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d, ptsphy, \
                             paphp1, papp1, pqm1, pqs, ptm1, pl, pi, plude, plu, pmfu, pmfd, ptent, pgtent,     \
                             ptenq, pgtenq, ptenl, pgtenl, pteni, pgteni, pclc, pfplsl, pfplsn, pfhpsl, pfhpsn, \
                             pcovptot, nthreads)

// CLOUD_S_LAYER (LEPCLD2) over all blocks of blocked host arrays: the CLOUD_S_LAYER_BLOCKED_ARGS order without
// the stream, then the thread count.
template <typename R>
int cloud_s_layer_blocked(const cloudsc_params_t* p, const cloudsc2_params_t* p2, const R* ceta, int ngptot,
                          int nproma, int ktdia, int klev, int ngpblks, int ldrain1d, R tsphy, const R* prs1,
                          const R* prsf1, const R* pq, const R* pt, const R* pa, const R* pl, const R* pi,
                          const R* plude, const R* plu, const R* pmfu, const R* pmfd, const R* ptendency_t,
                          const R* ptendency_q, const R* ptendency_l, const R* ptendency_i, R* ptendency_loc_t,
                          R* ptendency_loc_q, R* ptendency_loc_a, R* ptendency_loc_l, R* ptendency_loc_i,
                          R* ptendency_loc_r, R* ptendency_loc_s, R* pqsat, R* pclc, R* pfplsl, R* pfplsn,
                          R* pfhpsl, R* pfhpsn, R* pcovptot, R* pprectype, R* pfzra, R* pqiwp, R* pqlwp,
                          int nthreads) {
  int rc = cloudsc2_check<R>(p, p2, 1, nproma, nproma, ktdia, klev);
  if (rc) return rc;
  if (ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  if (ptendency_loc_a && !pa) return CLOUDSC_ENULL;
  if (p2->lepcld2 && (!pmfu || !pmfd || !ptendency_l || !ptendency_i)) return CLOUDSC_ENULL;
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1);
  auto off = [](auto* x, size_t n) { return x ? x + n : x; };
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh, o1 = (size_t)ibl * nproma;
    oracle::Cloudsc2Args<R> a{1,
                              kfdia,
                              nproma,
                              ktdia,
                              klev,
                              ldrain1d != 0,
                              tsphy,
                              ceta,
                              prs1 + oh,
                              prsf1 + o2,
                              pq + o2,
                              nullptr,
                              pt + o2,
                              pl + o2,
                              pi + o2,
                              plude + o2,
                              plu + o2,
                              off(pmfu, o2),
                              off(pmfd, o2),
                              ptendency_loc_t + o2,
                              ptendency_t + o2,
                              ptendency_loc_q + o2,
                              ptendency_q + o2,
                              ptendency_loc_l + o2,
                              off(ptendency_l, o2),
                              ptendency_loc_i + o2,
                              off(ptendency_i, o2),
                              off(pclc, o2),
                              pfplsl + oh,
                              pfplsn + oh,
                              pfhpsl + oh,
                              pfhpsn + oh,
                              pcovptot + o2};
    const oracle::LayerArgs<R> l{off(pa, o2),    off(ptendency_loc_a, o2), off(ptendency_loc_r, o2),
                                 off(ptendency_loc_s, o2), off(pqsat, o2), off(pprectype, o1),
                                 off(pfzra, o1),      off(pqiwp, o2),           off(pqlwp, o2)};
    oracle::cloud_s_layer<R>(*p, *p2, a, l);
  }
  return CLOUDSC_OK;
}

#define ORACLE_LAYER_ARGS(R)                                                                                   \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot, int nproma, \
      int ktdia, int klev, int ngpblks, int ldrain1d, R tsphy, const R *prs1, const R *prsf1, const R *pq,      \
      const R *pt, const R *pa, const R *pl, const R *pi, const R *plude, const R *plu, const R *pmfu,          \
      const R *pmfd, const R *ptendency_t, const R *ptendency_q, const R *ptendency_l, const R *ptendency_i,   \
      R *ptendency_loc_t, R *ptendency_loc_q, R *ptendency_loc_a, R *ptendency_loc_l, R *ptendency_loc_i,      \
      R *ptendency_loc_r, R *ptendency_loc_s, R *pqsat, R *pclc, R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn,    \
      R *pcovptot, R *pprectype, R *pfzra, R *pqiwp, R *pqlwp, int nthreads

#define ORACLE_LAYER_CALL(R)                                                                                     \
  return cloud_s_layer_blocked<R>(ydparams, ydparams2, ceta, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d,      \
                                  tsphy, prs1, prsf1, pq, pt, pa, pl, pi, plude, plu, pmfu, pmfd, ptendency_t,     \
                                  ptendency_q, ptendency_l, ptendency_i, ptendency_loc_t, ptendency_loc_q,         \
                                  ptendency_loc_a, ptendency_loc_l, ptendency_loc_i, ptendency_loc_r,              \
                                  ptendency_loc_s, pqsat, pclc, pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot,          \
                                  pprectype, pfzra, pqiwp, pqlwp, nthreads)

// CLOUDST over all blocks: the CLOUDST_BLOCKED_ARGS order without the stream, then the thread count.
template <typename R>
int cloudst_blocked(const cloudsc_params_t* p, const cloudsc2_params_t* p2, const R* ceta, int ngptot, int nproma,
                    int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R* paphp1, const R* papp1,
                    const R* pqm1, const R* pqs, const R* ptm1, const R* plude, const R* plu, R* ptent,
                    const R* pgtent, R* ptenq, const R* pgtenq, R* pclc, R* pqiwp, R* pqlwp, R* pfplsl, R* pfplsn,
                    R* pfhpsl, R* pfhpsn, R* pcovptot, int nthreads) {
  int rc = cloudsc2_check<R>(p, p2, 1, nproma, nproma, ktdia, klev);
  if (rc) return rc;
  if (ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma) return CLOUDSC_EDIMS;
  const size_t s2 = (size_t)nproma * klev, sh = (size_t)nproma * (klev + 1);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ibl = 0; ibl < ngpblks; ++ibl) {
    const int kfdia = std::min(nproma, ngptot - ibl * nproma);
    const size_t o2 = ibl * s2, oh = ibl * sh;
    const oracle::CloudstArgs<R> a{1,           kfdia,       nproma,      ktdia,       klev,        ldrain1d != 0,
                                   ptsphy,      ceta,        paphp1 + oh, papp1 + o2,  pqm1 + o2,   pqs + o2,
                                   ptm1 + o2,   plude + o2,  plu + o2,    ptent + o2,  pgtent + o2, ptenq + o2,
                                   pgtenq + o2, pclc + o2,   pqiwp + o2,  pqlwp + o2,  pfplsl + oh, pfplsn + oh,
                                   pfhpsl + oh, pfhpsn + oh, pcovptot + o2};
    oracle::cloudst<R>(*p, *p2, a);
  }
  return CLOUDSC_OK;
}

#define ORACLE_CLOUDST_ARGS(R)                                                                                 \
  const cloudsc_params_t *ydparams, const cloudsc2_params_t *ydparams2, const R *ceta, int ngptot, int nproma, \
      int ktdia, int klev, int ngpblks, int ldrain1d, R ptsphy, const R *paphp1, const R *papp1, const R *pqm1, \
      const R *pqs, const R *ptm1, const R *plude, const R *plu, R *ptent, const R *pgtent, R *ptenq,           \
      const R *pgtenq, R *pclc, R *pqiwp, R *pqlwp, R *pfplsl, R *pfplsn, R *pfhpsl, R *pfhpsn, R *pcovptot,      \
      int nthreads

#define ORACLE_CLOUDST_CALL(R)                                                                                \
  return cloudst_blocked<R>(ydparams, ydparams2, ceta, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d, ptsphy, \
                            paphp1, papp1, pqm1, pqs, ptm1, plude, plu, ptent, pgtent, ptenq, pgtenq, pclc, pqiwp,  \
                            pqlwp, pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot, nthreads)

}  // namespace

extern "C" {

#define ORACLE_CLOUDST_BLOCK_CALL(R)                                                                            \
  int rc = cloudsc2_check<R>(ydparams, ydparams2, kidia, kfdia, klon, ktdia, klev);                               \
  if (rc) return rc;                                                                                              \
  const oracle::CloudstArgs<R> a{kidia, kfdia, klon,  ktdia,  klev,   ldrain1d != 0, ptsphy, ceta,   paphp1, papp1, \
                                 pqm1,  pqs,   ptm1,  plude,  plu,    ptent,         pgtent, ptenq,  pgtenq, pclc,  \
                                 pqiwp, pqlwp, pfplsl, pfplsn, pfhpsl, pfhpsn,       pcovptot};                    \
  oracle::cloudst<R>(*ydparams, *ydparams2, a);                                                                   \
  return CLOUDSC_OK;

int oracle_cloudst_dp(CLOUDST_ARGS(double)) { ORACLE_CLOUDST_BLOCK_CALL(double) }
int oracle_cloudst_sp(CLOUDST_ARGS(float)) { ORACLE_CLOUDST_BLOCK_CALL(float) }
int oracle_cloudst_blocked_dp(ORACLE_CLOUDST_ARGS(double)) { ORACLE_CLOUDST_CALL(double); }
int oracle_cloudst_blocked_sp(ORACLE_CLOUDST_ARGS(float)) { ORACLE_CLOUDST_CALL(float); }
int oracle_cloud_s_layer_blocked_dp(ORACLE_LAYER_ARGS(double)) { ORACLE_LAYER_CALL(double); }
int oracle_cloud_s_layer_blocked_sp(ORACLE_LAYER_ARGS(float)) { ORACLE_LAYER_CALL(float); }

int oracle_cloudsc2_dp(CLOUDSC2_ARGS(double)) { ORACLE_CLOUDSC2_CALL(double) }
int oracle_cloudsc2_sp(CLOUDSC2_ARGS(float)) { ORACLE_CLOUDSC2_CALL(float) }
int oracle_cloudsc2_blocked_dp(ORACLE_CLOUDSC2_BLOCKED_ARGS(double)) { ORACLE_CLOUDSC2_BLOCKED_CALL(double); }
int oracle_cloudsc2_blocked_sp(ORACLE_CLOUDSC2_BLOCKED_ARGS(float)) { ORACLE_CLOUDSC2_BLOCKED_CALL(float); }

int oracle_cloudsc_dp(CLOUDSC_ARGS(double)) { ORACLE_BLOCK_CALL(double, double); }
void oracle_last_warnings(int64_t* n_negative_liquid, int64_t* n_negative_ice) {
  *n_negative_liquid = g_last_warn[0];
  *n_negative_ice = g_last_warn[1];
}
int oracle_cloudsc_sp(CLOUDSC_ARGS(float)) { ORACLE_BLOCK_CALL(float, float); }
int oracle_cloudsc_mixed(CLOUDSC_ARGS(float)) { ORACLE_BLOCK_CALL(float, double); }

int oracle_cloudsc_blocked_dp(ORACLE_BLOCKED_ARGS(double)) { ORACLE_BLOCKED_CALL(double, double); }
int oracle_cloudsc_blocked_sp(ORACLE_BLOCKED_ARGS(float)) { ORACLE_BLOCKED_CALL(float, float); }
int oracle_cloudsc_blocked_mixed(ORACLE_BLOCKED_ARGS(float)) { ORACLE_BLOCKED_CALL(float, double); }

int oracle_cloudsc_blocked_diag_dp(ORACLE_BLOCKED_DIAG_ARGS(double)) { ORACLE_BLOCKED_DIAG_CALL(double, double); }
int oracle_cloudsc_blocked_diag_sp(ORACLE_BLOCKED_DIAG_ARGS(float)) { ORACLE_BLOCKED_DIAG_CALL(float, float); }
int oracle_cloudsc_blocked_diag_mixed(ORACLE_BLOCKED_DIAG_ARGS(float)) {
  ORACLE_BLOCKED_DIAG_CALL(float, double);
}

// CLOUD_SUPERSATCHECK on n points (KIDIA = 1, KFDIA = n), outputs in RD.
int oracle_supersatcheck_dp(const cloudsc_params_t* p, int n, const double* t, const double* q, const double* a,
                            const double* pp, const double* psurf, double* t_adj, double* q_adj, double* a_adj,
                            double* l_adj, double* i_adj) {
  oracle::supersatcheck<double, double>(*p, 1, n, t, q, a, pp, psurf, t_adj, q_adj, a_adj, l_adj, i_adj);
  return 0;
}
int oracle_supersatcheck_sp(const cloudsc_params_t* p, int n, const float* t, const float* q, const float* a,
                            const float* pp, const float* psurf, float* t_adj, float* q_adj, float* a_adj,
                            float* l_adj, float* i_adj) {
  oracle::supersatcheck<float, float>(*p, 1, n, t, q, a, pp, psurf, t_adj, q_adj, a_adj, l_adj, i_adj);
  return 0;
}

// No-pivot LU solve of one 5x5 system (row-major lhs[a*5+b] = ZQLHS(a,b)).
int oracle_lu5_dp(const double* lhs_in, const double* rhs, double* x_out) {
  double lhs[5][5], x[5];
  for (int i = 0; i < 5; ++i) {
    x[i] = rhs[i];
    for (int j = 0; j < 5; ++j) lhs[i][j] = lhs_in[i * 5 + j];
  }
  oracle::lu5_solve<double>(lhs, x);
  for (int i = 0; i < 5; ++i) x_out[i] = x[i];
  return 0;
}

// Sink truncation on one column (row-major qa[m*5+n] = ZSOLQA(m,n)), in place.
int oracle_truncate_sinks_dp(double* qa_io, const double* qx) {
  double qa[5][5];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) qa[i][j] = qa_io[i * 5 + j];
  oracle::truncate_sinks<double>(qa, qx, 1.e-14);
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) qa_io[i * 5 + j] = qa[i][j];
  return 0;
}

// Thermodynamic statement functions on n temperatures:
// fn 0 FOEDELTA, 1 FOEEW, 2 FOEALFA, 3 FOEEWM, 4 FOEDEM, 5 FOELDCPM,
//    6 FOEELIQ, 7 FOEEICE, 8 FOEELIQ2ICE, 9 FOKOOP
int oracle_thermo_dp(const cloudsc_params_t* p, int fn, int n, const double* t, double* out) {
  const oracle::Thermo<double> th(*p);
  for (int i = 0; i < n; ++i) {
    const double x = t[i];
    double y;
    switch (fn) {
      case 0: y = th.foedelta(x); break;
      case 1: y = th.foeew(x); break;
      case 2: y = th.foealfa(x); break;
      case 3: y = th.foeewm(x); break;
      case 4: y = th.foedem(x); break;
      case 5: y = th.foeldcpm(x); break;
      case 6: y = th.foeeliq(x); break;
      case 7: y = th.foeeice(x); break;
      case 8: y = th.foeeliq2ice(x); break;
      case 9: y = th.fokoop(x); break;
      default: return -1;
    }
    out[i] = y;
  }
  return 0;
}

int oracle_satadj_dp(CLOUDSC_SATADJ_ARGS(double)) { ORACLE_SATADJ_CALL(double, double); }
int oracle_satadj_sp(CLOUDSC_SATADJ_ARGS(float)) { ORACLE_SATADJ_CALL(float, float); }
int oracle_satadj_mixed(CLOUDSC_SATADJ_ARGS(float)) { ORACLE_SATADJ_CALL(float, double); }

int oracle_satadj_blocked_dp(ORACLE_SATADJ_BLOCKED_ARGS(double)) { ORACLE_SATADJ_BLOCKED_CALL(double, double); }
int oracle_satadj_blocked_sp(ORACLE_SATADJ_BLOCKED_ARGS(float)) { ORACLE_SATADJ_BLOCKED_CALL(float, float); }
int oracle_satadj_blocked_mixed(ORACLE_SATADJ_BLOCKED_ARGS(float)) { ORACLE_SATADJ_BLOCKED_CALL(float, double); }

// CLOUD_SUPERSATCHECK on n points with an explicit KFTLIQICE (1 or 2).
int oracle_supersatcheck_k_dp(const cloudsc_params_t* p, int n, int kftliqice, const double* t, const double* q,
                              const double* a, const double* pp, const double* psurf, double* t_adj,
                              double* q_adj, double* a_adj, double* l_adj, double* i_adj) {
  oracle::supersatcheck<double, double>(*p, 1, n, t, q, a, pp, psurf, t_adj, q_adj, a_adj, l_adj, i_adj, kftliqice);
  return 0;
}

int oracle_sltend_supersat_blocked_dp(ORACLE_SLTEND_ARGS(double)) { ORACLE_SLTEND_CALL(double, double); }
int oracle_sltend_supersat_blocked_sp(ORACLE_SLTEND_ARGS(float)) { ORACLE_SLTEND_CALL(float, float); }
int oracle_sltend_supersat_blocked_mixed(ORACLE_SLTEND_ARGS(float)) { ORACLE_SLTEND_CALL(float, double); }

// QSATMIXADJ on n points: humidity after the two Newton iterations.
int oracle_qsatmixadj_dp(const cloudsc_params_t* p, int n, const double* t, const double* q, const double* pp,
                         double* out) {
  const oracle::Thermo<double> th(*p);
  for (int i = 0; i < n; ++i) out[i] = oracle::qsatmixadj<double>(th, p->retv, t[i], q[i], pp[i]);
  return 0;
}

}  // extern "C"
