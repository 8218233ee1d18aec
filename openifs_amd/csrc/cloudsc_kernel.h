// CLOUDSC column kernel for gfx950: one column per lane, levels swept top to
// bottom with all level-to-level state in registers.
//
// Restates the default path of ifs-source/arpifs/phys_ec/cloudsc.F90 and
// cloud_supersatcheck.F90 (hard-wired options at cloudsc.F90:727-793).
// Layout: IFS blocked fields (NPROMA, KLEV[+1], n, NGPBLKS), JL fastest, so
// lane JL of a wavefront reads consecutive words at every level.
//
// Register-level reformulation (same arithmetic as the reference):
//  * ZSOLQA is antisymmetric off the diagonal on every path the default
//    configuration takes (each process adds +x to (a,b) and -x to (b,a)), and
//    only 7 off-diagonal pairs and 4 diagonal entries are ever nonzero, so the
//    kernel keeps 11 scalars instead of 25 (cloudsc.F90:1335-3408).
//  * ZSOLQB has 5 nonzero entries (L,L), (I,I), (R,L), (S,I), (S,L), which
//    makes ZQLHS lower triangular: the no-pivot LU of cloudsc.F90:3600-3628
//    reduces to 3 multipliers and 5 divisions with the same operations on the
//    nonzero entries (the skipped ones only add or subtract exact zeros).
//  * Level JK-1 quantities (tidied ZA, ZQX(r,s), ZTP1, PAP) and level JK+1
//    inputs (PLU, PMFU, PMFD) are carried in registers instead of KLEV arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <limits>

#include "../../include/cloudsc_hip.h"
#include "cloudsc_budget.h"
#include "cloudsc_defaults.h"

namespace cloudsc_gpu {

// Constants in the kernel's working precision T.
template <typename T>
struct KConst {
  T rg, rd, rv_rd, retv, rtt, rlvtt, rlstt, rpi, r2es, r3les, r3ies, r4les, r4ies, r5les, r5ies;
  T r5alvcp, r5alscp, ralvdcp, ralsdcp, rtwat, rtice, rtwat_rtice_r, rkoop1, rkoop2, rv, rd_rv_dummy;
  T ramid, ramin, rlmin, rcldiff, rcldiff_convi, rkconv;
  T rsnowlin1, rsnowlin2, rlcritsnow, rvice, rvrain, rvsnow, rxfallspeedrain;
  T rcldtopcf, rdepliqrefrate, rdepliqrefdepth, rcl_overlapliqice, rcl_effrime, riceinit, rthomo;
  T rprecrhmax, rpecons, rvrfactor, rcovpmin, rtaumel, rssicefactor, rdenswat, rdensref;
  T rcl_dr, rcl_x1r, rcl_x2r, rcl_x1s, rcl_lam1r, rcl_lam2r, rcl_lam1s, rcl_lam2s;
  T rcl_cdenom1, rcl_cdenom2, rcl_cdenom3, rcl_const1r, rcl_const2r, rcl_const3r, rcl_const4r;
  T rcl_const5r, rcl_const6r, rcl_const7r, rcl_kkaac, rcl_kkbac, rcl_kkaau, rcl_kkbauq, rcl_kkbaun;
  T rcl_kk_cloud_num_sea, rcl_kk_cloud_num_land, rcl_fzrab, rcl_apb1, rcl_apb2, rcl_apb3;
  T rcl_const1s, rcl_const2s, rcl_const3s, rcl_const4s, rcl_const5s, rcl_const6s, rcl_const7s, rcl_const8s;
  T rmfadvw, rmfadvwdd;
  // derived scalars (cloudsc.F90:798-804), all evaluated as the reference does
  T ptsphy, zqtmst, zrdcp, zrg_r, zrldcp, zepsilon, rdtg_r;
  T c_aut, c_acc;  // RCL_INHOMOGAUT*PTSPHY*RCL_KKAAU, RCL_INHOMOGACC*PTSPHY*RCL_KKAAC
  int32_t ncldtop, lmfdsnow;
};

// The default configuration (cloudsc_defaults.h) as compile-time constants: the kernel body
// reads the same member names as in KConst, but every configuration constant becomes an
// instruction literal instead of a register.  Only the values that depend on the time step or
// on the vertical grid are run-time members.
template <typename T>
struct KConstDefault {
#define CLOUDSC_DEF(name) static constexpr T name = T(cloudsc_defaults::name);
  CLOUDSC_DEF(rg) CLOUDSC_DEF(rd) CLOUDSC_DEF(retv) CLOUDSC_DEF(rtt) CLOUDSC_DEF(rlvtt) CLOUDSC_DEF(rlstt)
  CLOUDSC_DEF(rpi) CLOUDSC_DEF(r2es) CLOUDSC_DEF(r3les) CLOUDSC_DEF(r3ies) CLOUDSC_DEF(r4les) CLOUDSC_DEF(r4ies)
  CLOUDSC_DEF(r5les) CLOUDSC_DEF(r5ies) CLOUDSC_DEF(r5alvcp) CLOUDSC_DEF(r5alscp) CLOUDSC_DEF(ralvdcp)
  CLOUDSC_DEF(ralsdcp) CLOUDSC_DEF(rtwat) CLOUDSC_DEF(rtice) CLOUDSC_DEF(rtwat_rtice_r) CLOUDSC_DEF(rkoop1)
  CLOUDSC_DEF(rkoop2) CLOUDSC_DEF(rv) CLOUDSC_DEF(ramid) CLOUDSC_DEF(ramin) CLOUDSC_DEF(rlmin) CLOUDSC_DEF(rcldiff)
  CLOUDSC_DEF(rcldiff_convi) CLOUDSC_DEF(rkconv) CLOUDSC_DEF(rsnowlin1) CLOUDSC_DEF(rsnowlin2)
  CLOUDSC_DEF(rlcritsnow) CLOUDSC_DEF(rvice) CLOUDSC_DEF(rvrain) CLOUDSC_DEF(rvsnow) CLOUDSC_DEF(rxfallspeedrain)
  CLOUDSC_DEF(rcldtopcf) CLOUDSC_DEF(rdepliqrefrate) CLOUDSC_DEF(rdepliqrefdepth) CLOUDSC_DEF(rcl_overlapliqice)
  CLOUDSC_DEF(rcl_effrime) CLOUDSC_DEF(riceinit) CLOUDSC_DEF(rthomo) CLOUDSC_DEF(rprecrhmax) CLOUDSC_DEF(rpecons)
  CLOUDSC_DEF(rvrfactor) CLOUDSC_DEF(rcovpmin) CLOUDSC_DEF(rtaumel) CLOUDSC_DEF(rdenswat) CLOUDSC_DEF(rdensref)
  CLOUDSC_DEF(rcl_dr) CLOUDSC_DEF(rcl_x1r) CLOUDSC_DEF(rcl_x2r) CLOUDSC_DEF(rcl_x1s) CLOUDSC_DEF(rcl_lam1r)
  CLOUDSC_DEF(rcl_lam2r) CLOUDSC_DEF(rcl_lam1s) CLOUDSC_DEF(rcl_lam2s) CLOUDSC_DEF(rcl_cdenom1)
  CLOUDSC_DEF(rcl_cdenom2) CLOUDSC_DEF(rcl_cdenom3) CLOUDSC_DEF(rcl_const1r) CLOUDSC_DEF(rcl_const2r)
  CLOUDSC_DEF(rcl_const3r) CLOUDSC_DEF(rcl_const4r) CLOUDSC_DEF(rcl_const5r) CLOUDSC_DEF(rcl_const6r)
  CLOUDSC_DEF(rcl_const7r) CLOUDSC_DEF(rcl_kkaac) CLOUDSC_DEF(rcl_kkbac) CLOUDSC_DEF(rcl_kkaau) CLOUDSC_DEF(rcl_kkbauq)
  CLOUDSC_DEF(rcl_kkbaun) CLOUDSC_DEF(rcl_kk_cloud_num_sea) CLOUDSC_DEF(rcl_kk_cloud_num_land)
  CLOUDSC_DEF(rcl_fzrab) CLOUDSC_DEF(rcl_apb1) CLOUDSC_DEF(rcl_apb2) CLOUDSC_DEF(rcl_apb3) CLOUDSC_DEF(rcl_const1s)
  CLOUDSC_DEF(rcl_const2s) CLOUDSC_DEF(rcl_const3s) CLOUDSC_DEF(rcl_const4s) CLOUDSC_DEF(rcl_const5s)
  CLOUDSC_DEF(rcl_const6s) CLOUDSC_DEF(rcl_const7s) CLOUDSC_DEF(rcl_const8s) CLOUDSC_DEF(rmfadvw)
  CLOUDSC_DEF(rmfadvwdd)
#undef CLOUDSC_DEF
  // derived as in KConst (cloudsc.F90:693, 798-804), in T arithmetic
  static constexpr T rv_rd = rv / rd;
  static constexpr T zrdcp = rd / T(cloudsc_defaults::rcpd);
  static constexpr T zrg_r = T(1) / rg;
  static constexpr T zrldcp = T(1) / (ralsdcp - ralvdcp);
  static constexpr T zepsilon = T(100) * std::numeric_limits<T>::epsilon();
  static constexpr int32_t lmfdsnow = cloudsc_defaults::lmfdsnow;
  // run-time values
  T ptsphy, zqtmst, rdtg_r, c_aut, c_acc, rssicefactor;
  int32_t ncldtop;
};

// Device pointers of one launch, blocked layout.
template <typename R>
struct KPtrs {
  const R *pt, *pq, *tcml_t, *tcml_q, *tcml_a, *tcml_cld;
  R *tloc_t, *tloc_q, *tloc_a, *tloc_cld;
  const R *pvfl, *pvfi, *pap, *paph, *plsm;
  const int32_t *ldcum, *ktype;
  const R *peis, *plu;
  R* plude;
  const R *pludeli, *psnde, *pmfu, *pmfd, *pa, *pclv;
  R *pcovptot, *prainfrac_toprfz;
  R *pfsqlf, *pfsqif, *pfcqnng, *pfcqlng, *pfsqrf, *pfsqsf, *pfcqrng, *pfcqsng, *pfsqltur, *pfsqitur;
  R *pfplsl, *pfplsn, *pfhpsl, *pfhpsn;
};

// Extra device pointers of the fused saturation adjustment + CLOUDSC launch (the CALLPAR sequence
// callpar.F90:1253-1293: CLOUD_SATADJ, STATE_INCREMENT of TENDENCY_CML, STATE_COPY to TENDENCY_SATADJ, CLOUDSC).
// tcml_* alias the CML tendencies of KPtrs and are updated in place; tsat_* (may be NULL) receive the adjustment's
// tendencies (TENDENCY_SATADJ; species 1 and 2 of tsat_cld).
template <typename R>
struct KSatPtrs {
  const R *phrsw, *phrlw, *tvdf_t, *tvdf_q, *pvervel;
  R *tcml_t, *tcml_q, *tcml_a, *tcml_cld;
  R *tsat_t, *tsat_q, *tsat_a, *tsat_cld;
};

// Host-side constant blocks for a launch (cloudsc_kernel.hip); is_default_config: every constant and switch
// equals the compiled-in default configuration, so KConstDefault may be used.
template <typename T>
KConst<T> make_kconst(const cloudsc_params_t& q, double ptsphy);
template <typename T>
KConstDefault<T> make_kconst_default(const cloudsc_params_t& q, double ptsphy);
bool is_default_config(const cloudsc_params_t& q);

// Launch configuration of the blocked kernel (tuned per arithmetic variant and KLEV).
struct LaunchCfg {
  int block = 256;  // columns per work-group: 64, 128, 256 or 512
  int wpe = 0;      // minimum waves per SIMD of the register allocation (0: compiler's choice)
  int kspec = 0;    // 1: use the instantiation with KLEV compiled in (137 and 91)
};
bool valid_launch(int variant, const LaunchCfg& l);
LaunchCfg tuned_launch(int variant, int klev);
void set_tuned_launch(int variant, int klev, const LaunchCfg* l);  // NULL removes the setting

struct KDims {
  int32_t ngptot;   // columns in total (blocked) or KFDIA (one block)
  int32_t nproma;   // KLON
  int32_t klev;
  int32_t ngpblks;
  int32_t kidia;    // first active column of every block (1-based)
  int32_t kfdia;    // last active column of a full block (1-based)
  // NULL, or two device counters the launch adds to: points where the reference writes its negative-condensate
  // warnings (cloudsc.F90:3794 liquid, :3802 ice; condensate below -RLMIN before the clean-up)
  uint32_t* nwarn = nullptr;
};

// Blocked launch over all blocks (cloudsc_kernel.hip): splits the grid so that every array of one launch spans
// less than 4 GiB.  ssc_off: KIDIA-1 offset of CLOUD_SUPERSATCHECK's pressure argument (per-block entry).
// sat: NULL for CLOUDSC alone, else the fused saturation adjustment + CLOUDSC.  bud: NULL, or the budget inputs and
// PEXTRA for the budget-instrumented kernel (cloudsc_budget.h; not combined with sat).
template <typename R, typename RD>
hipError_t launch_cloudsc(const cloudsc_params_t& q, double ptsphy, const KPtrs<R>& p, const KDims& d, int ssc_off,
                          hipStream_t stream, const KSatPtrs<R>* sat = nullptr, const KBudPtrs<R>* bud = nullptr);

}  // namespace cloudsc_gpu
