// CPU oracle of CLOUDSC -- TEST INFRASTRUCTURE ONLY.
//
// This file is the checker the HIP kernels are compared against.  Only the
// tests, the smoke check and the CPU-baseline leg of bench.py may load it;
// the product path (libcloudsc_hip.so) never links or calls it.
//
// It restates, block by block and in the reference's own loop structure
// (vector loops over JL inside a sequential level loop), the default path of
//   ifs-source/arpifs/phys_ec/cloudsc.F90:14-4365          SUBROUTINE CLOUDSC
//   ifs-source/arpifs/phys_ec/cloud_supersatcheck.F90:9-231
//   ifs-source/arpifs/function/fcttre.func.h:39-184, fccld.func.h:26-27
// Each block cites the lines it follows.  Hard-wired options are the
// reference's (cloudsc.F90:727-793): IWARMRAIN=3, IRAINACC=1, ISUBLSNOW=1,
// IDEPSNOW=1, IDEPICE=1, ISUBLICE=0, IP_SNOW_ACCRETES_RAIN=1,
// IP_ICE_ACCRETES_RAIN=1, IP_RAIN_ACCRETES_SNOW=0, IP_RAIN_ACCRETES_ICE=0,
// IVARFALL_RAIN=1, IVARFALL_ICE=0, ITURBEROSION=3.  The budget diagnostics
// (LCLDBUD*, LBUD23; cloudsc.F90:1271-1273 and the ZBUD* sites, PEXTRA writes
// :3826-4101) are restated when Args::bud is set.  Paths behind switches that
// are off by default (SPP, aerosol, inhomogeneity) are not restated; the
// driver rejects configurations that would need them.
//
// Template parameters: R = JPRB (storage and bulk arithmetic), RD = type of
// the water budget of a level that feeds the implicit solve -- the first-guess
// amounts ZQXFG, the explicit source/sink matrix ZSOLQA, the sink amounts
// limited by ZQXFG (MIN(rate, ZQXFG)), the sink truncation and the solve's
// right-hand side.  <double,double> is the ifs_dp build and <float,float> the
// ifs_sp build (RD = R: the reference's arithmetic unchanged); <float,double>
// is the mixed-precision variant.  In the mixed variant a species that a
// limited sink removes completely leaves a residue of FP64 size in the
// right-hand side, which the reference's ZEPSEC = 1E-14 clip removes as in the
// FP64 build; with FP32 sums the residue is ~1E-13 and survives (fully
// evaporated rain, for instance, keeps ZCOVPTOT alive below).  The elimination
// and CLOUD_SUPERSATCHECK run in R: in FP64 they change no output by more
// than FP32 rounding.  The stand-alone CLOUD_SUPERSATCHECK entry
// (supersatcheck<R, RD>) keeps RD.
//
// Notational choices (documented, not reference behaviour): a real exponent
// of 2 or 3 (e.g. ZTP1**3._JPRB) is evaluated as a product, a real exponent
// of 0.5 as sqrt.  Default-real literals of the reference (the 0.4 exponent
// at cloudsc.F90:2320,2980 and the 0.8 threshold at :2889) are kept as FP32
// constants widened to R, as the reference's FP64 build sees them.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/cloudsc_hip.h"

namespace oracle {

enum { NCLV = 5, QL = 0, QI = 1, QR = 2, QS = 3, QV = 4 };  // NCLDQ* - 1 (yoecldp.F90:102-107)

template <typename T>
struct Thermo {  // fcttre.func.h / fccld.func.h statement functions, constants in T
  T rtt, r2es, r3les, r3ies, r4les, r4ies, r5alvcp, r5alscp, ralvdcp, ralsdcp;
  T rtwat, rtice, rtwat_rtice_r, rkoop1, rkoop2;
  explicit Thermo(const cloudsc_params_t& p)
      : rtt(T(p.rtt)), r2es(T(p.r2es)), r3les(T(p.r3les)), r3ies(T(p.r3ies)), r4les(T(p.r4les)),
        r4ies(T(p.r4ies)), r5alvcp(T(p.r5alvcp)), r5alscp(T(p.r5alscp)), ralvdcp(T(p.ralvdcp)),
        ralsdcp(T(p.ralsdcp)), rtwat(T(p.rtwat)), rtice(T(p.rtice)), rtwat_rtice_r(T(p.rtwat_rtice_r)),
        rkoop1(T(p.rkoop1)), rkoop2(T(p.rkoop2)) {}
  // fcttre.func.h:41  MAX(0,SIGN(1,T-RTT))
  T foedelta(T t) const { return std::max(T(0), std::copysign(T(1), t - rtt)); }
  // fcttre.func.h:46-48
  T foeew(T t) const {
    const T d = foedelta(t);
    return r2es * std::exp((r3les * d + r3ies * (T(1) - d)) * (t - rtt) / (t - (r4les * d + r4ies * (T(1) - d))));
  }
  // fcttre.func.h:84-85
  T foealfa(T t) const {
    const T x = (std::max(rtice, std::min(rtwat, t)) - rtice) * rtwat_rtice_r;
    return std::min(T(1), x * x);
  }
  // fcttre.func.h:91-93
  T foeewm(T t) const {
    const T a = foealfa(t);
    return r2es * (a * std::exp(r3les * (t - rtt) / (t - r4les)) + (T(1) - a) * std::exp(r3ies * (t - rtt) / (t - r4ies)));
  }
  // fcttre.func.h:101-102
  T foedem(T t) const {
    const T a = foealfa(t);
    const T dl = t - r4les, di = t - r4ies;
    return a * r5alvcp * (T(1) / (dl * dl)) + (T(1) - a) * r5alscp * (T(1) / (di * di));
  }
  // fcttre.func.h:104-105
  T foeldcpm(T t) const {
    const T a = foealfa(t);
    return a * ralvdcp + (T(1) - a) * ralsdcp;
  }
  // fcttre.func.h:165-167
  T foeeliq(T t) const { return r2es * std::exp(r3les * (t - rtt) / (t - r4les)); }
  T foeeice(T t) const { return r2es * std::exp(r3ies * (t - rtt) / (t - r4ies)); }
  T foeeliq2ice(T t) const { return std::exp((t - rtt) * (r3les / (t - r4les) - r3ies / (t - r4ies))); }
  // fccld.func.h:27
  T fokoop(T t) const { return std::min(rkoop1 - rkoop2 * t, foeeliq2ice(t)); }
};

// ---------------------------------------------------------------------------
// CLOUD_SUPERSATCHECK (cloud_supersatcheck.F90:111-226), computed in RD.
// KFTLIQICE = 1 splits the condensate by FOEALFA, 2 makes it all liquid where
// PT+PT_ADJ >= RTHOMO and all ice below (:211-223).  PP and PSURF are indexed
// from 1 as the callee sees the sections PAP(KIDIA:KFDIA,JK),
// PAPH(KIDIA:KFDIA,KLEV+1) passed by CLOUDSC (cloudsc.F90:3774-3777): pp[jl-1]
// is the caller's column KIDIA+jl-1.
// ---------------------------------------------------------------------------
template <typename R, typename RD>
void supersatcheck(const cloudsc_params_t& p, int kidia, int kfdia, const RD* pt, const RD* pq, const RD* pa,
                   const R* pp, const R* psurf, RD* pt_adj, RD* pq_adj, RD* pa_adj, RD* pl_adj, RD* pi_adj,
                   int kftliqice = 1) {
  const Thermo<RD> th(p);
  const RD ramid = RD(p.ramid), retv = RD(p.retv), rssicefactor = RD(p.rssicefactor);
  const RD zepsec = RD(1.e-14);
  for (int jl = kidia; jl <= kfdia; ++jl) {
    const int i = jl - 1;
    RD za_upd = RD(0), zda = RD(0);
    RD zt = pt[i], zq = pq[i], za = pa[i];
    const RD ppv = RD(pp[i]);
    const RD zp_r = RD(1) / ppv;
    zt = std::max(zt, RD(160));
    za = std::max(std::min(za, RD(1)), RD(0));
    RD zrhc = ramid;  // computed by the reference, unused afterwards (:128-133)
    const RD zsigk = ppv / RD(psurf[i]);
    if (zsigk > RD(0.8)) {
      const RD s = (zsigk - RD(0.8)) / RD(0.2);
      zrhc = ramid + (RD(1) - ramid) * s * s;
    }
    (void)zrhc;
    const RD zfoeeliq = std::min(th.foeeliq(zt) * zp_r, RD(0.5));
    const RD zqsliq = zfoeeliq / (RD(1) - retv * zfoeeliq);
    (void)zqsliq;
    const RD zfoeeice = std::min(th.foeew(zt) * zp_r, RD(0.5));
    const RD zfokoop = std::min(zfoeeice * (th.rkoop1 - th.rkoop2 * zt), zfoeeliq);
    RD zqsliqk = zfokoop / (RD(1) - retv * zfokoop);
    const RD zfoealfa = th.foealfa(zt);
    const RD zfoeemix = std::min(th.foeewm(zt) * zp_r, RD(0.5));
    const RD zcor = RD(1) / (RD(1) - retv * zfoeemix);
    const RD zqsmix = zfoeemix * zcor;
    const RD zcorqsmix = RD(1) + zqsmix * zcor * th.foedem(zt);
    zqsliqk = RD(0.5) * (zqsliqk + zqsmix);
    RD zqslim = za * zqsmix + (RD(1) - za) * zqsliqk;
    RD zcond = std::max(zq - zqslim, RD(0));
    if (zcond > zepsec) {
      const RD zfaci = zfoealfa + (RD(1) - zfoealfa) * rssicefactor;
      zda = (RD(1) - za) * zfaci;
      za_upd = za + zda;
      zqslim = za_upd * zqsmix + (RD(1) - za_upd) * zqsliqk;
      zcond = (zq - zqslim) / zcorqsmix;
    }
    pt_adj[i] = th.foeldcpm(zt) * zcond;
    pq_adj[i] = -zcond;
    pa_adj[i] = zda;
    if (kftliqice == 2) {
      const bool liquid = pt[i] + pt_adj[i] >= RD(p.rthomo);
      pl_adj[i] = liquid ? zcond : RD(0);
      pi_adj[i] = liquid ? RD(0) : zcond;
    } else {
      pl_adj[i] = zcond * zfoealfa;
      pi_adj[i] = zcond * (RD(1) - zfoealfa);
    }
  }
}

// ---------------------------------------------------------------------------
// 5.2.1: scale explicit sinks so that no species is driven negative, species
// taken in the order in which they run out (cloudsc.F90:3455-3536).
// qa[m][n] = ZSOLQA(JL,m,n); qx = ZQX(JL,JK,:).  The insertion sort is stable
// (ties keep species order).  When a diagonal entry is negative it is scaled
// twice, as the two assignments at :3531-3532 both hit ZSOLQA(JL,JO,JO).
// ---------------------------------------------------------------------------
template <typename R>
void truncate_sinks(R qa[NCLV][NCLV], const R qx[NCLV], R zepsec) {
  R zratio[NCLV];
  for (int jm = 0; jm < NCLV; ++jm) {
    R zsinksum = R(0);
    for (int jn = 0; jn < NCLV; ++jn) zsinksum = zsinksum - qa[jm][jn];
    const R zmax = std::max(qx[jm], zepsec);
    const R zrat = std::max(zsinksum, zmax);
    zratio[jm] = zmax / zrat;
  }
  int iordv[NCLV];
  iordv[0] = 0;
  for (int jm = 1; jm < NCLV; ++jm) {
    int jn;
    for (jn = jm - 1; jn >= 0; --jn) {
      if (zratio[iordv[jn]] <= zratio[jm]) break;
      iordv[jn + 1] = iordv[jn];
    }
    iordv[jn + 1] = jm;
  }
  for (int jm = 0; jm < NCLV; ++jm) {
    const int jo = iordv[jm];
    bool llindex3[NCLV];
    R zsinksum = R(0);
    for (int jn = 0; jn < NCLV; ++jn) {
      llindex3[jn] = qa[jo][jn] < R(0);
      zsinksum = zsinksum - qa[jo][jn];
    }
    const R zmm = std::max(qx[jo], zepsec);
    const R zrr = std::max(zsinksum, zmm);
    const R zzratio = zmm / zrr;
    for (int jn = 0; jn < NCLV; ++jn) {
      if (llindex3[jn]) {
        qa[jo][jn] = qa[jo][jn] * zzratio;
        qa[jn][jo] = qa[jn][jo] * zzratio;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 5.2.2: LHS (:3546-3571), RHS (:3576-3587), no-pivot LU factorisation
// (:3600-3610) and forward/back substitution (:3614-3628), evaluated in RD.
// lhs[a][b] = ZQLHS(JL,a,b).
// ---------------------------------------------------------------------------
template <typename RD>
void lu5_solve(RD lhs[NCLV][NCLV], RD x[NCLV]) {
  for (int jn = 0; jn < NCLV - 1; ++jn)
    for (int jm = jn + 1; jm < NCLV; ++jm) {
      lhs[jm][jn] = lhs[jm][jn] / lhs[jn][jn];
      for (int ik = jn + 1; ik < NCLV; ++ik) lhs[jm][ik] = lhs[jm][ik] - lhs[jm][jn] * lhs[jn][ik];
    }
  for (int jn = 1; jn < NCLV; ++jn)
    for (int jm = 0; jm < jn; ++jm) x[jn] = x[jn] - lhs[jn][jm] * x[jm];
  x[NCLV - 1] = x[NCLV - 1] / lhs[NCLV - 1][NCLV - 1];
  for (int jn = NCLV - 2; jn >= 0; --jn) {
    for (int jm = jn + 1; jm < NCLV; ++jm) x[jn] = x[jn] - lhs[jn][jm] * x[jm];
    x[jn] = x[jn] / lhs[jn][jn];
  }
}

template <typename R, typename RD>
void implicit_solve(const RD qa[NCLV][NCLV], const R qb[NCLV][NCLV], const R fallsink[NCLV], const R qx[NCLV],
                    R qxn[NCLV]) {
  // the right-hand side (the level's water budget) is summed in RD, the elimination runs in R
  R lhs[NCLV][NCLV], x[NCLV];
  for (int jm = 0; jm < NCLV; ++jm)
    for (int jn = 0; jn < NCLV; ++jn) {
      if (jn == jm) {
        R d = R(1) + fallsink[jm];
        for (int jo = 0; jo < NCLV; ++jo) d = d + qb[jo][jn];
        lhs[jn][jm] = d;
      } else {
        lhs[jn][jm] = -qb[jn][jm];
      }
    }
  for (int jm = 0; jm < NCLV; ++jm) {
    RD zexplicit = RD(0);
    for (int jn = 0; jn < NCLV; ++jn) zexplicit = zexplicit + qa[jm][jn];
    x[jm] = R(RD(qx[jm]) + zexplicit);
  }
  lu5_solve<R>(lhs, x);
  for (int m = 0; m < NCLV; ++m) qxn[m] = x[m];
}

// Inputs and outputs of the diagnostics for one call, Fortran layout: the budgets (LCLDBUD*,
// cloudsc.F90:3826-4101) read (KLON,KLEV) arrays and KPBLTYPE/KCTOP (KLON) and accumulate into PEXTRA
// (KLON,KLEV,KFLDX); LRAD_CLOUD_INHOMOG reads PGAW (KLON) and writes PFSD (KLON,KLEV).
template <typename R>
struct BudArgs {
  const R *pvfa, *pdyna, *pdynl, *pdyni, *pdyns, *phrsw, *phrlw;
  const int32_t *kpbltype, *kctop;
  R* pextra;
  int kfldx;
  const R* pgaw;
  R* pfsd;
};

// Arguments of one CLOUDSC call, Fortran layout (JL fastest).
template <typename R>
struct Args {
  int kidia, kfdia, klon, klev;
  R ptsphy;
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
  const BudArgs<R>* bud = nullptr;  // diagnostics: switched by params.lcldbud and lrad_cloud_inhomog, NULL = none
  // NULL, or counts of the reference's negative-condensate warnings (liquid :3794, ice :3802) to add to
  int64_t* nwarn = nullptr;
};

// ---------------------------------------------------------------------------
// SUBROUTINE CLOUDSC, default configuration.
// ---------------------------------------------------------------------------
template <typename R, typename RD>
void cloudsc(const cloudsc_params_t& p, const Args<R>& a) {
  const int kidia = a.kidia, kfdia = a.kfdia, klon = a.klon, klev = a.klev;
  const int ncldtop = p.ncldtop;
  const R ptsphy = a.ptsphy;
  const Thermo<R> th(p);
  // 2D (KLON,KLEV) / (KLON,KLEV+1) and 3D (KLON,KLEV,N) Fortran indexing, 1-based
  auto i2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };
  auto i3 = [klon, klev](int jl, int jk, int jm) { return ((size_t)(jm - 1) * klev + (jk - 1)) * klon + (jl - 1); };
  auto ih3 = [klon, klev](int jl, int jk, int jm) {  // (KLON,KLEV+1,N)
    return ((size_t)(jm - 1) * (klev + 1) + (jk - 1)) * klon + (jl - 1);
  };

  // Constants (cloudsc.F90:693-804)
  const R RG = R(p.rg), RD_ = R(p.rd), RV = R(p.rv), RETV = R(p.retv), RTT = R(p.rtt), RLSTT = R(p.rlstt);
  const R RLVTT = R(p.rlvtt), RPI = R(p.rpi);
  const R RALVDCP = R(p.ralvdcp), RALSDCP = R(p.ralsdcp), R5LES = R(p.r5les), R5IES = R(p.r5ies);
  const R RALFDCP = R(p.ralfdcp);  // budget diagnostics only
  const R RA = R(p.ra), RCLOUD_FRAC_STD = R(p.rcloud_frac_std);  // LRAD_CLOUD_INHOMOG only
  const R ZR12 = R(1.3), ZU1 = R(0.446585), ZU2 = R(0.308061), ZU3 = R(0.395736), ZU4 = R(-0.744527);  // :587-592
  const R R4LES = R(p.r4les), R4IES = R(p.r4ies);
  const R RAMID = R(p.ramid), RAMIN = R(p.ramin), RLMIN = R(p.rlmin), RCLDIFF = R(p.rcldiff);
  const R RCLDIFF_CONVI = R(p.rcldiff_convi);  // ZLCRIT (RCLCRIT_*) is not used by IWARMRAIN=3
  const R RCL_INHOMOGAUT = R(p.rcl_inhomogaut), RCL_INHOMOGACC = R(p.rcl_inhomogacc);
  const R RSNOWLIN1 = R(p.rsnowlin1), RSNOWLIN2 = R(p.rsnowlin2), RLCRITSNOW = R(p.rlcritsnow);
  const R RVICE = R(p.rvice), RVRAIN = R(p.rvrain), RVSNOW = R(p.rvsnow), RXFALLSPEEDRAIN = R(p.rxfallspeedrain);
  const R RCLDTOPCF = R(p.rcldtopcf), RDEPLIQREFRATE = R(p.rdepliqrefrate), RDEPLIQREFDEPTH = R(p.rdepliqrefdepth);
  const R RCL_OVERLAPLIQICE = R(p.rcl_overlapliqice), RCL_EFFRIME = R(p.rcl_effrime), RICEINIT = R(p.riceinit);
  const R RTHOMO = R(p.rthomo), RPRECRHMAX = R(p.rprecrhmax), RPECONS = R(p.rpecons), RVRFACTOR = R(p.rvrfactor);
  const R RCOVPMIN = R(p.rcovpmin), RTAUMEL = R(p.rtaumel), RDENSWAT = R(p.rdenswat), RDENSREF = R(p.rdensref);
  const R RCL_DR = R(p.rcl_dr), RCL_X1R = R(p.rcl_x1r), RCL_X2R = R(p.rcl_x2r), RCL_X1S = R(p.rcl_x1s);
  const R RCL_LAM1R = R(p.rcl_lam1r), RCL_LAM2R = R(p.rcl_lam2r), RCL_LAM1S = R(p.rcl_lam1s), RCL_LAM2S = R(p.rcl_lam2s);
  const R RCL_CDENOM1 = R(p.rcl_cdenom1), RCL_CDENOM2 = R(p.rcl_cdenom2), RCL_CDENOM3 = R(p.rcl_cdenom3);
  const R RCL_KA273 = R(p.rcl_ka273), RCL_CONST1R = R(p.rcl_const1r), RCL_CONST2R = R(p.rcl_const2r);
  const R RCL_CONST3R = R(p.rcl_const3r), RCL_CONST4R = R(p.rcl_const4r), RCL_CONST5R = R(p.rcl_const5r);
  const R RCL_CONST6R = R(p.rcl_const6r), RCL_CONST7R = R(p.rcl_const7r);
  const R RCL_KKAAC = R(p.rcl_kkaac), RCL_KKBAC = R(p.rcl_kkbac), RCL_KKAAU = R(p.rcl_kkaau);
  const R RCL_KKBAUQ = R(p.rcl_kkbauq), RCL_KKBAUN = R(p.rcl_kkbaun);
  const R RCL_KK_CLOUD_NUM_SEA = R(p.rcl_kk_cloud_num_sea), RCL_KK_CLOUD_NUM_LAND = R(p.rcl_kk_cloud_num_land);
  const R RCL_FZRAB = R(p.rcl_fzrab), RCL_APB1 = R(p.rcl_apb1), RCL_APB2 = R(p.rcl_apb2), RCL_APB3 = R(p.rcl_apb3);
  const R RCL_CONST1S = R(p.rcl_const1s), RCL_CONST2S = R(p.rcl_const2s), RCL_CONST3S = R(p.rcl_const3s);
  const R RCL_CONST4S = R(p.rcl_const4s), RCL_CONST5S = R(p.rcl_const5s), RCL_CONST6S = R(p.rcl_const6s);
  const R RCL_CONST7S = R(p.rcl_const7s), RCL_CONST8S = R(p.rcl_const8s);
  const R RMFADVW = R(p.rmfadvw), RMFADVWDD = R(p.rmfadvwdd);
  const bool LMFDSNOW = p.lmfdsnow != 0;

  const R ZEPSILON = R(100) * std::numeric_limits<R>::epsilon();  // :693
  const R ZQTMST = R(1) / ptsphy;                                   // :798
  const R ZRDCP = RD_ / R(p.rcpd);
  const R ZEPSEC = R(1.e-14);
  const R ZRG_R = R(1) / RG;
  const R ZRLDCP = R(1) / (RALSDCP - RALVDCP);
  const R ZTW1 = R(1329.31), ZTW2 = R(0.0074615), ZTW3 = R(0.85E5), ZTW4 = R(40.637), ZTW5 = R(275.0);
  const R C04 = R(0.4f), C08 = R(0.8f);  // default-real literals, cloudsc.F90:2320,2980 and :2889

  // Species tables (:821-848)
  const int iphase[NCLV] = {1, 2, 1, 2, 0};
  const int imelt[NCLV] = {QI, QR, QS, QR, -99};
  R zvqx[NCLV] = {R(0), RVICE, RVRAIN, RVSNOW, R(0)};
  bool llfall[NCLV];
  for (int m = 0; m < NCLV; ++m) llfall[m] = zvqx[m] > R(0);

  const int KL = klon;
  std::vector<R> zadvw(KL), zadvwd(KL);
  for (int jl = kidia; jl <= kfdia; ++jl) {  // :696-703
    zadvw[jl - 1] = R(1);
    zadvwd[jl - 1] = R(1);
    if (a.ktype[jl - 1] == 1 && RMFADVW > R(0)) {
      zadvw[jl - 1] = R(1) - RMFADVW;
      zadvwd[jl - 1] = R(1) - RMFADVW * RMFADVWDD;
    }
  }

  // ---- 1. initial values (:963-1011) ----
  const size_t n2 = (size_t)klon * klev;
  std::vector<R> ztp1(n2), za(n2), zaorig(n2), zli(n2), zliqfrac(n2), zicefrac(n2);
  std::vector<R> zqx(n2 * NCLV), zqx0(n2 * NCLV), zlneg(n2 * NCLV, R(0)), zqxn2d(n2 * NCLV, R(0));
  std::vector<R> zpfplsx((size_t)klon * (klev + 1) * NCLV, R(0));
  std::vector<R> zfoealfa(n2), zfoeewm(n2), zfoeew(n2), zfoeeliqt(n2), zqsmix(n2), zqsice(n2), zqsliq(n2);

  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      a.tloc_t[i2(jl, jk)] = R(0);
      a.tloc_q[i2(jl, jk)] = R(0);
      a.tloc_a[i2(jl, jk)] = R(0);
    }
  for (int jm = 1; jm <= NCLV - 1; ++jm)
    for (int jk = 1; jk <= klev; ++jk)
      for (int jl = kidia; jl <= kfdia; ++jl) a.tloc_cld[i3(jl, jk, jm)] = R(0);

  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      ztp1[i2(jl, jk)] = a.pt[i2(jl, jk)] + ptsphy * a.tcml_t[i2(jl, jk)];
      zqx[i3(jl, jk, QV + 1)] = a.pq[i2(jl, jk)] + ptsphy * a.tcml_q[i2(jl, jk)];
      zqx0[i3(jl, jk, QV + 1)] = a.pq[i2(jl, jk)] + ptsphy * a.tcml_q[i2(jl, jk)];
      za[i2(jl, jk)] = a.pa[i2(jl, jk)] + ptsphy * a.tcml_a[i2(jl, jk)];
      zaorig[i2(jl, jk)] = a.pa[i2(jl, jk)] + ptsphy * a.tcml_a[i2(jl, jk)];
    }
  for (int jm = 1; jm <= NCLV - 1; ++jm)
    for (int jk = 1; jk <= klev; ++jk)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        zqx[i3(jl, jk, jm)] = a.pclv[i3(jl, jk, jm)] + ptsphy * a.tcml_cld[i3(jl, jk, jm)];
        zqx0[i3(jl, jk, jm)] = a.pclv[i3(jl, jk, jm)] + ptsphy * a.tcml_cld[i3(jl, jk, jm)];
      }
  for (int jl = 1; jl <= klon; ++jl) a.prainfrac_toprfz[jl - 1] = R(0);  // all KLON (:1010)
  std::vector<char> llrainliq(KL, 1);                                        // :1011
  // ZTRPAUS (:1016-1027) is computed by the reference and never used.

  // Saturation values (:1087-1115)
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      const R t = ztp1[k], pap = a.pap[k];
      zfoealfa[k] = th.foealfa(t);
      zfoeewm[k] = std::min(th.foeewm(t) / pap, R(0.5));
      zqsmix[k] = zfoeewm[k];
      zqsmix[k] = zqsmix[k] / (R(1) - RETV * zqsmix[k]);
      const R zalfa = th.foedelta(t);
      zfoeew[k] = std::min((zalfa * th.foeeliq(t) + (R(1) - zalfa) * th.foeeice(t)) / pap, R(0.5));
      zfoeew[k] = std::min(R(0.5), zfoeew[k]);
      zqsice[k] = zfoeew[k] / (R(1) - RETV * zfoeew[k]);
      zfoeeliqt[k] = std::min(th.foeeliq(t) / pap, R(0.5));
      zqsliq[k] = zfoeeliqt[k];
      zqsliq[k] = zqsliq[k] / (R(1) - RETV * zqsliq[k]);
    }

  // ---- 2. tidy up input values (:1127-1199) ----
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      if (zqx[i3(jl, jk, QL + 1)] + zqx[i3(jl, jk, QI + 1)] < RLMIN || za[k] < RAMIN) {
        zlneg[i3(jl, jk, QL + 1)] = zlneg[i3(jl, jk, QL + 1)] + zqx[i3(jl, jk, QL + 1)];
        R zqadj = zqx[i3(jl, jk, QL + 1)] * ZQTMST;
        a.tloc_q[k] = a.tloc_q[k] + zqadj;
        a.tloc_t[k] = a.tloc_t[k] - RALVDCP * zqadj;
        zqx[i3(jl, jk, QV + 1)] = zqx[i3(jl, jk, QV + 1)] + zqx[i3(jl, jk, QL + 1)];
        zqx[i3(jl, jk, QL + 1)] = R(0);
        zlneg[i3(jl, jk, QI + 1)] = zlneg[i3(jl, jk, QI + 1)] + zqx[i3(jl, jk, QI + 1)];
        zqadj = zqx[i3(jl, jk, QI + 1)] * ZQTMST;
        a.tloc_q[k] = a.tloc_q[k] + zqadj;
        a.tloc_t[k] = a.tloc_t[k] - RALSDCP * zqadj;
        zqx[i3(jl, jk, QV + 1)] = zqx[i3(jl, jk, QV + 1)] + zqx[i3(jl, jk, QI + 1)];
        zqx[i3(jl, jk, QI + 1)] = R(0);
        za[k] = R(0);
      }
    }
  for (int jm = 1; jm <= NCLV - 1; ++jm)
    for (int jk = 1; jk <= klev; ++jk)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        if (zqx[i3(jl, jk, jm)] < RLMIN) {
          const size_t k = i2(jl, jk);
          zlneg[i3(jl, jk, jm)] = zlneg[i3(jl, jk, jm)] + zqx[i3(jl, jk, jm)];
          const R zqadj = zqx[i3(jl, jk, jm)] * ZQTMST;
          a.tloc_q[k] = a.tloc_q[k] + zqadj;
          if (iphase[jm - 1] == 1) a.tloc_t[k] = a.tloc_t[k] - RALVDCP * zqadj;
          if (iphase[jm - 1] == 2) a.tloc_t[k] = a.tloc_t[k] - RALSDCP * zqadj;
          zqx[i3(jl, jk, QV + 1)] = zqx[i3(jl, jk, QV + 1)] + zqx[i3(jl, jk, jm)];
          zqx[i3(jl, jk, jm)] = R(0);
        }
      }
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      za[k] = std::max(R(0), std::min(R(1), za[k]));
      zli[k] = zqx[i3(jl, jk, QL + 1)] + zqx[i3(jl, jk, QI + 1)];
      if (zli[k] > RLMIN) {
        zliqfrac[k] = zqx[i3(jl, jk, QL + 1)] / zli[k];
        zicefrac[k] = R(1) - zliqfrac[k];
      } else {
        zliqfrac[k] = R(0);
        zicefrac[k] = R(0);
      }
    }

  // Single-level variables (:1205-1211)
  std::vector<R> zqxnm1(KL * NCLV, R(0)), zanewm1(KL, R(0)), zda(KL, R(0)), zcovpclr(KL, R(0));
  std::vector<R> zcovpmax(KL, R(0)), zcovptot(KL, R(0)), zcldtopdist(KL, R(0));
  // Per-level work arrays
  std::vector<RD> zqxfg(KL * NCLV), zsolqa(KL * NCLV * NCLV);  // RD: water budget of the level
  std::vector<R> zsolqb(KL * NCLV * NCLV);
  std::vector<R> zfallspeed(KL * NCLV), zfallsrce(KL * NCLV), zfallsink(KL * NCLV);
  std::vector<R> zconvsrce(KL * NCLV), zconvsink(KL * NCLV), zlcust(KL * NCLV), zqxn(KL * NCLV);
  std::vector<R> zfluxq(KL * NCLV);
  std::vector<R> zlicld(KL), zrainaut(KL), zrainacc(KL), zsnowaut(KL), zsnowrime(KL), zlfinalsum(KL);
  std::vector<R> zacust(KL), zqpretot(KL), zsolab(KL), zsolac(KL), zicetot(KL), zdp(KL), zgdp(KL), zrho(KL);
  std::vector<R> zdtgdp(KL), zrdtgdp(KL), zdqsliqdt(KL), zdqsicedt(KL), zdqsmixdt(KL), zcorqsliq(KL);
  std::vector<R> zcorqsice(KL), zcorqsmix(KL), zevaplimliq(KL), zevaplimice(KL), zevaplimmix(KL);
  std::vector<R> zliqcld(KL), zicecld(KL), zmf(KL), zdqs(KL), zldifdt(KL), zraincld(KL), zraincldm1(KL);
  std::vector<R> zsnowcld(KL), zsnowcldm1(KL), zmeltmax(KL), zfrzmax(KL), zpsdep(KL), zlamr(KL), zn0r(KL);
  std::vector<R> zlams(KL), zn0s(KL), zlami(KL), zn0i(KL), zpsacr(KL), zpiacr(KL);
  std::vector<R> zq_upd(KL), zt_upd(KL), za_upd(KL), zt_adj(KL), zq_adj(KL), za_adj(KL), zl_adj(KL), zi_adj(KL);
  auto QA = [&](int jl, int m, int n) -> RD& { return zsolqa[((size_t)n * NCLV + m) * KL + (jl - 1)]; };
  auto QB = [&](int jl, int m, int n) -> R& { return zsolqb[((size_t)n * NCLV + m) * KL + (jl - 1)]; };
  auto V = [&](auto& v, int jl, int m) -> auto& { return v[(size_t)m * KL + (jl - 1)]; };
  // Budget arrays ZBUDCC(KLON,12), ZBUDL(KLON,22), ZBUDI(KLON,18) (:557-559), 1-based second index.
  const int lcldbud = a.bud ? p.lcldbud : 0;
  std::vector<R> zbudcc(KL * 12), zbudl(KL * 22), zbudi(KL * 18);
  auto BCC = [&](int jl, int n) -> R& { return zbudcc[(size_t)(n - 1) * KL + (jl - 1)]; };
  auto BL = [&](int jl, int n) -> R& { return zbudl[(size_t)(n - 1) * KL + (jl - 1)]; };
  auto BI = [&](int jl, int n) -> R& { return zbudi[(size_t)(n - 1) * KL + (jl - 1)]; };
  // LRAD_CLOUD_INHOMOG arrays, initialised for all levels (:944-951); LCLOUD_INHOMOG uses ZLFSD too
  const bool lfsd = a.bud && p.lrad_cloud_inhomog;
  const bool linh = a.bud && p.lcloud_inhomog;
  std::vector<R> zifsd, zlfsd, zliqf, zicef, zratfsd, zanewp;
  if (lfsd || linh) {
    zifsd.assign(n2, RCLOUD_FRAC_STD);
    zlfsd.assign(n2, RCLOUD_FRAC_STD);
    zliqf.assign(n2, R(0));
    zicef.assign(n2, R(0));
    zratfsd.assign(n2, R(0));
    zanewp.assign(n2, R(0));
  }

  // ======================= level loop (:1219-4128) =======================
  for (int jk = ncldtop; jk <= klev; ++jk) {
    std::fill(zbudcc.begin(), zbudcc.end(), R(0));  // :1271-1273
    std::fill(zbudl.begin(), zbudl.end(), R(0));
    std::fill(zbudi.begin(), zbudi.end(), R(0));
    for (int m = 0; m < NCLV; ++m)
      for (int jl = kidia; jl <= kfdia; ++jl) V(zqxfg, jl, m) = zqx[i3(jl, jk, m + 1)];  // :1228-1232
    for (int jl = kidia; jl <= kfdia; ++jl) {  // :1238-1268
      zlicld[jl - 1] = zrainaut[jl - 1] = zrainacc[jl - 1] = zsnowaut[jl - 1] = zsnowrime[jl - 1] = R(0);
      zacust[jl - 1] = zqpretot[jl - 1] = zlfinalsum[jl - 1] = R(0);
      zsolab[jl - 1] = zsolac[jl - 1] = zicetot[jl - 1] = R(0);
      for (int m = 0; m < NCLV; ++m) {
        V(zfallspeed, jl, m) = V(zfallsrce, jl, m) = V(zfallsink, jl, m) = R(0);
        V(zconvsrce, jl, m) = V(zconvsink, jl, m) = R(0);
        for (int n = 0; n < NCLV; ++n) QA(jl, m, n) = QB(jl, m, n) = R(0);
      }
    }

    // derived variables, dqs/dT factors, evaporation limits, in-cloud condensate (:1275-1326)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdp[i] = a.paph[i2(jl, jk + 1)] - a.paph[i2(jl, jk)];
      zgdp[i] = RG / zdp[i];
      zrho[i] = a.pap[k] / (RD_ * ztp1[k]);
      zdtgdp[i] = ptsphy * zgdp[i];
      zrdtgdp[i] = zdp[i] * (R(1) / (ptsphy * RG));
      // ZDTGDPF (:1287) is computed by the reference and never used.
      const R tl = ztp1[k] - R4LES, ti = ztp1[k] - R4IES;
      const R zfacw = R5LES / (tl * tl);
      R zcor = R(1) / (R(1) - RETV * zfoeeliqt[k]);
      zdqsliqdt[i] = zfacw * zcor * zqsliq[k];
      zcorqsliq[i] = R(1) + RALVDCP * zdqsliqdt[i];
      const R zfaci = R5IES / (ti * ti);
      zcor = R(1) / (R(1) - RETV * zfoeew[k]);
      zdqsicedt[i] = zfaci * zcor * zqsice[k];
      zcorqsice[i] = R(1) + RALSDCP * zdqsicedt[i];
      const R zalfaw = zfoealfa[k];
      const R zfac = zalfaw * zfacw + (R(1) - zalfaw) * zfaci;
      zcor = R(1) / (R(1) - RETV * zfoeewm[k]);
      zdqsmixdt[i] = zfac * zcor * zqsmix[k];
      zcorqsmix[i] = R(1) + th.foeldcpm(ztp1[k]) * zdqsmixdt[i];
      const R qv = zqx[i3(jl, jk, QV + 1)];
      zevaplimmix[i] = std::max((zqsmix[k] - qv) / zcorqsmix[i], R(0));
      zevaplimliq[i] = std::max((zqsliq[k] - qv) / zcorqsliq[i], R(0));
      zevaplimice[i] = std::max((zqsice[k] - qv) / zcorqsice[i], R(0));
      const R ztmpa = R(1) / std::max(za[k], R(0.01));
      zliqcld[i] = zqx[i3(jl, jk, QL + 1)] * ztmpa;
      zicecld[i] = zqx[i3(jl, jk, QI + 1)] * ztmpa;
      zlicld[i] = zliqcld[i] + zicecld[i];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // evaporate very small amounts (:1332-1344)
      if (zqx[i3(jl, jk, QL + 1)] < RLMIN) {
        QA(jl, QV, QL) = zqx[i3(jl, jk, QL + 1)];
        QA(jl, QL, QV) = -zqx[i3(jl, jk, QL + 1)];
      }
      if (zqx[i3(jl, jk, QI + 1)] < RLMIN) {
        QA(jl, QV, QI) = zqx[i3(jl, jk, QI + 1)];
        QA(jl, QI, QV) = -zqx[i3(jl, jk, QI + 1)];
      }
    }

    // 3.2 detrainment (:1362-1398); PLUDE rescaled in place
    if (jk < klev && jk >= ncldtop) {
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        const size_t k = i2(jl, jk);
        a.plude[k] = a.plude[k] * zdtgdp[i];
        if (a.ldcum[i] && a.plude[k] > ZEPSEC && a.plu[i2(jl, jk + 1)] > ZEPSEC) {
          zsolac[i] = zsolac[i] + a.plude[k] / a.plu[i2(jl, jk + 1)];
          V(zconvsrce, jl, QL) = a.pludeli[i3(jl, jk, 1)] * zdtgdp[i];
          V(zconvsrce, jl, QI) = a.pludeli[i3(jl, jk, 2)] * zdtgdp[i];
          QA(jl, QL, QL) = QA(jl, QL, QL) + V(zconvsrce, jl, QL);
          QA(jl, QI, QI) = QA(jl, QI, QI) + V(zconvsrce, jl, QI);
          BL(jl, 3) = V(zconvsrce, jl, QL) * ZQTMST;  // :1378-1380
          BI(jl, 3) = V(zconvsrce, jl, QI) * ZQTMST;
          BCC(jl, 3) = ZQTMST * a.plude[k] / a.plu[i2(jl, jk + 1)];
        } else {
          a.plude[k] = R(0);
        }
        if (LMFDSNOW) {
          V(zconvsrce, jl, QR) = a.psnde[i3(jl, jk, 1)] * zdtgdp[i];
          V(zconvsrce, jl, QS) = a.psnde[i3(jl, jk, 2)] * zdtgdp[i];
          QA(jl, QR, QR) = QA(jl, QR, QR) + V(zconvsrce, jl, QR);
          QA(jl, QS, QS) = QA(jl, QS, QS) + V(zconvsrce, jl, QS);
        }
      }
    }

    // 3.3 subsidence source from the level above and its evaporation (:1419-1492)
    if (jk > ncldtop) {
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        zmf[i] = std::max(R(0), (a.pmfu[i2(jl, jk)] * zadvw[i] + a.pmfd[i2(jl, jk)] * zadvwd[i]) * zdtgdp[i]);
        zacust[i] = zmf[i] * zanewm1[i];
      }
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        V(zlcust, jl, QL) = zmf[i] * V(zqxnm1, jl, QL);
        V(zlcust, jl, QI) = zmf[i] * V(zqxnm1, jl, QI);
        V(zconvsrce, jl, QL) = V(zconvsrce, jl, QL) + V(zlcust, jl, QL);
        V(zconvsrce, jl, QI) = V(zconvsrce, jl, QI) + V(zlcust, jl, QI);
      }
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        const R zdtdp = ZRDCP * R(0.5) * (ztp1[i2(jl, jk - 1)] + ztp1[i2(jl, jk)]) / a.paph[i2(jl, jk)];
        const R zwtot = std::min(RG * (a.pmfu[i2(jl, jk)] * zadvw[i] + a.pmfd[i2(jl, jk)] * zadvwd[i]),
                                 (a.pap[i2(jl, jk)] - a.pap[i2(jl, jk - 1)]) / ptsphy);
        const R zdtforc = zdtdp * ptsphy * zwtot;
        zdqs[i] = zanewm1[i] * zdtforc * zdqsmixdt[i];
      }
      for (int m : {QL, QI}) {
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const int i = jl - 1;
          R zlfinal = std::max(R(0), V(zlcust, jl, m) - zdqs[i]);
          const R zevap = std::min(V(zlcust, jl, m) - zlfinal, zevaplimmix[i]);
          zlfinal = V(zlcust, jl, m) - zevap;
          zlfinalsum[i] = zlfinalsum[i] + zlfinal;
          QA(jl, m, m) = QA(jl, m, m) + V(zlcust, jl, m);
          QA(jl, QV, m) = QA(jl, QV, m) + zevap;
          QA(jl, m, QV) = QA(jl, m, QV) - zevap;
          if (m == QL) {  // :1459-1460
            BL(jl, 4) = V(zlcust, jl, m) * ZQTMST;
            BL(jl, 5) = -zevap * ZQTMST;
          } else {  // :1475-1476
            BI(jl, 4) = V(zlcust, jl, m) * ZQTMST;
            BI(jl, 5) = -zevap * ZQTMST;
          }
        }
      }
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        zsolac[i] = zsolac[i] + zacust[i];
        BCC(jl, 4) = zacust[i] * ZQTMST;  // :1484
        if (zlfinalsum[i] < ZEPSEC) {
          zsolac[i] = zsolac[i] - zacust[i];
          BCC(jl, 5) = -zacust[i] * ZQTMST;  // :1488
        }
      }
    }
    // subsidence sink to the level below, implicit (:1499-1514)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (jk < klev) {
        const R zmfdn = std::max(R(0), (a.pmfu[i2(jl, jk + 1)] * zadvw[i] + a.pmfd[i2(jl, jk + 1)] * zadvwd[i]) * zdtgdp[i]);
        zsolab[i] = zsolab[i] + zmfdn;
        QB(jl, QL, QL) = QB(jl, QL, QL) + zmfdn;
        QB(jl, QI, QI) = QB(jl, QI, QI) + zmfdn;
        V(zconvsink, jl, QL) = zmfdn;
        V(zconvsink, jl, QI) = zmfdn;
      }
    }

    // 3.4 turbulent erosion, ITURBEROSION = 3 (:1689-1754)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      zldifdt[i] = RCLDIFF * ptsphy;
      if (a.ktype[i] >= 2 && a.plude[i2(jl, jk)] > ZEPSEC && a.peis[i] < R(10.0)) {
        zldifdt[i] = R(20.0) * RCLDIFF_CONVI * zldifdt[i];
      } else if (a.ktype[i] == 1 && a.plude[i2(jl, jk)] > ZEPSEC) {
        zldifdt[i] = R(2.0) * RCLDIFF_CONVI * zldifdt[i];
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (zli[k] > ZEPSEC) {
        const R zqe = (zqx[i3(jl, jk, QV + 1)] - za[k] * zqsmix[k]) / std::max(ZEPSEC, R(1) - za[k]);
        const R ze = std::max(zqsmix[k] - zqe, R(0));
        R zleros = R(0.333) * za[k] * (R(1) - za[k]) * zldifdt[i] * ze;
        zleros = std::min(zleros, zevaplimmix[i]);
        zleros = std::min(zleros, zli[k]);
        const R zaeros = R(0.5) * zleros / zlicld[i];
        zsolac[i] = zsolac[i] - zaeros;
        QA(jl, QV, QL) = QA(jl, QV, QL) + zliqfrac[k] * zleros;
        QA(jl, QL, QV) = QA(jl, QL, QV) - zliqfrac[k] * zleros;
        QA(jl, QV, QI) = QA(jl, QV, QI) + zicefrac[k] * zleros;
        QA(jl, QI, QV) = QA(jl, QI, QV) - zicefrac[k] * zleros;
        BL(jl, 7) = -zliqfrac[k] * zleros * ZQTMST;  // :1749-1751
        BI(jl, 7) = -zicefrac[k] * zleros * ZQTMST;
        BCC(jl, 7) = -zaeros * ZQTMST;
      }
    }

    // 4.1 sedimentation: fall speeds (:1783-1844)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      V(zfallspeed, jl, QL) = zvqx[QL];
      V(zfallspeed, jl, QI) = zvqx[QI];
      V(zfallspeed, jl, QS) = zvqx[QS];
      const R zfallcorr = std::pow(RDENSREF / zrho[i], R(0.4));
      const R ztemp = R(0.5) * (zqx[i3(jl, jk, QR + 1)] + V(zqxnm1, jl, QR));
      const R zrlambda = std::pow(zrho[i] * ztemp / RCL_LAM1R, RCL_LAM2R);
      V(zfallspeed, jl, QR) = RXFALLSPEEDRAIN * zfallcorr * RCL_CONST7R * std::pow(zrlambda, RCL_DR);
    }
    // fall in (explicit) and out (implicit) (:1852-1877)
    for (int m = 0; m < NCLV; ++m) {
      if (!llfall[m]) continue;
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        if (jk > ncldtop) {
          V(zfallsrce, jl, m) = zpfplsx[ih3(jl, jk, m + 1)] * zdtgdp[i];
          QA(jl, m, m) = QA(jl, m, m) + V(zfallsrce, jl, m);
          if (m == QI) BI(jl, 12) = V(zfallsrce, jl, m) * ZQTMST;  // :1865
          V(zqxfg, jl, m) = V(zqxfg, jl, m) + V(zfallsrce, jl, m);
          zqpretot[i] = zqpretot[i] + V(zqxfg, jl, m);
        }
        V(zfallsink, jl, m) = zdtgdp[i] * zrho[i] * V(zfallspeed, jl, m);
      }
    }

    // 4.2 precipitation fraction, maximum-random overlap (:1898-1919)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (zqpretot[i] > ZEPSEC) {  // implies jk > ncldtop, so jk-1 is a valid level
        const R za_k = za[i2(jl, jk)], za_km1 = za[i2(jl, jk - 1)];
        zcovptot[i] = R(1) - ((R(1) - zcovptot[i]) * (R(1) - std::max(za_k, za_km1)) /
                              (R(1) - std::min(za_km1, R(1) - R(1.E-06))));
        zcovptot[i] = std::max(zcovptot[i], RCOVPMIN);
        zcovpclr[i] = std::max(R(0), zcovptot[i] - za_k);
        zraincld[i] = R(0.5) * (zqx[i3(jl, jk, QR + 1)] + zqx[i3(jl, jk - 1, QR + 1)]) / zcovptot[i];
        zraincldm1[i] = R(0.5) * (zqx[i3(jl, jk, QR + 1)] + V(zqxnm1, jl, QR)) / zcovptot[i];
        zsnowcld[i] = R(0.5) * (zqx[i3(jl, jk, QS + 1)] + zqx[i3(jl, jk - 1, QS + 1)]) / zcovptot[i];
        zsnowcldm1[i] = R(0.5) * (zqx[i3(jl, jk, QS + 1)] + V(zqxnm1, jl, QS)) / zcovptot[i];
        zcovpmax[i] = std::max(zcovptot[i], zcovpmax[i]);
      } else {
        zraincld[i] = zraincldm1[i] = zsnowcld[i] = zsnowcldm1[i] = R(0);
        zcovptot[i] = zcovpclr[i] = zcovpmax[i] = R(0);
      }
    }

    // 4.3 warm rain, Khairoutdinov-Kogan implicit (IWARMRAIN=3, IRAINACC=1) (:1930-2259)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (zliqcld[i] > ZEPSEC) {
        R zgridlen = R(0);
        if (lfsd || linh) {  // inhomogeneity (:1937-1971)
          const size_t k = i2(jl, jk);
          const R zqtot = std::min((zqx[i3(jl, jk, QV + 1)] + zqx[i3(jl, jk, QL + 1)]) * R(1000.), R(30.));
          zgridlen = R(2) * RA * std::sqrt(RPI * a.bud->pgaw[i]) * R(0.001);
          const R zphip1 = R(.2) + R(.01) * zqtot + R(.0027) * (zqtot * zqtot) - R(.00008) * (zqtot * zqtot * zqtot);
          const R zphip2 = zphip1 - R(0.2);
          const R zphip3 = R(0.123) * std::exp(-(zphip1 * R(0.55)));
          const R zanew2 = std::max(za[k], R(0.1));
          const R zphic = std::pow(zgridlen * zanew2, R(1) / R(3)) *
                          std::pow(std::pow(zphip3 * zgridlen * zanew2, R(1.5)) + R(3) * zphip3, R(-0.17));
          R zfracsdc = (zphip1 - zphip2 * zanew2) * zphic;
          if (za[k] > R(0.95)) zfracsdc = R(0.17) * zphic;
          zlfsd[k] = ZR12 * zfracsdc;
        }
        R zconst;
        if (a.plsm[i] > R(0.5)) {
          zconst = RCL_KK_CLOUD_NUM_LAND;
        } else {
          zconst = RCL_KK_CLOUD_NUM_SEA;
        }
        R zeaut = RCL_INHOMOGAUT, zeacc = RCL_INHOMOGACC;
        if (linh) {  // Boutle et al. (2013) enhancement factors (:2075-2098); ZCLDRAINCORR (:1947)
          const size_t k = i2(jl, jk);
          const R zfracsdc = zlfsd[k];
          const R zcldraincorr = R(1) - R(0.8) * za[k];
          zeaut = std::pow(R(1) + std::pow(zfracsdc, R(2)), (RCL_KKBAUQ - R(1)) * RCL_KKBAUQ / R(2));
          const R zphir = std::pow(zgridlen * za[k], R(1) / R(3)) *
                          std::pow(std::pow(R(0.11) * zgridlen * za[k], R(1.14)) + R(1), R(-0.22));
          R zfracsdr = (R(1.1) - R(0.8) * za[k]) * zphir;
          if (za[k] > R(0.95)) zfracsdr = R(0.3) * zphir;
          zeacc = std::pow(R(1) + std::pow(zfracsdc, R(2)), (RCL_KKBAC - R(1)) * RCL_KKBAC / R(2)) *
                  std::pow(R(1) + std::pow(zfracsdr, R(2)), (RCL_KKBAC - R(1)) * RCL_KKBAC / R(2)) *
                  std::exp(zcldraincorr * RCL_KKBAC * RCL_KKBAC *
                           std::sqrt(std::log(R(1) + std::pow(zfracsdc, R(2))) *
                                     std::log(R(1) + std::pow(zfracsdr, R(2)))));
        }
        zrainaut[i] = zeaut * ptsphy * RCL_KKAAU * std::pow(zliqcld[i], RCL_KKBAUQ - R(1)) * std::pow(zconst, RCL_KKBAUN);
        QB(jl, QR, QL) = QB(jl, QR, QL) + zrainaut[i];
        if (zraincld[i] > ZEPSEC) {
          zrainacc[i] = zeacc * ptsphy * RCL_KKAAC * std::pow(zliqcld[i], RCL_KKBAC - R(1)) * std::pow(zraincld[i], RCL_KKBAC);
        }
        QB(jl, QR, QL) = QB(jl, QR, QL) + zrainacc[i];
      }
    }

    // 4.4 rain evaporation (:2271-2378)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      R zzrh = RPRECRHMAX + (R(1) - RPRECRHMAX) * zcovpmax[i] / std::max(ZEPSEC, R(1) - za[k]);
      zzrh = std::min(std::max(zzrh, RPRECRHMAX), R(1));
      R zrhc = RAMID;
      const R zsigk = a.pap[k] / a.paph[i2(jl, klev + 1)];
      if (zsigk > R(0.8)) {
        const R s = (zsigk - R(0.8)) / R(0.2);
        zrhc = RAMID + (R(1) - RAMID) * (s * s);
      }
      zzrh = std::min(zrhc, zzrh);
      zzrh = std::min(R(0.9), zzrh);
      const R zqe = std::max(R(0), std::min(zqx[i3(jl, jk, QV + 1)], zqsliq[k]));
      const bool llo1 = zcovpclr[i] > ZEPSEC && zraincldm1[i] > ZEPSEC && zqe < zzrh * zqsliq[k];
      if (llo1) {
        const R zpreclr = zraincldm1[i];
        const R zfallcorr = std::pow(RDENSREF / zrho[i], C04);
        const R zesatliq = RV / RD_ * th.foeeliq(ztp1[k]);
        const R zlambda = std::pow(RCL_LAM1R / (zrho[i] * zpreclr), RCL_LAM2R);
        const R t = ztp1[k];
        const R zevap_denom = RCL_CDENOM1 * zesatliq - RCL_CDENOM2 * t * zesatliq + RCL_CDENOM3 * (t * t * t) * a.pap[k];
        const R zcorr2 = std::pow(t / R(273.), R(1.5)) * R(393.) / (t + R(120.));
        const R zka = RCL_KA273 * zcorr2;
        (void)zka;  // computed by the reference, unused (:2333)
        const R zsubsat = std::max(zzrh * zqsliq[k] - zqe, R(0));
        const R zbeta = (R(0.15) / zqsliq[k]) * (t * t) * zesatliq * RCL_CONST1R * (zcorr2 / zevap_denom) *
                        (R(0.78) / std::pow(zlambda, RCL_CONST4R) +
                         RCL_CONST2R * std::sqrt(zrho[i] * zfallcorr) / (std::sqrt(zcorr2) * std::pow(zlambda, RCL_CONST3R)));
        const R zdenom = R(1) + zbeta * ptsphy * zcorqsliq[i];
        const R zdpevap = zcovpclr[i] * zbeta * ptsphy * zsubsat / zdenom;
        const RD zevap = std::min<RD>(zdpevap, V(zqxfg, jl, QR));
        QA(jl, QV, QR) = QA(jl, QV, QR) + zevap;
        QA(jl, QR, QV) = QA(jl, QR, QV) - zevap;
        BL(jl, 20) = -zevap * ZQTMST;  // :2363
        V(zqxfg, jl, QR) = V(zqxfg, jl, QR) - zevap;
      }
    }

    // 4.5 ice deposition, Rotstayn et al. 2001 (IDEPICE=1) (:2404-2499)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (za[i2(jl, jk - 1)] < RCLDTOPCF && za[k] >= RCLDTOPCF) {
        zcldtopdist[i] = R(0);
      } else {
        zcldtopdist[i] = zcldtopdist[i] + zdp[i] / (zrho[i] * RG);
      }
      const R zoverlap_liqice = RCL_OVERLAPLIQICE;
      const R zsupersatice = (zqsliq[k] - zqsice[k]) / zqsice[k];
      const R t = ztp1[k];
      if (t < RTT - R(5.) && V(zqxfg, jl, QL) > RLMIN) {
        const R zvpice = th.foeeice(t) * RV / RD_;
        const R zvpliq = zvpice * th.fokoop(t);
        (void)zvpliq;  // computed by the reference, unused in IDEPICE=1 (:2439)
        const R zicenuclei = R(1000.0) * std::exp(R(12.96) * zsupersatice - R(0.639));
        const R zadd = RLSTT * (RLSTT / (RV * t) - R(1)) / (R(2.4E-2) * t);
        const R zbdd = RV * t * a.pap[k] / (R(2.21) * zvpice);
        const R zcvds = R(7.8) * std::pow(zicenuclei / zrho[i], R(0.666)) * zsupersatice / (R(8.87) * (zadd + zbdd));
        const R zice0 = std::max(zicecld[i], zicenuclei * RICEINIT / zrho[i]);
        const R zinew = std::pow(R(0.666) * zcvds * ptsphy + std::pow(zice0, R(0.666)), R(1.5));
        RD zdepos = std::max(zoverlap_liqice * za[k] * (zinew - zice0), R(0));
        zdepos = std::min<RD>(zdepos, V(zqxfg, jl, QL));
        const R zinfactor = std::min(zicenuclei / R(15000.), R(1));
        zdepos = zdepos * std::min(zinfactor + (R(1) - zinfactor) * (RDEPLIQREFRATE + zcldtopdist[i] / RDEPLIQREFDEPTH), R(1));
        QA(jl, QI, QL) = QA(jl, QI, QL) + zdepos;
        QA(jl, QL, QI) = QA(jl, QL, QI) - zdepos;
        V(zqxfg, jl, QI) = V(zqxfg, jl, QI) + zdepos;
        V(zqxfg, jl, QL) = V(zqxfg, jl, QL) - zdepos;
        BL(jl, 11) = -zdepos * ZQTMST;  // :2495-2496
        BI(jl, 11) = zdepos * ZQTMST;
      }
    }
    // 4.6 ice sublimation: ISUBLICE = 0, process rate stays zero (:2613-2615)

    // 4.7 deposition onto snow (IDEPSNOW=1) (:2698-2745)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zpsdep[i] = R(0);
      const R qv = zqx[i3(jl, jk, QV + 1)];
      const R zsupersatice = std::max((qv - zqsmix[k]) / zqsmix[k], R(0));
      const R t = ztp1[k];
      if (t < RTT && zsnowcld[i] > ZEPSEC && zsupersatice > R(0)) {
        const R zvpice = th.foeeice(t) * RV / RD_;
        const R ztcg = R(1), zfacx1s = R(1);
        const R zaplusb = RCL_APB1 * zvpice - RCL_APB2 * zvpice * t + a.pap[k] * RCL_APB3 * (t * t * t);
        const R zcorrfac = std::sqrt(R(1) / zrho[i]);
        const R zcorrfac2 = std::pow(t / R(273.0), R(1.5)) * (R(393.0) / (t + R(120.0)));
        const R zpr02 = zrho[i] * zsnowcld[i] * RCL_CONST1S / (ztcg * zfacx1s);
        const R zterm1 = zsupersatice * (t * t) * zvpice * zcorrfac2 * ztcg * RCL_CONST2S * zfacx1s / (zrho[i] * zaplusb);
        const R zterm2 = R(0.65) * RCL_CONST6S * std::pow(zpr02, RCL_CONST4S) +
                         RCL_CONST3S * std::sqrt(zcorrfac) * std::sqrt(zrho[i]) * std::pow(zpr02, RCL_CONST5S) / std::sqrt(zcorrfac2);
        zpsdep[i] = std::max(zcovptot[i] * zterm1 * zterm2 * ptsphy, R(0));
        zpsdep[i] = std::min(zpsdep[i], qv - zqsmix[k]) / zcorqsmix[i];  // division applies to the MIN (:2733)
        QA(jl, QS, QV) = QA(jl, QS, QV) + zpsdep[i];
        QA(jl, QV, QS) = QA(jl, QV, QS) - zpsdep[i];
      }
    }

    // revise in-cloud condensate (:2750-2754)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const R ztmpa = R(1) / std::max(za[i2(jl, jk)], R(0.01));
      zliqcld[i] = V(zqxfg, jl, QL) * ztmpa;
      zicecld[i] = V(zqxfg, jl, QI) * ztmpa;
    }

    // 4.8 ice -> snow autoconversion (:2769-2800)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const R t = ztp1[i2(jl, jk)];
      if (t <= RTT) {
        if (zicecld[i] > ZEPSEC) {
          const R zzco = ptsphy * RSNOWLIN1 * std::exp(RSNOWLIN2 * (t - RTT));
          const R zlcrit = RLCRITSNOW;
          const R r = zicecld[i] / zlcrit;
          const R ztemp = R(-1.) * (r * r);
          if (ztemp < R(50.0)) {
            zsnowaut[i] = zzco * (R(1) - std::exp(ztemp));
          } else {
            zsnowaut[i] = zzco;
          }
          QB(jl, QS, QI) = QB(jl, QS, QI) + zsnowaut[i];
        }
      }
    }

    // 4.12 melting of snow and ice (:2815-2862)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zicetot[i] = V(zqxfg, jl, QI) + V(zqxfg, jl, QS);
      zmeltmax[i] = R(0);
      const R t = ztp1[k];
      if (zicetot[i] > ZEPSEC && t > RTT) {
        const R zsubsat = std::max(zqsice[k] - zqx[i3(jl, jk, QV + 1)], R(0));
        const R ztdmtw0 = t - RTT - zsubsat * (ZTW1 + ZTW2 * (a.pap[k] - ZTW3) - ZTW4 * (t - ZTW5));
        const R zcons1 = std::fabs(ptsphy * (R(1) + R(0.5) * ztdmtw0) / RTAUMEL);
        zmeltmax[i] = std::max(ztdmtw0 * zcons1 * ZRLDCP, R(0));
      }
    }
    for (int jm = 0; jm < NCLV; ++jm) {
      if (iphase[jm] != 2) continue;
      const int jn = imelt[jm];
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        if (zmeltmax[i] > ZEPSEC && zicetot[i] > ZEPSEC) {
          const R zalfa = V(zqxfg, jl, jm) / zicetot[i];
          const RD zmelt = std::min<RD>(V(zqxfg, jl, jm), zalfa * zmeltmax[i]);
          V(zqxfg, jl, jm) = V(zqxfg, jl, jm) - zmelt;
          V(zqxfg, jl, jn) = V(zqxfg, jl, jn) + zmelt;
          QA(jl, jn, jm) = QA(jl, jn, jm) + zmelt;
          QA(jl, jm, jn) = QA(jl, jm, jn) - zmelt;
          if (jm == QI) BI(jl, 15) = -zmelt * ZQTMST;  // :2856-2858
          if (jm == QI) BL(jl, 17) = zmelt * ZQTMST;
          if (jm == QS) BI(jl, 16) = -zmelt * ZQTMST;
        }
      }
    }

    // 4.13 freezing of rain (:2877-2933)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      const R qr = zqx[i3(jl, jk, QR + 1)];
      if (qr > ZEPSEC) {
        const R t = ztp1[k];
        if (t <= RTT && ztp1[i2(jl, jk - 1)] > RTT) {
          zqpretot[i] = std::max(zqx[i3(jl, jk, QS + 1)] + qr, ZEPSEC);
          a.prainfrac_toprfz[i] = qr / zqpretot[i];
          llrainliq[i] = a.prainfrac_toprfz[i] > C08;
        }
        if (t < RTT) {
          if (llrainliq[i]) {
            const R zlambda = std::pow(RCL_LAM1R / (zrho[i] * qr), RCL_LAM2R);
            const R ztemp = std::min(RCL_FZRAB * (t - RTT), R(50.));
            const R zfrz = ptsphy * (RCL_CONST5R / zrho[i]) * (std::exp(ztemp) - R(1.)) * std::pow(zlambda, RCL_CONST6R);
            zfrzmax[i] = std::max(zfrz, R(0));
          } else {
            const R zcons1 = std::fabs(ptsphy * (R(1) + R(0.5) * (RTT - t)) / RTAUMEL);
            zfrzmax[i] = std::max((RTT - t) * zcons1 * ZRLDCP, R(0));
          }
          if (zfrzmax[i] > ZEPSEC) {
            const R zfrz = std::min(qr, zfrzmax[i]);
            QA(jl, QS, QR) = QA(jl, QS, QR) + zfrz;
            QA(jl, QR, QS) = QA(jl, QR, QS) - zfrz;
            BL(jl, 18) = -zfrz * ZQTMST;  // :2927
          }
        }
      }
    }

    // 4.14 homogeneous freezing of cloud liquid (:2948-2960)
    for (int jl = kidia; jl <= kfdia; ++jl) zfrzmax[jl - 1] = std::max((RTHOMO - ztp1[i2(jl, jk)]) * ZRLDCP, R(0));
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (zfrzmax[i] > ZEPSEC && V(zqxfg, jl, QL) > ZEPSEC) {
        const RD zfrz = std::min<RD>(V(zqxfg, jl, QL), zfrzmax[i]);
        QA(jl, QI, QL) = QA(jl, QI, QL) + zfrz;
        QA(jl, QL, QI) = QA(jl, QL, QI) - zfrz;
        BL(jl, 19) = -zfrz * ZQTMST;  // :2958
      }
    }

    // 4.9 riming of snow (:2975-3017)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (ztp1[i2(jl, jk)] <= RTT && zliqcld[i] > ZEPSEC) {
        const R zfallcorr = std::pow(RDENSREF / zrho[i], C04);
        if (zsnowcld[i] > ZEPSEC && zcovptot[i] > R(0.01)) {
          zsnowrime[i] = RCL_EFFRIME * zcovptot[i] * ptsphy * RCL_CONST7S * zfallcorr *
                         std::pow(zrho[i] * zsnowcld[i] * RCL_CONST1S, RCL_CONST8S);
          zsnowrime[i] = std::min(zsnowrime[i], R(1));
          QB(jl, QS, QL) = QB(jl, QS, QL) + zsnowrime[i];
        }
      }
    }
    // PSD slopes (:3029-3067)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      if (zraincld[i] > ZEPSEC) {
        zlamr[i] = std::pow(RCL_LAM1R / (zrho[i] * zraincld[i]), RCL_LAM2R);
        zn0r[i] = RCL_X1R * std::pow(zlamr[i], RCL_X2R);
      } else {
        zlamr[i] = zn0r[i] = R(0);
      }
      if (zsnowcld[i] > ZEPSEC) {
        zlams[i] = std::pow(RCL_LAM1S / (zrho[i] * zsnowcld[i]), RCL_LAM2S);
        zn0s[i] = RCL_X1S;
      } else {
        zlams[i] = zn0s[i] = R(0);
      }
      if (zicecld[i] > ZEPSEC) {
        zlami[i] = std::pow(RCL_LAM1S / (zrho[i] * zicecld[i]), RCL_LAM2S);
        zn0i[i] = RCL_X1S;
      } else {
        zlami[i] = zn0i[i] = R(0);
      }
    }
    // PSACR: snow accretes rain (:3084-3115)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      zpsacr[i] = R(0);
      if (ztp1[i2(jl, jk)] <= RTT && a.prainfrac_toprfz[i] < ZEPSEC && zraincld[i] > ZEPSEC && zsnowcld[i] > ZEPSEC) {
        const R vr = V(zfallspeed, jl, QR), vs = V(zfallspeed, jl, QS);
        const R zfalldiff = std::max(std::fabs(vr - vs), (vr + vs) / R(8.));
        const R lr3 = zlamr[i] * zlamr[i] * zlamr[i];
        const R zfac = RPI * RPI * RDENSWAT * R(1.0) * zrho[i] * zn0r[i] * zn0s[i] / (zlams[i] * lr3);
        const R zrat = zlamr[i] / zlams[i];
        zpsacr[i] = zcovptot[i] * ptsphy * zfac * zfalldiff * ((R(0.5) * zrat + R(2.)) * zrat + R(5.)) / lr3;
        QA(jl, QS, QR) = QA(jl, QS, QR) + zpsacr[i];
        QA(jl, QR, QS) = QA(jl, QR, QS) - zpsacr[i];
      }
    }
    // PIACR: ice accretes rain (:3171-3198)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      zpiacr[i] = R(0);
      if (ztp1[i2(jl, jk)] <= RTT && a.prainfrac_toprfz[i] < ZEPSEC && zraincld[i] > ZEPSEC && zicecld[i] > ZEPSEC) {
        const R vr = V(zfallspeed, jl, QR), vi = V(zfallspeed, jl, QI);
        const R zfalldiff = std::max(std::fabs(vr - vi), (vr + vi) / R(8.));
        const R lr3 = zlamr[i] * zlamr[i] * zlamr[i];
        const R zfac = RPI * RPI * RDENSWAT * R(0.1) * zrho[i] * zn0r[i] * zn0i[i] / (zlami[i] * lr3);
        const R zrat = zlamr[i] / zlami[i];
        zpiacr[i] = zcovptot[i] * ptsphy * zfac * zfalldiff * ((R(0.5) * zrat + R(2.)) * zrat + R(5.)) / lr3;
        QA(jl, QI, QR) = QA(jl, QI, QR) + zpiacr[i];
        QA(jl, QR, QI) = QA(jl, QR, QI) - zpiacr[i];
      }
    }

    // 4.15 snow sublimation (ISUBLSNOW=1) (:3254-3318)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      R zzrh = RPRECRHMAX + (R(1) - RPRECRHMAX) * zcovpmax[i] / std::max(ZEPSEC, R(1) - za[k]);
      zzrh = std::min(std::max(zzrh, RPRECRHMAX), R(1));
      R zqe = (zqx[i3(jl, jk, QV + 1)] - za[k] * zqsice[k]) / std::max(ZEPSEC, R(1) - za[k]);
      zqe = std::max(R(0), std::min(zqe, zqsice[k]));
      const bool llo1 = zcovpclr[i] > ZEPSEC && V(zqxfg, jl, QS) > ZEPSEC && zqe < zzrh * zqsice[k];
      if (llo1) {
        const R d = zcovptot[i] * zdtgdp[i];
        const R zpreclr = V(zqxfg, jl, QS) * zcovpclr[i] / std::copysign(std::max(std::fabs(d), ZEPSILON), d);
        const R zbeta1 = std::sqrt(a.pap[k] / a.paph[i2(jl, klev + 1)]) / RVRFACTOR * zpreclr / std::max(zcovpclr[i], ZEPSEC);
        const R zbeta = RG * RPECONS * std::pow(zbeta1, R(0.5777));
        const R zdenom = R(1) + zbeta * ptsphy * zcorqsice[i];
        const R zdpr = zcovpclr[i] * zbeta * (zqsice[k] - zqe) / zdenom * zdp[i] * ZRG_R;
        const R zdpevap = zdpr * zdtgdp[i];
        const RD zevap = std::min<RD>(zdpevap, V(zqxfg, jl, QS));
        QA(jl, QV, QS) = QA(jl, QV, QS) + zevap;
        QA(jl, QS, QV) = QA(jl, QS, QV) - zevap;
        BI(jl, 17) = -zevap * ZQTMST;  // :3304
        zcovptot[i] = std::max(RCOVPMIN, zcovptot[i] - std::max(R(0), R((zcovptot[i] - za[k]) * zevap / V(zqxfg, jl, QS))));
        V(zqxfg, jl, QS) = V(zqxfg, jl, QS) - zevap;
      }
    }

    // evaporate small precipitation amounts (:3402-3412)
    for (int m = 0; m < NCLV; ++m) {
      if (!llfall[m]) continue;
      for (int jl = kidia; jl <= kfdia; ++jl) {
        if (V(zqxfg, jl, m) < RLMIN) {
          QA(jl, QV, m) = QA(jl, QV, m) + V(zqxfg, jl, m);
          QA(jl, m, QV) = QA(jl, m, QV) - V(zqxfg, jl, m);
        }
      }
    }

    // 5.1 cloud cover (:3429-3438)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      R zanew = (za[i2(jl, jk)] + zsolac[i]) / (R(1) + zsolab[i]);
      zanew = std::min(zanew, R(1));
      if (zanew < RAMIN) zanew = R(0);
      zda[i] = zanew - zaorig[i2(jl, jk)];
      zanewm1[i] = zanew;
    }

    // 5.2.1 truncate explicit sinks (:3455-3536) and 5.2.2 solve (:3546-3628), per column
    for (int jl = kidia; jl <= kfdia; ++jl) {
      RD qa[NCLV][NCLV], qxd[NCLV];
      R qb[NCLV][NCLV], qx[NCLV], fs[NCLV], qxn[NCLV];
      for (int m = 0; m < NCLV; ++m) {
        qx[m] = zqx[i3(jl, jk, m + 1)];
        fs[m] = V(zfallsink, jl, m);
        for (int n = 0; n < NCLV; ++n) {
          qa[m][n] = QA(jl, m, n);
          qb[m][n] = QB(jl, m, n);
        }
      }
      for (int m = 0; m < NCLV; ++m) qxd[m] = RD(qx[m]);
      truncate_sinks<RD>(qa, qxd, RD(ZEPSEC));
      implicit_solve<R, RD>(qa, qb, fs, qx, qxn);
      for (int m = 0; m < NCLV; ++m) {
        V(zqxn, jl, m) = qxn[m];
        for (int n = 0; n < NCLV; ++n) QA(jl, m, n) = qa[m][n];
      }
    }
    // clip small values (:3633-3640)
    for (int jn = 0; jn < NCLV - 1; ++jn)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        if (V(zqxn, jl, jn) < ZEPSEC) {
          V(zqxn, jl, QV) = V(zqxn, jl, QV) + V(zqxn, jl, jn);
          V(zqxn, jl, jn) = R(0);
        }
      }
    // carry to next level (:3645-3650)
    for (int m = 0; m < NCLV; ++m)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        V(zqxnm1, jl, m) = V(zqxn, jl, m);
        zqxn2d[i3(jl, jk, m + 1)] = V(zqxn, jl, m);
      }
    // 5.3 fluxes to the next level (:3658-3672)
    for (int m = 0; m < NCLV; ++m)
      for (int jl = kidia; jl <= kfdia; ++jl)
        zpfplsx[ih3(jl, jk + 1, m + 1)] = V(zfallsink, jl, m) * V(zqxn, jl, m) * zrdtgdp[jl - 1];
    for (int jl = kidia; jl <= kfdia; ++jl) {
      zqpretot[jl - 1] = zpfplsx[ih3(jl, jk + 1, QS + 1)] + zpfplsx[ih3(jl, jk + 1, QR + 1)];
      if (zqpretot[jl - 1] < ZEPSEC) zcovptot[jl - 1] = R(0);
    }
    // diagnosed rates of the implicit processes (:3674-3698); IWARMRAIN = 3 with IP_SNOW_ACCRETES_RAIN = 1
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      BL(jl, 14) = -zrainaut[i] * V(zqxn, jl, QL) * ZQTMST;
      BL(jl, 15) = -zrainacc[i] * V(zqxn, jl, QL) * ZQTMST;
      BCC(jl, 6) = -V(zconvsink, jl, QL) * zanewm1[i] * ZQTMST;
      BL(jl, 6) = -V(zconvsink, jl, QL) * V(zqxn, jl, QL) * ZQTMST;
      BI(jl, 6) = -V(zconvsink, jl, QI) * V(zqxn, jl, QI) * ZQTMST;
      BI(jl, 13) = -V(zfallsink, jl, QI) * V(zqxn, jl, QI) * ZQTMST;
      BL(jl, 21) = -zsnowrime[i] * V(zqxn, jl, QL) * ZQTMST;
      BI(jl, 14) = -zsnowaut[i] * V(zqxn, jl, QI) * ZQTMST;
    }

    // 6.1 temperature and condensate tendencies (:3713-3745)
    for (int jm = 0; jm < NCLV - 1; ++jm) {
      for (int jl = kidia; jl <= kfdia; ++jl)
        V(zfluxq, jl, jm) = V(zconvsrce, jl, jm) + V(zfallsrce, jl, jm) -
                            (V(zfallsink, jl, jm) + V(zconvsink, jl, jm)) * V(zqxn, jl, jm);
      if (iphase[jm] == 1)
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const size_t k = i2(jl, jk);
          a.tloc_t[k] = a.tloc_t[k] + RALVDCP * (V(zqxn, jl, jm) - zqx[i3(jl, jk, jm + 1)] - V(zfluxq, jl, jm)) * ZQTMST;
        }
      if (iphase[jm] == 2)
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const size_t k = i2(jl, jk);
          a.tloc_t[k] = a.tloc_t[k] + RALSDCP * (V(zqxn, jl, jm) - zqx[i3(jl, jk, jm + 1)] - V(zfluxq, jl, jm)) * ZQTMST;
        }
      for (int jl = kidia; jl <= kfdia; ++jl)
        a.tloc_cld[i3(jl, jk, jm + 1)] = a.tloc_cld[i3(jl, jk, jm + 1)] + (V(zqxn, jl, jm) - zqx0[i3(jl, jk, jm + 1)]) * ZQTMST;
    }
    // 6.2-6.3 humidity and cloud-cover tendencies; updated state (:3748-3769)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      a.tloc_q[k] = a.tloc_q[k] + (V(zqxn, jl, QV) - zqx[i3(jl, jk, QV + 1)]) * ZQTMST;
      a.tloc_a[k] = a.tloc_a[k] + zda[i] * ZQTMST;
      zq_upd[i] = a.pq[k] + ptsphy * (a.tloc_q[k] + a.tcml_q[k]);
      zt_upd[i] = a.pt[k] + ptsphy * (a.tloc_t[k] + a.tcml_t[k]);
      za_upd[i] = a.pa[k] + ptsphy * (a.tloc_a[k] + a.tcml_a[k]);
    }
    // 6.4 CLOUD_SUPERSATCHECK (:3772-3777).  The sections PAP(KIDIA:KFDIA,JK) and
    // PAPH(KIDIA:KFDIA,KLEV+1) are indexed from KIDIA by the callee, so column JL
    // reads the pressures of column KIDIA+JL-1 (identical when KIDIA = 1).
    supersatcheck<R, R>(p, kidia, kfdia, zt_upd.data(), zq_upd.data(), za_upd.data(), a.pap + i2(kidia, jk),
                         a.paph + i2(kidia, klev + 1), zt_adj.data(), zq_adj.data(), za_adj.data(), zl_adj.data(),
                         zi_adj.data());
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:3779-3809)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      a.tloc_q[k] = a.tloc_q[k] + R(zq_adj[i]) * ZQTMST;
      a.tloc_t[k] = a.tloc_t[k] + R(zt_adj[i]) * ZQTMST;
      a.tloc_a[k] = a.tloc_a[k] + R(za_adj[i]) * ZQTMST;
      a.tloc_cld[i3(jl, jk, QL + 1)] = a.tloc_cld[i3(jl, jk, QL + 1)] + R(zl_adj[i]) * ZQTMST;
      a.tloc_cld[i3(jl, jk, QI + 1)] = a.tloc_cld[i3(jl, jk, QI + 1)] + R(zi_adj[i]) * ZQTMST;
      BCC(jl, 2) = R(za_adj[i]) * ZQTMST;  // :3787-3789
      BL(jl, 2) = R(zl_adj[i]) * ZQTMST;
      BI(jl, 2) = R(zi_adj[i]) * ZQTMST;
      const R zl = zqx0[i3(jl, jk, QL + 1)] + a.tloc_cld[i3(jl, jk, QL + 1)] * ptsphy;
      if (a.nwarn && zl < -RLMIN) ++a.nwarn[0];  // WRITE(NULOUT,*) 'CLOUDSC: WARNING Negative cloud liquid!' (:3794)
      if (zl < R(0)) {
        a.tloc_cld[i3(jl, jk, QL + 1)] = a.tloc_cld[i3(jl, jk, QL + 1)] - zl * ZQTMST;
        a.tloc_q[k] = a.tloc_q[k] + zl * ZQTMST;
        a.tloc_t[k] = a.tloc_t[k] - RALVDCP * zl * ZQTMST;
      }
      const R zi = zqx0[i3(jl, jk, QI + 1)] + a.tloc_cld[i3(jl, jk, QI + 1)] * ptsphy;
      if (a.nwarn && zi < -RLMIN) ++a.nwarn[1];  // 'CLOUDSC: WARNING Negative cloud ice!' (:3802)
      if (zi < R(0)) {
        a.tloc_cld[i3(jl, jk, QI + 1)] = a.tloc_cld[i3(jl, jk, QI + 1)] - zi * ZQTMST;
        a.tloc_q[k] = a.tloc_q[k] + zi * ZQTMST;
        a.tloc_t[k] = a.tloc_t[k] - RALSDCP * zi * ZQTMST;
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) a.pcovptot[i2(jl, jk)] = zcovptot[jl - 1];  // (:3814-3816)
    if (lfsd)  // cloud fraction/precipitation fraction at the end of the step (:3818-3822)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const size_t k = i2(jl, jk);
        if (a.pa[k] + zda[jl - 1] > R(0.001)) zanewp[k] = std::max(std::max(a.pa[k] + zda[jl - 1], a.pcovptot[k]), R(0.1));
      }

    // ---- 7. cloud budget diagnostics (:3826-4101) ----
    if (lcldbud) {
      const BudArgs<R>& b = *a.bud;
      auto PX = [&](int jl, int k, int f) -> R& {  // PEXTRA(JL,K,F)
        return b.pextra[((size_t)(f - 1) * klev + (k - 1)) * klon + (jl - 1)];
      };
      int is = (lcldbud & CLOUDSC_BUD_LBUD23) ? 26 : 0;
      const R zdt = (lcldbud & CLOUDSC_BUD_TIMEINT) ? ptsphy : R(1);
      // Rates that the default configuration never sets keep their zero from :1271-1273 (ZBUDCC 1, 10-12;
      // ZBUDL 1, 8-10, 12-13, 16, 22; ZBUDI 1, 8-10, 18) and are read as such below.  ZPRACS, ZPRACI and
      // ZPIEVAP are zero likewise (:3130, :3207, :2613 with IP_RAIN_ACCRETES_SNOW = 0, ISUBLICE = 0).
      const R zpracs = R(0), zpraci = R(0), zpievap = R(0);
      if (lcldbud & CLOUDSC_BUD_VERTINT) {  // :3849-3943
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const int i = jl - 1;
          const size_t k = i2(jl, jk);
          const R zdz = zdp[i] / (zrho[i] * RG);
          const int f = is + 1;
          int ik = 0;
          PX(jl, ik + 4, f) = PX(jl, ik + 4, f) + BCC(jl, 2) * zdz * zdt;
          PX(jl, ik + 6, f) = PX(jl, ik + 6, f) + BCC(jl, 3) * zdz * zdt;
          PX(jl, ik + 7, f) = PX(jl, ik + 7, f) + (BCC(jl, 4) + BCC(jl, 6)) * zdz * zdt;
          PX(jl, ik + 8, f) = PX(jl, ik + 8, f) + BCC(jl, 5) * zdz * zdt;
          PX(jl, ik + 9, f) = PX(jl, ik + 9, f) + BCC(jl, 7) * zdz * zdt;
          PX(jl, ik + 10, f) = PX(jl, ik + 10, f) + BCC(jl, 12) * zdz * zdt;
          PX(jl, ik + 11, f) = PX(jl, ik + 11, f) + b.pvfa[k] * zdz * zdt;
          PX(jl, ik + 12, f) = PX(jl, ik + 12, f) + b.pdyna[k] * zdz * zdt;
          ik = 14;
          PX(jl, ik + 5, f) = PX(jl, ik + 5, f) + BL(jl, 2) * zdz * zdt;
          PX(jl, ik + 7, f) = PX(jl, ik + 7, f) + BL(jl, 3) * zdz * zdt;
          PX(jl, ik + 8, f) = PX(jl, ik + 8, f) + (BL(jl, 4) + BL(jl, 6)) * zdz * zdt;
          PX(jl, ik + 9, f) = PX(jl, ik + 9, f) + BL(jl, 5) * zdz * zdt;
          PX(jl, ik + 10, f) = PX(jl, ik + 10, f) + BL(jl, 7) * zdz * zdt;
          PX(jl, ik + 11, f) = PX(jl, ik + 11, f) + BL(jl, 11) * zdz * zdt;
          PX(jl, ik + 12, f) = PX(jl, ik + 12, f) + BL(jl, 14) * zdz * zdt;
          PX(jl, ik + 13, f) = PX(jl, ik + 13, f) + BL(jl, 15) * zdz * zdt;
          PX(jl, ik + 14, f) = PX(jl, ik + 14, f) + BL(jl, 12) * zdz * zdt;
          PX(jl, ik + 15, f) = PX(jl, ik + 15, f) + BL(jl, 13) * zdz * zdt;
          PX(jl, ik + 16, f) = PX(jl, ik + 16, f) + BL(jl, 17) * zdz * zdt;
          PX(jl, ik + 17, f) = PX(jl, ik + 17, f) + (BL(jl, 18) + BL(jl, 19)) * zdz * zdt;
          PX(jl, ik + 18, f) = PX(jl, ik + 18, f) + BL(jl, 20) * zdz * zdt;
          PX(jl, ik + 19, f) = PX(jl, ik + 19, f) + BL(jl, 21) * zdz * zdt;
          PX(jl, ik + 20, f) = PX(jl, ik + 20, f) + zpracs * zdz * zdt * ZQTMST;
          PX(jl, ik + 21, f) = PX(jl, ik + 21, f) + zpsacr[i] * zdz * zdt * ZQTMST;
          PX(jl, ik + 22, f) = PX(jl, ik + 22, f) + zpraci * zdz * zdt * ZQTMST;
          PX(jl, ik + 23, f) = PX(jl, ik + 23, f) + zpiacr[i] * zdz * zdt * ZQTMST;
          PX(jl, ik + 24, f) = PX(jl, ik + 24, f) + a.pvfl[k] * zdz * zdt;
          PX(jl, ik + 25, f) = PX(jl, ik + 25, f) + b.pdynl[k] * zdz * zdt;
          ik = 39;
          PX(jl, ik + 5, f) = PX(jl, ik + 5, f) + BI(jl, 2) * zdz * zdt;
          PX(jl, ik + 7, f) = PX(jl, ik + 7, f) + BI(jl, 3) * zdz * zdt;
          PX(jl, ik + 8, f) = PX(jl, ik + 8, f) + (BI(jl, 4) + BI(jl, 6)) * zdz * zdt;
          PX(jl, ik + 9, f) = PX(jl, ik + 9, f) + BI(jl, 5) * zdz * zdt;
          PX(jl, ik + 10, f) = PX(jl, ik + 10, f) + BI(jl, 7) * zdz * zdt;
          PX(jl, ik + 11, f) = PX(jl, ik + 11, f) + BI(jl, 11) * zdz * zdt;
          PX(jl, ik + 12, f) = PX(jl, ik + 12, f) + (BI(jl, 12) + BI(jl, 13)) * zdz * zdt;
          PX(jl, ik + 13, f) = PX(jl, ik + 13, f) + BI(jl, 14) * zdz * zdt;
          PX(jl, ik + 14, f) = PX(jl, ik + 14, f) + BI(jl, 16) * zdz * zdt;
          PX(jl, ik + 15, f) = PX(jl, ik + 15, f) + BI(jl, 17) * zdz * zdt;
          PX(jl, ik + 16, f) = PX(jl, ik + 16, f) + zpsdep[i] * ZQTMST * zdz * zdt;
          PX(jl, ik + 17, f) = PX(jl, ik + 17, f) + a.pvfi[k] * zdz * zdt;
          PX(jl, ik + 18, f) = PX(jl, ik + 18, f) + b.pdyni[k] * zdz * zdt;
          PX(jl, ik + 19, f) = PX(jl, ik + 19, f) + b.pdyns[k] * zdz * zdt;
          PX(jl, 60, f) = a.prainfrac_toprfz[i];
          ik = 70;
          const int kt = a.ktype[i], kp = b.kpbltype[i];
          const R eis = a.peis[i];
          if (kt == 0) PX(jl, ik, f) = PX(jl, ik, f) + R(1);
          if (kt == is + 1) PX(jl, ik + 1, f) = PX(jl, ik + 1, f) + R(1);  // IS+1 as in the reference
          if (kt == 2) PX(jl, ik + 2, f) = PX(jl, ik + 2, f) + R(1);
          if (kt == 3) PX(jl, ik + 3, f) = PX(jl, ik + 3, f) + R(1);
          if (kp == 0) PX(jl, ik + 4, f) = PX(jl, ik + 4, f) + R(1);
          if (kp == is + 1) PX(jl, ik + 5, f) = PX(jl, ik + 5, f) + R(1);
          if (kp == 2) PX(jl, ik + 6, f) = PX(jl, ik + 6, f) + R(1);
          if (kp == 3) PX(jl, ik + 7, f) = PX(jl, ik + 7, f) + R(1);
          PX(jl, ik + 8, f) = PX(jl, ik + 8, f) + eis;
          if (eis > R(6.0)) PX(jl, ik + 9, f) = PX(jl, ik + 9, f) + R(1);
          if (eis > R(8.0)) PX(jl, ik + 10, f) = PX(jl, ik + 10, f) + R(1);
          if (eis > R(10.0)) PX(jl, ik + 11, f) = PX(jl, ik + 11, f) + R(1);
          if (kt >= 2 && eis < R(p.reisthsc)) PX(jl, ik + 12, f) = PX(jl, ik + 12, f) + R(1);
          if (kt > 0) PX(jl, ik + 13, f) = R(b.kctop[i]);
        }
        is = is + 1;
      }
      if (lcldbud & CLOUDSC_BUD_C) {  // :3948-3966
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const size_t k = i2(jl, jk);
          PX(jl, jk, is + 4) = PX(jl, jk, is + 4) + BCC(jl, 2) * zdt;
          PX(jl, jk, is + 6) = PX(jl, jk, is + 6) + BCC(jl, 3) * zdt;
          PX(jl, jk, is + 7) = PX(jl, jk, is + 7) + (BCC(jl, 4) + BCC(jl, 6)) * zdt;
          PX(jl, jk, is + 8) = PX(jl, jk, is + 8) + BCC(jl, 5) * zdt;
          PX(jl, jk, is + 9) = PX(jl, jk, is + 9) + BCC(jl, 7) * zdt;
          PX(jl, jk, is + 11) = PX(jl, jk, is + 11) + b.pvfa[k] * zdt;
          PX(jl, jk, is + 12) = PX(jl, jk, is + 12) + b.pdyna[k] * zdt;
        }
        is = is + 12;
      }
      if (lcldbud & CLOUDSC_BUD_L) {  // :3971-4003
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const size_t k = i2(jl, jk);
          PX(jl, jk, is + 5) = PX(jl, jk, is + 5) + BL(jl, 2) * zdt;
          PX(jl, jk, is + 7) = PX(jl, jk, is + 7) + BL(jl, 3) * zdt;
          PX(jl, jk, is + 8) = PX(jl, jk, is + 8) + (BL(jl, 4) + BL(jl, 6)) * zdt;
          PX(jl, jk, is + 9) = PX(jl, jk, is + 9) + BL(jl, 5) * zdt;
          PX(jl, jk, is + 10) = PX(jl, jk, is + 10) + BL(jl, 7) * zdt;
          PX(jl, jk, is + 11) = PX(jl, jk, is + 11) + BL(jl, 11) * zdt;
          PX(jl, jk, is + 12) = PX(jl, jk, is + 12) + BL(jl, 12) * zdt;
          PX(jl, jk, is + 13) = PX(jl, jk, is + 13) + BL(jl, 13) * zdt;
          PX(jl, jk, is + 14) = PX(jl, jk, is + 14) + BL(jl, 14) * zdt;
          PX(jl, jk, is + 15) = PX(jl, jk, is + 15) + BL(jl, 15) * zdt;
          PX(jl, jk, is + 16) = PX(jl, jk, is + 16) + BL(jl, 17) * zdt;
          PX(jl, jk, is + 17) = PX(jl, jk, is + 17) + (BL(jl, 18) + BL(jl, 19)) * zdt;
          PX(jl, jk, is + 18) = PX(jl, jk, is + 18) + BL(jl, 20) * zdt;
          PX(jl, jk, is + 19) = PX(jl, jk, is + 19) + BL(jl, 21) * zdt;
          PX(jl, jk, is + 20) = PX(jl, jk, is + 20) + a.pvfl[k] * zdt;
          PX(jl, jk, is + 21) = PX(jl, jk, is + 21) + b.pdynl[k] * zdt;
          PX(jl, jk, is + 22) = PX(jl, jk, is + 22) + V(zqxn, jl, QL);
        }
        is = is + 22;
      }
      if (lcldbud & CLOUDSC_BUD_I) {  // :4008-4041
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const int i = jl - 1;
          const size_t k = i2(jl, jk);
          PX(jl, jk, is + 5) = PX(jl, jk, is + 5) + BI(jl, 2) * zdt;
          PX(jl, jk, is + 7) = PX(jl, jk, is + 7) + BI(jl, 3) * zdt;
          PX(jl, jk, is + 8) = PX(jl, jk, is + 8) + (BI(jl, 4) + BI(jl, 6)) * zdt;
          PX(jl, jk, is + 9) = PX(jl, jk, is + 9) + BI(jl, 5) * zdt;
          PX(jl, jk, is + 10) = PX(jl, jk, is + 10) + BI(jl, 7) * zdt;
          PX(jl, jk, is + 11) = PX(jl, jk, is + 11) + BI(jl, 11) * zdt;
          PX(jl, jk, is + 12) = PX(jl, jk, is + 12) + (BI(jl, 12) + BI(jl, 13)) * zdt;
          PX(jl, jk, is + 13) = PX(jl, jk, is + 13) + BI(jl, 14) * zdt;
          PX(jl, jk, is + 14) = PX(jl, jk, is + 14) + (BI(jl, 15) + BI(jl, 16)) * zdt;
          PX(jl, jk, is + 15) = PX(jl, jk, is + 15) + zpsdep[i] * ZQTMST * zdt;
          PX(jl, jk, is + 16) = PX(jl, jk, is + 16) + BI(jl, 17) * zdt;
          PX(jl, jk, is + 17) = PX(jl, jk, is + 17) + a.pvfi[k] * zdt;
          PX(jl, jk, is + 18) = PX(jl, jk, is + 18) + zpievap * ZQTMST * zdt;
          PX(jl, jk, is + 19) = PX(jl, jk, is + 19) + b.pdyni[k] * zdt;
          PX(jl, jk, is + 20) = PX(jl, jk, is + 20) + V(zqxn, jl, QI);
        }
        is = is + 20;
      }
      if (lcldbud & CLOUDSC_BUD_T) {  // :4052-4101
        for (int jl = kidia; jl <= kfdia; ++jl) {
          const size_t k = i2(jl, jk);
          PX(jl, jk, is + 1) = PX(jl, jk, is + 1) + b.phrsw[k];
          PX(jl, jk, is + 2) = PX(jl, jk, is + 2) + b.phrlw[k];
          PX(jl, jk, is + 3) = PX(jl, jk, is + 3) +
                               ((BL(jl, 1) + BL(jl, 2) + BL(jl, 9) + BL(jl, 10)) * RALVDCP - BI(jl, 11) * RALVDCP) * zdt;
          PX(jl, jk, is + 4) = PX(jl, jk, is + 4) +
                               (BI(jl, 1) + BI(jl, 2) + BI(jl, 9) + BI(jl, 10) + BI(jl, 11)) * RALSDCP * zdt;
          PX(jl, jk, is + 5) = PX(jl, jk, is + 5) + BI(jl, 11) * RALFDCP * zdt;
          PX(jl, jk, is + 6) = PX(jl, jk, is + 6) + (BL(jl, 7) + BL(jl, 8) + BL(jl, 5)) * RALVDCP * zdt;
          PX(jl, jk, is + 7) = PX(jl, jk, is + 7) + (BI(jl, 7) + BI(jl, 8) + BI(jl, 5)) * RALSDCP * zdt;
          PX(jl, jk, is + 8) = PX(jl, jk, is + 8) + BL(jl, 20) * RALVDCP * zdt;
          PX(jl, jk, is + 9) = PX(jl, jk, is + 9) + BI(jl, 17) * RALSDCP * zdt;
          PX(jl, jk, is + 10) = PX(jl, jk, is + 10) + BI(jl, 15) * RALFDCP * zdt;
          PX(jl, jk, is + 11) = PX(jl, jk, is + 11) + BI(jl, 16) * RALFDCP * zdt;
          PX(jl, jk, is + 12) = PX(jl, jk, is + 12) - (BL(jl, 12) + BL(jl, 13) + BL(jl, 18) + BL(jl, 19)) * RALFDCP * zdt;
          PX(jl, jk, is + 13) = PX(jl, jk, is + 13) + BL(jl, 21) * RALFDCP * zdt;
        }
      }
    }
    if (lfsd)  // for the ice FSD (:4112-4126)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const size_t k = i2(jl, jk);
        const R q0 = zqx0[i3(jl, jk, QL + 1)] + zqx0[i3(jl, jk, QI + 1)];
        if (q0 + a.plude[k] > RLMIN) {
          zratfsd[k] = a.plude[k] / (a.plude[k] + zqx0[i3(jl, jk, QL + 1)] + zqx0[i3(jl, jk, QI + 1)]);
          zratfsd[k] = std::max(R(0), std::min(R(1), zratfsd[k]));
        }
        zicef[k] = V(zqxn, jl, QI);
        zliqf[k] = V(zqxn, jl, QL);
      }
  }  // level loop

  // ---- 8. fluxes (:4218-4303) ----
  for (int jk = 1; jk <= klev + 1; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      a.pfplsl[i2(jl, jk)] = zpfplsx[ih3(jl, jk, QR + 1)] + zpfplsx[ih3(jl, jk, QL + 1)];
      a.pfplsn[i2(jl, jk)] = zpfplsx[ih3(jl, jk, QS + 1)] + zpfplsx[ih3(jl, jk, QI + 1)];
    }
  for (int jl = kidia; jl <= kfdia; ++jl) {
    const size_t k = i2(jl, 1);
    a.pfsqlf[k] = a.pfsqif[k] = a.pfsqrf[k] = a.pfsqsf[k] = R(0);
    a.pfcqlng[k] = a.pfcqnng[k] = a.pfcqrng[k] = a.pfcqsng[k] = R(0);
    a.pfsqltur[k] = a.pfsqitur[k] = R(0);
  }
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk), k1 = i2(jl, jk + 1);
      const R zgdph_r = -ZRG_R * (a.paph[k1] - a.paph[k]) * ZQTMST;
      a.pfsqlf[k1] = a.pfsqlf[k];
      a.pfsqif[k1] = a.pfsqif[k];
      a.pfsqrf[k1] = a.pfsqlf[k];
      a.pfsqsf[k1] = a.pfsqif[k];
      a.pfcqlng[k1] = a.pfcqlng[k];
      a.pfcqnng[k1] = a.pfcqnng[k];
      a.pfcqrng[k1] = a.pfcqlng[k];
      a.pfcqsng[k1] = a.pfcqnng[k];
      a.pfsqltur[k1] = a.pfsqltur[k];
      a.pfsqitur[k1] = a.pfsqitur[k];
      a.pfsqlf[k1] = a.pfsqlf[k1] + (zqxn2d[i3(jl, jk, QL + 1)] - zqx0[i3(jl, jk, QL + 1)] + a.pvfl[k] * ptsphy) * zgdph_r +
                     a.pludeli[i3(jl, jk, 1)];
      a.pfcqlng[k1] = a.pfcqlng[k1] + zlneg[i3(jl, jk, QL + 1)] * zgdph_r;
      a.pfsqltur[k1] = a.pfsqltur[k1] + a.pvfl[k] * ptsphy * zgdph_r;
      a.pfsqrf[k1] = a.pfsqrf[k1] + (zqxn2d[i3(jl, jk, QR + 1)] - zqx0[i3(jl, jk, QR + 1)]) * zgdph_r;
      a.pfcqrng[k1] = a.pfcqrng[k1] + zlneg[i3(jl, jk, QR + 1)] * zgdph_r;
      a.pfsqif[k1] = a.pfsqif[k1] + (zqxn2d[i3(jl, jk, QI + 1)] - zqx0[i3(jl, jk, QI + 1)] + a.pvfi[k] * ptsphy) * zgdph_r +
                     a.pludeli[i3(jl, jk, 2)];
      a.pfcqnng[k1] = a.pfcqnng[k1] + zlneg[i3(jl, jk, QI + 1)] * zgdph_r;
      a.pfsqitur[k1] = a.pfsqitur[k1] + a.pvfi[k] * ptsphy * zgdph_r;
      a.pfsqsf[k1] = a.pfsqsf[k1] + (zqxn2d[i3(jl, jk, QS + 1)] - zqx0[i3(jl, jk, QS + 1)]) * zgdph_r;
      a.pfcqsng[k1] = a.pfcqsng[k1] + zlneg[i3(jl, jk, QS + 1)] * zgdph_r;
    }
  for (int jk = 1; jk <= klev + 1; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      a.pfhpsl[i2(jl, jk)] = -RLVTT * a.pfplsl[i2(jl, jk)];
      a.pfhpsn[i2(jl, jk)] = -RLSTT * a.pfplsn[i2(jl, jk)];
    }

  // Ice FSD and PFSD (:4307-4359).  The literals .5, .001, .24, .11, 10., .03 and 0.00001 are default real in the
  // reference: FP32 values, widened.
  if (lfsd) {
    for (int jk = 1; jk <= klev; ++jk)
      for (int jl = kidia; jl <= kfdia; ++jl) {
        const int i = jl - 1;
        const size_t k = i2(jl, jk);
        const R zgridlen = R(2) * RA * std::sqrt(RPI * a.bud->pgaw[i]) * R(0.001);
        const R zifsdback = ZU1 * std::pow(zgridlen, R(1) / R(3)) * std::pow(std::pow(ZU2 * zgridlen, ZU3) + R(1), ZU4);
        const R zqtot = std::max(std::min((zqx[i3(jl, jk, QV + 1)] + zqx[i3(jl, jk, QI + 1)]) * R(1000.), R(30.)),
                                 R(0.00001f));
        const R zdp_k = a.paph[i2(jl, jk + 1)] - a.paph[k];
        const R zrho_k = a.pap[k] / (RD_ * ztp1[k]);
        R zdelz = R(1) / (RG * zrho_k) * zdp_k * R(0.001);
        zdelz = std::max(std::min(zdelz, R(.5f)), R(.001f));
        const R zphi = std::pow(zdelz / R(.24f), R(.11f)) * std::pow(zqtot / R(10.f), R(.03f));
        if (zanewp[k] > R(0.95)) {
          zifsd[k] = zifsdback * zphi;
          zifsd[k] = ZR12 * (zifsd[k] + zratfsd[k] * R(1.5));
        } else if (zanewp[k] > R(0.001) && zanewp[k] <= R(0.95)) {
          zifsd[k] = zifsdback * zphi * R(1.5);
          zifsd[k] = ZR12 * (zifsd[k] + zratfsd[k] * R(1.5));
        }
        R zfsd = RCLOUD_FRAC_STD;
        if (zicef[k] > ZEPSEC) zfsd = zifsd[k];
        if (zliqf[k] > R(0) && zicef[k] <= ZEPSEC) zfsd = zlfsd[k];
        a.bud->pfsd[k] = std::min(std::max(R(0.1), zfsd), R(3.575));
      }
  }
}

}  // namespace oracle
