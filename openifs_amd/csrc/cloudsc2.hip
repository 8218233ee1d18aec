// CLOUDSC2 for gfx950: the simplified prognostic cloud scheme of the linearised physics
// (ifs-source/arpifs/phys_ec/cloudsc2.F90:10-709) with its final saturation adjustment (CUADJTQ / CUADJTQS,
// KCALL = 0).  One column per lane: the precipitation fluxes and the precipitation fraction are carried from level
// to level in registers, like CLOUDSC's sweep.  C-ABI in include/cloudsc2.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "../../include/cloudsc2.h"
#include "capi_util.h"
#include "cloudsc_device.h"

namespace cloudsc2_gpu {

using cloudsc_capi::fail;
using cloudsc_gpu::m_div;
using cloudsc_gpu::m_exp;
using cloudsc_gpu::m_max;
using cloudsc_gpu::m_min;
using cloudsc_gpu::m_pow;
using cloudsc_gpu::m_sqrt;

using cloudsc_gpu::LD;
using cloudsc_gpu::ST;

// Constants of one launch in the storage precision (cloudsc2.F90:243-253 and the TCST/TTHF/TECLDP/TEPHLI values).
template <typename R>
struct C2Const {
  R rg, rd, rcpd, retv, rtt, rlvtt, rlstt, rlmlt, r2es, r3les, r3ies, r4les, r4ies, r5les, r5ies, r5alvcp, r5alscp;
  R ralvdcp, ralsdcp, rtwat, rtice, rtwat_rtice_r, rlmin, rpecons, rclcrit, rlptrc, rvtmp2;
  R ptsphy, zckcodtl, zckcodti, zcons2, zcons3, zmeltp2, zqtmst;
  int32_t lin;      // LPHYLIN or LDRAIN1D: tanh phase function and CUADJTQS
  int32_t evap;     // LEVAPLS2 or LDRAIN1D: precipitation evaporation, the 1.9 RCLCRIT / 1e-4 thresholds
  int32_t phylin;   // LPHYLIN alone: SATUR's form of the saturation humidity (CLOUD_S_LAYER)
};

template <typename R>
struct C2Ptrs {
  const R *ceta, *paphp1, *papp1, *pqm1, *pqs, *ptm1, *pl, *pi, *plude, *plu, *pmfu, *pmfd;
  R* ptent;
  const R* pgtent;
  R* ptenq;
  const R* pgtenq;
  R* ptenl;
  const R* pgtenl;
  R* pteni;
  const R* pgteni;
  R *pclc, *pfplsl, *pfplsn, *pfhpsl, *pfhpsn, *pcovptot;
  // CLOUD_S_LAYER only (LAYER): the cloud fraction and the wrap-up outputs, NULL where the caller has no such field
  const R* pa;
  R *ptena, *ptenr, *ptens, *pqsat, *pprectype, *pfzra;
  R *pqiwp, *pqlwp;  // CLOUDST: diagnostic ice and liquid water (AUXL%ZPLSMP, ZPLTMP in CLOUD_S_LAYER)
};

struct C2Dims {
  int ngptot, nproma, klev, ngpblks, ktdia, kidia, kfdia;  // 1-based KIDIA/KFDIA bound the columns of each block
};

// fcttre.func.h statement functions on the mixed phase (FOEALFA, FOEEWM, FOEDEM, FOELDCPM)
template <typename R>
__device__ __forceinline__ R foealfa(const C2Const<R>& c, R t) {
  const R x = (m_max(c.rtice, m_min(c.rtwat, t)) - c.rtice) * c.rtwat_rtice_r;
  return m_min(R(1), x * x);
}
template <typename R>
__device__ __forceinline__ R foeewm(const C2Const<R>& c, R t, R a) {
  return c.r2es * (a * m_exp(m_div(c.r3les * (t - c.rtt), t - c.r4les)) +
                   (R(1) - a) * m_exp(m_div(c.r3ies * (t - c.rtt), t - c.r4ies)));
}

// CUADJTQ, KCALL = 0 (cuadjtq.F90:256-272), or CUADJTQS, KCALL = 0 (cuadjtqs.F90:123-134, 217-250): two
// Newton steps of the isobaric saturation adjustment of (pt, pq) at pressure psp.
template <typename R>
__device__ __forceinline__ void adjust(const C2Const<R>& c, R psp, R& pt, R& pq) {
  const R zqp = m_div(R(1), psp);
  if (!c.lin) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const R a = foealfa(c, pt);
      R zqsat = foeewm(c, pt, a) * zqp;
      zqsat = m_min(R(0.5), zqsat);
      const R zcor = m_div(R(1), R(1) - c.retv * zqsat);
      zqsat = zqsat * zcor;
      const R dl = pt - c.r4les, di = pt - c.r4ies;
      const R foedem = a * c.r5alvcp * m_div(R(1), dl * dl) + (R(1) - a) * c.r5alscp * m_div(R(1), di * di);
      const R zcond1 = m_div(pq - zqsat, R(1) + zqsat * zcor * foedem);
      pt = pt + (a * c.ralvdcp + (R(1) - a) * c.ralsdcp) * zcond1;
      pq = pq - zcond1;
    }
  } else {
    const bool liq = pt > c.rtt;  // phase from the entry temperature, kept for both steps
    const R z3es = liq ? c.r3les : c.r3ies, z4es = liq ? c.r4les : c.r4ies;
    const R z5alcp = liq ? c.r5alvcp : c.r5alscp, zaldcp = liq ? c.ralvdcp : c.ralsdcp;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const R ztarg = pt;
      R zqsat = zqp * (c.r2es * m_exp(m_div(z3es * (ztarg - c.rtt), ztarg - z4es)));
      if (zqsat > R(0.5)) zqsat = R(0.5);
      const R zcor = m_div(R(1), R(1) - c.retv * zqsat);
      zqsat = zqsat * zcor;
      const R d = ztarg - z4es;
      const R z2s = m_div(z5alcp, d * d);
      const R zcond1 = m_div(pq - zqsat, R(1) + zqsat * zcor * z2s);
      pt = pt + zaldcp * zcond1;
      pq = pq - zcond1;
    }
  }
}

// Work-group size: 64, 128 and 256 run within 1 % of each other at 655,360 x L137 in FP32.
constexpr int CLOUDSC2_BLOCK = 128;

// SATUR with KFLAG = 2 (satur.F90:110-141): the saturation specific humidity of the mixed phase at (t, p), in the
// LPHYLIN form (separate water and ice pressures weighted by FOEALFA) or as FOEEWM.
template <typename R>
__device__ __forceinline__ R satur(const C2Const<R>& c, R t, R paprsf) {
  const R a = foealfa(c, t);
  if (c.phylin) {
    const R el = c.r2es * m_exp(m_div(c.r3les * (t - c.rtt), t - c.r4les));
    const R ei = c.r2es * m_exp(m_div(c.r3ies * (t - c.rtt), t - c.r4ies));
    R qs = m_div(a * el + (R(1.0) - a) * ei, paprsf);
    if (qs > R(0.5)) qs = R(0.5);
    return qs * m_div(R(1.0), R(1.0) - c.retv * qs);
  }
  const R qs = m_min(R(0.5), m_div(foeewm(c, t, a), paprsf));
  return qs * m_div(R(1.0), R(1.0) - c.retv * qs);
}

// UNI: NPROMA is a multiple of 64, so every wave lies in one block and the block's offsets are wave-uniform.
// Arrays are addressed as a scalar base plus a 32-bit byte offset (LD/ST); the launcher keeps every array of one
// launch below 4 GiB.
// LAYER: CLOUD_S_LAYER around CLOUDSC2 (cloud_s_layer.F90:112-176): PQS from SATUR at T + TSPHY*TENDENCY%T in the
// sweep instead of an input, the cloud-fraction tendency (PCLC - PA)/TSPHY, zero rain and snow tendencies and
// zero PPRECTYPE/PFZRA.
template <typename R, bool UNI, bool LAYER>
__global__ __launch_bounds__(CLOUDSC2_BLOCK) void cloudsc2_kernel(const C2Const<R> c, const C2Ptrs<R> p,
                                                                  const C2Dims d) {
  const uint32_t g = blockIdx.x * CLOUDSC2_BLOCK + threadIdx.x;
  const uint32_t nproma = (uint32_t)d.nproma;
  const int klev = d.klev;
  if (g >= (uint32_t)d.ngpblks * nproma) return;
  uint32_t b = g / nproma;
  if (UNI) b = __builtin_amdgcn_readfirstlane(b);
  const uint32_t jl = g - b * nproma;
  constexpr uint32_t E = sizeof(R);
  const uint32_t jo = jl * E;                          // the lane's byte offset within a level
  const uint32_t lstep = nproma * E;                   // one level
  const uint32_t b2 = b * nproma * (uint32_t)klev * E;  // level 0 of a (NPROMA,KLEV) array
  const uint32_t bh = b2 + b * lstep;                  // level 0 of a (NPROMA,KLEV+1) array
  auto at2 = [&](int k) { return b2 + (uint32_t)k * lstep; };
  auto ath = [&](int k) { return bh + (uint32_t)k * lstep; };
  if constexpr (LAYER) {  // TENDENCY_LOC%CLD(:,:,NCLDQR/NCLDQS) = 0 on every column of the block (:164-165)
    for (int k = 0; k < klev; ++k) {
      if (p.ptenr) ST(p.ptenr, at2(k), jo, R(0));
      if (p.ptens) ST(p.ptens, at2(k), jo, R(0));
    }
  }
  const int kfdia = std::min(d.kfdia, d.ngptot - (int)(b * nproma));
  if ((int)jl < d.kidia - 1 || (int)jl >= kfdia) return;
  const R ptsphy = c.ptsphy;
  if constexpr (LAYER) {  // (:168-169)
    if (p.pprectype) ST(p.pprectype, b * lstep, jo, R(0));
    if (p.pfzra) ST(p.pfzra, b * lstep, jo, R(0));
  }
  const R zeps1 = R(1.e-12), zeps2 = R(1.e-10), zqmax = R(0.5);

  // eta of the tropopause (:323-334): the last level with 0.1 < eta < 0.4 and T above the level below
  R ztrpaus = R(0.1);
  for (int k = 0; k < klev - 1; ++k) {
    const R ceta = p.ceta[k];
    if (ceta > R(0.1) && ceta < R(0.4)) {
      const R t0 = LD(p.ptm1, at2(k), jo) + ptsphy * LD(p.pgtent, at2(k), jo);
      const R t1 = LD(p.ptm1, at2(k + 1), jo) + ptsphy * LD(p.pgtent, at2(k + 1), jo);
      if (t0 > t1) ztrpaus = ceta;
    }
  }
  // outputs the level loop does not reach (:297-320): PCLC, PCOVPTOT zero, PFPLSL/PFPLSN zero at the top
  for (int k = 0; k < d.ktdia - 1; ++k) {
    if (!LAYER || p.pclc) ST(p.pclc, at2(k), jo, R(0));
    ST(p.pcovptot, at2(k), jo, R(0));
    if constexpr (LAYER) {  // (PCLC - PA)/TSPHY with PCLC = 0 above KTDIA (:155-158)
      if (p.ptena) ST(p.ptena, at2(k), jo, c.zqtmst * (R(0) - LD(p.pa, at2(k), jo)));
    }
  }
  ST(p.pfplsl, ath(0), jo, R(0));
  ST(p.pfplsn, ath(0), jo, R(0));
  ST(p.pfhpsl, ath(0), jo, -R(0) * c.rlvtt);  // (:694-699), as -PFPLSL*RLVTT
  ST(p.pfhpsn, ath(0), jo, -R(0) * c.rlstt);
  for (int k = 1; k < d.ktdia; ++k) {  // half levels the loop leaves as they are
    ST(p.pfhpsl, ath(k), jo, -LD(p.pfplsl, ath(k), jo) * c.rlvtt);
    ST(p.pfhpsn, ath(k), jo, -LD(p.pfplsn, ath(k), jo) * c.rlstt);
  }
  const R paph_surf = LD(p.paphp1, ath(klev), jo);

  // The inputs of one level, loaded one level ahead of the sweep so that the loads of level k+1 are in flight
  // while level k computes.
  struct Level {
    R tm1, gtent, qm1, gtenq, l, gtenl, i, gteni, paph1, qs, app1, lude, lu_next, mfu, mfd, a;
  };
  auto load = [&](int k) {
    const uint32_t j = at2(k);
    Level v;
    v.tm1 = LD(p.ptm1, j, jo); v.gtent = LD(p.pgtent, j, jo); v.qm1 = LD(p.pqm1, j, jo);
    v.gtenq = LD(p.pgtenq, j, jo); v.l = LD(p.pl, j, jo); v.gtenl = LD(p.pgtenl, j, jo); v.i = LD(p.pi, j, jo);
    v.gteni = LD(p.pgteni, j, jo); v.paph1 = LD(p.paphp1, ath(k + 1), jo);
    v.qs = LAYER ? R(0) : LD(p.pqs, j, jo);
    v.a = (LAYER && p.ptena) ? LD(p.pa, j, jo) : R(0);
    v.app1 = LD(p.papp1, j, jo); v.lude = LD(p.plude, j, jo);
    v.lu_next = k < klev - 1 ? LD(p.plu, j + lstep, jo) : R(0);
    v.mfu = LD(p.pmfu, j, jo); v.mfd = LD(p.pmfd, j, jo);
    return v;
  };

  R zrfl = R(0), zsfl = R(0), zcovptot = R(0);
  R paph0 = LD(p.paphp1, ath(d.ktdia - 1), jo);
  Level nxt = load(d.ktdia - 1);
  for (int k = d.ktdia - 1; k < klev; ++k) {
    const Level cur = nxt;
    if (k + 1 < klev) nxt = load(k + 1);
    const uint32_t i2 = at2(k);
    const R ceta = p.ceta[k];
    const R zscalm = R(0.9) * m_pow(m_max(ceta - R(0.2), zeps1), R(0.2));  // (:276)
    R ztp1 = cur.tm1 + ptsphy * cur.gtent;
    R zqp1 = cur.qm1 + ptsphy * cur.gtenq;
    const R zl = cur.l + ptsphy * cur.gtenl;
    const R zi = cur.i + ptsphy * cur.gteni;
    const R paph1 = cur.paph1;
    const R zdp = paph1 - paph0;
    const R zzz = m_div(R(1), c.rcpd + c.rcpd * c.rvtmp2 * zqp1);
    const R zlfdcp = c.rlmlt * zzz, zlsdcp = c.rlstt * zzz, zlvdcp = c.rlvtt * zzz;
    const R papp1 = cur.app1, plude = cur.lude;
    // CLOUD_S_LAYER's SATUR at the first-guess temperature (cloud_s_layer.F90:112-118), else the PQS input
    const R pqs = LAYER ? satur(c, ztp1, papp1) : cur.qs;
    if constexpr (LAYER) {
      if (p.pqsat) ST(p.pqsat, i2, jo, pqs);
    }

    // 3.1 (:352-416)
    R zfwat, zfoeew, zesdp;
    if (c.lin) {
      const R zoealfaw = R(0.545) * (tanh(R(0.17) * (ztp1 - c.rlptrc)) + R(1.0));
      const bool cold = ztp1 < c.rtt;
      zfwat = cold ? zoealfaw : R(1.0);
      const R z3es = cold ? c.r3ies : c.r3les, z4es = cold ? c.r4ies : c.r4les;
      zfoeew = c.r2es * m_exp(m_div(z3es * (ztp1 - c.rtt), ztp1 - z4es));
      zesdp = m_div(zfoeew, papp1);
      if (zesdp > zqmax) zesdp = zqmax;
    } else {
      zfwat = foealfa(c, ztp1);
      zfoeew = foeewm(c, ztp1, zfwat);
      zesdp = m_div(zfoeew, papp1);
    }
    const R dl = ztp1 - c.r4les, di = ztp1 - c.r4ies;
    const R zfac = zfwat * m_div(c.r5les, dl * dl) + (R(1.0) - zfwat) * m_div(c.r5ies, di * di);
    const R zcor = m_div(R(1.0), R(1.0) - c.retv * zesdp);
    const R zdqsdtemp = zfac * zcor * pqs;
    const R zcorqs = R(1.0) + c.zcons3 * zdqsdtemp;
    const R zqlim = zqp1 > pqs ? pqs : zqp1;
    // critical relative humidity (:386-407)
    const R zeta3 = ztrpaus;
    const R sr = m_div(zeta3 - R(0.25), R(0.15));
    const R zrh2 = R(0.35) + R(0.14) * (sr * sr) + m_div(R(0.04) * m_min(zeta3 - R(0.25), R(0.0)), R(0.15));
    const R zdeta2 = R(0.3);
    const R zdeta1 = R(0.09) + m_div(R(0.16) * (R(0.4) - zeta3), R(0.3));
    R zcrh2 = R(0);
    if (ceta < zeta3) {
      zcrh2 = R(1.0);
    } else if (ceta >= zeta3 && ceta < (zeta3 + zdeta2)) {
      zcrh2 = R(1.0) + (zrh2 - R(1.0)) * m_div(ceta - zeta3, zdeta2);
    } else if (ceta >= (zeta3 + zdeta2) && ceta < (R(1.0) - zdeta1)) {
      zcrh2 = zrh2;
    } else if (ceta >= (R(1.0) - zdeta1)) {
      zcrh2 = R(1.0) + (zrh2 - R(1.0)) * m_sqrt(m_div(R(1.0) - ceta, zdeta1));
    }
    const R zsupsat = ztp1 < c.rtice ? R(1.8) - R(3.E-03) * ztp1 : R(1.0);
    const R zqsat = pqs * zsupsat;
    const R zqcrit = zcrh2 * zqsat;

    // uniform total-water distribution (:419-436)
    R pclc, zqc;
    const R zqt = zqp1 + zl + zi;
    if (zqt <= zqcrit) {
      pclc = R(0.0);
      zqc = R(0.0);
    } else if (zqt >= zqsat) {
      pclc = R(1.0);
      zqc = (R(1.0) - zscalm) * (zqsat - zqcrit);
    } else {
      const R zqpd = zqsat - zqt, zqcd = zqsat - zqcrit;
      pclc = R(1.0) - m_sqrt(m_div(zqpd, zqcd - zscalm * (zqt - zqcrit)));
      zqc = (zscalm * zqpd + (R(1.0) - zscalm) * zqcd) * (pclc * pclc);
    }
    // convective component (:440-453)
    const R zgdp = m_div(c.rg, paph1 - paph0);
    const R zlude = plude * ptsphy * zgdp;
    if (k < klev - 1) {
      const R plu_n = cur.lu_next;
      if (zlude >= c.rlmin && plu_n >= zeps2) {
        pclc = pclc + (R(1.0) - pclc) * (R(1.0) - m_exp(-m_div(zlude, plu_n)));
        zqc = zqc + zlude;
      }
    }
    // compensating subsidence (:457-471)
    {
      const R zfac1 = m_div(R(1.0), c.rd * ztp1);
      const R zrho = papp1 * zfac1;
      const R zfac2 = m_div(R(1.0), papp1 - c.retv * zfoeew);
      const R zrodqsdp = -zrho * pqs * zfac2;
      const R zldcp = zfwat * zlvdcp + (R(1.0) - zfwat) * zlsdcp;
      const R zfac3 = m_div(R(1.0), R(1.0) + zldcp * zdqsdtemp);
      const R dtdzmo = c.rg * (m_div(R(1.0), c.rcpd) - zldcp * zrodqsdp) * zfac3;
      const R zdqsdz = zdqsdtemp * dtdzmo - c.rg * zrodqsdp;
      const R zfac4 = m_div(R(1.0), zrho);
      const R zdqc = m_min(zdqsdz * (cur.mfu + cur.mfd) * ptsphy * zfac4, zqc);
      zqc = zqc - zdqc;
    }
    // new cloud contents and condensation (:475-481)
    R zqlwc = zqc * zfwat, zqiwc = zqc * (R(1.0) - zfwat);
    R zcondl = (zqlwc - zl) * c.zqtmst, zcondi = (zqiwc - zi) * c.zqtmst;
    // maximum overlap of precipitation (:487-493)
    if (pclc > zcovptot) zcovptot = pclc;
    const R zcovpclr = m_max(zcovptot - pclc, R(0.0));
    // melting of incoming snow (:499-511)
    R zrfln, zsfln;
    if (zsfl != R(0.0)) {
      const R zcons = m_div(c.zcons2 * zdp, zlfdcp);
      const R zsnmlt = m_min(zsfl, zcons * m_max(R(0.0), ztp1 - c.zmeltp2));
      zrfln = zrfl + zsnmlt;
      zsfln = zsfl - zsnmlt;
      ztp1 = ztp1 - m_div(zsnmlt, zcons);
    } else {
      zrfln = zrfl;
      zsfln = zsfl;
    }
    // rain and snow production (:513-545)
    R zprr = R(0), zprs = R(0);
    if (pclc > zeps2) {
      const R zlcrit = c.evap ? R(1.9) * c.rclcrit : c.rclcrit * R(2.);
      const R zcldl = m_div(zqlwc, pclc);
      const R r = m_div(zcldl, zlcrit);
      const R zd = c.zckcodtl * (R(1.0) - m_exp(-(r * r)));
      const R zlnew = pclc * zcldl * m_exp(-zd);
      zprr = zqlwc - zlnew;
      zqlwc = zqlwc - zprr;
      const R zlcriti = c.evap ? R(1.E-04) : c.rclcrit * R(2.);
      const R zcldi = m_div(zqiwc, pclc);
      const R ri = m_div(zcldi, zlcriti);
      const R zdi = c.zckcodti * m_exp(R(0.025) * (ztp1 - c.rtt)) * (R(1.0) - m_exp(-(ri * ri)));
      const R zinew = pclc * zcldi * m_exp(-zdi);
      zprs = zqiwc - zinew;
      zqiwc = zqiwc - zprs;
    }
    const R zdr = c.zcons2 * zdp * (zprr + zprs);
    R zrfreeze = R(0), zfwatr;
    if (ztp1 < c.rtt) {
      zrfreeze = c.zcons2 * zdp * zprr;
      zfwatr = R(0.0);
    } else {
      zfwatr = R(1.0);
    }
    zrfln = zrfln + zfwatr * zdr;
    zsfln = zsfln + (R(1.0) - zfwatr) * zdr;
    // precipitation evaporation (:568-597)
    R zevapr = R(0), zevaps = R(0), pcovptot = R(0);
    const R zprtot = zrfln + zsfln;
    if (zprtot > zeps2 && zcovpclr > zeps2 && c.evap) {
      R zpreclr = m_div(zprtot * zcovpclr, zcovptot);
      const R om = R(1.0) - pclc;
      const R zqe = pqs - m_div((pqs - zqlim) * zcovpclr, om * om);
      const R zbeta = c.rg * c.rpecons *
                      m_pow(m_div(m_div(m_sqrt(m_div(papp1, paph_surf)), R(5.09E-3)) * zpreclr, zcovpclr), R(0.5777));
      const R zb = m_div(ptsphy * zbeta * (pqs - zqe), R(1.0) + zbeta * ptsphy * zcorqs);
      const R zdtgdp = m_div(ptsphy * c.rg, paph1 - paph0);
      R zdpr = m_div(zcovpclr * zb, zdtgdp);
      zdpr = m_min(zdpr, zpreclr);
      zpreclr = zpreclr - zdpr;
      if (zpreclr <= R(0.0)) zcovptot = pclc;
      pcovptot = zcovptot;
      zevapr = m_div(zdpr * zrfln, zprtot);
      zrfln = zrfln - zevapr;
      zevaps = m_div(zdpr * zsfln, zprtot);
      zsfln = zsfln - zevaps;
    }
    // tendencies, first guess and the final adjustment (:608-665)
    auto tendency_t = [&](R condl, R condi, R rfreeze) {
      return zlvdcp * condl + zlsdcp * condi -
             (zlvdcp * zevapr + zlsdcp * zevaps + plude * (zfwat * zlvdcp + (R(1.0) - zfwat) * zlsdcp) -
              (zlsdcp - zlvdcp) * rfreeze) *
                 zgdp;
    };
    {
      const R zdqdt = -(zcondl + zcondi) + (plude + zevapr + zevaps) * zgdp;
      const R zdtdt = tendency_t(zcondl, zcondi, zrfreeze);
      ztp1 = ztp1 + ptsphy * zdtdt;
      zqp1 = zqp1 + ptsphy * zdqdt;
    }
    const R zqold = zqp1;
    adjust(c, papp1, ztp1, zqp1);
    const R zdq = m_max(R(0.0), zqold - zqp1);
    const R zdr2 = c.zcons2 * zdp * zdq;
    R zrfreeze2;
    if (ztp1 < c.rtt) {
      zrfreeze2 = zfwat * zdr2;
      zfwatr = R(0.0);
    } else {
      zrfreeze2 = R(0.0);
      zfwatr = R(1.0);
    }
    zcondl = zcondl + zfwatr * zdq * c.zqtmst;
    zcondi = zcondi + (R(1.0) - zfwatr) * zdq * c.zqtmst;
    zrfln = zrfln + zfwatr * zdr2;
    zsfln = zsfln + (R(1.0) - zfwatr) * zdr2;
    zrfreeze = zrfreeze + zrfreeze2;

    ST(p.ptenq, i2, jo, -(zcondl + zcondi) + (plude + zevapr + zevaps) * zgdp);
    ST(p.ptent, i2, jo, tendency_t(zcondl, zcondi, zrfreeze));
    ST(p.ptenl, i2, jo, (zqlwc - zl) * c.zqtmst);
    ST(p.pteni, i2, jo, (zqiwc - zi) * c.zqtmst);
    if (!LAYER || p.pclc) ST(p.pclc, i2, jo, pclc);
    if constexpr (LAYER) {  // (:155-158)
      if (p.ptena) ST(p.ptena, i2, jo, c.zqtmst * (pclc - cur.a));
    }
    ST(p.pcovptot, i2, jo, pcovptot);
    const uint32_t ih = ath(k + 1);
    ST(p.pfplsl, ih, jo, zrfln);
    ST(p.pfplsn, ih, jo, zsfln);
    ST(p.pfhpsl, ih, jo, -zrfln * c.rlvtt);
    ST(p.pfhpsn, ih, jo, -zsfln * c.rlstt);
    zrfl = zrfln;
    zsfl = zsfln;
    paph0 = paph1;
  }
}

// CLOUDST (cloudst.F90:9-641), the diagnostic scheme: cover from the vapour alone with its own critical humidity
// profile, no subsidence term, liquid autoconversion only, tendencies from the precipitation flux divergence and the
// diagnostic water paths PQLWP/PQIWP.  One column per lane, as cloudsc2_kernel.  LAYER: CLOUD_S_LAYER with
// LEPCLD2 false (cloud_s_layer.F90:112-176): PQS from SATUR, and the A/L/I tendencies (ZPATMP/ZPLTMP/ZPLSMP minus
// the state)/TSPHY, zero rain and snow tendencies and zero PPRECTYPE/PFZRA.
template <typename R, bool UNI, bool LAYER>
__global__ __launch_bounds__(CLOUDSC2_BLOCK) void cloudst_kernel(const C2Const<R> c, const C2Ptrs<R> p,
                                                                 const C2Dims d) {
  const uint32_t g = blockIdx.x * CLOUDSC2_BLOCK + threadIdx.x;
  const uint32_t nproma = (uint32_t)d.nproma;
  const int klev = d.klev;
  if (g >= (uint32_t)d.ngpblks * nproma) return;
  uint32_t b = g / nproma;
  if (UNI) b = __builtin_amdgcn_readfirstlane(b);
  const uint32_t jl = g - b * nproma;
  constexpr uint32_t E = sizeof(R);
  const uint32_t jo = jl * E;
  const uint32_t lstep = nproma * E;
  const uint32_t b2 = b * nproma * (uint32_t)klev * E;
  const uint32_t bh = b2 + b * lstep;
  auto at2 = [&](int k) { return b2 + (uint32_t)k * lstep; };
  auto ath = [&](int k) { return bh + (uint32_t)k * lstep; };
  if constexpr (LAYER) {  // (:164-165), every column of the block
    for (int k = 0; k < klev; ++k) {
      if (p.ptenr) ST(p.ptenr, at2(k), jo, R(0));
      if (p.ptens) ST(p.ptens, at2(k), jo, R(0));
    }
  }
  const int kfdia = std::min(d.kfdia, d.ngptot - (int)(b * nproma));
  if ((int)jl < d.kidia - 1 || (int)jl >= kfdia) return;
  const R ptsphy = c.ptsphy;
  const R zeps1 = R(1.e-12), zeps2 = R(1.e-10), zqmax = R(0.5);
  const R zcons3 = R(2.5E6) / R(1005.);  // (:215)
  const R zpecons = R(1.E-5);           // (:217)
  if constexpr (LAYER) {
    if (p.pprectype) ST(p.pprectype, b * lstep, jo, R(0));
    if (p.pfzra) ST(p.pfzra, b * lstep, jo, R(0));
  }
  // eta of the tropopause (:234-247)
  R ztrpaus = R(0.1);
  for (int k = 0; k < klev - 1; ++k) {
    const R ceta = p.ceta[k];
    if (ceta > R(0.1) && ceta < R(0.4)) {
      const R t0 = LD(p.ptm1, at2(k), jo) + ptsphy * LD(p.pgtent, at2(k), jo);
      const R t1 = LD(p.ptm1, at2(k + 1), jo) + ptsphy * LD(p.pgtent, at2(k + 1), jo);
      if (t0 > t1) ztrpaus = ceta;
    }
  }
  // outputs above KTDIA (:311-330): cover, water paths and PCOVPTOT zero; fluxes zero at the top
  // CLOUD_S_LAYER's tendencies from ZPATMP/ZPLTMP/ZPLSMP and the state PA/PL/PI of level k
  auto wrap_state = [&](int k, R clc, R lwp, R iwp, R a, R l, R i) {
    if constexpr (LAYER) {
      const uint32_t j = at2(k);
      if (p.ptena) ST(p.ptena, j, jo, c.zqtmst * (clc - a));
      ST(p.ptenl, j, jo, c.zqtmst * (lwp - l));
      ST(p.pteni, j, jo, c.zqtmst * (iwp - i));
    }
  };
  for (int k = 0; k < d.ktdia - 1; ++k) {
    if (!LAYER || p.pclc) ST(p.pclc, at2(k), jo, R(0));
    if (!LAYER || p.pqiwp) ST(p.pqiwp, at2(k), jo, R(0));
    if (!LAYER || p.pqlwp) ST(p.pqlwp, at2(k), jo, R(0));
    ST(p.pcovptot, at2(k), jo, R(0));
    if constexpr (LAYER)
      wrap_state(k, R(0), R(0), R(0), p.ptena ? LD(p.pa, at2(k), jo) : R(0), LD(p.pl, at2(k), jo),
                 LD(p.pi, at2(k), jo));
  }
  ST(p.pfplsl, ath(0), jo, R(0));
  ST(p.pfplsn, ath(0), jo, R(0));
  ST(p.pfhpsl, ath(0), jo, -R(0) * c.rlvtt);  // (:629-634)
  ST(p.pfhpsn, ath(0), jo, -R(0) * c.rlstt);
  for (int k = 1; k < d.ktdia; ++k) {
    ST(p.pfhpsl, ath(k), jo, -LD(p.pfplsl, ath(k), jo) * c.rlvtt);
    ST(p.pfhpsn, ath(k), jo, -LD(p.pfplsn, ath(k), jo) * c.rlstt);
  }
  const R paph_surf = LD(p.paphp1, ath(klev), jo);

  // the inputs of level k+1 are loaded while level k computes
  struct Level {
    R tm1, gtent, qm1, gtenq, paph1, app1, lude, qs, lu_next, a, l, i;
  };
  auto load = [&](int k) {
    const uint32_t j = at2(k);
    Level v;
    v.tm1 = LD(p.ptm1, j, jo); v.gtent = LD(p.pgtent, j, jo); v.qm1 = LD(p.pqm1, j, jo);
    v.gtenq = LD(p.pgtenq, j, jo); v.paph1 = LD(p.paphp1, ath(k + 1), jo); v.app1 = LD(p.papp1, j, jo);
    v.lude = LD(p.plude, j, jo);
    v.qs = LAYER ? R(0) : LD(p.pqs, j, jo);
    v.lu_next = k < klev - 1 ? LD(p.plu, j + lstep, jo) : R(0);
    v.a = (LAYER && p.ptena) ? LD(p.pa, j, jo) : R(0);
    v.l = LAYER ? LD(p.pl, j, jo) : R(0);
    v.i = LAYER ? LD(p.pi, j, jo) : R(0);
    return v;
  };
  R zrfl = R(0), zsfl = R(0), zcovptot = R(0);
  R paph0 = LD(p.paphp1, ath(d.ktdia - 1), jo);
  Level nxt = load(d.ktdia - 1);
  for (int k = d.ktdia - 1; k < klev; ++k) {
    const Level cur = nxt;
    if (k + 1 < klev) nxt = load(k + 1);
    const uint32_t i2 = at2(k);
    const R ceta = p.ceta[k];
    R ztp1 = cur.tm1 + ptsphy * cur.gtent;
    R zqp1 = cur.qm1 + ptsphy * cur.gtenq;
    const R paph1 = cur.paph1;
    const R zdp = paph1 - paph0;
    const R papp1 = cur.app1, plude = cur.lude;
    const R pqs = LAYER ? satur(c, ztp1, papp1) : cur.qs;
    if constexpr (LAYER) {
      if (p.pqsat) ST(p.pqsat, i2, jo, pqs);
    }
    // critical relative humidity (:263-290)
    R zcrh = R(0);
    {
      const R zeta3 = ztrpaus, zrh1 = R(0.9), zrh2 = R(0.5), zrh3 = R(1.0), zeta1 = R(0.75);
      const R zdeta1 = R(1.0) - zeta1;
      if (ceta < zeta3) {
        zcrh = zrh3 + m_div((zrh2 - zrh3) * ceta, zeta3);
      } else if (ceta >= zeta3 && ceta < zeta1) {
        zcrh = zrh2;
      } else if (ceta >= zeta1) {
        zcrh = zrh1 + (zrh2 - zrh1) * m_sqrt(m_div(R(1.0) - ceta, zdeta1));
      }
    }
    const R zscalm = R(0.9) * m_pow(m_max(ceta - R(0.2), zeps1), R(0.2));
    const R zzz = m_div(R(1), c.rcpd + c.rcpd * c.rvtmp2 * zqp1);
    const R zlfdcp = c.rlmlt * zzz, zlsdcp = c.rlstt * zzz, zlvdcp = c.rlvtt * zzz;
    // (:342-385)
    R zalfaw, zesdp;
    if (c.lin) {
      const R zoealfaw = R(0.545) * (tanh(R(0.17) * (ztp1 - c.rlptrc)) + R(1.0));
      const bool cold = ztp1 < c.rtt;
      zalfaw = cold ? zoealfaw : R(1.0);
      const R z3es = cold ? c.r3ies : c.r3les, z4es = cold ? c.r4ies : c.r4les;
      zesdp = m_div(c.r2es * m_exp(m_div(z3es * (ztp1 - c.rtt), ztp1 - z4es)), papp1);
      if (zesdp > zqmax) zesdp = zqmax;
    } else {
      zalfaw = foealfa(c, ztp1);
      zesdp = m_div(foeewm(c, ztp1, zalfaw), papp1);
    }
    const R dl = ztp1 - c.r4les, di = ztp1 - c.r4ies;
    const R zfac = zalfaw * m_div(c.r5les, dl * dl) + (R(1.0) - zalfaw) * m_div(c.r5ies, di * di);
    const R zcor = m_div(R(1.0), R(1.0) - c.retv * zesdp);
    const R zdqsdtemp = zfac * zcor * pqs;
    const R zcorqs = R(1.0) + zcons3 * zdqsdtemp;
    const R zqlim = zqp1 > pqs ? pqs : zqp1;
    const R zqcrit = zcrh * pqs;
    // layer cloud amounts (:387-403)
    R pclc, zqlwc;
    if (zqp1 <= zqcrit) {
      pclc = R(0.0);
      zqlwc = R(0.0);
    } else if (zqp1 >= pqs) {
      pclc = R(1.0);
      zqlwc = (R(1.0) - zscalm) * (pqs - zqcrit);
    } else {
      const R zqpd = pqs - zqp1, zqcd = pqs - zqcrit;
      pclc = R(1.0) - m_sqrt(m_div(zqpd, zqcd - zscalm * (zqp1 - zqcrit)));
      zqlwc = (zscalm * zqpd + (R(1.0) - zscalm) * zqcd) * (pclc * pclc);
    }
    // convective cloud (:407-424)
    const R zgdp = m_div(c.rg, paph1 - paph0);
    const R zlude = plude * ptsphy * zgdp;
    if (k < klev - 1) {
      const R plu_n = cur.lu_next;
      if (zlude >= c.rlmin && plu_n >= zeps2) {
        pclc = pclc + (R(1.0) - pclc) * (R(1.0) - m_exp(-m_div(zlude, plu_n)));
        zqlwc = zqlwc + zlude;
      }
    }
    zcovptot = m_max(zcovptot, pclc);  // (:428-434)
    const R zcovpclr = m_max(zcovptot - pclc, R(0.0));
    // melting of incoming snow (:448-460)
    R zrfln, zsfln;
    if (zsfl != R(0.0)) {
      const R zcons = m_div(c.zcons2 * zdp, zlfdcp);
      const R zsnmlt = m_min(zsfl, zcons * m_max(R(0.0), ztp1 - c.zmeltp2));
      zrfln = zrfl + zsnmlt;
      zsfln = zsfl - zsnmlt;
      ztp1 = ztp1 - m_div(zsnmlt, zcons);
    } else {
      zrfln = zrfl;
      zsfln = zsfl;
    }
    // precipitation (:462-542)
    R zpr = R(0.0);
    if (pclc > zeps2) {
      const R zlcrit = c.evap ? c.rclcrit : c.rclcrit * R(2.);
      const R zcld = m_div(zqlwc, pclc);
      const R r = m_div(zcld, zlcrit);
      const R zd = c.zckcodtl * (R(1.0) - m_exp(-(r * r)));
      const R zlnew = pclc * zcld * m_exp(-zd);
      zpr = zqlwc - zlnew;
      zqlwc = zqlwc - zpr;
    }
    R zdr = c.zcons2 * zdp * zpr;
    R zfwat;
    if (c.lin) {
      const R zoealfa = R(0.545) * (tanh(R(0.17) * (ztp1 - c.rlptrc)) + R(1.0));
      zfwat = ztp1 < c.rtt ? zoealfa : R(1.0);
    } else {
      zfwat = foealfa(c, ztp1);
    }
    R zrfreeze = R(0), zfwatr;
    if (ztp1 < c.rtt) {
      zrfreeze = zfwat * zdr;
      zfwatr = R(0.0);
    } else {
      zfwatr = R(1.0);
    }
    zrfln = zrfln + zfwatr * zdr;
    zsfln = zsfln + (R(1.0) - zfwatr) * zdr;
    R pcovptot = R(0);
    const R zprtot = zrfln + zsfln;
    if (zprtot > zeps2 && zcovpclr > zeps2 && c.evap) {
      R zpreclr = m_div(zprtot * zcovpclr, zcovptot);
      const R om = R(1.0) - pclc;
      const R zqe = pqs - m_div((pqs - zqlim) * zcovpclr, om * om);
      const R zbeta = c.rg * zpecons *
                      m_pow(m_div(m_div(m_sqrt(m_div(papp1, paph_surf)), R(5.09E-3)) * zpreclr, zcovpclr), R(0.5777));
      const R zb = m_div(ptsphy * zbeta * (pqs - zqe), R(1.0) + zbeta * ptsphy * zcorqs);
      const R zdtgdp = m_div(ptsphy * c.rg, paph1 - paph0);
      R zdpr = m_div(zcovpclr * zb, zdtgdp);
      zdpr = m_min(zdpr, zpreclr);
      zpreclr = zpreclr - zdpr;
      if (zpreclr <= R(0.0)) zcovptot = pclc;
      pcovptot = zcovptot;
      const R zdprsfl = m_div(zdpr * zrfln, zprtot);
      zrfln = zrfln - zdprsfl;
      const R zdpssfl = m_div(zdpr * zsfln, zprtot);
      zsfln = zsfln - zdpssfl;
    }
    const R pqlwp = zfwat * zqlwc, pqiwp = (R(1.0) - zfwat) * zqlwc;
    // tendencies from the flux divergence, first guess and the final adjustment (:546-616)
    auto tendency_q = [&](R rfln, R sfln) {
      return -((rfln - zrfl) + (sfln - zsfl)) * m_div(c.rg, zdp) + plude * zgdp;
    };
    auto tendency_t = [&](R rfln, R sfln, R rfreeze) {
      return (zlvdcp * (rfln - zrfl) + zlsdcp * (sfln - zsfl)) * m_div(c.rg, zdp) -
             plude * zgdp * (zfwat * zlvdcp + (R(1.0) - zfwat) * zlsdcp) + (zlsdcp - zlvdcp) * rfreeze * zgdp;
    };
    ztp1 = ztp1 + ptsphy * tendency_t(zrfln, zsfln, zrfreeze);
    zqp1 = zqp1 + ptsphy * tendency_q(zrfln, zsfln);
    const R zqold = zqp1;
    adjust(c, papp1, ztp1, zqp1);
    const R zdq = m_max(R(0.0), zqold - zqp1);
    zdr = c.zcons2 * zdp * zdq;
    if (ztp1 < c.rtt) {
      zrfreeze = zrfreeze + zfwat * zdr;
      zfwatr = R(0.0);
    } else {
      zfwatr = R(1.0);
    }
    zrfln = zrfln + zfwatr * zdr;
    zsfln = zsfln + (R(1.0) - zfwatr) * zdr;

    ST(p.ptenq, i2, jo, tendency_q(zrfln, zsfln));
    ST(p.ptent, i2, jo, tendency_t(zrfln, zsfln, zrfreeze));
    if (!LAYER || p.pclc) ST(p.pclc, i2, jo, pclc);
    if (!LAYER || p.pqlwp) ST(p.pqlwp, i2, jo, pqlwp);
    if (!LAYER || p.pqiwp) ST(p.pqiwp, i2, jo, pqiwp);
    ST(p.pcovptot, i2, jo, pcovptot);
    wrap_state(k, pclc, pqlwp, pqiwp, cur.a, cur.l, cur.i);
    const uint32_t ih = ath(k + 1);
    ST(p.pfplsl, ih, jo, zrfln);
    ST(p.pfplsn, ih, jo, zsfln);
    ST(p.pfhpsl, ih, jo, -zrfln * c.rlvtt);
    ST(p.pfhpsn, ih, jo, -zsfln * c.rlstt);
    zrfl = zrfln;
    zsfl = zsfln;
    paph0 = paph1;
  }
}

template <typename R>
C2Const<R> make_const(const cloudsc_params_t& q, const cloudsc2_params_t& q2, double ptsphy, bool ldrain1d) {
  C2Const<R> c{};
  c.rg = R(q.rg); c.rd = R(q.rd); c.rcpd = R(q.rcpd); c.retv = R(q.retv); c.rtt = R(q.rtt); c.rlvtt = R(q.rlvtt);
  c.rlstt = R(q.rlstt); c.rlmlt = R(q.rlmlt); c.r2es = R(q.r2es); c.r3les = R(q.r3les); c.r3ies = R(q.r3ies);
  c.r4les = R(q.r4les); c.r4ies = R(q.r4ies); c.r5les = R(q.r5les); c.r5ies = R(q.r5ies);
  c.r5alvcp = R(q.r5alvcp); c.r5alscp = R(q.r5alscp); c.ralvdcp = R(q.ralvdcp); c.ralsdcp = R(q.ralsdcp);
  c.rtwat = R(q.rtwat); c.rtice = R(q.rtice); c.rtwat_rtice_r = R(q.rtwat_rtice_r); c.rlmin = R(q.rlmin);
  c.rpecons = R(q.rpecons); c.rclcrit = R(q2.rclcrit); c.rlptrc = R(q2.rlptrc); c.rvtmp2 = R(q2.rvtmp2);
  // (:243-249), evaluated in R as the reference does in JPRB
  const R ts = R(ptsphy), rkconv = R(q.rkconv);
  c.ptsphy = ts;
  c.zckcodtl = R(2.0) * rkconv * ts;
  c.zckcodti = R(5.0) * rkconv * ts;
  c.zcons2 = R(1) / (ts * c.rg);
  c.zcons3 = c.rlvtt / c.rcpd;
  c.zmeltp2 = c.rtt + R(2.0);
  c.zqtmst = R(1) / ts;
  c.lin = (q2.lphylin || ldrain1d) ? 1 : 0;
  c.phylin = q2.lphylin ? 1 : 0;
  c.evap = (q2.levapls2 || ldrain1d) ? 1 : 0;
  return c;
}

int check(const cloudsc_params_t* q, const cloudsc2_params_t* q2, int ktdia, int klev) {
  if (!q || !q2) return fail(CLOUDSC_ENULL, "cloudsc2: NULL parameter block");
  if (klev < 2 || ktdia < 1 || ktdia > klev) return fail(CLOUDSC_EDIMS, "cloudsc2: need KLEV >= 2, 1 <= KTDIA <= KLEV");
  return CLOUDSC_OK;
}

// The blocked fields of blocks b0.. of one launch: (NPROMA,KLEV) arrays advance by NPROMA*KLEV words per block,
// (NPROMA,KLEV+1) arrays by NPROMA*(KLEV+1); CETA is shared.
template <typename R>
C2Ptrs<R> offset_ptrs(const C2Ptrs<R>& p, int64_t b0, int nproma, int klev) {
  const int64_t s2 = b0 * nproma * klev, sh = b0 * nproma * (klev + 1);
  C2Ptrs<R> o = p;
  auto adv0 = [](auto* x, int64_t n) { return x ? x + n : x; };
  o.paphp1 += sh; o.papp1 += s2; o.pqm1 += s2; o.pqs = adv0(o.pqs, s2); o.ptm1 += s2; o.pl = adv0(o.pl, s2);
  o.pi = adv0(o.pi, s2); o.plude += s2; o.plu += s2; o.pmfu = adv0(o.pmfu, s2); o.pmfd = adv0(o.pmfd, s2);
  o.ptent += s2; o.pgtent += s2; o.ptenq += s2; o.pgtenq += s2; o.ptenl = adv0(o.ptenl, s2);
  o.pgtenl = adv0(o.pgtenl, s2); o.pteni = adv0(o.pteni, s2); o.pgteni = adv0(o.pgteni, s2); o.pfplsl += sh;
  o.pfplsn += sh; o.pfhpsl += sh; o.pfhpsn += sh; o.pcovptot += s2;
  auto adv = [](auto* x, int64_t n) { return x ? x + n : x; };  // optional arrays stay NULL
  o.pa = adv(o.pa, s2); o.ptena = adv(o.ptena, s2); o.ptenr = adv(o.ptenr, s2); o.ptens = adv(o.ptens, s2);
  o.pqsat = adv(o.pqsat, s2); o.pclc = adv(o.pclc, s2);
  o.pprectype = adv(o.pprectype, b0 * nproma); o.pfzra = adv(o.pfzra, b0 * nproma);
  o.pqiwp = adv(o.pqiwp, s2); o.pqlwp = adv(o.pqlwp, s2);
  return o;
}

template <typename R>
hipError_t launch(const cloudsc_params_t& q, const cloudsc2_params_t& q2, bool ldrain1d, double ptsphy,
                  const C2Ptrs<R>& p, const C2Dims& d, hipStream_t stream, bool layer = false, bool st = false) {
  const C2Const<R> c = make_const<R>(q, q2, ptsphy, ldrain1d);
  const bool uni = d.nproma % 64 == 0;
  // blocks per launch: every (NPROMA,KLEV+1) array of one launch stays below 4 GiB (32-bit byte offsets)
  const int64_t per_block = (int64_t)d.nproma * (d.klev + 1) * (int64_t)sizeof(R);
  const int64_t max_blocks = std::max<int64_t>(1, ((int64_t)1 << 32) / per_block - 1);
  // launches of equal size: a short last launch would leave most of the GPU idle while it runs
  const int64_t nlaunch = (d.ngpblks + max_blocks - 1) / max_blocks;
  const int64_t step = (d.ngpblks + nlaunch - 1) / nlaunch;
  for (int64_t b0 = 0; b0 < d.ngpblks; b0 += step) {
    const int64_t nb = std::min<int64_t>(step, d.ngpblks - b0);
    C2Dims dc = d;
    dc.ngpblks = (int)nb;
    dc.ngptot = (int)(d.ngptot - b0 * d.nproma);
    const C2Ptrs<R> pc = offset_ptrs(p, b0, d.nproma, d.klev);
    const dim3 grid((unsigned)((nb * d.nproma + CLOUDSC2_BLOCK - 1) / CLOUDSC2_BLOCK));
    if (st && layer && uni)
      hipLaunchKernelGGL((cloudst_kernel<R, true, true>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (st && layer)
      hipLaunchKernelGGL((cloudst_kernel<R, false, true>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (st && uni)
      hipLaunchKernelGGL((cloudst_kernel<R, true, false>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (st)
      hipLaunchKernelGGL((cloudst_kernel<R, false, false>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (layer && uni)
      hipLaunchKernelGGL((cloudsc2_kernel<R, true, true>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (layer)
      hipLaunchKernelGGL((cloudsc2_kernel<R, false, true>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else if (uni)
      hipLaunchKernelGGL((cloudsc2_kernel<R, true, false>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    else
      hipLaunchKernelGGL((cloudsc2_kernel<R, false, false>), grid, dim3(CLOUDSC2_BLOCK), 0, stream, c, pc, dc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// One host array of a per-block call: copied to the workspace, and back to `out` when that is set.
struct HostArr {
  const void* host;
  void* out;
  size_t n;
};

// Copies the arrays into this thread's device workspace, runs `launch_fn(D, stream)` with D(i) the device copy of
// array i, copies the outputs back and waits.
template <typename R, typename F>
int run_on_device(const char* what, const HostArr* arrs, int narr, F launch_fn) {
  size_t off[64], total = 0;
  if (narr > 64) return fail(CLOUDSC_EDIMS, what);
  for (int i = 0; i < narr; ++i) {
    off[i] = total;
    total += (arrs[i].n * sizeof(R) + 255) & ~(size_t)255;
  }
  // the calling thread's workspace (shared with the other per-block entries)
  cloudsc_capi::HostWorkspace* wp = nullptr;
  const int rc = cloudsc_capi::thread_workspace(total, "cloudsc2", &wp);
  if (rc) return rc;
  cloudsc_capi::HostWorkspace& w = *wp;
  hipError_t e = hipSuccess;
  char* base = (char*)w.dev;
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    if (arrs[i].host)
      e = hipMemcpyAsync(base + off[i], arrs[i].host, arrs[i].n * sizeof(R), hipMemcpyHostToDevice, w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc2: copy to device", e);
  auto D = [&](int i) { return arrs[i].host ? (R*)(base + off[i]) : (R*)nullptr; };  // NULL stays NULL
  e = launch_fn(D, w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, what, e);
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    if (arrs[i].out && arrs[i].host)
      e = hipMemcpyAsync(arrs[i].out, base + off[i], arrs[i].n * sizeof(R), hipMemcpyDeviceToHost, w.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc2: copy to host", e);
  return CLOUDSC_OK;
}

template <typename R>
int per_block(const cloudsc_params_t* q, const cloudsc2_params_t* q2, const R* ceta, int kidia, int kfdia, int klon,
              int ktdia, int klev, int ldrain1d, R ptsphy, const R* paphp1, const R* papp1, const R* pqm1,
              const R* pqs, const R* ptm1, const R* pl, const R* pi, const R* plude, const R* plu, const R* pmfu,
              const R* pmfd, R* ptent, const R* pgtent, R* ptenq, const R* pgtenq, R* ptenl, const R* pgtenl,
              R* pteni, const R* pgteni, R* pclc, R* pfplsl, R* pfplsn, R* pfhpsl, R* pfhpsn, R* pcovptot) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (klon < 1 || kidia < 1 || kfdia < kidia || kfdia > klon)
    return fail(CLOUDSC_EDIMS, "cloudsc2: need 1 <= KIDIA <= KFDIA <= KLON");
  const void* req[] = {ceta,  paphp1, papp1, pqm1,  pqs,    ptm1,   pl,    pi,     plude, plu,   pmfu,  pmfd,
                       ptent, pgtent, ptenq, pgtenq, ptenl, pgtenl, pteni, pgteni, pclc,  pfplsl, pfplsn, pfhpsl,
                       pfhpsn, pcovptot};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudsc2: NULL array");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const size_t n2 = (size_t)klon * klev, nh = (size_t)klon * (klev + 1);
  // outputs are copied in too, so that what CLOUDSC2 leaves untouched comes back unchanged
  HostArr arrs[] = {{ceta, nullptr, (size_t)klev}, {paphp1, nullptr, nh}, {papp1, nullptr, n2}, {pqm1, nullptr, n2},
                    {pqs, nullptr, n2},   {ptm1, nullptr, n2},  {pl, nullptr, n2},     {pi, nullptr, n2},
                    {plude, nullptr, n2}, {plu, nullptr, n2},   {pmfu, nullptr, n2},   {pmfd, nullptr, n2},
                    {ptent, ptent, n2},   {pgtent, nullptr, n2}, {ptenq, ptenq, n2},   {pgtenq, nullptr, n2},
                    {ptenl, ptenl, n2},   {pgtenl, nullptr, n2}, {pteni, pteni, n2},   {pgteni, nullptr, n2},
                    {pclc, pclc, n2},     {pfplsl, pfplsl, nh},  {pfplsn, pfplsn, nh}, {pfhpsl, pfhpsl, nh},
                    {pfhpsn, pfhpsn, nh}, {pcovptot, pcovptot, n2}};
  return run_on_device<R>("cloudsc2", arrs, sizeof(arrs) / sizeof(arrs[0]), [&](auto D, hipStream_t st) {
    C2Ptrs<R> p{D(0),  D(1),  D(2),  D(3),  D(4),  D(5),  D(6),  D(7),  D(8),  D(9),  D(10), D(11), D(12),
                D(13), D(14), D(15), D(16), D(17), D(18), D(19), D(20), D(21), D(22), D(23), D(24), D(25)};
    C2Dims d{kfdia, klon, klev, 1, ktdia, kidia, kfdia};
    return launch<R>(*q, *q2, ldrain1d != 0, (double)ptsphy, p, d, st);
  });
}

// CLOUDST on one block of host arrays (the per-block drop-in, as cloudsc2_*_c).
template <typename R>
int per_block_st(const cloudsc_params_t* q, const cloudsc2_params_t* q2, const R* ceta, int kidia, int kfdia,
                 int klon, int ktdia, int klev, int ldrain1d, R ptsphy, const R* paphp1, const R* papp1,
                 const R* pqm1, const R* pqs, const R* ptm1, const R* plude, const R* plu, R* ptent,
                 const R* pgtent, R* ptenq, const R* pgtenq, R* pclc, R* pqiwp, R* pqlwp, R* pfplsl, R* pfplsn,
                 R* pfhpsl, R* pfhpsn, R* pcovptot) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (klon < 1 || kidia < 1 || kfdia < kidia || kfdia > klon)
    return fail(CLOUDSC_EDIMS, "cloudst: need 1 <= KIDIA <= KFDIA <= KLON");
  const void* req[] = {ceta,  paphp1, papp1,  pqm1,  pqs,   ptm1,   plude,  plu,    ptent,  pgtent,
                       ptenq, pgtenq, pclc,   pqiwp, pqlwp, pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudst: NULL array");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const size_t n2 = (size_t)klon * klev, nh = (size_t)klon * (klev + 1);
  HostArr arrs[] = {{ceta, nullptr, (size_t)klev}, {paphp1, nullptr, nh}, {papp1, nullptr, n2}, {pqm1, nullptr, n2},
                    {pqs, nullptr, n2},     {ptm1, nullptr, n2},   {plude, nullptr, n2},  {plu, nullptr, n2},
                    {ptent, ptent, n2},     {pgtent, nullptr, n2}, {ptenq, ptenq, n2},    {pgtenq, nullptr, n2},
                    {pclc, pclc, n2},       {pqiwp, pqiwp, n2},    {pqlwp, pqlwp, n2},    {pfplsl, pfplsl, nh},
                    {pfplsn, pfplsn, nh},   {pfhpsl, pfhpsl, nh},  {pfhpsn, pfhpsn, nh},  {pcovptot, pcovptot, n2}};
  return run_on_device<R>("cloudst", arrs, sizeof(arrs) / sizeof(arrs[0]), [&](auto D, hipStream_t st) {
    C2Ptrs<R> p{};
    p.ceta = D(0); p.paphp1 = D(1); p.papp1 = D(2); p.pqm1 = D(3); p.pqs = D(4); p.ptm1 = D(5); p.plude = D(6);
    p.plu = D(7); p.ptent = D(8); p.pgtent = D(9); p.ptenq = D(10); p.pgtenq = D(11); p.pclc = D(12);
    p.pqiwp = D(13); p.pqlwp = D(14); p.pfplsl = D(15); p.pfplsn = D(16); p.pfhpsl = D(17); p.pfhpsn = D(18);
    p.pcovptot = D(19);
    C2Dims d{kfdia, klon, klev, 1, ktdia, kidia, kfdia};
    return launch<R>(*q, *q2, ldrain1d != 0, (double)ptsphy, p, d, st, false, true);
  });
}

template <typename R>
int blocked(const cloudsc_params_t* q, const cloudsc2_params_t* q2, const R* ceta, int ngptot, int nproma, int ktdia,
            int klev, int ngpblks, int ldrain1d, R ptsphy, const C2Ptrs<R>& p, void* stream) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (nproma < 1 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloudsc2_blocked: need NGPBLKS = ceil(NGPTOT/NPROMA)");
  const void* req[] = {p.ceta,  p.paphp1, p.papp1, p.pqm1,   p.pqs,   p.ptm1,   p.pl,    p.pi,     p.plude,
                       p.plu,   p.pmfu,   p.pmfd,  p.ptent,  p.pgtent, p.ptenq, p.pgtenq, p.ptenl, p.pgtenl,
                       p.pteni, p.pgteni, p.pclc,  p.pfplsl, p.pfplsn, p.pfhpsl, p.pfhpsn, p.pcovptot};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudsc2_blocked: NULL array");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  C2Dims d{ngptot, nproma, klev, ngpblks, ktdia, 1, nproma};
  hipError_t e = launch<R>(*q, *q2, ldrain1d != 0, (double)ptsphy, p, d, (hipStream_t)stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc2_blocked: kernel launch", e);
  return CLOUDSC_OK;
}

// CLOUD_S_LAYER on one block of host arrays (Fortran order), the arguments of CLOUD_S_LAYER_ARGS; NULL optional
// arrays stay NULL on the device.
template <typename R>
int layer_per_block(const cloudsc_params_t* q, const cloudsc2_params_t* q2, const R* ceta, int kidia, int kfdia,
                    int klon, int ktdia, int klev, int ldrain1d, R tsphy, const R* const* in, R* const* out) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (klon < 1 || kidia < 1 || kfdia < kidia || kfdia > klon)
    return fail(CLOUDSC_EDIMS, "cloud_s_layer: need 1 <= KIDIA <= KFDIA <= KLON");
  const bool c2 = q2->lepcld2 != 0;
  for (int i : {0, 1, 2, 3, 5, 6, 7, 8, 11, 12})
    if (!in[i]) return fail(CLOUDSC_ENULL, "cloud_s_layer: NULL array");
  for (int i : {0, 1, 3, 4, 9, 10, 11, 12, 13})
    if (!out[i]) return fail(CLOUDSC_ENULL, "cloud_s_layer: NULL array");
  if (c2 && (!in[9] || !in[10] || !in[13] || !in[14]))
    return fail(CLOUDSC_ENULL, "cloud_s_layer: LEPCLD2 needs PMFU, PMFD and the condensate tendencies");
  if (out[2] && !in[4]) return fail(CLOUDSC_ENULL, "cloud_s_layer: PTENDENCY_LOC_A needs PA");
  if (!ceta) return fail(CLOUDSC_ENULL, "cloud_s_layer: NULL CETA");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const size_t n2 = (size_t)klon * klev, nh = (size_t)klon * (klev + 1), n1 = (size_t)klon;
  // in: prs1 prsf1 pq pt pa pl pi plude plu pmfu pmfd ptendency_t/q/l/i (15);
  // out: loc_t loc_q loc_a loc_l loc_i loc_r loc_s pqsat pclc pfplsl pfplsn pfhpsl pfhpsn pcovptot pprectype pfzra
  //      pqiwp pqlwp (18)
  HostArr arrs[1 + 15 + 18];
  arrs[0] = {ceta, nullptr, (size_t)klev};
  for (int i = 0; i < 15; ++i) arrs[1 + i] = {in[i], nullptr, i == 0 ? nh : n2};
  for (int i = 0; i < 18; ++i) {
    const size_t n = (i >= 9 && i <= 12) ? nh : (i == 14 || i == 15) ? n1 : n2;
    arrs[16 + i] = {out[i], out[i], n};
  }
  return run_on_device<R>("cloud_s_layer", arrs, 34, [&](auto D, hipStream_t st) {
    C2Ptrs<R> p{};
    p.ceta = D(0); p.paphp1 = D(1); p.papp1 = D(2); p.pqm1 = D(3); p.ptm1 = D(4); p.pa = D(5); p.pl = D(6);
    p.pi = D(7); p.plude = D(8); p.plu = D(9); p.pmfu = D(10); p.pmfd = D(11); p.pgtent = D(12);
    p.pgtenq = D(13); p.pgtenl = D(14); p.pgteni = D(15);
    p.ptent = D(16); p.ptenq = D(17); p.ptena = D(18); p.ptenl = D(19); p.pteni = D(20); p.ptenr = D(21);
    p.ptens = D(22); p.pqsat = D(23); p.pclc = D(24); p.pfplsl = D(25); p.pfplsn = D(26); p.pfhpsl = D(27);
    p.pfhpsn = D(28); p.pcovptot = D(29); p.pprectype = D(30); p.pfzra = D(31); p.pqiwp = D(32); p.pqlwp = D(33);
    C2Dims d{kfdia, klon, klev, 1, ktdia, kidia, kfdia};
    return launch<R>(*q, *q2, ldrain1d != 0, (double)tsphy, p, d, st, true, !c2);
  });
}

// CLOUD_S_LAYER with LEPCLD2 (cloud_s_layer.F90:112-176) over all blocks: SATUR, CLOUDSC2 and the wrap-up in one
// launch.  PTENDENCY_LOC_A/R/S, PQSAT, PCLC (AUXL%ZPATMP), PPRECTYPE and PFZRA may be NULL.
template <typename R>
int layer_blocked(const cloudsc_params_t* q, const cloudsc2_params_t* q2, const R* ceta, int ngptot, int nproma,
                  int ktdia, int klev, int ngpblks, int ldrain1d, R tsphy, const C2Ptrs<R>& p, void* stream) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (nproma < 1 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloud_s_layer_blocked: need NGPBLKS = ceil(NGPTOT/NPROMA)");
  const bool c2 = q2->lepcld2 != 0;  // CLOUDSC2 (prognostic) or CLOUDST (diagnostic) (:121-150)
  const void* req[] = {p.ceta,  p.paphp1, p.papp1,  p.pqm1,   p.ptm1,   p.pl,     p.pi,     p.plude,
                       p.plu,   p.ptent,  p.pgtent, p.ptenq,  p.pgtenq, p.ptenl,  p.pteni,  p.pfplsl,
                       p.pfplsn, p.pfhpsl, p.pfhpsn, p.pcovptot};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloud_s_layer_blocked: NULL array");
  if (c2 && (!p.pmfu || !p.pmfd || !p.pgtenl || !p.pgteni))
    return fail(CLOUDSC_ENULL, "cloud_s_layer_blocked: LEPCLD2 needs PMFU, PMFD and the condensate tendencies");
  if (p.ptena && !p.pa) return fail(CLOUDSC_ENULL, "cloud_s_layer_blocked: PTENDENCY_LOC_A needs PA");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  C2Dims d{ngptot, nproma, klev, ngpblks, ktdia, 1, nproma};
  hipError_t e = launch<R>(*q, *q2, ldrain1d != 0, (double)tsphy, p, d, (hipStream_t)stream, true, !c2);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_s_layer_blocked: kernel launch", e);
  return CLOUDSC_OK;
}

// CLOUDST over all blocks of device-resident blocked fields.
template <typename R>
int cloudst_blocked(const cloudsc_params_t* q, const cloudsc2_params_t* q2, int ngptot, int nproma, int ktdia,
                    int klev, int ngpblks, int ldrain1d, R ptsphy, const C2Ptrs<R>& p, void* stream) {
  int rc = check(q, q2, ktdia, klev);
  if (rc) return rc;
  if (nproma < 1 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloudst_blocked: need NGPBLKS = ceil(NGPTOT/NPROMA)");
  const void* req[] = {p.ceta,  p.paphp1, p.papp1, p.pqm1,   p.pqs,   p.ptm1,   p.plude,  p.plu,  p.ptent,
                       p.pgtent, p.ptenq, p.pgtenq, p.pclc,   p.pqiwp, p.pqlwp,  p.pfplsl, p.pfplsn,
                       p.pfhpsl, p.pfhpsn, p.pcovptot};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudst_blocked: NULL array");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  C2Dims d{ngptot, nproma, klev, ngpblks, ktdia, 1, nproma};
  hipError_t e = launch<R>(*q, *q2, ldrain1d != 0, (double)ptsphy, p, d, (hipStream_t)stream, false, true);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudst_blocked: kernel launch", e);
  return CLOUDSC_OK;
}

}  // namespace cloudsc2_gpu

#define CLOUDSC2_PER_BLOCK(R)                                                                                     \
  return cloudsc2_gpu::per_block<R>(ydparams, ydparams2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, \
                                    paphp1, papp1, pqm1, pqs, ptm1, pl, pi, plude, plu, pmfu, pmfd, ptent, pgtent, \
                                    ptenq, pgtenq, ptenl, pgtenl, pteni, pgteni, pclc, pfplsl, pfplsn, pfhpsl,     \
                                    pfhpsn, pcovptot)

#define CLOUDSC2_BLOCKED(R)                                                                                        \
  const cloudsc2_gpu::C2Ptrs<R> p{ceta,  paphp1, papp1, pqm1,   pqs,   ptm1,   pl,    pi,     plude,  plu,     \
                                  pmfu,  pmfd,   ptent, pgtent, ptenq, pgtenq, ptenl, pgtenl, pteni,  pgteni,  \
                                  pclc,  pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot};                              \
  return cloudsc2_gpu::blocked<R>(ydparams, ydparams2, ceta, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d,     \
                                  ptsphy, p, stream)

#define CLOUD_S_LAYER_BLOCKED(R)                                                                                \
  cloudsc2_gpu::C2Ptrs<R> p{ceta,  prs1,   prsf1,  pq,    nullptr, pt,    pl,    pi,     plude,   plu,         \
                            pmfu,  pmfd,   ptendency_loc_t, ptendency_t, ptendency_loc_q, ptendency_q,            \
                            ptendency_loc_l, ptendency_l, ptendency_loc_i, ptendency_i, pclc, pfplsl, pfplsn,    \
                            pfhpsl, pfhpsn, pcovptot};                                                             \
  p.pa = pa; p.ptena = ptendency_loc_a; p.ptenr = ptendency_loc_r; p.ptens = ptendency_loc_s; p.pqsat = pqsat;   \
  p.pprectype = pprectype; p.pfzra = pfzra; p.pqiwp = pqiwp; p.pqlwp = pqlwp;                                    \
  return cloudsc2_gpu::layer_blocked<R>(ydparams, ydparams2, ceta, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d, \
                                        tsphy, p, stream)

#define CLOUD_S_LAYER_PER_BLOCK(R)                                                                             \
  const R* in[15] = {prs1, prsf1, pq, pt, pa, pl, pi, plude, plu, pmfu, pmfd, ptendency_t, ptendency_q,          \
                     ptendency_l, ptendency_i};                                                                   \
  R* out[18] = {ptendency_loc_t, ptendency_loc_q, ptendency_loc_a, ptendency_loc_l, ptendency_loc_i,             \
                ptendency_loc_r, ptendency_loc_s, pqsat, pclc, pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot,         \
                pprectype, pfzra, pqiwp, pqlwp};                                                                  \
  return cloudsc2_gpu::layer_per_block<R>(ydparams, ydparams2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d,   \
                                          tsphy, in, out)

#define CLOUDST_PER_BLOCK(R)                                                                                   \
  return cloudsc2_gpu::per_block_st<R>(ydparams, ydparams2, ceta, kidia, kfdia, klon, ktdia, klev, ldrain1d, ptsphy, \
                                       paphp1, papp1, pqm1, pqs, ptm1, plude, plu, ptent, pgtent, ptenq, pgtenq, pclc,  \
                                       pqiwp, pqlwp, pfplsl, pfplsn, pfhpsl, pfhpsn, pcovptot)

#define CLOUDST_BLOCKED(R)                                                                                      \
  cloudsc2_gpu::C2Ptrs<R> p{};                                                                                     \
  p.ceta = ceta; p.paphp1 = paphp1; p.papp1 = papp1; p.pqm1 = pqm1; p.pqs = pqs; p.ptm1 = ptm1; p.plude = plude;   \
  p.plu = plu; p.ptent = ptent; p.pgtent = pgtent; p.ptenq = ptenq; p.pgtenq = pgtenq; p.pclc = pclc;               \
  p.pqiwp = pqiwp; p.pqlwp = pqlwp; p.pfplsl = pfplsl; p.pfplsn = pfplsn; p.pfhpsl = pfhpsl; p.pfhpsn = pfhpsn;     \
  p.pcovptot = pcovptot;                                                                                          \
  return cloudsc2_gpu::cloudst_blocked<R>(ydparams, ydparams2, ngptot, nproma, ktdia, klev, ngpblks, ldrain1d,     \
                                          ptsphy, p, stream)

extern "C" {

int cloudsc2_params_default(cloudsc2_params_t* p2, const cloudsc_params_t* p) {
  if (!p2 || !p) return cloudsc_capi::fail(CLOUDSC_ENULL, "cloudsc2_params_default: NULL");
  p2->rclcrit = 4.0E-4;                                    // sucldp.F90:264
  p2->rlptrc = p->rtice + (p->rtwat - p->rtice) / std::sqrt(2.0);  // suphli.F90:125
  p2->rlpal1 = 0.15;                                        // suphli.F90:126
  p2->rlpal2 = 20.0;                                        // suphli.F90:127
  p2->rvtmp2 = 0.0;                                         // suphec.F90:229
  p2->lphylin = 0;                                          // suphli.F90:107
  p2->levapls2 = 0;                                         // su0phy.F90:1361
  p2->lepcld2 = 0;                                          // su0phy.F90:1360
  p2->pad_ = 0;
  return CLOUDSC_OK;
}

int cloudsc2_dp_c(CLOUDSC2_ARGS(double)) { CLOUDSC2_PER_BLOCK(double); }
int cloudsc2_sp_c(CLOUDSC2_ARGS(float)) { CLOUDSC2_PER_BLOCK(float); }
int cloudsc2_blocked_dp_c(CLOUDSC2_BLOCKED_ARGS(double)) { CLOUDSC2_BLOCKED(double); }
int cloudsc2_blocked_sp_c(CLOUDSC2_BLOCKED_ARGS(float)) { CLOUDSC2_BLOCKED(float); }
int cloud_s_layer_dp_c(CLOUD_S_LAYER_ARGS(double)) { CLOUD_S_LAYER_PER_BLOCK(double); }
int cloud_s_layer_sp_c(CLOUD_S_LAYER_ARGS(float)) { CLOUD_S_LAYER_PER_BLOCK(float); }
int cloudst_dp_c(CLOUDST_ARGS(double)) { CLOUDST_PER_BLOCK(double); }
int cloudst_sp_c(CLOUDST_ARGS(float)) { CLOUDST_PER_BLOCK(float); }
int cloudst_blocked_dp_c(CLOUDST_BLOCKED_ARGS(double)) { CLOUDST_BLOCKED(double); }
int cloudst_blocked_sp_c(CLOUDST_BLOCKED_ARGS(float)) { CLOUDST_BLOCKED(float); }
int cloud_s_layer_blocked_dp_c(CLOUD_S_LAYER_BLOCKED_ARGS(double)) { CLOUD_S_LAYER_BLOCKED(double); }
int cloud_s_layer_blocked_sp_c(CLOUD_S_LAYER_BLOCKED_ARGS(float)) { CLOUD_S_LAYER_BLOCKED(float); }

}  // extern "C"
