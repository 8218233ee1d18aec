// CPU oracle of CLOUDSC2 -- TEST INFRASTRUCTURE ONLY (see cloudsc_oracle.hpp).
//
// Restates, in the reference's loop structure (vector loops over JL inside the level loop),
//   ifs-source/arpifs/phys_ec/cloudsc2.F90:10-709   SUBROUTINE CLOUDSC2 (the simplified scheme of the
//                                                   linearised physics, called from CLOUD_SOLVER's CLOUD_S_LAYER)
//   ifs-source/arpifs/phys_ec/cuadjtq.F90:250-274    CUADJTQ, KCALL = 0, LPHYLIN = .FALSE.
//   ifs-source/arpifs/phys_ec/cuadjtqs.F90:123-134,217-250  CUADJTQS, KCALL = 0 (LPHYLIN or LDRAIN1D)
// Each block cites the lines it follows.  Integer powers (**2) are products; a real exponent of 0.5 is sqrt.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../include/cloudsc2.h"
#include "cloudsc_oracle.hpp"

namespace oracle {

template <typename R>
struct Cloudsc2Args {
  int kidia, kfdia, klon, ktdia, klev;
  bool ldrain1d;
  R ptsphy;
  const R* ceta;
  const R *paphp1, *papp1, *pqm1, *pqs, *ptm1, *pl, *pi, *plude, *plu, *pmfu, *pmfd;
  R* ptent;
  const R* pgtent;
  R* ptenq;
  const R* pgtenq;
  R* ptenl;
  const R* pgtenl;
  R* pteni;
  const R* pgteni;
  R *pclc, *pfplsl, *pfplsn, *pfhpsl, *pfhpsn, *pcovptot;
};

// CUADJTQ, KCALL = 0, default thermodynamics (cuadjtq.F90:256-272): two Newton steps of the isobaric
// saturation adjustment on FOEEWM / FOEDEM / FOELDCPM.
template <typename R>
void cuadjtq0(const Thermo<R>& th, R retv, R psp, R& pt, R& pq) {
  const R zqp = R(1) / psp;
  for (int it = 0; it < 2; ++it) {
    R zqsat = th.foeewm(pt) * zqp;
    zqsat = std::min(R(0.5), zqsat);
    const R zcor = R(1) / (R(1) - retv * zqsat);
    zqsat = zqsat * zcor;
    const R zcond1 = (pq - zqsat) / (R(1) + zqsat * zcor * th.foedem(pt));
    pt = pt + th.foeldcpm(pt) * zcond1;
    pq = pq - zcond1;
  }
}

// CUADJTQS, KCALL = 0 (cuadjtqs.F90:123-134, 217-250): the phase (liquid above RTT, ice at or below) is chosen from
// the temperature on entry and kept for both steps.
template <typename R>
void cuadjtqs0(const Thermo<R>& th, R retv, R psp, R& pt, R& pq) {
  const bool liq = pt > th.rtt;
  const R z3es = liq ? th.r3les : th.r3ies, z4es = liq ? th.r4les : th.r4ies;
  const R z5alcp = liq ? th.r5alvcp : th.r5alscp, zaldcp = liq ? th.ralvdcp : th.ralsdcp;
  const R zqp = R(1) / psp;
  for (int it = 0; it < 2; ++it) {
    const R ztarg = pt;
    const R zfoeew = th.r2es * std::exp(z3es * (ztarg - th.rtt) / (ztarg - z4es));
    R zqsat = zqp * zfoeew;
    if (zqsat > R(0.5)) zqsat = R(0.5);
    const R zcor = R(1) / (R(1) - retv * zqsat);
    zqsat = zqsat * zcor;
    const R d = ztarg - z4es;
    const R z2s = z5alcp / (d * d);
    const R zcond1 = (pq - zqsat) / (R(1) + zqsat * zcor * z2s);
    pt = pt + zaldcp * zcond1;
    pq = pq - zcond1;
  }
}

template <typename R>
void cloudsc2(const cloudsc_params_t& p, const cloudsc2_params_t& p2, const Cloudsc2Args<R>& a) {
  const int kidia = a.kidia, kfdia = a.kfdia, klon = a.klon, klev = a.klev, ktdia = a.ktdia;
  const R ptsphy = a.ptsphy;
  const Thermo<R> th(p);
  auto i2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };
  const R RG = R(p.rg), RD = R(p.rd), RCPD = R(p.rcpd), RETV = R(p.retv), RTT = R(p.rtt), RLVTT = R(p.rlvtt);
  const R RLSTT = R(p.rlstt), RLMLT = R(p.rlmlt), R2ES = R(p.r2es), R3LES = R(p.r3les), R3IES = R(p.r3ies);
  const R R4LES = R(p.r4les), R4IES = R(p.r4ies), R5LES = R(p.r5les), R5IES = R(p.r5ies), RTICE = R(p.rtice);
  const R RKCONV = R(p.rkconv), RLMIN = R(p.rlmin), RPECONS = R(p.rpecons);
  const R RCLCRIT = R(p2.rclcrit), RLPTRC = R(p2.rlptrc), RVTMP2 = R(p2.rvtmp2);
  const bool LPHYLIN = p2.lphylin != 0, LEVAPLS2 = p2.levapls2 != 0, LDRAIN1D = a.ldrain1d;
  const R ZSCAL = R(0.9);  // :167

  // 1.1 constants (:243-253)
  const R ZCKCODTL = R(2.0) * RKCONV * ptsphy;
  const R ZCKCODTI = R(5.0) * RKCONV * ptsphy;
  const R ZCONS2 = R(1) / (ptsphy * RG);
  const R ZCONS3 = RLVTT / RCPD;
  const R ZMELTP2 = RTT + R(2.0);
  const R ZQTMST = R(1) / ptsphy;
  const R ZQMAX = R(0.5), ZEPS1 = R(1.E-12), ZEPS2 = R(1.E-10);

  const size_t n2 = (size_t)klon * klev;
  std::vector<R> ztp1(n2), zqp1(n2), zl(n2), zi(n2), zqc(n2), zqlwc(n2), zqiwc(n2), zdp(n2);
  std::vector<R> zlsdcp(n2), zlfdcp(n2), zlvdcp(n2), zrfreeze(n2), zcondl(n2), zcondi(n2), zevapr(n2), zevaps(n2);
  std::vector<R> zscalm(klev);
  const int KL = klon;
  std::vector<R> zrfl(KL), zsfl(KL), zrfln(KL), zsfln(KL), zgdp(KL), zdqdt(KL), zdtdt(KL), zdldt(KL), zdidt(KL);
  std::vector<R> zqcrit(KL), zcovpclr(KL), zcovptot(KL), zdqsdtemp(KL), zcorqs(KL), zdtgdp(KL), zqold(KL), zpp(KL);
  std::vector<R> zdq(KL), zqlim(KL), zqsat(KL), zfoeew(KL), zfwat(KL), ztrpaus(KL), zlude(KL);

  // 2.1 first guess (:263-270)
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      ztp1[k] = a.ptm1[k] + ptsphy * a.pgtent[k];
      zqp1[k] = a.pqm1[k] + ptsphy * a.pgtenq[k];
      zl[k] = a.pl[k] + ptsphy * a.pgtenl[k];
      zi[k] = a.pi[k] + ptsphy * a.pgteni[k];
    }
  for (int jk = 1; jk <= klev; ++jk) {  // (:272-289)
    zscalm[jk - 1] = ZSCAL * std::pow(std::max(a.ceta[jk - 1] - R(0.2), ZEPS1), R(0.2));
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      zdp[k] = a.paphp1[i2(jl, jk + 1)] - a.paphp1[k];
      const R zzz = R(1) / (RCPD + RCPD * RVTMP2 * zqp1[k]);
      zlfdcp[k] = RLMLT * zzz;
      zlsdcp[k] = RLSTT * zzz;
      zlvdcp[k] = RLVTT * zzz;
    }
  }
  // 2.2 initialisation (:297-320)
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      a.pclc[k] = R(0);
      zqc[k] = zqlwc[k] = zqiwc[k] = zrfreeze[k] = zcondl[k] = zcondi[k] = zevapr[k] = zevaps[k] = R(0);
      a.pcovptot[k] = R(0);
    }
  for (int jl = kidia; jl <= kfdia; ++jl) {
    const int i = jl - 1;
    zrfl[i] = zsfl[i] = R(0);
    a.pfplsl[i2(jl, 1)] = R(0);
    a.pfplsn[i2(jl, 1)] = R(0);
    zcovptot[i] = zcovpclr[i] = R(0);
  }
  // eta of the tropopause (:323-334)
  for (int jl = kidia; jl <= kfdia; ++jl) ztrpaus[jl - 1] = R(0.1);
  for (int jk = 1; jk <= klev - 1; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const bool llo1 = a.ceta[jk - 1] > R(0.1) && a.ceta[jk - 1] < R(0.4) && ztp1[i2(jl, jk)] > ztp1[i2(jl, jk + 1)];
      if (llo1) ztrpaus[jl - 1] = a.ceta[jk - 1];
    }

  // 3. level loop (:348-685)
  for (int jk = ktdia; jk <= klev; ++jk) {
    const R ceta = a.ceta[jk - 1];
    for (int jl = kidia; jl <= kfdia; ++jl) {  // 3.1 (:352-416)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      const R t = ztp1[k];
      R zesdp;
      if (LPHYLIN || LDRAIN1D) {
        const R zoealfaw = R(0.545) * (std::tanh(R(0.17) * (t - RLPTRC)) + R(1.0));
        R z3es, z4es;
        if (t < RTT) {
          zfwat[i] = zoealfaw;
          z3es = R3IES;
          z4es = R4IES;
        } else {
          zfwat[i] = R(1.0);
          z3es = R3LES;
          z4es = R4LES;
        }
        zfoeew[i] = R2ES * std::exp(z3es * (t - RTT) / (t - z4es));
        zesdp = zfoeew[i] / a.papp1[k];
        if (zesdp > ZQMAX) zesdp = ZQMAX;
      } else {
        zfwat[i] = th.foealfa(t);
        zfoeew[i] = th.foeewm(t);
        zesdp = zfoeew[i] / a.papp1[k];
      }
      const R dl = t - R4LES, di = t - R4IES;
      const R zfacw = R5LES / (dl * dl);
      const R zfaci = R5IES / (di * di);
      const R zfac = zfwat[i] * zfacw + (R(1.0) - zfwat[i]) * zfaci;
      const R zcor = R(1.0) / (R(1.0) - RETV * zesdp);
      zdqsdtemp[i] = zfac * zcor * a.pqs[k];
      zcorqs[i] = R(1.0) + ZCONS3 * zdqsdtemp[i];
      zqlim[i] = zqp1[k];
      if (zqp1[k] > a.pqs[k]) zqlim[i] = a.pqs[k];
      // critical relative humidity (:386-407)
      const R zeta3 = ztrpaus[i];
      const R zrh1 = R(1.0);
      const R s = (zeta3 - R(0.25)) / R(0.15);
      const R zrh2 = R(0.35) + R(0.14) * (s * s) + R(0.04) * std::min(zeta3 - R(0.25), R(0.0)) / R(0.15);
      const R zrh3 = R(1.0);
      const R zdeta2 = R(0.3);
      const R zdeta1 = R(0.09) + R(0.16) * (R(0.4) - zeta3) / R(0.3);
      R zcrh2 = R(0);
      if (ceta < zeta3) {
        zcrh2 = zrh3;
      } else if (ceta >= zeta3 && ceta < (zeta3 + zdeta2)) {
        zcrh2 = zrh3 + (zrh2 - zrh3) * ((ceta - zeta3) / zdeta2);
      } else if (ceta >= (zeta3 + zdeta2) && ceta < (R(1.0) - zdeta1)) {
        zcrh2 = zrh2;
      } else if (ceta >= (R(1.0) - zdeta1)) {
        zcrh2 = zrh1 + (zrh2 - zrh1) * std::sqrt((R(1.0) - ceta) / zdeta1);
      }
      const R zsupsat = t < RTICE ? R(1.8) - R(3.E-03) * t : R(1.0);  // (:409-413)
      zqsat[i] = a.pqs[k] * zsupsat;
      zqcrit[i] = zcrh2 * zqsat[i];
    }
    // uniform total-water distribution, Le Treut & Li (1990) (:419-436)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      const R zqt = zqp1[k] + zl[k] + zi[k];
      if (zqt <= zqcrit[i]) {
        a.pclc[k] = R(0.0);
        zqc[k] = R(0.0);
      } else if (zqt >= zqsat[i]) {
        a.pclc[k] = R(1.0);
        zqc[k] = (R(1.0) - zscalm[jk - 1]) * (zqsat[i] - zqcrit[i]);
      } else {
        const R zqpd = zqsat[i] - zqt;
        const R zqcd = zqsat[i] - zqcrit[i];
        a.pclc[k] = R(1.0) - std::sqrt(zqpd / (zqcd - zscalm[jk - 1] * (zqt - zqcrit[i])));
        zqc[k] = (zscalm[jk - 1] * zqpd + (R(1.0) - zscalm[jk - 1]) * zqcd) * (a.pclc[k] * a.pclc[k]);
      }
    }
    // convective component (:440-453)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zgdp[i] = RG / (a.paphp1[i2(jl, jk + 1)] - a.paphp1[k]);
      zlude[i] = a.plude[k] * ptsphy * zgdp[i];
      bool llo1 = false;
      if (jk < klev) llo1 = zlude[i] >= RLMIN && a.plu[i2(jl, jk + 1)] >= ZEPS2;
      if (llo1) {
        a.pclc[k] = a.pclc[k] + (R(1.0) - a.pclc[k]) * (R(1.0) - std::exp(-zlude[i] / a.plu[i2(jl, jk + 1)]));
        zqc[k] = zqc[k] + zlude[i];
      }
    }
    // compensating subsidence (:457-471)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      const R zfac1 = R(1.0) / (RD * ztp1[k]);
      const R zrho = a.papp1[k] * zfac1;
      const R zfac2 = R(1.0) / (a.papp1[k] - RETV * zfoeew[i]);
      const R zrodqsdp = -zrho * a.pqs[k] * zfac2;
      const R zldcp = zfwat[i] * zlvdcp[k] + (R(1.0) - zfwat[i]) * zlsdcp[k];
      const R zfac3 = R(1.0) / (R(1.0) + zldcp * zdqsdtemp[i]);
      const R dtdzmo = RG * (R(1.0) / RCPD - zldcp * zrodqsdp) * zfac3;
      const R zdqsdz = zdqsdtemp[i] * dtdzmo - RG * zrodqsdp;
      const R zfac4 = R(1.0) / zrho;
      const R zdqc = std::min(zdqsdz * (a.pmfu[k] + a.pmfd[k]) * ptsphy * zfac4, zqc[k]);
      zqc[k] = zqc[k] - zdqc;
    }
    // new liquid/ice and condensation rates (:475-481)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zqlwc[k] = zqc[k] * zfwat[i];
      zqiwc[k] = zqc[k] * (R(1.0) - zfwat[i]);
      zcondl[k] = (zqlwc[k] - zl[k]) * ZQTMST;
      zcondi[k] = (zqiwc[k] - zi[k]) * ZQTMST;
    }
    // precipitation overlap, maximum (:487-493)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (a.pclc[k] > zcovptot[i]) zcovptot[i] = a.pclc[k];
      zcovpclr[i] = zcovptot[i] - a.pclc[k];
      zcovpclr[i] = std::max(zcovpclr[i], R(0.0));
    }
    // melting of incoming snow (:499-511)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (zsfl[i] != R(0.0)) {
        const R zcons = ZCONS2 * zdp[k] / zlfdcp[k];
        const R zsnmlt = std::min(zsfl[i], zcons * std::max(R(0.0), (ztp1[k] - ZMELTP2)));
        zrfln[i] = zrfl[i] + zsnmlt;
        zsfln[i] = zsfl[i] - zsnmlt;
        ztp1[k] = ztp1[k] - zsnmlt / zcons;
      } else {
        zrfln[i] = zrfl[i];
        zsfln[i] = zsfl[i];
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:513-599)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      R zprr = R(0), zprs = R(0);
      if (a.pclc[k] > ZEPS2) {  // rain from cloud liquid (:517-529)
        const R zlcrit = (LEVAPLS2 || LDRAIN1D) ? R(1.9) * RCLCRIT : RCLCRIT * R(2.);
        const R zcldl = zqlwc[k] / a.pclc[k];
        const R r = zcldl / zlcrit;
        const R zd = ZCKCODTL * (R(1.0) - std::exp(-(r * r)));
        const R zlnew = a.pclc[k] * zcldl * std::exp(-zd);
        zprr = zqlwc[k] - zlnew;
        zqlwc[k] = zqlwc[k] - zprr;
      }
      if (a.pclc[k] > ZEPS2) {  // snow from cloud ice (:533-545)
        const R zlcrit = (LEVAPLS2 || LDRAIN1D) ? R(1.E-04) : RCLCRIT * R(2.);
        const R zcldi = zqiwc[k] / a.pclc[k];
        const R r = zcldi / zlcrit;
        const R zd = ZCKCODTI * std::exp(R(0.025) * (ztp1[k] - RTT)) * (R(1.0) - std::exp(-(r * r)));
        const R zinew = a.pclc[k] * zcldi * std::exp(-zd);
        zprs = zqiwc[k] - zinew;
        zqiwc[k] = zqiwc[k] - zprs;
      }
      const R zdr = ZCONS2 * zdp[k] * (zprr + zprs);  // (:549)
      R zfwatr;
      if (ztp1[k] < RTT) {  // (:553-559)
        zrfreeze[k] = ZCONS2 * zdp[k] * zprr;
        zfwatr = R(0.0);
      } else {
        zfwatr = R(1.0);
      }
      const R zrn = zfwatr * zdr;
      const R zsn = (R(1.0) - zfwatr) * zdr;
      zrfln[i] = zrfln[i] + zrn;
      zsfln[i] = zsfln[i] + zsn;
      // precipitation evaporation (:568-597)
      const R zprtot = zrfln[i] + zsfln[i];
      const bool llo2 = zprtot > ZEPS2 && zcovpclr[i] > ZEPS2 && (LEVAPLS2 || LDRAIN1D);
      if (llo2) {
        R zpreclr = zprtot * zcovpclr[i] / zcovptot[i];
        const R om = R(1.0) - a.pclc[k];
        const R zqe = a.pqs[k] - (a.pqs[k] - zqlim[i]) * zcovpclr[i] / (om * om);
        const R zbeta = RG * RPECONS *
                        std::pow(std::sqrt(a.papp1[k] / a.paphp1[i2(jl, klev + 1)]) / R(5.09E-3) * zpreclr / zcovpclr[i],
                                 R(0.5777));
        const R zb = ptsphy * zbeta * (a.pqs[k] - zqe) / (R(1.0) + zbeta * ptsphy * zcorqs[i]);
        zdtgdp[i] = ptsphy * RG / (a.paphp1[i2(jl, jk + 1)] - a.paphp1[k]);
        R zdpr = zcovpclr[i] * zb / zdtgdp[i];
        zdpr = std::min(zdpr, zpreclr);
        zpreclr = zpreclr - zdpr;
        if (zpreclr <= R(0.0)) zcovptot[i] = a.pclc[k];
        a.pcovptot[k] = zcovptot[i];
        zevapr[k] = zdpr * zrfln[i] / zprtot;
        zrfln[i] = zrfln[i] - zevapr[k];
        zevaps[k] = zdpr * zsfln[i] / zprtot;
        zsfln[i] = zsfln[i] - zevaps[k];
      }
    }
    // T and q tendencies, first guess, CUADJTQ (:608-632)
    auto tendencies = [&](int jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdqdt[i] = -(zcondl[k] + zcondi[k]) + (a.plude[k] + zevapr[k] + zevaps[k]) * zgdp[i];
      zdtdt[i] = zlvdcp[k] * zcondl[k] + zlsdcp[k] * zcondi[k] -
                 (zlvdcp[k] * zevapr[k] + zlsdcp[k] * zevaps[k] +
                  a.plude[k] * (zfwat[i] * zlvdcp[k] + (R(1.0) - zfwat[i]) * zlsdcp[k]) -
                  (zlsdcp[k] - zlvdcp[k]) * zrfreeze[k]) *
                     zgdp[i];
    };
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      tendencies(jl);
      ztp1[k] = ztp1[k] + ptsphy * zdtdt[i];
      zqp1[k] = zqp1[k] + ptsphy * zdqdt[i];
      zpp[i] = a.papp1[k];
      zqold[i] = zqp1[k];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // clipping of final qv (:636-643)
      const size_t k = i2(jl, jk);
      if (LPHYLIN || LDRAIN1D)
        cuadjtqs0(th, RETV, zpp[jl - 1], ztp1[k], zqp1[k]);
      else
        cuadjtq0(th, RETV, zpp[jl - 1], ztp1[k], zqp1[k]);
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:645-665)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdq[i] = std::max(R(0.0), zqold[i] - zqp1[k]);
      const R zdr2 = ZCONS2 * zdp[k] * zdq[i];
      R zrfreeze2, zfwatr;
      if (ztp1[k] < RTT) {
        zrfreeze2 = zfwat[i] * zdr2;
        zfwatr = R(0.0);
      } else {
        zrfreeze2 = R(0.0);
        zfwatr = R(1.0);
      }
      const R zrn = zfwatr * zdr2;
      const R zsn = (R(1.0) - zfwatr) * zdr2;
      zcondl[k] = zcondl[k] + zfwatr * zdq[i] * ZQTMST;
      zcondi[k] = zcondi[k] + (R(1.0) - zfwatr) * zdq[i] * ZQTMST;
      zrfln[i] = zrfln[i] + zrn;
      zsfln[i] = zsfln[i] + zsn;
      zrfreeze[k] = zrfreeze[k] + zrfreeze2;
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:667-686)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      tendencies(jl);
      zdldt[i] = (zqlwc[k] - zl[k]) * ZQTMST;
      zdidt[i] = (zqiwc[k] - zi[k]) * ZQTMST;
      a.ptenq[k] = zdqdt[i];
      a.ptent[k] = zdtdt[i];
      a.ptenl[k] = zdldt[i];
      a.pteni[k] = zdidt[i];
      a.pfplsl[i2(jl, jk + 1)] = zrfln[i];
      a.pfplsn[i2(jl, jk + 1)] = zsfln[i];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {
      zrfl[jl - 1] = zrfln[jl - 1];
      zsfl[jl - 1] = zsfln[jl - 1];
    }
  }
  // enthalpy fluxes (:694-699)
  for (int jk = 1; jk <= klev + 1; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      a.pfhpsl[i2(jl, jk)] = -a.pfplsl[i2(jl, jk)] * RLVTT;
      a.pfhpsn[i2(jl, jk)] = -a.pfplsn[i2(jl, jk)] * RLSTT;
    }
}

// CLOUDST (cloudst.F90:9-641), the diagnostic scheme of CLOUD_S_LAYER when LEPCLD2 is false.  The arguments are
// CLOUDSC2's without PL/PI/PMFU/PMFD/PTENL/PGTENL/PTENI/PGTENI, plus the diagnostic water paths.
template <typename R>
struct CloudstArgs {
  int kidia, kfdia, klon, ktdia, klev;
  bool ldrain1d;
  R ptsphy;
  const R* ceta;
  const R *paphp1, *papp1, *pqm1, *pqs, *ptm1, *plude, *plu;
  R* ptent;
  const R* pgtent;
  R* ptenq;
  const R* pgtenq;
  R *pclc, *pqiwp, *pqlwp, *pfplsl, *pfplsn, *pfhpsl, *pfhpsn, *pcovptot;
};

template <typename R>
void cloudst(const cloudsc_params_t& p, const cloudsc2_params_t& p2, const CloudstArgs<R>& a) {
  const int kidia = a.kidia, kfdia = a.kfdia, klon = a.klon, klev = a.klev, ktdia = a.ktdia;
  const R ptsphy = a.ptsphy;
  const Thermo<R> th(p);
  auto i2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };
  const R RG = R(p.rg), RCPD = R(p.rcpd), RETV = R(p.retv), RTT = R(p.rtt), RLVTT = R(p.rlvtt);
  const R RLSTT = R(p.rlstt), RLMLT = R(p.rlmlt), R2ES = R(p.r2es), R3LES = R(p.r3les), R3IES = R(p.r3ies);
  const R R4LES = R(p.r4les), R4IES = R(p.r4ies), R5LES = R(p.r5les), R5IES = R(p.r5ies);
  const R RKCONV = R(p.rkconv), RLMIN = R(p.rlmin);
  const R RCLCRIT = R(p2.rclcrit), RLPTRC = R(p2.rlptrc), RVTMP2 = R(p2.rvtmp2);
  const bool LPHYLIN = p2.lphylin != 0, LEVAPLS2 = p2.levapls2 != 0, LDRAIN1D = a.ldrain1d;
  const R ZSCAL = R(0.9);  // :165

  // 1.1 (:213-221)
  const R ZCKCODT = R(2.0) * RKCONV * ptsphy;
  const R ZCONS2 = R(1.0) / (ptsphy * RG);
  const R ZCONS3 = R(2.5E6) / R(1005.);
  const R ZMELTP2 = RTT + R(2.0);
  const R ZPECONS = R(1.E-5);
  const R ZQMAX = R(0.5), ZEPS1 = R(1.E-12), ZEPS2 = R(1.E-10);

  const size_t n2 = (size_t)klon * klev;
  std::vector<R> ztp1(n2), zqp1(n2), zcrh(n2), zdp(n2), zlfdcp(n2), zlsdcp(n2), zlvdcp(n2), zqlwc(n2);
  std::vector<R> zrfreeze(n2), zfwat(n2), zlude(n2), zscalm(klev);
  const int KL = klon;
  std::vector<R> ztrpaus(KL), zrfl(KL), zsfl(KL), zrfln(KL), zsfln(KL), zcovptot(KL), zcovpclr(KL), zdr(KL);
  std::vector<R> zdqsdtemp(KL), zcorqs(KL), zqlim(KL), zqcrit(KL), zgdp(KL), zdtgdp(KL), zdqdt(KL), zdtdt(KL);
  std::vector<R> zpp(KL), zqold(KL), zdq(KL);

  for (int jk = 1; jk <= klev; ++jk)  // (:225-230)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      ztp1[i2(jl, jk)] = a.ptm1[i2(jl, jk)] + ptsphy * a.pgtent[i2(jl, jk)];
      zqp1[i2(jl, jk)] = a.pqm1[i2(jl, jk)] + ptsphy * a.pgtenq[i2(jl, jk)];
    }
  for (int jl = kidia; jl <= kfdia; ++jl) ztrpaus[jl - 1] = R(0.1);  // (:234-247)
  for (int jk = 1; jk <= klev - 1; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl)
      if (a.ceta[jk - 1] > R(0.1) && a.ceta[jk - 1] < R(0.4) && ztp1[i2(jl, jk)] > ztp1[i2(jl, jk + 1)])
        ztrpaus[jl - 1] = a.ceta[jk - 1];
  // 2.1 critical relative humidity (:263-306)
  for (int jk = 1; jk <= klev; ++jk) {
    const R ceta = a.ceta[jk - 1];
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const R zeta3 = ztrpaus[jl - 1];
      const R zrh1 = R(0.9), zrh2 = R(0.5), zrh3 = R(1.0), zeta1 = R(0.75);
      const R zdeta1 = R(1.0) - zeta1;
      R& c = zcrh[i2(jl, jk)];
      if (ceta < zeta3) {
        c = zrh3 + (zrh2 - zrh3) * ceta / zeta3;
      } else if (ceta >= zeta3 && ceta < zeta1) {
        c = zrh2;
      } else if (ceta >= zeta1) {
        c = zrh1 + (zrh2 - zrh1) * std::sqrt((R(1.0) - ceta) / zdeta1);
      }
    }
    zscalm[jk - 1] = ZSCAL * std::pow(std::max(ceta - R(0.2), ZEPS1), R(0.2));
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      zdp[k] = a.paphp1[i2(jl, jk + 1)] - a.paphp1[k];
      const R zzz = R(1.0) / (RCPD + RCPD * RVTMP2 * zqp1[k]);
      zlfdcp[k] = RLMLT * zzz;
      zlsdcp[k] = RLSTT * zzz;
      zlvdcp[k] = RLVTT * zzz;
    }
  }
  // 2.2 (:311-330)
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      a.pclc[k] = a.pqiwp[k] = a.pqlwp[k] = R(0.0);
      zqlwc[k] = zrfreeze[k] = R(0.0);
      a.pcovptot[k] = R(0.0);
    }
  for (int jl = kidia; jl <= kfdia; ++jl) {
    const int i = jl - 1;
    zrfl[i] = zsfl[i] = R(0.0);
    a.pfplsl[i2(jl, 1)] = a.pfplsn[i2(jl, 1)] = R(0.0);
    zcovptot[i] = zcovpclr[i] = R(0.0);
  }
  // 3. level loop (:340-623)
  for (int jk = ktdia; jk <= klev; ++jk) {
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:342-385)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdr[i] = R(0.0);
      const R t = ztp1[k];
      R zalfaw, zesdp;
      if (LPHYLIN || LDRAIN1D) {
        const R zoealfaw = R(0.545) * (std::tanh(R(0.17) * (t - RLPTRC)) + R(1.0));
        R z3es, z4es;
        if (t < RTT) {
          zalfaw = zoealfaw;
          z3es = R3IES;
          z4es = R4IES;
        } else {
          zalfaw = R(1.0);
          z3es = R3LES;
          z4es = R4LES;
        }
        const R zfoeew = R2ES * std::exp(z3es * (t - RTT) / (t - z4es));
        zesdp = zfoeew / a.papp1[k];
        if (zesdp > ZQMAX) zesdp = ZQMAX;
      } else {
        zalfaw = th.foealfa(t);
        zesdp = th.foeewm(t) / a.papp1[k];
      }
      const R dl = t - R4LES, di = t - R4IES;
      const R zfacw = R5LES / (dl * dl);
      const R zfaci = R5IES / (di * di);
      const R zfac = zalfaw * zfacw + (R(1.0) - zalfaw) * zfaci;
      const R zcor = R(1.0) / (R(1.0) - RETV * zesdp);
      zdqsdtemp[i] = zfac * zcor * a.pqs[k];
      zcorqs[i] = R(1.0) + ZCONS3 * zdqsdtemp[i];
      zqlim[i] = zqp1[k];
      if (zqp1[k] > a.pqs[k]) zqlim[i] = a.pqs[k];
      zqcrit[i] = zcrh[k] * a.pqs[k];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // layer cloud amounts (:387-403)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (zqp1[k] <= zqcrit[i]) {
        a.pclc[k] = R(0.0);
        zqlwc[k] = R(0.0);
      } else if (zqp1[k] >= a.pqs[k]) {
        a.pclc[k] = R(1.0);
        zqlwc[k] = (R(1.0) - zscalm[jk - 1]) * (a.pqs[k] - zqcrit[i]);
      } else {
        const R zqpd = a.pqs[k] - zqp1[k];
        const R zqcd = a.pqs[k] - zqcrit[i];
        a.pclc[k] = R(1.0) - std::sqrt(zqpd / (zqcd - zscalm[jk - 1] * (zqp1[k] - zqcrit[i])));
        zqlwc[k] = (zscalm[jk - 1] * zqpd + (R(1.0) - zscalm[jk - 1]) * zqcd) * (a.pclc[k] * a.pclc[k]);
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // convective cloud (:407-424)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zgdp[i] = RG / (a.paphp1[i2(jl, jk + 1)] - a.paphp1[k]);
      zlude[k] = a.plude[k] * ptsphy * zgdp[i];
      bool llo1 = false;
      if (jk < klev) llo1 = zlude[k] >= RLMIN && a.plu[i2(jl, jk + 1)] >= ZEPS2;
      if (llo1) {
        a.pclc[k] = a.pclc[k] + (R(1.0) - a.pclc[k]) * (R(1.0) - std::exp(-zlude[k] / a.plu[i2(jl, jk + 1)]));
        zqlwc[k] = zqlwc[k] + zlude[k];
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:428-434)
      const int i = jl - 1;
      zcovptot[i] = std::max(zcovptot[i], a.pclc[i2(jl, jk)]);
      zcovpclr[i] = zcovptot[i] - a.pclc[i2(jl, jk)];
      zcovpclr[i] = std::max(zcovpclr[i], R(0.0));
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // melting of incoming snow (:448-460)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      if (zsfl[i] != R(0.0)) {
        const R zcons = ZCONS2 * zdp[k] / zlfdcp[k];
        const R zsnmlt = std::min(zsfl[i], zcons * std::max(R(0.0), (ztp1[k] - ZMELTP2)));
        zrfln[i] = zrfl[i] + zsnmlt;
        zsfln[i] = zsfl[i] - zsnmlt;
        ztp1[k] = ztp1[k] - zsnmlt / zcons;
      } else {
        zrfln[i] = zrfl[i];
        zsfln[i] = zsfl[i];
      }
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // precipitation (:462-542)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      R zpr;
      if (a.pclc[k] > ZEPS2) {
        const R zlcrit = (LEVAPLS2 || LDRAIN1D) ? RCLCRIT : RCLCRIT * R(2.);
        const R zcld = zqlwc[k] / a.pclc[k];
        const R r = zcld / zlcrit;
        const R zd = ZCKCODT * (R(1.0) - std::exp(-(r * r)));
        const R zlnew = a.pclc[k] * zcld * std::exp(-zd);
        zpr = zqlwc[k] - zlnew;
        zqlwc[k] = zqlwc[k] - zpr;
      } else {
        zpr = R(0.0);
      }
      zdr[i] = zdr[i] + ZCONS2 * zdp[k] * zpr;
      if (LPHYLIN || LDRAIN1D) {
        const R zoealfa = R(0.545) * (std::tanh(R(0.17) * (ztp1[k] - RLPTRC)) + R(1.0));
        zfwat[k] = ztp1[k] < RTT ? zoealfa : R(1.0);
      } else {
        zfwat[k] = th.foealfa(ztp1[k]);
      }
      R zfwatr;
      if (ztp1[k] < RTT) {
        zrfreeze[k] = zfwat[k] * zdr[i];
        zfwatr = R(0.0);
      } else {
        zfwatr = R(1.0);
      }
      zrfln[i] = zrfln[i] + zfwatr * zdr[i];
      zsfln[i] = zsfln[i] + (R(1.0) - zfwatr) * zdr[i];
      const R zprtot = zrfln[i] + zsfln[i];
      const bool llo2 = zprtot > ZEPS2 && zcovpclr[i] > ZEPS2 && (LEVAPLS2 || LDRAIN1D);
      if (llo2) {
        R zpreclr = zprtot * zcovpclr[i] / zcovptot[i];
        const R om = R(1.0) - a.pclc[k];
        const R zqe = a.pqs[k] - (a.pqs[k] - zqlim[i]) * zcovpclr[i] / (om * om);
        const R zbeta = RG * ZPECONS *
                        std::pow(std::sqrt(a.papp1[k] / a.paphp1[i2(jl, klev + 1)]) / R(5.09E-3) * zpreclr / zcovpclr[i],
                                 R(0.5777));
        const R zb = ptsphy * zbeta * (a.pqs[k] - zqe) / (R(1.0) + zbeta * ptsphy * zcorqs[i]);
        zdtgdp[i] = ptsphy * RG / (a.paphp1[i2(jl, jk + 1)] - a.paphp1[k]);
        R zdpr = zcovpclr[i] * zb / zdtgdp[i];
        zdpr = std::min(zdpr, zpreclr);
        zpreclr = zpreclr - zdpr;
        if (zpreclr <= R(0.0)) zcovptot[i] = a.pclc[k];
        a.pcovptot[k] = zcovptot[i];
        const R zdprsfl = zdpr * zrfln[i] / zprtot;
        zrfln[i] = zrfln[i] - zdprsfl;
        const R zdpssfl = zdpr * zsfln[i] / zprtot;
        zsfln[i] = zsfln[i] - zdpssfl;
      }
      a.pqlwp[k] = zfwat[k] * zqlwc[k];
      a.pqiwp[k] = (R(1.0) - zfwat[k]) * zqlwc[k];
    }
    // tendencies from the flux divergence, first guess, CUADJTQ (:546-583)
    auto tendencies = [&](int jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdqdt[i] = -((zrfln[i] - zrfl[i]) + (zsfln[i] - zsfl[i])) * (RG / zdp[k]) + a.plude[k] * zgdp[i];
      zdtdt[i] = (zlvdcp[k] * (zrfln[i] - zrfl[i]) + zlsdcp[k] * (zsfln[i] - zsfl[i])) * (RG / zdp[k]) -
                 a.plude[k] * zgdp[i] * (zfwat[k] * zlvdcp[k] + (R(1.0) - zfwat[k]) * zlsdcp[k]) +
                 (zlsdcp[k] - zlvdcp[k]) * zrfreeze[k] * zgdp[i];
    };
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      tendencies(jl);
      ztp1[k] = ztp1[k] + ptsphy * zdtdt[i];
      zqp1[k] = zqp1[k] + ptsphy * zdqdt[i];
      zpp[i] = a.papp1[k];
      zqold[i] = zqp1[k];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      if (LPHYLIN || LDRAIN1D)
        cuadjtqs0(th, RETV, zpp[jl - 1], ztp1[k], zqp1[k]);
      else
        cuadjtq0(th, RETV, zpp[jl - 1], ztp1[k], zqp1[k]);
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:585-600)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      zdq[i] = std::max(R(0.0), zqold[i] - zqp1[k]);
      zdr[i] = ZCONS2 * zdp[k] * zdq[i];
      R zrfreeze2, zfwatr;
      if (ztp1[k] < RTT) {
        zrfreeze2 = zfwat[k] * zdr[i];
        zfwatr = R(0.0);
      } else {
        zrfreeze2 = R(0.0);
        zfwatr = R(1.0);
      }
      zrfln[i] = zrfln[i] + zfwatr * zdr[i];
      zsfln[i] = zsfln[i] + (R(1.0) - zfwatr) * zdr[i];
      zrfreeze[k] = zrfreeze[k] + zrfreeze2;
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {  // (:602-616)
      const int i = jl - 1;
      const size_t k = i2(jl, jk);
      tendencies(jl);
      a.ptenq[k] = zdqdt[i];
      a.ptent[k] = zdtdt[i];
      a.pfplsl[i2(jl, jk + 1)] = zrfln[i];
      a.pfplsn[i2(jl, jk + 1)] = zsfln[i];
    }
    for (int jl = kidia; jl <= kfdia; ++jl) {
      zrfl[jl - 1] = zrfln[jl - 1];
      zsfl[jl - 1] = zsfln[jl - 1];
    }
  }
  for (int jk = 1; jk <= klev + 1; ++jk)  // enthalpy fluxes (:629-634)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      a.pfhpsl[i2(jl, jk)] = -a.pfplsl[i2(jl, jk)] * RLVTT;
      a.pfhpsn[i2(jl, jk)] = -a.pfplsn[i2(jl, jk)] * RLSTT;
    }
}

// SATUR, KFLAG = 2 (satur.F90:110-141), levels KTDIA..KLEV.
template <typename R>
void satur2(const cloudsc_params_t& p, bool ldphylin, int kidia, int kfdia, int klon, int ktdia, int klev,
            const R* paprsf, const R* pt, R* pqsat) {
  const Thermo<R> th(p);
  const R RETV = R(p.retv), RTT = R(p.rtt), R2ES = R(p.r2es), R3LES = R(p.r3les), R3IES = R(p.r3ies);
  const R R4LES = R(p.r4les), R4IES = R(p.r4ies);
  const R ZQMAX = R(0.5);
  for (int jk = ktdia; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = (size_t)(jk - 1) * klon + (jl - 1);
      if (ldphylin) {
        const R ztarg = pt[k];
        const R zalfa = th.foealfa(ztarg);
        const R zfoeewl = R2ES * std::exp(R3LES * (ztarg - RTT) / (ztarg - R4LES));
        const R zfoeewi = R2ES * std::exp(R3IES * (ztarg - RTT) / (ztarg - R4IES));
        const R zfoeew = zalfa * zfoeewl + (R(1.0) - zalfa) * zfoeewi;
        R zqs = zfoeew / paprsf[k];
        if (zqs > ZQMAX) zqs = ZQMAX;
        const R zcor = R(1.0) / (R(1.0) - RETV * zqs);
        pqsat[k] = zqs * zcor;
      } else {
        const R zew = th.foeewm(pt[k]);
        R zqs = zew / paprsf[k];
        zqs = std::min(ZQMAX, zqs);
        const R zcor = R(1.0) / (R(1.0) - RETV * zqs);
        pqsat[k] = zqs * zcor;
      }
    }
}

// The arrays of CLOUD_S_LAYER's LEPCLD2 branch beyond CLOUDSC2's (NULL where the caller has no such field).
template <typename R>
struct LayerArgs {
  const R* pa;
  R *ptena, *ptenr, *ptens, *pqsat, *pprectype, *pfzra, *pqiwp, *pqlwp;
};

// CLOUD_S_LAYER (cloud_s_layer.F90:112-176): SATUR at T + TSPHY*TENDENCY%T into PQSAT, CLOUDSC2 (LEPCLD2) or
// CLOUDST with it, then the cloud-fraction tendency (and with CLOUDST the liquid and ice tendencies from the
// diagnostic water), zero rain and snow tendencies and zero PPRECTYPE/PFZRA.  `a` carries the CLOUDSC2 arguments
// (CLOUDST reads its subset); its PQS is the PQSAT array filled here and its PCLC is ZPATMP.
template <typename R>
void cloud_s_layer(const cloudsc_params_t& p, const cloudsc2_params_t& p2, Cloudsc2Args<R> a, const LayerArgs<R>& l) {
  const int kidia = a.kidia, kfdia = a.kfdia, klon = a.klon, klev = a.klev;
  auto i2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };
  std::vector<R> z_t1((size_t)klon * klev, R(0)), zpqsat((size_t)klon * klev, R(0));
  std::vector<R> zpatmp((size_t)klon * klev, R(0));
  for (int jk = 1; jk <= klev; ++jk)  // (:113-115)
    for (int jl = kidia; jl <= kfdia; ++jl) z_t1[i2(jl, jk)] = a.ptm1[i2(jl, jk)] + a.ptsphy * a.pgtent[i2(jl, jk)];
  satur2<R>(p, p2.lphylin != 0, kidia, kfdia, klon, a.ktdia, klev, a.papp1, z_t1.data(), zpqsat.data());
  a.pqs = zpqsat.data();
  R* clc_out = a.pclc;
  a.pclc = zpatmp.data();
  std::vector<R> zplsmp((size_t)klon * klev, R(0)), zpltmp((size_t)klon * klev, R(0));
  if (p2.lepcld2) {
    cloudsc2<R>(p, p2, a);
  } else {
    const CloudstArgs<R> s{a.kidia,  a.kfdia, a.klon,  a.ktdia, a.klev,  a.ldrain1d, a.ptsphy, a.ceta,
                           a.paphp1, a.papp1, a.pqm1,  a.pqs,   a.ptm1,  a.plude,    a.plu,    a.ptent,
                           a.pgtent, a.ptenq, a.pgtenq, a.pclc, zplsmp.data(), zpltmp.data(), a.pfplsl, a.pfplsn,
                           a.pfhpsl, a.pfhpsn, a.pcovptot};
    cloudst<R>(p, p2, s);
  }
  const R zcons = R(1) / a.ptsphy;  // (:154)
  for (int jk = 1; jk <= klev; ++jk)
    for (int jl = kidia; jl <= kfdia; ++jl) {
      const size_t k = i2(jl, jk);
      if (l.ptena) l.ptena[k] = zcons * (zpatmp[k] - l.pa[k]);
      if (clc_out) clc_out[k] = zpatmp[k];
      if (l.pqsat && jk >= a.ktdia) l.pqsat[k] = zpqsat[k];
      if (!p2.lepcld2) {  // (:159-163)
        a.ptenl[k] = zcons * (zpltmp[k] - a.pl[k]);
        a.pteni[k] = zcons * (zplsmp[k] - a.pi[k]);
        if (l.pqlwp) l.pqlwp[k] = zpltmp[k];
        if (l.pqiwp) l.pqiwp[k] = zplsmp[k];
      }
    }
  for (int jk = 1; jk <= klev; ++jk)  // (:164-165), every column
    for (int jl = 1; jl <= klon; ++jl) {
      if (l.ptenr) l.ptenr[i2(jl, jk)] = R(0);
      if (l.ptens) l.ptens[i2(jl, jk)] = R(0);
    }
  for (int jl = kidia; jl <= kfdia; ++jl) {  // (:168-169)
    if (l.pprectype) l.pprectype[jl - 1] = R(0);
    if (l.pfzra) l.pfzra[jl - 1] = R(0);
  }
}

}  // namespace oracle
