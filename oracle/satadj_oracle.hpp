// CPU oracle of CLOUD_SATADJ -- TEST INFRASTRUCTURE ONLY.
//
// The checker of the gfx950 saturation-adjustment kernel.  Only the tests load it (through
// liboracle_cloudsc.so); the product never links or calls it.
//
// Restates the default path (LSPP = .FALSE., budgets off) of
//   ifs-source/arpifs/phys_ec/cloud_satadj.F90:13-945     SUBROUTINE CLOUD_SATADJ
//   ifs-source/arpifs/phys_ec/cloud_satadj.F90:948-1017   FUNCTION QSATMIXADJ
//   ifs-source/arpifs/phys_ec/cloud_supersatcheck.F90     (KFTLIQICE = 2; only L+I is used)
// in the reference's loop structure (vector loops over JL inside the level loop).  Template parameters:
// R = JPRB, RD = type of the Newton saturation adjustment (QSATMIXADJ) and of CLOUD_SUPERSATCHECK.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "cloudsc_oracle.hpp"

namespace oracle {

// QSATMIXADJ: two Newton iterations towards saturation over the mixed phase (cloud_satadj.F90:988-1014).
template <typename RD>
RD qsatmixadj(const Thermo<RD>& th, RD retv, RD pt, RD pq, RD pp) {
  RD zt = pt, zq = pq;
  const RD zp_r = RD(1) / pp;
  for (int ji = 0; ji < 2; ++ji) {
    const RD zfoeet = std::min(th.foeewm(zt) * zp_r, RD(0.5));
    const RD zcor = RD(1) / (RD(1) - retv * zfoeet);
    const RD zqsat = zfoeet * zcor;
    const RD zcorqs = RD(1) + zqsat * zcor * th.foedem(zt);
    const RD zcond = (zq - zqsat) / zcorqs;
    zt = zt + th.foeldcpm(zt) * zcond;
    zq = zq - zcond;
  }
  return zq;
}

// One block, Fortran order: (KLON,KLEV) arrays, PAPH (KLON,KLEV+1), PLUDELI (KLON,KLEV,4).
template <typename R>
struct SatadjArgs {
  int kidia, kfdia, klon, klev;
  R ptsphy;
  const R *pap, *paph, *pt, *pq, *pa, *pl, *pi;
  const R *tcml_t, *tcml_q, *tcml_a, *tcml_l, *tcml_i;
  const R *phrsw, *phrlw, *tvdf_t, *tvdf_q, *pludeli, *pvervel;
  R *tloc_t, *tloc_q, *tloc_a, *tloc_l, *tloc_i;
};

template <typename R, typename RD>
void satadj(const cloudsc_params_t& p, const SatadjArgs<R>& a) {
  const int klon = a.klon, klev = a.klev;
  auto I2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };                  // (JL,JK)
  auto IH = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };                  // PAPH
  auto I3 = [klon, klev](int jl, int jk, int s) { return ((size_t)(s - 1) * klev + (jk - 1)) * klon + (jl - 1); };
  const Thermo<R> th(p);
  const Thermo<RD> thd(p);
  const R rg = R(p.rg), rd = R(p.rd), rcpd = R(p.rcpd), retv = R(p.retv);
  const R ralvdcp = R(p.ralvdcp), ralsdcp = R(p.ralsdcp);
  const R ramid = R(p.ramid), ramin = R(p.ramin), rlmin = R(p.rlmin), rthomo = R(p.rthomo);
  const R ptsphy = a.ptsphy;
  const int ncldtop = p.ncldtop;
  // :291-293 (ZEPSILON at :287 is unused on this path)
  const R zqtmst = R(1) / ptsphy;
  const R zrdcp = rd / rcpd;
  const R zepsec = R(1.e-14);
  const int n = klon;
  std::vector<R> zt(n + 1), zq(n + 1), za(n + 1), zl(n + 1), zi(n + 1), zli(n + 1), zliqfrac(n + 1), zicefrac(n + 1);
  std::vector<R> zrhc(n + 1), zqsmix(n + 1), zqsliqk(n + 1), zqsmix_upd(n + 1), zcorqsmix_upd(n + 1);
  std::vector<R> zdtforce(n + 1), zdqforce(n + 1), zdqs(n + 1), zttmp(n + 1), zqtmp(n + 1);
  std::vector<R> env_l(n + 1), env_i(n + 1), zacond(n + 1), cld_l(n + 1), cld_i(n + 1), ev_l(n + 1), ev_i(n + 1);
  std::vector<R> zaevap(n + 1);
  std::vector<RD> t_upd(n + 1), q_upd(n + 1), a_upd(n + 1), t_adj(n + 1), q_adj(n + 1), a_adj(n + 1), l_adj(n + 1),
      i_adj(n + 1);
  std::vector<R> pp(n + 1), psurf(n + 1);

  // :340-348 zero tendencies above NCLDTOP
  for (int jk = 1; jk <= ncldtop - 1; ++jk)
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      a.tloc_t[I2(jl, jk)] = R(0);
      a.tloc_q[I2(jl, jk)] = R(0);
      a.tloc_a[I2(jl, jk)] = R(0);
      a.tloc_l[I2(jl, jk)] = R(0);
      a.tloc_i[I2(jl, jk)] = R(0);
    }

  for (int jk = ncldtop; jk <= klev; ++jk) {
    // :360-368 updated state
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      zt[jl] = a.pt[I2(jl, jk)] + ptsphy * a.tcml_t[I2(jl, jk)];
      zq[jl] = a.pq[I2(jl, jk)] + ptsphy * a.tcml_q[I2(jl, jk)];
      za[jl] = a.pa[I2(jl, jk)] + ptsphy * a.tcml_a[I2(jl, jk)];
      zl[jl] = a.pl[I2(jl, jk)] + ptsphy * a.tcml_l[I2(jl, jk)];
      zi[jl] = a.pi[I2(jl, jk)] + ptsphy * a.tcml_i[I2(jl, jk)];
      za[jl] = std::max(R(0), std::min(R(1), za[jl]));
    }
    // :384-416 forcing terms (ZPVERVEL = PVERVEL without SPP, :378)
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      const R zdp = a.paph[IH(jl, jk + 1)] - a.paph[IH(jl, jk)];
      const R zgdp = rg / zdp;
      const R zdtgdp = ptsphy * zgdp;
      const R zdtdp = zrdcp * a.pt[I2(jl, jk)] / a.pap[I2(jl, jk)];
      const R zdpmxdt = zdp * zqtmst;
      R zwtot = a.pvervel[I2(jl, jk)];
      zwtot = std::min(zdpmxdt, std::max(-zdpmxdt, zwtot));
      const R zzzdt = a.phrsw[I2(jl, jk)] + a.phrlw[I2(jl, jk)];
      const R zdtdiab = std::min(zdpmxdt * zdtdp, std::max(-zdpmxdt * zdtdp, zzzdt)) * ptsphy;
      zdtforce[jl] = zdtdp * zwtot * ptsphy + zdtdiab + a.pludeli[I3(jl, jk, 4)] * zdtgdp +
                     a.tvdf_t[I2(jl, jk)] * ptsphy;
      zdqforce[jl] = a.pludeli[I3(jl, jk, 3)] * zdtgdp + a.tvdf_q[I2(jl, jk)] * ptsphy;
    }
    // :418-539 derived variables and saturation functions
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      const R t0 = a.pt[I2(jl, jk)], pap = a.pap[I2(jl, jk)];
      const R zp_r = R(1) / pap;
      zli[jl] = zl[jl] + zi[jl];
      if (zli[jl] > rlmin) {
        zliqfrac[jl] = zl[jl] / zli[jl];
        zicefrac[jl] = R(1) - zliqfrac[jl];
      } else {
        zliqfrac[jl] = R(0);
        zicefrac[jl] = R(0);
      }
      // :464-469 critical relative humidity (ZXRAMID = RAMID without SPP)
      zrhc[jl] = ramid;
      const R zsigk = pap / a.paph[IH(jl, klev + 1)];
      if (zsigk > R(0.8)) {
        const R s = (zsigk - R(0.8)) / R(0.2);
        zrhc[jl] = ramid + (R(1) - ramid) * (s * s);
      }
      // start-of-step state (:476-506)
      const R zfoeeliq = std::min(th.foeeliq(t0) * zp_r, R(0.5));
      const R zfoeeice = std::min(th.foeew(t0) * zp_r, R(0.5));
      const R zfokoop = std::min(zfoeeice * (th.rkoop1 - th.rkoop2 * t0), zfoeeliq);
      R qsliqk = zfokoop / (R(1) - retv * zfokoop);
      const R zfoeemix = std::min(th.foeewm(t0) * zp_r, R(0.5));
      const R zcor = R(1) / (R(1) - retv * zfoeemix);
      zqsmix[jl] = zfoeemix * zcor;
      zqsliqk[jl] = R(0.5) * (qsliqk + zqsmix[jl]);
      // updated state (:513-535): only QSMIX_UPD and CORQSMIX_UPD are used further on
      const R zfoeemix_u = std::min(th.foeewm(zt[jl]) * zp_r, R(0.5));
      const R zcor_u = R(1) / (R(1) - retv * zfoeemix_u);
      zqsmix_upd[jl] = zfoeemix_u * zcor_u;
      zcorqsmix_upd[jl] = R(1) + zqsmix_upd[jl] * zcor_u * th.foedem(zt[jl]);
    }
    // :561-576 dqs/dt by the Newton adjustment (in RD)
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      zttmp[jl] = a.pt[I2(jl, jk)] + zdtforce[jl];
      zttmp[jl] = std::max(zttmp[jl], R(160));
      zqtmp[jl] = zqsmix[jl] + zdqforce[jl];
      const RD zqsatnew = qsatmixadj<RD>(thd, RD(p.retv), RD(zttmp[jl]), RD(zqtmp[jl]), RD(a.pap[I2(jl, jk)]));
      zdqs[jl] = R(zqsatnew - RD(zqtmp[jl]));
    }
    // :586-653 condensation in the clear part
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      const R pa0 = a.pa[I2(jl, jk)], pq0 = a.pq[I2(jl, jk)];
      env_l[jl] = R(0);
      env_i[jl] = R(0);
      zacond[jl] = R(0);
      if (zdqs[jl] <= -rlmin && pa0 < R(1) - zepsec) {
        R zqe = (pq0 - pa0 * zqsmix[jl]) / std::max(zepsec, R(1) - pa0);
        zqe = std::max(R(0), zqe);
        if (zqe >= zrhc[jl] * zqsliqk[jl] && zqe < zqsliqk[jl]) {
          zacond[jl] = -(R(1) - pa0) * zdqs[jl] / std::max(R(2) * (zqsliqk[jl] - zqe), zepsec);
          zacond[jl] = std::min(zacond[jl], R(1) - pa0);
          R zlcond_env = -zdqs[jl] * R(0.5) * zacond[jl];
          const R zzdl = R(2) * (zqsliqk[jl] - zqe) / std::max(zepsec, R(1) - pa0);
          if (zdqs[jl] < -zzdl) {
            const R zlcondlim = (pa0 - R(1)) * zdqs[jl] - zqsliqk[jl] + pq0;
            zlcond_env = std::min(zlcond_env, zlcondlim);
          }
          zlcond_env = std::max(zlcond_env, R(0));
          if (zlcond_env < rlmin || (R(1) - pa0) < zepsec) {
            zlcond_env = R(0);
            zacond[jl] = R(0);
          }
          if (zlcond_env == R(0)) zacond[jl] = R(0);
          if (a.pt[I2(jl, jk)] > rthomo)
            env_l[jl] = zlcond_env;
          else
            env_i[jl] = zlcond_env;
        }
      }
    }
    // :661-697 condensation in the cloudy part
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      const R pa0 = a.pa[I2(jl, jk)];
      cld_l[jl] = R(0);
      cld_i[jl] = R(0);
      if (pa0 > zepsec && zdqs[jl] <= -rlmin) {
        R zlcond_cld = std::max(-zdqs[jl], R(0));
        const R zcdmax = std::max((zq[jl] - zqsmix_upd[jl]) / zcorqsmix_upd[jl], R(0));
        zlcond_cld = std::min(zlcond_cld, zcdmax);
        zlcond_cld = pa0 * zlcond_cld;
        if (zlcond_cld < rlmin) zlcond_cld = R(0);
        if (a.pt[I2(jl, jk)] > rthomo)
          cld_l[jl] = zlcond_cld;
        else
          cld_i[jl] = zlcond_cld;
      }
    }
    // :703-724 evaporation in the cloudy part
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      if (zdqs[jl] > R(0) && zli[jl] > R(0)) {
        const R zlevap = std::min(a.pa[I2(jl, jk)] * zdqs[jl], zli[jl]);
        ev_l[jl] = zliqfrac[jl] * zlevap;
        ev_i[jl] = zicefrac[jl] * zlevap;
      } else {
        ev_l[jl] = R(0);
        ev_i[jl] = R(0);
      }
      zaevap[jl] = R(0);
    }
    // :741-758 state for the supersaturation check
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      const R tu = zt[jl] + ralvdcp * (cld_l[jl] + env_l[jl] - ev_l[jl]) + ralsdcp * (cld_i[jl] + env_i[jl] - ev_i[jl]);
      const R qu = zq[jl] + (ev_l[jl] - cld_l[jl] - env_l[jl]) + (ev_i[jl] - cld_i[jl] - env_i[jl]);
      R zanew = za[jl] + zacond[jl] - zaevap[jl];
      zanew = std::min(zanew, R(1));
      if (zanew < ramin) zanew = R(0);
      t_upd[jl] = RD(tu);
      q_upd[jl] = RD(qu);
      a_upd[jl] = RD(zanew);
      pp[jl] = a.pap[I2(jl, jk)];
      psurf[jl] = a.paph[IH(jl, klev + 1)];
    }
    // :770-773 CLOUD_SUPERSATCHECK on the full level (PAP(:,JK), PAPH(:,KLEV+1)); only L+I and A are used
    supersatcheck<R, RD>(p, a.kidia, a.kfdia, &t_upd[1], &q_upd[1], &a_upd[1], &pp[1], &psurf[1], &t_adj[1],
                         &q_adj[1], &a_adj[1], &l_adj[1], &i_adj[1]);
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {
      // :775-793
      R supl = R(0), supi = R(0);
      const R sup = R(l_adj[jl] + i_adj[jl]);
      if (R(t_upd[jl]) > rthomo)
        supl = sup;
      else
        supi = sup;
      const R supa = R(a_adj[jl]);
      // :800-825 tendencies
      R tt = ralvdcp * (cld_l[jl] + env_l[jl] + supl - ev_l[jl]) * zqtmst +
             ralsdcp * (cld_i[jl] + env_i[jl] + supi - ev_i[jl]) * zqtmst;
      R tq = (ev_l[jl] - cld_l[jl] - env_l[jl] - supl + ev_i[jl] - cld_i[jl] - env_i[jl] - supi) * zqtmst;
      R zanew = za[jl] + zacond[jl] - zaevap[jl] + supa;
      zanew = std::min(zanew, R(1));
      if (zanew < ramin) zanew = R(0);
      const R ta = (zanew - za[jl]) * zqtmst;
      R tl = (cld_l[jl] + env_l[jl] + supl - ev_l[jl]) * zqtmst;
      R ti = (cld_i[jl] + env_i[jl] + supi - ev_i[jl]) * zqtmst;
      // :827-842 negative clean-up
      const R zl_adj = zl[jl] + tl * ptsphy;
      if (zl_adj < R(0)) {
        tl = tl - zl_adj;
        tq = tq + zl_adj;
        tt = tt - ralvdcp * zl_adj;
      }
      const R zi_adj = zi[jl] + ti * ptsphy;
      if (zi_adj < R(0)) {
        ti = ti - zi_adj;
        tq = tq + zi_adj;
        tt = tt - ralsdcp * zi_adj;
      }
      a.tloc_t[I2(jl, jk)] = tt;
      a.tloc_q[I2(jl, jk)] = tq;
      a.tloc_a[I2(jl, jk)] = ta;
      a.tloc_l[I2(jl, jk)] = tl;
      a.tloc_i[I2(jl, jk)] = ti;
    }
  }
}

}  // namespace oracle
