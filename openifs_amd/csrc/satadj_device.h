// CLOUD_SATADJ for one (column, level) point: the default path of ifs-source/arpifs/phys_ec/cloud_satadj.F90
// (LSPP = .FALSE., budgets off), shared by the stand-alone kernel (satadj.hip) and the fused saturation
// adjustment + CLOUDSC kernel.  Line citations refer to cloud_satadj.F90.
#pragma once

#include "cloudsc_device.h"

namespace cloudsc_gpu {

// Inputs of one point.  psurf = PAPH(JL,KLEV+1); tc* = TENDENCY_CML; lude3/lude4 = PLUDELI(JL,JK,3:4).
template <typename R>
struct SatIn {
  R pt, pq, pa, pl, pi, pap, paph0, paph1, psurf;
  R tct, tcq, tca, tcl, tci;
  R hrsw, hrlw, vdft, vdfq, lude3, lude4, vervel;
};
template <typename R>
struct SatOut {
  R t, q, a, l, i;  // TENDENCY_LOC of T, Q, A, CLD(:,:,NCLDQL), CLD(:,:,NCLDQI)
};

// saturation() evaluating only the exponentials that FOEALFA weights: above RTWAT (alfa = 1) the ice term and
// below RTICE (alfa = 0) the liquid term is multiplied by exactly zero, so skipping it gives the same bits.
// Levels of one block row lie in one temperature regime, so the branch is almost always uniform over a wave.
template <typename T, typename KCT>
__device__ __forceinline__ Sat<T> saturation_sel(const KCT& c, T t) {
  Sat<T> s;
  T x = (m_max(c.rtice, m_min(c.rtwat, t)) - c.rtice) * c.rtwat_rtice_r;
  s.alfa = m_min(T(1), x * x);
  s.el = T(0);
  s.ei = T(0);
  if (s.alfa > T(0)) s.el = m_exp(m_div(c.r3les * (t - c.rtt), t - c.r4les));
  if (s.alfa < T(1)) s.ei = m_exp(m_div(c.r3ies * (t - c.rtt), t - c.r4ies));
  return s;
}

// QSATMIXADJ (:948-1017): two Newton iterations towards mixed-phase saturation; returns the adjusted humidity.
template <typename T, typename KCT>
__device__ __forceinline__ T qsatmixadj(const KCT& c, T pt, T pq, T pp) {
  T zt = pt, zq = pq;
  const T zp_r = m_div(T(1), pp);
#pragma unroll
  for (int ji = 0; ji < 2; ++ji) {
    const Sat<T> s = saturation_sel<T, KCT>(c, zt);
    const T zfoeet = m_min(c.r2es * (s.alfa * s.el + (T(1) - s.alfa) * s.ei) * zp_r, T(0.5));
    const T zcor = m_div(T(1), T(1) - c.retv * zfoeet);
    const T zqsat = zfoeet * zcor;
    const T dl = zt - c.r4les, di = zt - c.r4ies;
    const T foedem = s.alfa * c.r5alvcp * m_div(T(1), dl * dl) + (T(1) - s.alfa) * c.r5alscp * m_div(T(1), di * di);
    const T zcorqs = T(1) + zqsat * zcor * foedem;
    const T zcond = m_div(zq - zqsat, zcorqs);
    const T foeldcpm = s.alfa * c.ralvdcp + (T(1) - s.alfa) * c.ralsdcp;
    zt = zt + foeldcpm * zcond;
    zq = zq - zcond;
  }
  return zq;
}

// One point on a level at or below NCLDTOP.  R: storage and bulk arithmetic; RD: the Newton adjustment and
// CLOUD_SUPERSATCHECK (FP64 in the mixed variant).
template <typename R, typename RD, typename KC, typename KCD>
__device__ __forceinline__ SatOut<R> satadj_point(const KC& c, const KCD& cd, const SatIn<R>& x) {
  const R ptsphy = c.ptsphy, zqtmst = c.zqtmst, zepsec = R(1.e-14), rlmin = c.rlmin;
  const R pt = x.pt, pq = x.pq, pa = x.pa, pap = x.pap;
  // :360-368 updated state
  const R zt = pt + ptsphy * x.tct;
  const R zq = pq + ptsphy * x.tcq;
  const R za = m_max(R(0), m_min(R(1), pa + ptsphy * x.tca));
  const R zl = x.pl + ptsphy * x.tcl;
  const R zi = x.pi + ptsphy * x.tci;
  // :384-416 forcing terms (ZPVERVEL = PVERVEL without SPP)
  const R zdp = x.paph1 - x.paph0;
  const R zdtgdp = ptsphy * m_div(c.rg, zdp);
  const R zdtdp = m_div(c.zrdcp * pt, pap);
  const R zdpmxdt = zdp * zqtmst;
  const R zwtot = m_min(zdpmxdt, m_max(-zdpmxdt, x.vervel));
  const R zzzdt = x.hrsw + x.hrlw;
  const R zdtdiab = m_min(zdpmxdt * zdtdp, m_max(-zdpmxdt * zdtdp, zzzdt)) * ptsphy;
  const R zdtforce = zdtdp * zwtot * ptsphy + zdtdiab + x.lude4 * zdtgdp + x.vdft * ptsphy;
  const R zdqforce = x.lude3 * zdtgdp + x.vdfq * ptsphy;
  // :430-442 liquid / ice fractions
  const R zli = zl + zi;
  R zliqfrac = R(0), zicefrac = R(0);
  if (zli > rlmin) {
    zliqfrac = m_div(zl, zli);
    zicefrac = R(1) - zliqfrac;
  }
  // :464-469 critical relative humidity
  R zrhc = c.ramid;
  const R zsigk = m_div(pap, x.psurf);
  if (zsigk > R(0.8)) {
    const R s = m_div(zsigk - R(0.8), R(0.2));
    zrhc = c.ramid + (R(1) - c.ramid) * (s * s);
  }
  // :476-506 saturation of the start-of-step state
  const R zp_r = m_div(R(1), pap);
  // EXP for ice only where FOEEW or the FOEALFA mix uses it (below RTT or below RTWAT)
  Sat<R> s0;
  {
    const R x = (m_max(c.rtice, m_min(c.rtwat, pt)) - c.rtice) * c.rtwat_rtice_r;
    s0.alfa = m_min(R(1), x * x);
    s0.el = m_exp(m_div(c.r3les * (pt - c.rtt), pt - c.r4les));
    s0.ei = R(0);
    if (pt - c.rtt < R(0) || s0.alfa < R(1)) s0.ei = m_exp(m_div(c.r3ies * (pt - c.rtt), pt - c.r4ies));
  }
  const R zfoeeliq = m_min(c.r2es * s0.el * zp_r, R(0.5));
  const R foeew = (pt - c.rtt >= R(0)) ? c.r2es * s0.el : c.r2es * s0.ei;
  const R zfoeeice = m_min(foeew * zp_r, R(0.5));
  const R zfokoop = m_min(zfoeeice * (c.rkoop1 - c.rkoop2 * pt), zfoeeliq);
  const R qsliqk0 = m_div(zfokoop, R(1) - c.retv * zfokoop);
  const R zfoeemix = m_min(c.r2es * (s0.alfa * s0.el + (R(1) - s0.alfa) * s0.ei) * zp_r, R(0.5));
  const R zqsmix = zfoeemix * m_div(R(1), R(1) - c.retv * zfoeemix);
  const R zqsliqk = R(0.5) * (qsliqk0 + zqsmix);
  // :513-535 updated state: only QSMIX_UPD and CORQSMIX_UPD are used further on
  const Sat<R> s1 = saturation_sel<R, KC>(c, zt);
  const R zfoeemix_u = m_min(c.r2es * (s1.alfa * s1.el + (R(1) - s1.alfa) * s1.ei) * zp_r, R(0.5));
  const R zcor_u = m_div(R(1), R(1) - c.retv * zfoeemix_u);
  const R zqsmix_upd = zfoeemix_u * zcor_u;
  const R dl = zt - c.r4les, di = zt - c.r4ies;
  const R foedem_u = s1.alfa * c.r5alvcp * m_div(R(1), dl * dl) + (R(1) - s1.alfa) * c.r5alscp * m_div(R(1), di * di);
  const R zcorqsmix_upd = R(1) + zqsmix_upd * zcor_u * foedem_u;
  // :561-576 dqs from the Newton adjustment (RD)
  const R zttmp = m_max(pt + zdtforce, R(160));
  const R zqtmp = zqsmix + zdqforce;
  const R zdqs = R(qsatmixadj<RD, KCD>(cd, RD(zttmp), RD(zqtmp), RD(pap)) - RD(zqtmp));
  // :586-653 condensation in the clear part
  R env_l = R(0), env_i = R(0), zacond = R(0);
  if (zdqs <= -rlmin && pa < R(1) - zepsec) {
    const R zqe = m_max(R(0), m_div(pq - pa * zqsmix, m_max(zepsec, R(1) - pa)));
    if (zqe >= zrhc * zqsliqk && zqe < zqsliqk) {
      zacond = m_div(-(R(1) - pa) * zdqs, m_max(R(2) * (zqsliqk - zqe), zepsec));
      zacond = m_min(zacond, R(1) - pa);
      R zlcond_env = -zdqs * R(0.5) * zacond;
      const R zzdl = m_div(R(2) * (zqsliqk - zqe), m_max(zepsec, R(1) - pa));
      if (zdqs < -zzdl) zlcond_env = m_min(zlcond_env, (pa - R(1)) * zdqs - zqsliqk + pq);
      zlcond_env = m_max(zlcond_env, R(0));
      if (zlcond_env < rlmin || (R(1) - pa) < zepsec) {
        zlcond_env = R(0);
        zacond = R(0);
      }
      if (zlcond_env == R(0)) zacond = R(0);
      if (pt > c.rthomo)
        env_l = zlcond_env;
      else
        env_i = zlcond_env;
    }
  }
  // :661-697 condensation in the cloudy part
  R cld_l = R(0), cld_i = R(0);
  if (pa > zepsec && zdqs <= -rlmin) {
    R zlcond_cld = m_max(-zdqs, R(0));
    const R zcdmax = m_max(m_div(zq - zqsmix_upd, zcorqsmix_upd), R(0));
    zlcond_cld = pa * m_min(zlcond_cld, zcdmax);
    if (zlcond_cld < rlmin) zlcond_cld = R(0);
    if (pt > c.rthomo)
      cld_l = zlcond_cld;
    else
      cld_i = zlcond_cld;
  }
  // :703-724 evaporation in the cloudy part (ZAEVAP = 0)
  R ev_l = R(0), ev_i = R(0);
  if (zdqs > R(0) && zli > R(0)) {
    const R zlevap = m_min(pa * zdqs, zli);
    ev_l = zliqfrac * zlevap;
    ev_i = zicefrac * zlevap;
  }
  // :741-758 state for the supersaturation check
  const R t_upd = zt + c.ralvdcp * (cld_l + env_l - ev_l) + c.ralsdcp * (cld_i + env_i - ev_i);
  const R q_upd = zq + (ev_l - cld_l - env_l) + (ev_i - cld_i - env_i);
  R zanew = m_min(za + zacond, R(1));
  if (zanew < c.ramin) zanew = R(0);
  // :770-793 CLOUD_SUPERSATCHECK on PAP(:,JK) (RD); the supersaturation becomes liquid above RTHOMO, else ice
  RD t_adj, q_adj, a_adj, l_adj, i_adj;
  supersatcheck<RD, KCD>(cd, RD(t_upd), RD(q_upd), RD(zanew), RD(pap), t_adj, q_adj, a_adj, l_adj, i_adj);
  const R sup = R(l_adj + i_adj);
  const R supl = t_upd > c.rthomo ? sup : R(0), supi = t_upd > c.rthomo ? R(0) : sup;
  // :800-825 tendencies
  SatOut<R> o;
  o.t = c.ralvdcp * (cld_l + env_l + supl - ev_l) * zqtmst + c.ralsdcp * (cld_i + env_i + supi - ev_i) * zqtmst;
  o.q = (ev_l - cld_l - env_l - supl + ev_i - cld_i - env_i - supi) * zqtmst;
  R zanew2 = m_min(za + zacond + R(a_adj), R(1));
  if (zanew2 < c.ramin) zanew2 = R(0);
  o.a = (zanew2 - za) * zqtmst;
  o.l = (cld_l + env_l + supl - ev_l) * zqtmst;
  o.i = (cld_i + env_i + supi - ev_i) * zqtmst;
  // :827-842 negative clean-up
  const R zl_adj = zl + o.l * ptsphy;
  if (zl_adj < R(0)) {
    o.l = o.l - zl_adj;
    o.q = o.q + zl_adj;
    o.t = o.t - c.ralvdcp * zl_adj;
  }
  const R zi_adj = zi + o.i * ptsphy;
  if (zi_adj < R(0)) {
    o.i = o.i - zi_adj;
    o.q = o.q + zi_adj;
    o.t = o.t - c.ralsdcp * zi_adj;
  }
  return o;
}

}  // namespace cloudsc_gpu
