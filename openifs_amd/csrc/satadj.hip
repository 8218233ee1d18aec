// CLOUD_SATADJ for gfx950 (declared in include/cloudsc_satadj.h).
//
// Restates the default path of ifs-source/arpifs/phys_ec/cloud_satadj.F90 (LSPP = .FALSE., budgets off).
// Unlike CLOUDSC, nothing is carried from level to level: every (column, level) point is independent, so the
// kernel runs one lane per point over the whole blocked field, lanes of a wave on consecutive columns of one
// level -- every load and store is a contiguous row, and the grid holds NGPTOT x KLEV lanes instead of NGPTOT.
// Line citations refer to cloud_satadj.F90 unless another file is named.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/cloudsc_satadj.h"
#include "capi_util.h"
#include "cloudsc_device.h"
#include "cloudsc_kernel.h"
#include "synth.h"

namespace cloudsc_gpu {

template <typename R>
struct SatPtrs {
  const R *pap, *paph, *pt, *pq, *pa, *pl, *pi;
  const R *tcml_t, *tcml_q, *tcml_a, *tcml_l, *tcml_i;
  const R *phrsw, *phrlw, *tvdf_t, *tvdf_q, *pludeli, *pvervel;
  R *tloc_t, *tloc_q, *tloc_a, *tloc_l, *tloc_i;
};

struct SatDims {
  int32_t ngptot, nproma, klev, ngpblks;
  int32_t ncs;     // species stride of the liquid/ice arrays: 5 (slices of PCLV-like fields) or 1 (own arrays)
  int32_t kidia, kfdia;  // active columns of a full block (1-based)
};

// QSATMIXADJ (:948-1017): two Newton iterations towards mixed-phase saturation; returns the adjusted humidity.
template <typename T, typename KCT>
__device__ __forceinline__ T qsatmixadj(const KCT& c, T pt, T pq, T pp) {
  T zt = pt, zq = pq;
  const T zp_r = m_div(T(1), pp);
#pragma unroll
  for (int ji = 0; ji < 2; ++ji) {
    const Sat<T> s = saturation<T, KCT>(c, zt);
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

template <typename R, typename RD, typename KC, typename KCD>
__global__ __launch_bounds__(256) void satadj_kernel(const KC c, const KCD cd, const SatPtrs<R> p, const SatDims d) {
  const uint32_t np = (uint32_t)d.nproma, klev = (uint32_t)d.klev;
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= (uint32_t)d.ngpblks * klev * np) return;
  const uint32_t jl = e % np, r = e / np, k = r % klev, b = r / klev;
  const int kfdia = min(d.kfdia, d.ngptot - (int)(b * np));
  if ((int)jl < d.kidia - 1 || (int)jl >= kfdia) return;  // columns outside KIDIA:KFDIA are not touched
  constexpr uint32_t E = sizeof(R);
  const uint32_t o2 = e * E;                                                    // (NPROMA,KLEV) arrays
  const uint32_t oc = ((b * (uint32_t)d.ncs * klev + k) * np + jl) * E;         // liquid / ice arrays
  const uint32_t o4 = ((b * 4u * klev + k) * np + jl) * E;                     // PLUDELI(:,:,1)
  const uint32_t oh = ((b * (klev + 1u) + k) * np + jl) * E;                   // PAPH(:,JK)
  const uint32_t lstep = np * E, s4 = klev * np * E;
  if ((int)k < c.ncldtop - 1) {  // :338-348
    ST(p.tloc_t, o2, 0u, R(0));
    ST(p.tloc_q, o2, 0u, R(0));
    ST(p.tloc_a, o2, 0u, R(0));
    ST(p.tloc_l, oc, 0u, R(0));
    ST(p.tloc_i, oc, 0u, R(0));
    return;
  }
  const R ptsphy = c.ptsphy, zqtmst = c.zqtmst, zepsec = R(1.e-14), rlmin = c.rlmin;
  const R pt = LD(p.pt, o2, 0u), pq = LD(p.pq, o2, 0u), pa = LD(p.pa, o2, 0u), pap = LD(p.pap, o2, 0u);
  const R paph0 = LD(p.paph, oh, 0u), paph1 = LD(p.paph, oh, lstep);
  const R psurf = LD(p.paph, ((b * (klev + 1u) + klev) * np + jl) * E, 0u);
  // :360-368 updated state
  const R zt = pt + ptsphy * LD(p.tcml_t, o2, 0u);
  const R zq = pq + ptsphy * LD(p.tcml_q, o2, 0u);
  const R za = m_max(R(0), m_min(R(1), pa + ptsphy * LD(p.tcml_a, o2, 0u)));
  const R zl = LD(p.pl, oc, 0u) + ptsphy * LD(p.tcml_l, oc, 0u);
  const R zi = LD(p.pi, oc, 0u) + ptsphy * LD(p.tcml_i, oc, 0u);
  // :384-416 forcing terms (ZPVERVEL = PVERVEL without SPP)
  const R zdp = paph1 - paph0;
  const R zdtgdp = ptsphy * m_div(c.rg, zdp);
  const R zdtdp = m_div(c.zrdcp * pt, pap);
  const R zdpmxdt = zdp * zqtmst;
  const R zwtot = m_min(zdpmxdt, m_max(-zdpmxdt, LD(p.pvervel, o2, 0u)));
  const R zzzdt = LD(p.phrsw, o2, 0u) + LD(p.phrlw, o2, 0u);
  const R zdtdiab = m_min(zdpmxdt * zdtdp, m_max(-zdpmxdt * zdtdp, zzzdt)) * ptsphy;
  const R zdtforce = zdtdp * zwtot * ptsphy + zdtdiab + LD(p.pludeli, o4, 3u * s4) * zdtgdp +
                     LD(p.tvdf_t, o2, 0u) * ptsphy;
  const R zdqforce = LD(p.pludeli, o4, 2u * s4) * zdtgdp + LD(p.tvdf_q, o2, 0u) * ptsphy;
  // :430-442 liquid / ice fractions
  const R zli = zl + zi;
  R zliqfrac = R(0), zicefrac = R(0);
  if (zli > rlmin) {
    zliqfrac = m_div(zl, zli);
    zicefrac = R(1) - zliqfrac;
  }
  // :464-469 critical relative humidity
  R zrhc = c.ramid;
  const R zsigk = m_div(pap, psurf);
  if (zsigk > R(0.8)) {
    const R s = m_div(zsigk - R(0.8), R(0.2));
    zrhc = c.ramid + (R(1) - c.ramid) * (s * s);
  }
  // :476-506 saturation of the start-of-step state
  const R zp_r = m_div(R(1), pap);
  const Sat<R> s0 = saturation<R, KC>(c, pt);
  const R zfoeeliq = m_min(c.r2es * s0.el * zp_r, R(0.5));
  const R foeew = (pt - c.rtt >= R(0)) ? c.r2es * s0.el : c.r2es * s0.ei;
  const R zfoeeice = m_min(foeew * zp_r, R(0.5));
  const R zfokoop = m_min(zfoeeice * (c.rkoop1 - c.rkoop2 * pt), zfoeeliq);
  const R qsliqk0 = m_div(zfokoop, R(1) - c.retv * zfokoop);
  const R zfoeemix = m_min(c.r2es * (s0.alfa * s0.el + (R(1) - s0.alfa) * s0.ei) * zp_r, R(0.5));
  const R zqsmix = zfoeemix * m_div(R(1), R(1) - c.retv * zfoeemix);
  const R zqsliqk = R(0.5) * (qsliqk0 + zqsmix);
  // :513-535 updated state: only QSMIX_UPD and CORQSMIX_UPD are used further on
  const Sat<R> s1 = saturation<R, KC>(c, zt);
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
  R tt = c.ralvdcp * (cld_l + env_l + supl - ev_l) * zqtmst + c.ralsdcp * (cld_i + env_i + supi - ev_i) * zqtmst;
  R tq = (ev_l - cld_l - env_l - supl + ev_i - cld_i - env_i - supi) * zqtmst;
  R zanew2 = m_min(za + zacond + R(a_adj), R(1));
  if (zanew2 < c.ramin) zanew2 = R(0);
  const R ta = (zanew2 - za) * zqtmst;
  R tl = (cld_l + env_l + supl - ev_l) * zqtmst;
  R ti = (cld_i + env_i + supi - ev_i) * zqtmst;
  // :827-842 negative clean-up
  const R zl_adj = zl + tl * ptsphy;
  if (zl_adj < R(0)) {
    tl = tl - zl_adj;
    tq = tq + zl_adj;
    tt = tt - c.ralvdcp * zl_adj;
  }
  const R zi_adj = zi + ti * ptsphy;
  if (zi_adj < R(0)) {
    ti = ti - zi_adj;
    tq = tq + zi_adj;
    tt = tt - c.ralsdcp * zi_adj;
  }
  ST(p.tloc_t, o2, 0u, tt);
  ST(p.tloc_q, o2, 0u, tq);
  ST(p.tloc_a, o2, 0u, ta);
  ST(p.tloc_l, oc, 0u, tl);
  ST(p.tloc_i, oc, 0u, ti);
}

// Launcher: splits the blocks so that the largest array of one launch, (NPROMA,KLEV,NCS) or PLUDELI
// (NPROMA,KLEV,4), spans less than 4 GiB (32-bit byte offsets).
template <typename R, typename RD>
hipError_t launch_satadj(const cloudsc_params_t& q, double ptsphy, const SatPtrs<R>& p, const SatDims& d,
                         hipStream_t stream) {
  const int64_t per_block = (int64_t)d.nproma * d.klev * std::max(4, d.ncs) * (int64_t)sizeof(R);
  const int64_t max_blocks = std::max<int64_t>(1, ((int64_t)1 << 32) / per_block - 1);
  const bool def = is_default_config(q);
  for (int64_t b0 = 0; b0 < d.ngpblks; b0 += max_blocks) {
    const int64_t nb = std::min<int64_t>(max_blocks, d.ngpblks - b0);
    SatDims dc = d;
    dc.ngpblks = (int32_t)nb;
    dc.ngptot = (int32_t)(d.ngptot - b0 * d.nproma);
    const size_t s2 = (size_t)d.nproma * d.klev, sh = (size_t)d.nproma * (d.klev + 1);
    const size_t sc = s2 * d.ncs, s4 = s2 * 4;
    SatPtrs<R> pc = p;
    pc.pap += b0 * s2; pc.paph += b0 * sh; pc.pt += b0 * s2; pc.pq += b0 * s2; pc.pa += b0 * s2;
    pc.pl += b0 * sc; pc.pi += b0 * sc; pc.tcml_t += b0 * s2; pc.tcml_q += b0 * s2; pc.tcml_a += b0 * s2;
    pc.tcml_l += b0 * sc; pc.tcml_i += b0 * sc; pc.phrsw += b0 * s2; pc.phrlw += b0 * s2;
    pc.tvdf_t += b0 * s2; pc.tvdf_q += b0 * s2; pc.pludeli += b0 * s4; pc.pvervel += b0 * s2;
    pc.tloc_t += b0 * s2; pc.tloc_q += b0 * s2; pc.tloc_a += b0 * s2; pc.tloc_l += b0 * sc; pc.tloc_i += b0 * sc;
    const unsigned grid = (unsigned)((nb * (int64_t)d.klev * d.nproma + 255) / 256);
    if (def) {
      const KConstDefault<R> c = make_kconst_default<R>(q, ptsphy);
      const KConstDefault<RD> cd = make_kconst_default<RD>(q, ptsphy);
      hipLaunchKernelGGL((satadj_kernel<R, RD, KConstDefault<R>, KConstDefault<RD>>), dim3(grid), dim3(256), 0,
                         stream, c, cd, pc, dc);
    } else {
      const KConst<R> c = make_kconst<R>(q, ptsphy);
      const KConst<RD> cd = make_kconst<RD>(q, ptsphy);
      hipLaunchKernelGGL((satadj_kernel<R, RD, KConst<R>, KConst<RD>>), dim3(grid), dim3(256), 0, stream, c, cd,
                         pc, dc);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Synthetic forcings (include/cloudsc_satadj.h): one lane per (column, level).
template <typename R>
__global__ __launch_bounds__(256) void satadj_synth_kernel(uint64_t seed, int ngptot, int nproma, int klev,
                                                           int ngpblks, const R* plude, R* hrsw, R* hrlw, R* vdft,
                                                           R* vdfq, R* vervel, R* pludeli) {
  const uint32_t np = (uint32_t)nproma, kl = (uint32_t)klev;
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= (uint32_t)ngpblks * kl * np) return;
  const uint32_t jl = e % np, r = e / np, k = r % kl, b = r / kl;
  const uint32_t col = b * np + jl;
  // padding columns get the same draws as real ones, so every array is initialised
  const synth::Draw d1 = synth::draw(seed, 201u, col, k), d2 = synth::draw(seed, 202u, col, k);
  const synth::Draw d3 = synth::draw(seed, 203u, col, k), d4 = synth::draw(seed, 204u, col, k);
  hrsw[e] = (R)(2e-5 * d1.u0);
  hrlw[e] = (R)(-2e-5 + 1e-5 * d1.n);
  vdft[e] = (R)(1e-5 * d2.n);
  vdfq[e] = (R)(1e-8 * d3.n);
  vervel[e] = (R)(0.3 * d4.n);
  const size_t o4 = ((size_t)b * 4u * kl + k) * np + jl, s4 = (size_t)kl * np;
  const double lude = (double)plude[e];
  pludeli[o4 + 2 * s4] = (R)(1e-2 * lude);
  pludeli[o4 + 3 * s4] = (R)(1e3 * lude);
}

}  // namespace cloudsc_gpu

using cloudsc_gpu::SatDims;
using cloudsc_gpu::SatPtrs;

namespace {

using cloudsc_capi::fail;

int check_satadj_config(const cloudsc_params_t* p) {
  if (!p) return fail(CLOUDSC_ENULL, "cloud_satadj: NULL parameter block");
  if (p->lspp || p->lcldbud)
    return fail(CLOUDSC_EUNSUPPORTED, "cloud_satadj: SPP perturbations and budget diagnostics are not implemented");
  return CLOUDSC_OK;
}

template <typename R, typename RD>
int blocked(const cloudsc_params_t* q, int ngptot, int nproma, int klev, int ngpblks, R ptsphy, const SatPtrs<R>& p,
            void* stream) {
  int rc = check_satadj_config(q);
  if (rc) return rc;
  if (nproma < 1 || klev < 2 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloud_satadj_blocked: need NGPBLKS = ceil(NGPTOT/NPROMA), KLEV >= 2");
  if (q->ncldtop < 2 || q->ncldtop > klev) return fail(CLOUDSC_EDIMS, "cloud_satadj: NCLDTOP outside [2,KLEV]");
  if ((int64_t)ngpblks * klev * nproma >= ((int64_t)1 << 32))
    return fail(CLOUDSC_EDIMS, "cloud_satadj_blocked: NGPBLKS*KLEV*NPROMA must be below 2^32");
  const void* req[] = {p.pap, p.paph, p.pt, p.pq, p.pa, p.pl, p.tcml_t, p.tcml_q, p.tcml_a, p.tcml_l, p.phrsw,
                       p.phrlw, p.tvdf_t, p.tvdf_q, p.pludeli, p.pvervel, p.tloc_t, p.tloc_q, p.tloc_a, p.tloc_l};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloud_satadj_blocked: NULL array on the default path");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  SatDims d{ngptot, nproma, klev, ngpblks, 5, 1, nproma};
  const hipError_t e = cloudsc_gpu::launch_satadj<R, RD>(*q, (double)ptsphy, p, d, (hipStream_t)stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj_blocked: kernel launch", e);
  return CLOUDSC_OK;
}

// Per-block drop-in: a per-thread device workspace and stream, as the CLOUDSC per-block entry.
struct SatWorkspace {
  void* dev = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  int device = -1;
};
thread_local SatWorkspace g_sws;

template <typename R, typename RD>
int per_block(const cloudsc_params_t* q, int kidia, int kfdia, int klon, int klev, R ptsphy, const R* pap,
              const R* paph, const R* pt, const R* pq, const R* pa, const R* pl, const R* pi, const R* tt,
              const R* tq, const R* ta, const R* tl, const R* ti, const R* hrsw, const R* hrlw, const R* vdft,
              const R* vdfq, const R* pludeli, const R* vervel, R* ot, R* oq, R* oa, R* ol, R* oi) {
  int rc = check_satadj_config(q);
  if (rc) return rc;
  if (klon < 1 || klev < 2 || kidia < 1 || kfdia < kidia || kfdia > klon)
    return fail(CLOUDSC_EDIMS, "cloud_satadj: need 1 <= KIDIA <= KFDIA <= KLON, KLEV >= 2");
  if (q->ncldtop < 2 || q->ncldtop > klev) return fail(CLOUDSC_EDIMS, "cloud_satadj: NCLDTOP outside [2,KLEV]");
  const void* req[] = {pap, paph, pt, pq, pa, pl, pi, tt, tq, ta, tl, ti, hrsw, hrlw, vdft, vdfq, pludeli, vervel,
                       ot, oq, oa, ol, oi};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloud_satadj: NULL array on the default path");
  rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const size_t n2 = (size_t)klon * klev, nh = (size_t)klon * (klev + 1);
  struct Arr {
    const void* host;
    void* out;
    size_t n;
  };
  Arr arrs[] = {{pap, nullptr, n2}, {paph, nullptr, nh}, {pt, nullptr, n2},  {pq, nullptr, n2},  {pa, nullptr, n2},
                {pl, nullptr, n2},  {pi, nullptr, n2},   {tt, nullptr, n2},  {tq, nullptr, n2},  {ta, nullptr, n2},
                {tl, nullptr, n2},  {ti, nullptr, n2},   {hrsw, nullptr, n2}, {hrlw, nullptr, n2}, {vdft, nullptr, n2},
                {vdfq, nullptr, n2}, {pludeli, nullptr, 4 * n2}, {vervel, nullptr, n2}, {ot, ot, n2}, {oq, oq, n2},
                {oa, oa, n2}, {ol, ol, n2}, {oi, oi, n2}};
  const int narr = sizeof(arrs) / sizeof(arrs[0]);
  size_t off[sizeof(arrs) / sizeof(arrs[0])], total = 0;
  for (int i = 0; i < narr; ++i) {
    off[i] = total;
    total += (arrs[i].n * sizeof(R) + 255) & ~(size_t)255;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj: hipGetDevice", e);
  SatWorkspace& w = g_sws;
  if (w.device != dev) {
    w = SatWorkspace{};
    w.device = dev;
    e = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj: hipStreamCreate", e);
  }
  if (w.bytes < total) {
    if (w.dev) (void)hipFree(w.dev);
    w.dev = nullptr;
    w.bytes = 0;
    e = hipMalloc(&w.dev, total);
    if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloud_satadj: hipMalloc workspace", e);
    w.bytes = total;
  }
  char* base = (char*)w.dev;
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    e = hipMemcpyAsync(base + off[i], arrs[i].host, arrs[i].n * sizeof(R), hipMemcpyHostToDevice, w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj: copy to device", e);
  auto D = [&](int i) { return (R*)(base + off[i]); };
  SatPtrs<R> p{D(0),  D(1),  D(2),  D(3),  D(4),  D(5),  D(6),  D(7),  D(8),  D(9),  D(10), D(11),
               D(12), D(13), D(14), D(15), D(16), D(17), D(18), D(19), D(20), D(21), D(22)};
  SatDims d{kfdia, klon, klev, 1, 1, kidia, kfdia};
  e = cloudsc_gpu::launch_satadj<R, RD>(*q, (double)ptsphy, p, d, w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj: kernel launch", e);
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    if (arrs[i].out)
      e = hipMemcpyAsync(arrs[i].out, base + off[i], arrs[i].n * sizeof(R), hipMemcpyDeviceToHost, w.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj: copy to host", e);
  return CLOUDSC_OK;
}

template <typename R>
int run_synth(uint64_t seed, int ngptot, int nproma, int klev, int ngpblks, const R* plude, R* hrsw, R* hrlw, R* vdft,
          R* vdfq, R* vervel, R* pludeli, void* stream) {
  if (nproma < 1 || klev < 2 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma ||
      (int64_t)ngpblks * klev * nproma >= ((int64_t)1 << 32))
    return fail(CLOUDSC_EDIMS, "cloud_satadj_synth: inconsistent dimensions");
  if (!plude || !hrsw || !hrlw || !vdft || !vdfq || !vervel || !pludeli)
    return fail(CLOUDSC_ENULL, "cloud_satadj_synth: NULL array");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const unsigned grid = (unsigned)(((int64_t)ngpblks * klev * nproma + 255) / 256);
  hipLaunchKernelGGL(cloudsc_gpu::satadj_synth_kernel<R>, dim3(grid), dim3(256), 0, (hipStream_t)stream, seed,
                     ngptot, nproma, klev, ngpblks, plude, hrsw, hrlw, vdft, vdfq, vervel, pludeli);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_satadj_synth: kernel launch", e);
  return CLOUDSC_OK;
}

}  // namespace

#define SATADJ_PER_BLOCK(R, RD)                                                                               \
  (void)ktype, (void)lddiag, (void)pmfu, (void)pmfd, (void)pgp2dspp, (void)pextra, (void)kfldx;                 \
  return per_block<R, RD>(ydparams, kidia, kfdia, klon, klev, ptsphy, pap, paph, pt, pq, pa, pl, pi,          \
                          ptendency_cml_t, ptendency_cml_q, ptendency_cml_a, ptendency_cml_l, ptendency_cml_i,   \
                          phrsw, phrlw, ptendency_vdf_t, ptendency_vdf_q, pludeli, pvervel, ptendency_loc_t,      \
                          ptendency_loc_q, ptendency_loc_a, ptendency_loc_l, ptendency_loc_i)

#define SATADJ_BLOCKED(R, RD)                                                                                 \
  const size_t sl = (size_t)nproma * klev;                                                                    \
  SatPtrs<R> p{pap, paph, pt, pq, pa, pclv, pclv ? pclv + sl : nullptr, ptendency_cml_t, ptendency_cml_q,    \
               ptendency_cml_a, ptendency_cml_cld, ptendency_cml_cld ? ptendency_cml_cld + sl : nullptr, phrsw, \
               phrlw, ptendency_vdf_t, ptendency_vdf_q, pludeli, pvervel, ptendency_loc_t, ptendency_loc_q,    \
               ptendency_loc_a, ptendency_loc_cld, ptendency_loc_cld ? ptendency_loc_cld + sl : nullptr};       \
  return blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy, p, stream)

extern "C" {

int cloudsc_satadj_dp_c(CLOUDSC_SATADJ_ARGS(double)) { SATADJ_PER_BLOCK(double, double); }
int cloudsc_satadj_sp_c(CLOUDSC_SATADJ_ARGS(float)) { SATADJ_PER_BLOCK(float, float); }
int cloudsc_satadj_mixed_c(CLOUDSC_SATADJ_ARGS(float)) { SATADJ_PER_BLOCK(float, double); }

int cloudsc_satadj_blocked_dp_c(CLOUDSC_SATADJ_BLOCKED_ARGS(double)) { SATADJ_BLOCKED(double, double); }
int cloudsc_satadj_blocked_sp_c(CLOUDSC_SATADJ_BLOCKED_ARGS(float)) { SATADJ_BLOCKED(float, float); }
int cloudsc_satadj_blocked_mixed_c(CLOUDSC_SATADJ_BLOCKED_ARGS(float)) { SATADJ_BLOCKED(float, double); }

int cloudsc_satadj_synth_sp(uint64_t seed, int ngptot, int nproma, int klev, int ngpblks, const float* plude,
                            float* phrsw, float* phrlw, float* ptendency_vdf_t, float* ptendency_vdf_q,
                            float* pvervel, float* pludeli, void* stream) {
  return run_synth<float>(seed, ngptot, nproma, klev, ngpblks, plude, phrsw, phrlw, ptendency_vdf_t, ptendency_vdf_q,
                      pvervel, pludeli, stream);
}
int cloudsc_satadj_synth_dp(uint64_t seed, int ngptot, int nproma, int klev, int ngpblks, const double* plude,
                            double* phrsw, double* phrlw, double* ptendency_vdf_t, double* ptendency_vdf_q,
                            double* pvervel, double* pludeli, void* stream) {
  return run_synth<double>(seed, ngptot, nproma, klev, ngpblks, plude, phrsw, phrlw, ptendency_vdf_t, ptendency_vdf_q,
                       pvervel, pludeli, stream);
}

}  // extern "C"
