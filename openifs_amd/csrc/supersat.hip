// CLOUD_SUPERSATCHECK as a drop-in of its own and SLTEND's supersaturation adjustment on gfx950
// (include/cloudsc_supersat.h).  The per-point arithmetic is supersatcheck() of cloudsc_device.h, shared with the
// CLOUDSC and CLOUD_SATADJ kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/cloudsc_supersat.h"
#include "capi_util.h"
#include "cloudsc_device.h"
#include "cloudsc_kernel.h"

namespace cloudsc_gpu {

// KFTLIQICE = 2 (cloud_supersatcheck.F90:215-223): all liquid where PT + PT_ADJ >= RTHOMO, all ice below; the
// condensate is ZCOND = -PQ_ADJ exactly.
template <typename T, typename KCT>
__device__ __forceinline__ void supersat_point(const KCT& c, int kftliqice, T pt, T pq, T pa, T pp, T& t_adj,
                                               T& q_adj, T& a_adj, T& l_adj, T& i_adj) {
  supersatcheck<T, KCT>(c, pt, pq, pa, pp, t_adj, q_adj, a_adj, l_adj, i_adj);
  if (kftliqice == 2) {
    const T zcond = -q_adj;
    const bool liquid = pt + t_adj >= T(c.rthomo);
    l_adj = liquid ? zcond : T(0);
    i_adj = liquid ? T(0) : zcond;
  }
}

template <typename R, typename RD, typename KCD>
__global__ __launch_bounds__(256) void supersat_points_kernel(const KCD cd, int n, int kftliqice, const R* pt,
                                                              const R* pq, const R* pa, const R* pp, R* t_adj,
                                                              R* q_adj, R* a_adj, R* l_adj, R* i_adj) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  RD t, q, a, l, ic;
  supersat_point<RD, KCD>(cd, kftliqice, RD(pt[i]), RD(pq[i]), RD(pa[i]), RD(pp[i]), t, q, a, l, ic);
  t_adj[i] = R(t);
  q_adj[i] = R(q);
  a_adj[i] = R(a);
  l_adj[i] = R(l);
  i_adj[i] = R(ic);
}

// SLTEND section 3 (sltend.F90:370-409), one lane per (column, level) point of levels NCLDTOP..KLEV.
template <typename R>
struct SltPtrs {
  const R *pt, *pq, *pa, *prsf1, *pcent, *pcenq, *pcena;
  R *ptent, *ptenq, *ptena, *ptencld, *adja, *adjl, *adji;
};

template <typename R, typename RD, typename KC, typename KCD>
__global__ __launch_bounds__(256) void sltend_supersat_kernel(const KC c, const KCD cd, const SltPtrs<R> p,
                                                              int ngptot, int nproma, int klev, int ngpblks,
                                                              int k0) {
  const uint32_t np = (uint32_t)nproma, nk = (uint32_t)(klev - k0);
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= (uint32_t)ngpblks * nk * np) return;
  const uint32_t jl = e % np, r = e / np, k = (uint32_t)k0 + r % nk, b = r / nk;
  if ((int)(b * np + jl) >= ngptot) return;
  constexpr uint32_t E = sizeof(R);
  const uint32_t o2 = ((b * (uint32_t)klev + k) * np + jl) * E;
  const uint32_t o5 = ((b * 5u * (uint32_t)klev + k) * np + jl) * E, s5 = (uint32_t)klev * np * E;
  const R tsphy = c.ptsphy, zrtsphy = c.zqtmst;
  R tt = LD(p.ptent, o2, 0u), tq = LD(p.ptenq, o2, 0u), ta = LD(p.ptena, o2, 0u);
  const R zt = LD(p.pt, o2, 0u) + tsphy * (LD(p.pcent, o2, 0u) + tt);
  const R zq = LD(p.pq, o2, 0u) + tsphy * (LD(p.pcenq, o2, 0u) + tq);
  const R za = LD(p.pa, o2, 0u) + tsphy * (LD(p.pcena, o2, 0u) + ta);
  RD t_adj, q_adj, a_adj, l_adj, i_adj;
  supersat_point<RD, KCD>(cd, 1, RD(zt), RD(zq), RD(za), RD(LD(p.prsf1, o2, 0u)), t_adj, q_adj, a_adj, l_adj,
                          i_adj);
  const R ra = R(a_adj) * zrtsphy, rl = R(l_adj) * zrtsphy, ri = R(i_adj) * zrtsphy;
  ST(p.ptent, o2, 0u, tt + R(t_adj) * zrtsphy);
  ST(p.ptenq, o2, 0u, tq + R(q_adj) * zrtsphy);
  ST(p.ptena, o2, 0u, ta + ra);
  ST(p.ptencld, o5, 0u, LD(p.ptencld, o5, 0u) + rl);
  ST(p.ptencld, o5, s5, LD(p.ptencld, o5, s5) + ri);
  if (p.adja) {
    ST(p.adja, o2, 0u, ra);
    ST(p.adjl, o2, 0u, rl);
    ST(p.adji, o2, 0u, ri);
  }
}

}  // namespace cloudsc_gpu

namespace {

using cloudsc_capi::fail;
using cloudsc_gpu::SltPtrs;

template <typename R, typename RD>
int points(const cloudsc_params_t* q, int kidia, int kfdia, int klon, int klev, int kftliqice, const R* pt,
           const R* pq, const R* pa, const R* pp, const R* psurf, R* t_adj, R* q_adj, R* a_adj, R* l_adj,
           R* i_adj) {
  (void)klev, (void)psurf;
  if (!q) return fail(CLOUDSC_ENULL, "cloud_supersatcheck: NULL parameter block");
  if (klon < 1 || kidia < 1 || kfdia < kidia || kfdia > klon)
    return fail(CLOUDSC_EDIMS, "cloud_supersatcheck: need 1 <= KIDIA <= KFDIA <= KLON");
  if (kftliqice != 1 && kftliqice != 2) return fail(CLOUDSC_EDIMS, "cloud_supersatcheck: KFTLIQICE must be 1 or 2");
  if (!pt || !pq || !pa || !pp || !t_adj || !q_adj || !a_adj || !l_adj || !i_adj)
    return fail(CLOUDSC_ENULL, "cloud_supersatcheck: NULL array");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const int n = kfdia - kidia + 1, off = kidia - 1;
  const size_t nb = (size_t)n * sizeof(R), slot = (nb + 255) & ~(size_t)255;
  // the calling thread's workspace (shared with the other per-block entries)
  cloudsc_capi::HostWorkspace* wp = nullptr;
  rc = cloudsc_capi::thread_workspace(9 * slot, "cloud_supersatcheck", &wp);
  if (rc) return rc;
  cloudsc_capi::HostWorkspace& w = *wp;
  hipError_t e = hipSuccess;
  char* base = (char*)w.dev;
  R* d[9];
  for (int i = 0; i < 9; ++i) d[i] = (R*)(base + i * slot);
  const R* in[4] = {pt + off, pq + off, pa + off, pp + off};
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipMemcpyAsync(d[i], in[i], nb, hipMemcpyHostToDevice, w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_supersatcheck: copy to device", e);
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (cloudsc_gpu::is_default_config(*q)) {
    const auto cd = cloudsc_gpu::make_kconst_default<RD>(*q, 1.0);
    hipLaunchKernelGGL((cloudsc_gpu::supersat_points_kernel<R, RD, decltype(cd)>), dim3(grid), dim3(256), 0,
                       w.stream, cd, n, kftliqice, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8]);
  } else {
    const auto cd = cloudsc_gpu::make_kconst<RD>(*q, 1.0);
    hipLaunchKernelGGL((cloudsc_gpu::supersat_points_kernel<R, RD, decltype(cd)>), dim3(grid), dim3(256), 0,
                       w.stream, cd, n, kftliqice, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8]);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_supersatcheck: kernel launch", e);
  R* out[5] = {t_adj + off, q_adj + off, a_adj + off, l_adj + off, i_adj + off};
  for (int i = 0; i < 5 && e == hipSuccess; ++i)
    e = hipMemcpyAsync(out[i], d[4 + i], nb, hipMemcpyDeviceToHost, w.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w.stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloud_supersatcheck: copy to host", e);
  return CLOUDSC_OK;
}

template <typename R, typename RD>
int sltend(const cloudsc_params_t* q, int ngptot, int nproma, int klev, int ngpblks, R tsphy, const SltPtrs<R>& p,
           void* stream) {
  if (!q) return fail(CLOUDSC_ENULL, "sltend_supersat: NULL parameter block");
  if (nproma < 1 || klev < 2 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "sltend_supersat: need NGPBLKS = ceil(NGPTOT/NPROMA), KLEV >= 2");
  if (q->ncldtop < 1 || q->ncldtop > klev) return fail(CLOUDSC_EDIMS, "sltend_supersat: NCLDTOP outside [1,KLEV]");
  if ((int64_t)ngpblks * nproma * klev * 5 * (int64_t)sizeof(R) >= ((int64_t)1 << 32))
    return fail(CLOUDSC_EDIMS, "sltend_supersat: the five-species field must stay below 4 GiB per call");
  const void* req[] = {p.pt, p.pq, p.pa, p.prsf1, p.pcent, p.pcenq, p.pcena, p.ptent, p.ptenq, p.ptena, p.ptencld};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "sltend_supersat: NULL array");
  if ((p.adja || p.adjl || p.adji) && !(p.adja && p.adjl && p.adji))
    return fail(CLOUDSC_ENULL, "sltend_supersat: PADJA, PADJL, PADJI must be all set or all NULL");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const int k0 = q->ncldtop - 1;
  const unsigned grid = (unsigned)(((int64_t)ngpblks * (klev - k0) * nproma + 255) / 256);
  if (cloudsc_gpu::is_default_config(*q)) {
    const auto c = cloudsc_gpu::make_kconst_default<R>(*q, (double)tsphy);
    const auto cd = cloudsc_gpu::make_kconst_default<RD>(*q, (double)tsphy);
    hipLaunchKernelGGL((cloudsc_gpu::sltend_supersat_kernel<R, RD, decltype(c), decltype(cd)>), dim3(grid),
                       dim3(256), 0, (hipStream_t)stream, c, cd, p, ngptot, nproma, klev, ngpblks, k0);
  } else {
    const auto c = cloudsc_gpu::make_kconst<R>(*q, (double)tsphy);
    const auto cd = cloudsc_gpu::make_kconst<RD>(*q, (double)tsphy);
    hipLaunchKernelGGL((cloudsc_gpu::sltend_supersat_kernel<R, RD, decltype(c), decltype(cd)>), dim3(grid),
                       dim3(256), 0, (hipStream_t)stream, c, cd, p, ngptot, nproma, klev, ngpblks, k0);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "sltend_supersat: kernel launch", e);
  return CLOUDSC_OK;
}

}  // namespace

#define SS_POINTS(R, RD)                                                                                       \
  return points<R, RD>(ydparams, kidia, kfdia, klon, klev, kftliqice, pt, pq, pa, pp, psurf, pt_adj, pq_adj,    \
                       pa_adj, pl_adj, pi_adj)
#define SS_SLTEND(R, RD)                                                                                       \
  return sltend<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, tsphy,                                          \
                       SltPtrs<R>{pt, pq, pa, prsf1, pcent, pcenq, pcena, ptent, ptenq, ptena, ptencld, padja,   \
                                  padjl, padji},                                                               \
                       stream)

extern "C" {

int cloudsc_supersatcheck_dp_c(CLOUDSC_SUPERSATCHECK_ARGS(double)) { SS_POINTS(double, double); }
int cloudsc_supersatcheck_sp_c(CLOUDSC_SUPERSATCHECK_ARGS(float)) { SS_POINTS(float, float); }
int cloudsc_supersatcheck_mixed_c(CLOUDSC_SUPERSATCHECK_ARGS(float)) { SS_POINTS(float, double); }

int cloudsc_sltend_supersat_blocked_dp_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(double)) { SS_SLTEND(double, double); }
int cloudsc_sltend_supersat_blocked_sp_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(float)) { SS_SLTEND(float, float); }
int cloudsc_sltend_supersat_blocked_mixed_c(CLOUDSC_SLTEND_SUPERSAT_ARGS(float)) { SS_SLTEND(float, double); }

}  // extern "C"
