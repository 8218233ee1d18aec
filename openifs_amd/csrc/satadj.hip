// CLOUD_SATADJ for gfx950 (declared in include/cloudsc_satadj.h).
//
// Restates the default path of ifs-source/arpifs/phys_ec/cloud_satadj.F90 (LSPP = .FALSE., budgets off); the
// per-point arithmetic is satadj_point (satadj_device.h).  Unlike CLOUDSC, nothing is carried from level to
// level: every (column, level) point is independent, so the kernel spreads the whole blocked field over the
// grid, lanes of a wave on consecutive columns of one level -- every load and store is a contiguous row, and
// the grid holds NGPTOT x KLEV points instead of NGPTOT columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/cloudsc_satadj.h"
#include "capi_util.h"
#include "cloudsc_device.h"
#include "cloudsc_kernel.h"
#include "satadj_device.h"
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

template <typename R, int V>
struct Vec {
  typedef R type __attribute__((ext_vector_type(V)));
};
template <typename R>
struct Vec<R, 1> {
  typedef R type;
};
template <int V, typename R>
__device__ __forceinline__ typename Vec<R, V>::type LDV(const R* base, uint32_t u, uint32_t v) {
  return *reinterpret_cast<const typename Vec<R, V>::type*>(reinterpret_cast<const char*>(base) + (u + v));
}
template <int V, typename R>
__device__ __forceinline__ void STV(R* base, uint32_t u, uint32_t v, typename Vec<R, V>::type x) {
  *reinterpret_cast<typename Vec<R, V>::type*>(reinterpret_cast<char*>(base) + (u + v)) = x;
}
template <int V, typename R>
__device__ __forceinline__ R lane(const typename Vec<R, V>::type& x, int j) {
  if constexpr (V == 1)
    return x;
  else
    return x[j];
}
template <int V, typename R>
__device__ __forceinline__ void set_lane(typename Vec<R, V>::type& x, int j, R y) {
  if constexpr (V == 1)
    x = y;
  else
    x[j] = y;
}

// One lane handles V consecutive columns of one level (V = 2 when NPROMA is even and every array is aligned to
// 2 words, else 1); lanes of a wave cover consecutive columns, so each load and store instruction moves one
// contiguous row segment of 64 V words.  Measured at 655,360 columns x L137: V = 2 is 9 % faster than V = 1 in
// FP32 (1.40 vs 1.53 ms) and 7 % in FP64; V = 4 is no faster than V = 1 (its 93 VGPRs halve the waves per SIMD).
template <typename R, typename RD, typename KC, typename KCD, int V>
__global__ __launch_bounds__(256) void satadj_kernel(const KC c, const KCD cd, const SatPtrs<R> p, const SatDims d) {
  typedef typename Vec<R, V>::type VT;
  const uint32_t np = (uint32_t)d.nproma, klev = (uint32_t)d.klev, npv = np / V;
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint32_t)d.ngpblks * klev * npv) return;
  const uint32_t jv = t % npv, r = t / npv, k = r % klev, b = r / klev;
  const uint32_t jl0 = jv * V;
  const int kfdia = min(d.kfdia, d.ngptot - (int)(b * np));
  const int lo = d.kidia - 1 - (int)jl0, hi = kfdia - (int)jl0;  // valid j: lo <= j < hi
  if (hi <= 0 || lo >= V) return;  // columns outside KIDIA:KFDIA are not touched
  constexpr uint32_t E = sizeof(R);
  const uint32_t o2 = ((b * klev + k) * np + jl0) * E;                         // (NPROMA,KLEV) arrays
  const uint32_t oc = ((b * (uint32_t)d.ncs * klev + k) * np + jl0) * E;       // liquid / ice arrays
  const uint32_t o4 = ((b * 4u * klev + k) * np + jl0) * E;                    // PLUDELI(:,:,1)
  const uint32_t oh = ((b * (klev + 1u) + k) * np + jl0) * E;                  // PAPH(:,JK)
  const uint32_t os = ((b * (klev + 1u) + klev) * np + jl0) * E;               // PAPH(:,KLEV+1)
  const uint32_t lstep = np * E, s4 = klev * np * E;
  VT ot, oq, oa, ol, oi;
  if ((int)k < c.ncldtop - 1) {  // :338-348
    ot = oq = oa = ol = oi = VT(R(0));
  } else {
    const VT pt = LDV<V>(p.pt, o2, 0u), pq = LDV<V>(p.pq, o2, 0u), pa = LDV<V>(p.pa, o2, 0u);
    const VT pl = LDV<V>(p.pl, oc, 0u), pi = LDV<V>(p.pi, oc, 0u), pap = LDV<V>(p.pap, o2, 0u);
    const VT ph0 = LDV<V>(p.paph, oh, 0u), ph1 = LDV<V>(p.paph, oh, lstep), ps = LDV<V>(p.paph, os, 0u);
    const VT tct = LDV<V>(p.tcml_t, o2, 0u), tcq = LDV<V>(p.tcml_q, o2, 0u), tca = LDV<V>(p.tcml_a, o2, 0u);
    const VT tcl = LDV<V>(p.tcml_l, oc, 0u), tci = LDV<V>(p.tcml_i, oc, 0u);
    const VT hrsw = LDV<V>(p.phrsw, o2, 0u), hrlw = LDV<V>(p.phrlw, o2, 0u);
    const VT vdft = LDV<V>(p.tvdf_t, o2, 0u), vdfq = LDV<V>(p.tvdf_q, o2, 0u);
    const VT l3 = LDV<V>(p.pludeli, o4, 2u * s4), l4 = LDV<V>(p.pludeli, o4, 3u * s4);
    const VT vv = LDV<V>(p.pvervel, o2, 0u);
#pragma unroll
    for (int j = 0; j < V; ++j) {
      SatIn<R> x;
      x.pt = lane<V, R>(pt, j), x.pq = lane<V, R>(pq, j), x.pa = lane<V, R>(pa, j);
      x.pl = lane<V, R>(pl, j), x.pi = lane<V, R>(pi, j), x.pap = lane<V, R>(pap, j);
      x.paph0 = lane<V, R>(ph0, j), x.paph1 = lane<V, R>(ph1, j), x.psurf = lane<V, R>(ps, j);
      x.tct = lane<V, R>(tct, j), x.tcq = lane<V, R>(tcq, j), x.tca = lane<V, R>(tca, j);
      x.tcl = lane<V, R>(tcl, j), x.tci = lane<V, R>(tci, j);
      x.hrsw = lane<V, R>(hrsw, j), x.hrlw = lane<V, R>(hrlw, j);
      x.vdft = lane<V, R>(vdft, j), x.vdfq = lane<V, R>(vdfq, j);
      x.lude3 = lane<V, R>(l3, j), x.lude4 = lane<V, R>(l4, j), x.vervel = lane<V, R>(vv, j);
      const SatOut<R> y = satadj_point<R, RD, KC, KCD>(c, cd, x);
      set_lane<V, R>(ot, j, y.t), set_lane<V, R>(oq, j, y.q), set_lane<V, R>(oa, j, y.a);
      set_lane<V, R>(ol, j, y.l), set_lane<V, R>(oi, j, y.i);
    }
  }
  if (lo <= 0 && hi >= V) {
    STV<V>(p.tloc_t, o2, 0u, ot);
    STV<V>(p.tloc_q, o2, 0u, oq);
    STV<V>(p.tloc_a, o2, 0u, oa);
    STV<V>(p.tloc_l, oc, 0u, ol);
    STV<V>(p.tloc_i, oc, 0u, oi);
  } else {  // a vector that straddles KIDIA or KFDIA: element stores inside the window only
#pragma unroll
    for (int j = 0; j < V; ++j)
      if (j >= lo && j < hi) {
        ST(p.tloc_t, o2, j * E, lane<V, R>(ot, j));
        ST(p.tloc_q, o2, j * E, lane<V, R>(oq, j));
        ST(p.tloc_a, o2, j * E, lane<V, R>(oa, j));
        ST(p.tloc_l, oc, j * E, lane<V, R>(ol, j));
        ST(p.tloc_i, oc, j * E, lane<V, R>(oi, j));
      }
  }
}

template <typename R>
bool aligned_to(const SatPtrs<R>& p, size_t bytes) {
  const void* a[] = {p.pap, p.paph, p.pt, p.pq, p.pa, p.pl, p.pi, p.tcml_t, p.tcml_q, p.tcml_a, p.tcml_l, p.tcml_i,
                     p.phrsw, p.phrlw, p.tvdf_t, p.tvdf_q, p.pludeli, p.pvervel, p.tloc_t, p.tloc_q, p.tloc_a,
                     p.tloc_l, p.tloc_i};
  for (const void* x : a)
    if ((uintptr_t)x % bytes) return false;
  return true;
}

// Launcher: splits the blocks so that the largest array of one launch, (NPROMA,KLEV,NCS) or PLUDELI
// (NPROMA,KLEV,4), spans less than 4 GiB (32-bit byte offsets).
template <typename R, typename RD, typename KC, typename KCD>
hipError_t launch_satadj_c(const KC& c, const KCD& cd, const SatPtrs<R>& p, const SatDims& d, hipStream_t stream) {
  constexpr int VW = 2;
  const bool vec = d.nproma % VW == 0 && aligned_to(p, VW * sizeof(R));
  const int64_t per_block = (int64_t)d.nproma * d.klev * std::max(4, d.ncs) * (int64_t)sizeof(R);
  const int64_t max_blocks = std::max<int64_t>(1, ((int64_t)1 << 32) / per_block - 1);
  // launches of equal size: a short last launch would leave most of the GPU idle while it runs
  const int64_t nlaunch = (d.ngpblks + max_blocks - 1) / max_blocks;
  const int64_t step = (d.ngpblks + nlaunch - 1) / nlaunch;
  for (int64_t b0 = 0; b0 < d.ngpblks; b0 += step) {
    const int64_t nb = std::min<int64_t>(step, d.ngpblks - b0);
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
    const int64_t lanes = nb * (int64_t)d.klev * (d.nproma / (vec ? VW : 1));
    const unsigned grid = (unsigned)((lanes + 255) / 256);
    if (vec)
      hipLaunchKernelGGL((satadj_kernel<R, RD, KC, KCD, VW>), dim3(grid), dim3(256), 0, stream, c, cd, pc, dc);
    else
      hipLaunchKernelGGL((satadj_kernel<R, RD, KC, KCD, 1>), dim3(grid), dim3(256), 0, stream, c, cd, pc, dc);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename R, typename RD>
hipError_t launch_satadj(const cloudsc_params_t& q, double ptsphy, const SatPtrs<R>& p, const SatDims& d,
                         hipStream_t stream) {
  if (is_default_config(q))
    return launch_satadj_c<R, RD>(make_kconst_default<R>(q, ptsphy), make_kconst_default<RD>(q, ptsphy), p, d,
                                  stream);
  return launch_satadj_c<R, RD>(make_kconst<R>(q, ptsphy), make_kconst<RD>(q, ptsphy), p, d, stream);
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
  // the calling thread's workspace (shared with the other per-block entries)
  cloudsc_capi::HostWorkspace* wp = nullptr;
  const int wrc = cloudsc_capi::thread_workspace(total, "cloud_satadj", &wp);
  if (wrc) return wrc;
  cloudsc_capi::HostWorkspace& w = *wp;
  hipError_t e = hipSuccess;
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
