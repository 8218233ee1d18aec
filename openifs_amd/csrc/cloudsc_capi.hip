// extern "C" entry points of libcloudsc_hip.so (declared in include/cloudsc_hip.h).
//
//   cloudsc_blocked_{dp,sp,mixed}_c : device pointers, one launch over all
//       NGPBLKS blocks (the production path).
//   cloudsc_{dp,sp,mixed}_c         : host pointers, one NPROMA block, the
//       positional drop-in for SUBROUTINE CLOUDSC (cloudsc.F90:14-44).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cloudsc_hip.h"
#include "../../include/cloudsc_satadj.h"
#include "../../include/cloudsc_tune.h"
#include "capi_util.h"
#include "cloudsc_kernel.h"

using cloudsc_gpu::KDims;
using cloudsc_gpu::KPtrs;

namespace cloudsc_capi {

thread_local std::string g_last_error;
thread_local int64_t g_last_warn[2] = {0, 0};

int fail(int code, const char* what, hipError_t e) {
  g_last_error = what;
  if (e != hipSuccess) {
    g_last_error += ": ";
    g_last_error += hipGetErrorString(e);
  }
  return code;
}

int check_config(const cloudsc_params_t* p, bool diag) {
  if (!p) return fail(CLOUDSC_ENULL, "cloudsc: NULL parameter block");
  if (p->laeraware || p->lspp) return fail(CLOUDSC_EUNSUPPORTED, "cloudsc: aerosol/SPP are not implemented");
  if (p->lcloud_inhomog && !diag)
    return fail(CLOUDSC_EUNSUPPORTED,
                "cloudsc: LCLOUD_INHOMOG reads PGAW: use the per-block entry or cloudsc_blocked_diag_*_c");
  const int32_t out = CLOUDSC_BUD_C | CLOUDSC_BUD_L | CLOUDSC_BUD_I | CLOUDSC_BUD_T | CLOUDSC_BUD_VERTINT;
  if ((p->lcldbud & out) && !diag)
    return fail(CLOUDSC_EUNSUPPORTED,
                "cloudsc: the budget diagnostics write PEXTRA: use the per-block entry or cloudsc_blocked_diag_*_c");
  if (p->lrad_cloud_inhomog && !diag)
    return fail(CLOUDSC_EUNSUPPORTED,
                "cloudsc: LRAD_CLOUD_INHOMOG writes PFSD: use the per-block entry or cloudsc_blocked_diag_*_c");
  if (p->lcldbud & ~255) return fail(CLOUDSC_EUNSUPPORTED, "cloudsc: unknown budget switch bits");
  return CLOUDSC_OK;
}

// PGAW and PFSD where LRAD_CLOUD_INHOMOG asks for the condensate FSD.
int check_fsd(const cloudsc_params_t* p, const void* pgaw, const void* pfsd) {
  if (p->lrad_cloud_inhomog && (!pgaw || !pfsd))
    return fail(CLOUDSC_ENULL, "cloudsc: LRAD_CLOUD_INHOMOG needs PGAW and PFSD");
  if (p->lcloud_inhomog && !pgaw) return fail(CLOUDSC_ENULL, "cloudsc: LCLOUD_INHOMOG needs PGAW");
  return CLOUDSC_OK;
}

// KFLDX and KLEV for the budget switches in force; NULL arrays only when no budget is written.
int check_budget(const cloudsc_params_t* p, int klev, int kfldx, const void* const* arrays, int narrays) {
  const int32_t out = CLOUDSC_BUD_C | CLOUDSC_BUD_L | CLOUDSC_BUD_I | CLOUDSC_BUD_T | CLOUDSC_BUD_VERTINT;
  if (!(p->lcldbud & out)) return CLOUDSC_OK;
  for (int i = 0; i < narrays; ++i)
    if (!arrays[i]) return fail(CLOUDSC_ENULL, "cloudsc: NULL array the budget diagnostics read or write");
  if (kfldx < cloudsc_gpu::budget_fields(p->lcldbud))
    return fail(CLOUDSC_EDIMS, "cloudsc: KFLDX is smaller than the PEXTRA fields the budget switches fill");
  if ((p->lcldbud & CLOUDSC_BUD_VERTINT) && klev < 83)
    return fail(CLOUDSC_EDIMS, "cloudsc: the vertical-integral budget needs KLEV >= 83");
  return CLOUDSC_OK;
}

// The kernel is built for gfx950 only; refuse any other device up front.
int check_device() {
  static std::mutex m;
  static std::vector<int> ok(64, -1);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CLOUDSC_ENODEVICE, "cloudsc: hipGetDevice", e);
  std::lock_guard<std::mutex> lk(m);
  if (dev < 0 || dev >= (int)ok.size()) return fail(CLOUDSC_ENODEVICE, "cloudsc: device index out of range");
  if (ok[dev] < 0) {
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail(CLOUDSC_ENODEVICE, "cloudsc: hipGetDeviceProperties", e);
    ok[dev] = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
  }
  if (!ok[dev]) return fail(CLOUDSC_ENODEVICE, "cloudsc: device is not gfx950 (MI355X)");
  return CLOUDSC_OK;
}

}  // namespace cloudsc_capi

namespace {

using cloudsc_capi::check_device;
using cloudsc_capi::check_budget;
using cloudsc_capi::check_config;
using cloudsc_capi::check_fsd;
using cloudsc_capi::fail;

template <typename R, typename RD>
int blocked(const cloudsc_params_t* q, int ngptot, int nproma, int klev, int ngpblks, R ptsphy, const KPtrs<R>& p,
            void* stream, const cloudsc_gpu::KSatPtrs<R>* sat = nullptr,
            const cloudsc_gpu::KBudPtrs<R>* bud = nullptr) {
  int rc = check_config(q, bud != nullptr);
  if (rc) return rc;
  if (nproma < 1 || klev < 2 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloudsc_blocked: need NGPBLKS = ceil(NGPTOT/NPROMA), KLEV >= 2");
  if (q->ncldtop < 2 || q->ncldtop > klev) return fail(CLOUDSC_EDIMS, "cloudsc_blocked: NCLDTOP outside [2,KLEV]");
  const void* req[] = {p.pt, p.pq, p.tcml_t, p.tcml_q, p.tcml_a, p.tcml_cld, p.tloc_t, p.tloc_q, p.tloc_a,
                       p.tloc_cld, p.pvfl, p.pvfi, p.pap, p.paph, p.plsm, p.ldcum, p.ktype, p.peis, p.plu,
                       p.plude, p.pludeli, p.psnde, p.pmfu, p.pmfd, p.pa, p.pclv, p.pcovptot,
                       p.prainfrac_toprfz, p.pfsqlf, p.pfsqif, p.pfcqnng, p.pfcqlng, p.pfsqrf, p.pfsqsf,
                       p.pfcqrng, p.pfcqsng, p.pfsqltur, p.pfsqitur, p.pfplsl, p.pfplsn, p.pfhpsl, p.pfhpsn};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudsc_blocked: NULL array on the default path");
  if (sat) {
    const void* sreq[] = {sat->phrsw, sat->phrlw, sat->tvdf_t, sat->tvdf_q, sat->pvervel};
    for (const void* x : sreq)
      if (!x) return fail(CLOUDSC_ENULL, "cloudsc_satadj_cloudsc_blocked: NULL CLOUD_SATADJ input");
    const bool any = sat->tsat_t || sat->tsat_q || sat->tsat_a || sat->tsat_cld;
    const bool all = sat->tsat_t && sat->tsat_q && sat->tsat_a && sat->tsat_cld;
    if (any && !all)
      return fail(CLOUDSC_ENULL, "cloudsc_satadj_cloudsc_blocked: TENDENCY_SATADJ arrays must be all set or all NULL");
  }
  const bool budgets = bud && (q->lcldbud & (CLOUDSC_BUD_C | CLOUDSC_BUD_L | CLOUDSC_BUD_I | CLOUDSC_BUD_T |
                                             CLOUDSC_BUD_VERTINT));
  if (bud) {
    const void* breq[] = {bud->pvfa, bud->pdyna, bud->pdynl, bud->pdyni, bud->pdyns, bud->phrsw, bud->phrlw,
                          bud->kpbltype, bud->kctop, bud->pextra};
    rc = check_budget(q, klev, bud->kfldx, breq, 10);
    if (rc) return rc;
    rc = check_fsd(q, bud->pgaw, bud->pfsd);
    if (rc) return rc;
  }
  const bool diag = budgets || (bud && (bud->fsd || bud->inhomog));
  rc = check_device();
  if (rc) return rc;
  KDims d{ngptot, nproma, klev, ngpblks, 1, nproma};
  hipError_t e = cloudsc_gpu::launch_cloudsc<R, RD>(*q, (double)ptsphy, p, d, 0, (hipStream_t)stream, sat,
                                                    diag ? bud : nullptr);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_blocked: kernel launch", e);
  return CLOUDSC_OK;
}

}  // namespace

namespace cloudsc_capi {
// ---- per-block host entries: a per-thread device workspace and stream ----
thread_local HostWorkspace g_ws;

int thread_workspace(size_t bytes, const char* who, HostWorkspace** out) {
  std::string msg(who);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, (msg + ": hipGetDevice").c_str(), e);
  HostWorkspace& w = g_ws;
  if (w.device != dev) {  // first call on this thread, or the thread switched devices
    if (w.device >= 0 && hipSetDevice(w.device) == hipSuccess) {
      if (w.stream) (void)hipStreamDestroy(w.stream);
      if (w.dev) (void)hipFree(w.dev);
    }
    w = HostWorkspace{};
    e = hipSetDevice(dev);
    if (e != hipSuccess) return fail(CLOUDSC_EHIP, (msg + ": hipSetDevice").c_str(), e);
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(CLOUDSC_EHIP, (msg + ": hipStreamCreate").c_str(), e);
    w.stream = st;
    w.device = dev;
  }
  if (w.bytes < bytes) {
    if (w.dev) (void)hipFree(w.dev);
    w.dev = nullptr;
    w.bytes = 0;
    e = hipMalloc(&w.dev, bytes);
    if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, (msg + ": hipMalloc workspace").c_str(), e);
    w.bytes = bytes;
  }
  *out = &w;
  return CLOUDSC_OK;
}
}  // namespace cloudsc_capi

// ---- device memory of the allocation-free per-block entry (cloudsc_mixed_noalloc_c) ----
// For callers that may not allocate device memory (hosts that forbid hipMalloc, tight serving loops): the staging
// buffer of one block lives in this fixed array instead of the per-thread workspace.  Each GPU the library runs on
// holds its own copy until the process ends.  Every call on a GPU shares that GPU's copy and leaves its contents
// behind (the warning counters are cleared at the start of each call): calls on one GPU that overlap race.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
using cloudsc_capi::HostWorkspace;
using cloudsc_capi::thread_workspace;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

// The current device's copy of `scratch`, looked up once per device with that device current.  Host code uses only
// this address: the array's name is a host-side stand-in, not a device address.
int scratch_on_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_mixed_noalloc_c: hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return fail(CLOUDSC_EDIMS, "cloudsc_mixed_noalloc_c: device index above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CLOUDSC_EHIP, "cloudsc_mixed_noalloc_c: hipGetSymbolAddress(scratch)", e);
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return CLOUDSC_OK;
}

// Clears the two warning counters: a kernel rather than hipMemsetAsync, which stays correct in captured graphs.
__global__ void zero_counters(uint32_t* n) {
  if (threadIdx.x < 2) n[threadIdx.x] = 0;
}

// Where per_block stages the block: the calling thread's workspace (nullptr) or the fixed array on `stream`.
struct NoAlloc {
  hipStream_t stream;
};

template <typename R, typename RD>
int per_block(const cloudsc_params_t* q, int kidia, int kfdia, int klon, int klev, R ptsphy, const R* pt,
              const R* pq, const R* tcml_t, const R* tcml_q, const R* tcml_a, const R* tcml_cld, R* tloc_t, R* tloc_q,
              R* tloc_a, R* tloc_cld, const R* pvfl, const R* pvfi, const R* pap, const R* paph, const R* plsm,
              const int32_t* ldcum, const int32_t* ktype, R* peis, const R* plu, R* plude, const R* pludeli,
              const R* psnde, const R* pmfu, const R* pmfd, const R* pa, const R* pclv, R* pcovptot,
              R* prainfrac_toprfz, R* pfsqlf, R* pfsqif, R* pfcqnng, R* pfcqlng, R* pfsqrf, R* pfsqsf, R* pfcqrng,
              R* pfcqsng, R* pfsqltur, R* pfsqitur, R* pfplsl, R* pfplsn, R* pfhpsl, R* pfhpsn, const R* pvfa,
              const R* pdyna, const R* pdynl, const R* pdyni, const R* pdyns, const R* phrsw, const R* phrlw,
              const int32_t* kpbltype, const int32_t* kctop, R* pextra, int kfldx, const R* pgaw, R* pfsd,
              const NoAlloc* na = nullptr) {
  int rc = check_config(q, true);
  if (rc) return rc;
  if (klon < 1 || klev < 2 || kidia < 1 || kfdia < kidia || kfdia > klon || kidia + kfdia - 1 > klon)
    return fail(CLOUDSC_EDIMS, "cloudsc: need 1 <= KIDIA <= KFDIA <= KLON and KIDIA+KFDIA-1 <= KLON");
  if (q->ncldtop < 2 || q->ncldtop > klev) return fail(CLOUDSC_EDIMS, "cloudsc: NCLDTOP outside [2,KLEV]");
  const void* req[] = {pt, pq, tcml_t, tcml_q, tcml_a, tcml_cld, tloc_t, tloc_q, tloc_a, tloc_cld, pvfl, pvfi,
                       pap, paph, plsm, ldcum, ktype, peis, plu, plude, pludeli, psnde, pmfu, pmfd, pa, pclv,
                       pcovptot, prainfrac_toprfz, pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng,
                       pfcqsng, pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudsc: NULL array on the default path");
  const bool budgets = q->lcldbud & (CLOUDSC_BUD_C | CLOUDSC_BUD_L | CLOUDSC_BUD_I | CLOUDSC_BUD_T |
                                     CLOUDSC_BUD_VERTINT);
  const bool fsd = q->lrad_cloud_inhomog != 0;
  const bool inh = q->lcloud_inhomog != 0;
  {
    const void* breq[] = {pvfa, pdyna, pdynl, pdyni, pdyns, phrsw, phrlw, kpbltype, kctop, pextra};
    rc = check_budget(q, klev, kfldx, breq, 10);
    if (rc) return rc;
    rc = check_fsd(q, pgaw, pfsd);
    if (rc) return rc;
  }
  rc = check_device();
  if (rc) return rc;

  const size_t n2 = (size_t)klon * klev, nh = (size_t)klon * (klev + 1), n1 = klon;
  // every array the kernel touches, with its element count; outputs are copied in too so that
  // columns outside KIDIA:KFDIA and levels the reference leaves untouched come back unchanged
  struct Arr {
    const void* host;
    void* host_out;
    size_t n, esz;
  };
  Arr arrs[] = {{pt, nullptr, n2, sizeof(R)},          {pq, nullptr, n2, sizeof(R)},
                {tcml_t, nullptr, n2, sizeof(R)},      {tcml_q, nullptr, n2, sizeof(R)},
                {tcml_a, nullptr, n2, sizeof(R)},      {tcml_cld, nullptr, 5 * n2, sizeof(R)},
                {tloc_t, tloc_t, n2, sizeof(R)},       {tloc_q, tloc_q, n2, sizeof(R)},
                {tloc_a, tloc_a, n2, sizeof(R)},       {tloc_cld, tloc_cld, 5 * n2, sizeof(R)},
                {pvfl, nullptr, n2, sizeof(R)},        {pvfi, nullptr, n2, sizeof(R)},
                {pap, nullptr, n2, sizeof(R)},         {paph, nullptr, nh, sizeof(R)},
                {plsm, nullptr, n1, sizeof(R)},        {ldcum, nullptr, n1, sizeof(int32_t)},
                {ktype, nullptr, n1, sizeof(int32_t)}, {peis, nullptr, n1, sizeof(R)},
                {plu, nullptr, n2, sizeof(R)},         {plude, plude, n2, sizeof(R)},
                {pludeli, nullptr, 4 * n2, sizeof(R)}, {psnde, nullptr, 2 * n2, sizeof(R)},
                {pmfu, nullptr, n2, sizeof(R)},        {pmfd, nullptr, n2, sizeof(R)},
                {pa, nullptr, n2, sizeof(R)},          {pclv, nullptr, 5 * n2, sizeof(R)},
                {pcovptot, pcovptot, n2, sizeof(R)},   {prainfrac_toprfz, prainfrac_toprfz, n1, sizeof(R)},
                {pfsqlf, pfsqlf, nh, sizeof(R)},       {pfsqif, pfsqif, nh, sizeof(R)},
                {pfcqnng, pfcqnng, nh, sizeof(R)},     {pfcqlng, pfcqlng, nh, sizeof(R)},
                {pfsqrf, pfsqrf, nh, sizeof(R)},       {pfsqsf, pfsqsf, nh, sizeof(R)},
                {pfcqrng, pfcqrng, nh, sizeof(R)},     {pfcqsng, pfcqsng, nh, sizeof(R)},
                {pfsqltur, pfsqltur, nh, sizeof(R)},   {pfsqitur, pfsqitur, nh, sizeof(R)},
                {pfplsl, pfplsl, nh, sizeof(R)},       {pfplsn, pfplsn, nh, sizeof(R)},
                {pfhpsl, pfhpsl, nh, sizeof(R)},       {pfhpsn, pfhpsn, nh, sizeof(R)},
                // budget diagnostics (copied only when a budget is written)
                {pvfa, nullptr, budgets ? n2 : 0, sizeof(R)},   {pdyna, nullptr, budgets ? n2 : 0, sizeof(R)},
                {pdynl, nullptr, budgets ? n2 : 0, sizeof(R)},  {pdyni, nullptr, budgets ? n2 : 0, sizeof(R)},
                {pdyns, nullptr, budgets ? n2 : 0, sizeof(R)},  {phrsw, nullptr, budgets ? n2 : 0, sizeof(R)},
                {phrlw, nullptr, budgets ? n2 : 0, sizeof(R)},
                {kpbltype, nullptr, budgets ? n1 : 0, sizeof(int32_t)},
                {kctop, nullptr, budgets ? n1 : 0, sizeof(int32_t)},
                {pextra, budgets ? pextra : nullptr, budgets ? n2 * (size_t)kfldx : 0, sizeof(R)},
                // condensate FSD (LRAD_CLOUD_INHOMOG)
                {pgaw, nullptr, (fsd || inh) ? n1 : 0, sizeof(R)},       {pfsd, fsd ? pfsd : nullptr, fsd ? n2 : 0, sizeof(R)}};
  const int narr = sizeof(arrs) / sizeof(arrs[0]);
  size_t off[sizeof(arrs) / sizeof(arrs[0])];
  size_t total = 0;
  for (int i = 0; i < narr; ++i) {
    off[i] = total;
    total += (arrs[i].n * arrs[i].esz + 255) & ~(size_t)255;
  }
  const size_t warn_off = total;  // two counters of the negative-condensate warnings
  char* base = nullptr;
  hipStream_t stream = nullptr;
  if (na) {
    if (total + 256 > kScratchBytes)
      return fail(CLOUDSC_EDIMS, "cloudsc_mixed_noalloc_c: the block needs more than the 32 MiB fixed device array "
                                 "(KLON x KLEV too large, or too many PEXTRA fields)");
    void* fixed = nullptr;
    rc = scratch_on_current_device(&fixed);
    if (rc) return rc;
    base = (char*)fixed;
    stream = na->stream;
  } else {
    HostWorkspace* w = nullptr;
    rc = thread_workspace(total + 256, "cloudsc", &w);
    if (rc) return rc;
    base = (char*)w->dev;
    stream = w->stream;
  }
  uint32_t* d_warn = (uint32_t*)(base + warn_off);
  (void)hipGetLastError();  // a stale error of an earlier call is not this launch's
  zero_counters<<<1, 64, 0, stream>>>(d_warn);
  hipError_t e = hipGetLastError();
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    if (arrs[i].n)
      e = hipMemcpyAsync(base + off[i], arrs[i].host, arrs[i].n * arrs[i].esz, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc: copy to device", e);
  auto D = [&](int i) { return (R*)(base + off[i]); };
  KPtrs<R> p;
  p.pt = D(0); p.pq = D(1); p.tcml_t = D(2); p.tcml_q = D(3); p.tcml_a = D(4); p.tcml_cld = D(5);
  p.tloc_t = D(6); p.tloc_q = D(7); p.tloc_a = D(8); p.tloc_cld = D(9); p.pvfl = D(10); p.pvfi = D(11);
  p.pap = D(12); p.paph = D(13); p.plsm = D(14); p.ldcum = (const int32_t*)D(15); p.ktype = (const int32_t*)D(16);
  p.peis = D(17); p.plu = D(18); p.plude = D(19); p.pludeli = D(20); p.psnde = D(21); p.pmfu = D(22);
  p.pmfd = D(23); p.pa = D(24); p.pclv = D(25); p.pcovptot = D(26); p.prainfrac_toprfz = D(27);
  p.pfsqlf = D(28); p.pfsqif = D(29); p.pfcqnng = D(30); p.pfcqlng = D(31); p.pfsqrf = D(32); p.pfsqsf = D(33);
  p.pfcqrng = D(34); p.pfcqsng = D(35); p.pfsqltur = D(36); p.pfsqitur = D(37); p.pfplsl = D(38);
  p.pfplsn = D(39); p.pfhpsl = D(40); p.pfhpsn = D(41);
  // one block: KFDIA bounds the active columns, KIDIA-1 shifts the pressure section of the supersaturation check
  KDims d{kfdia, klon, klev, 1, kidia, kfdia};
  d.nwarn = d_warn;
  cloudsc_gpu::KBudPtrs<R> bp{};
  if (budgets) {
    bp.pvfa = D(42); bp.pdyna = D(43); bp.pdynl = D(44); bp.pdyni = D(45); bp.pdyns = D(46); bp.phrsw = D(47);
    bp.phrlw = D(48); bp.kpbltype = (const int32_t*)D(49); bp.kctop = (const int32_t*)D(50); bp.pextra = D(51);
    bp.kfldx = kfldx; bp.flags = q->lcldbud;
  }
  if (fsd) {
    bp.pfsd = D(53); bp.fsd = 1;
  }
  if (fsd || inh) bp.pgaw = D(52);
  bp.inhomog = inh ? 1 : 0;
  bp.pvfl = p.pvfl; bp.pvfi = p.pvfi;
  bp.reisthsc = R(q->reisthsc); bp.ralfdcp = R(q->ralfdcp);
  bp.ra = R(q->ra); bp.rpi = R(q->rpi); bp.rcloud_frac_std = R(q->rcloud_frac_std);
  e = cloudsc_gpu::launch_cloudsc<R, RD>(*q, (double)ptsphy, p, d, kidia - 1, stream, nullptr,
                                         (budgets || fsd || inh) ? &bp : nullptr);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc: kernel launch", e);
  for (int i = 0; i < narr && e == hipSuccess; ++i)
    if (arrs[i].host_out && arrs[i].n)
      e = hipMemcpyAsync(arrs[i].host_out, base + off[i], arrs[i].n * arrs[i].esz, hipMemcpyDeviceToHost, stream);
  uint32_t warn[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(warn, d_warn, sizeof(warn), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc: copy to host", e);
  cloudsc_capi::g_last_warn[0] = warn[0];
  cloudsc_capi::g_last_warn[1] = warn[1];
  // non-finite outputs in the reference's write set: columns KIDIA:KFDIA of the tendencies (species 1-4 of
  // PTENDENCY_LOC_CLD), PCOVPTOT from NCLDTOP and the 14 flux profiles
  int64_t nbad = 0;
  auto scan = [&](const R* a, size_t ld, int k0, int k1) {
    for (int k = k0; k < k1; ++k)
      for (int jl = kidia - 1; jl < kfdia; ++jl) nbad += !std::isfinite(a[(size_t)k * ld + jl]);
  };
  const R* tend[] = {tloc_t, tloc_q, tloc_a};
  for (const R* a : tend) scan(a, klon, 0, klev);
  for (int m = 0; m < 4; ++m) scan(tloc_cld + (size_t)m * n2, klon, 0, klev);
  scan(pcovptot, klon, q->ncldtop - 1, klev);
  const R* flux[] = {pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng,
                     pfcqsng, pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn};
  for (const R* a : flux) scan(a, klon, 0, klev + 1);
  if (nbad) {
    const std::string msg = "cloudsc: " + std::to_string(nbad) + " non-finite outputs";
    return fail(CLOUDSC_ENONFINITE, msg.c_str());
  }
  return CLOUDSC_OK;
}

template <typename R>
KPtrs<R> blocked_ptrs(const R* pt, const R* pq, const R* tcml_t, const R* tcml_q, const R* tcml_a, const R* tcml_cld,
                      R* tloc_t, R* tloc_q, R* tloc_a, R* tloc_cld, const R* pvfl, const R* pvfi, const R* pap,
                      const R* paph, const R* plsm, const int32_t* ldcum, const int32_t* ktype, R* peis, const R* plu,
                      R* plude, const R* pludeli, const R* psnde, const R* pmfu, const R* pmfd, const R* pa,
                      const R* pclv, R* pcovptot, R* prainfrac_toprfz, R* pfsqlf, R* pfsqif, R* pfcqnng, R* pfcqlng,
                      R* pfsqrf, R* pfsqsf, R* pfcqrng, R* pfcqsng, R* pfsqltur, R* pfsqitur, R* pfplsl, R* pfplsn,
                      R* pfhpsl, R* pfhpsn) {
  KPtrs<R> p;
  p.pt = pt; p.pq = pq; p.tcml_t = tcml_t; p.tcml_q = tcml_q; p.tcml_a = tcml_a; p.tcml_cld = tcml_cld;
  p.tloc_t = tloc_t; p.tloc_q = tloc_q; p.tloc_a = tloc_a; p.tloc_cld = tloc_cld; p.pvfl = pvfl; p.pvfi = pvfi;
  p.pap = pap; p.paph = paph; p.plsm = plsm; p.ldcum = ldcum; p.ktype = ktype; p.peis = peis; p.plu = plu;
  p.plude = plude; p.pludeli = pludeli; p.psnde = psnde; p.pmfu = pmfu; p.pmfd = pmfd; p.pa = pa; p.pclv = pclv;
  p.pcovptot = pcovptot; p.prainfrac_toprfz = prainfrac_toprfz; p.pfsqlf = pfsqlf; p.pfsqif = pfsqif;
  p.pfcqnng = pfcqnng; p.pfcqlng = pfcqlng; p.pfsqrf = pfsqrf; p.pfsqsf = pfsqsf; p.pfcqrng = pfcqrng;
  p.pfcqsng = pfcqsng; p.pfsqltur = pfsqltur; p.pfsqitur = pfsqitur; p.pfplsl = pfplsl; p.pfplsn = pfplsn;
  p.pfhpsl = pfhpsl; p.pfhpsn = pfhpsn;
  return p;
}

}  // namespace

#define PER_BLOCK_CALL(R, RD)                                                                                   \
  (void)lddiag, (void)pdynr, (void)pvervel, (void)pgeom, (void)pgaw, (void)pgp2dspp, (void)ldslphy,              \
      (void)plcrit_aer, (void)picrit_aer, (void)pre_ice, (void)pccn, (void)pnice;                               \
  return per_block<R, RD>(ydparams, kidia, kfdia, klon, klev, ptsphy, pt, pq, ptendency_cml_t, ptendency_cml_q,  \
                          ptendency_cml_a, ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, ptendency_loc_a,   \
                          ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis, plu, plude, pludeli, \
                          psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz, pfsqlf, pfsqif, pfcqnng,        \
                          pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng, pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl,   \
                          pfhpsn, pvfa, pdyna, pdynl, pdyni, pdyns, phrsw, phrlw, kpbltype, kctop, pextra, kfldx,   \
                          pgaw, pfsd)

#define DIAG_CALL(R, RD)                                                                                         \
  const cloudsc_params_t* qd = ydparams;                                                                          \
  cloudsc_gpu::KBudPtrs<R> bud{pvfa, pdyna, pdynl, pdyni, pdyns, phrsw, phrlw, pvfl, pvfi, kpbltype, kctop,        \
                               pextra, kfldx, qd ? qd->lcldbud : 0, R(qd ? qd->reisthsc : 0.0),                   \
                               R(qd ? qd->ralfdcp : 0.0), pgaw, pfsd, qd ? qd->lrad_cloud_inhomog : 0,            \
                               R(qd ? qd->ra : 0.0), R(qd ? qd->rpi : 0.0), R(qd ? qd->rcloud_frac_std : 0.0),    \
                               qd ? qd->lcloud_inhomog : 0};                                                    \
  return blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy,                                          \
                        blocked_ptrs<R>(pt, pq, ptendency_cml_t, ptendency_cml_q, ptendency_cml_a,                \
                                        ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, ptendency_loc_a,     \
                                        ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis, plu,  \
                                        plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz,  \
                                        pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng,       \
                                        pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn),                      \
                        stream, nullptr, &bud)

#define BLOCKED_CALL(R, RD)                                                                                      \
  return blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy,                                          \
                        blocked_ptrs<R>(pt, pq, ptendency_cml_t, ptendency_cml_q, ptendency_cml_a,                \
                                        ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, ptendency_loc_a,     \
                                        ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis, plu,  \
                                        plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz,  \
                                        pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng,       \
                                        pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn),                      \
                        stream)

#define FUSED_CALL(R, RD)                                                                                        \
  cloudsc_gpu::KSatPtrs<R> sat{phrsw, phrlw, ptendency_vdf_t, ptendency_vdf_q, pvervel, ptendency_cml_t,          \
                               ptendency_cml_q, ptendency_cml_a, ptendency_cml_cld, ptendency_satadj_t,           \
                               ptendency_satadj_q, ptendency_satadj_a, ptendency_satadj_cld};                     \
  return blocked<R, RD>(ydparams, ngptot, nproma, klev, ngpblks, ptsphy,                                          \
                        blocked_ptrs<R>(pt, pq, ptendency_cml_t, ptendency_cml_q, ptendency_cml_a,                \
                                        ptendency_cml_cld, ptendency_loc_t, ptendency_loc_q, ptendency_loc_a,     \
                                        ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm, ldcum, ktype, peis, plu,  \
                                        plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot, prainfrac_toprfz,  \
                                        pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng, pfcqsng,       \
                                        pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn),                      \
                        stream, &sat)

extern "C" {

int cloudsc_satadj_cloudsc_blocked_dp_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(double)) { FUSED_CALL(double, double); }
int cloudsc_satadj_cloudsc_blocked_sp_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(float)) { FUSED_CALL(float, float); }
int cloudsc_satadj_cloudsc_blocked_mixed_c(CLOUDSC_SATADJ_CLOUDSC_BLOCKED_ARGS(float)) { FUSED_CALL(float, double); }

int cloudsc_blocked_diag_dp_c(CLOUDSC_BLOCKED_DIAG_ARGS(double)) { DIAG_CALL(double, double); }
int cloudsc_blocked_diag_sp_c(CLOUDSC_BLOCKED_DIAG_ARGS(float)) { DIAG_CALL(float, float); }
int cloudsc_blocked_diag_mixed_c(CLOUDSC_BLOCKED_DIAG_ARGS(float)) { DIAG_CALL(float, double); }

int cloudsc_dp_c(CLOUDSC_ARGS(double)) { PER_BLOCK_CALL(double, double); }
int cloudsc_sp_c(CLOUDSC_ARGS(float)) { PER_BLOCK_CALL(float, float); }
int cloudsc_mixed_c(CLOUDSC_ARGS(float)) { PER_BLOCK_CALL(float, double); }

int cloudsc_mixed_noalloc_c(CLOUDSC_ARGS(float), void* stream) {
  (void)lddiag, (void)pdynr, (void)pvervel, (void)pgeom, (void)pgp2dspp, (void)ldslphy, (void)plcrit_aer,
      (void)picrit_aer, (void)pre_ice, (void)pccn, (void)pnice;
  const NoAlloc na{(hipStream_t)stream};
  return per_block<float, double>(ydparams, kidia, kfdia, klon, klev, ptsphy, pt, pq, ptendency_cml_t,
                                  ptendency_cml_q, ptendency_cml_a, ptendency_cml_cld, ptendency_loc_t,
                                  ptendency_loc_q, ptendency_loc_a, ptendency_loc_cld, pvfl, pvfi, pap, paph, plsm,
                                  ldcum, ktype, peis, plu, plude, pludeli, psnde, pmfu, pmfd, pa, pclv, pcovptot,
                                  prainfrac_toprfz, pfsqlf, pfsqif, pfcqnng, pfcqlng, pfsqrf, pfsqsf, pfcqrng,
                                  pfcqsng, pfsqltur, pfsqitur, pfplsl, pfplsn, pfhpsl, pfhpsn, pvfa, pdyna, pdynl,
                                  pdyni, pdyns, phrsw, phrlw, kpbltype, kctop, pextra, kfldx, pgaw, pfsd, &na);
}

int cloudsc_blocked_dp_c(CLOUDSC_BLOCKED_ARGS(double)) { BLOCKED_CALL(double, double); }
int cloudsc_blocked_sp_c(CLOUDSC_BLOCKED_ARGS(float)) { BLOCKED_CALL(float, float); }
int cloudsc_blocked_mixed_c(CLOUDSC_BLOCKED_ARGS(float)) { BLOCKED_CALL(float, double); }

const char* cloudsc_last_error(void) { return cloudsc_capi::g_last_error.c_str(); }

int cloudsc_last_warnings(int64_t* n_negative_liquid, int64_t* n_negative_ice) {
  if (!n_negative_liquid || !n_negative_ice)
    return cloudsc_capi::fail(CLOUDSC_ENULL, "cloudsc_last_warnings: NULL output pointer");
  *n_negative_liquid = cloudsc_capi::g_last_warn[0];
  *n_negative_ice = cloudsc_capi::g_last_warn[1];
  return CLOUDSC_OK;
}

}  // extern "C"

extern "C" int cloudsc_tune_set_block(int variant, int klev, int block) {
  if (variant < 0 || variant > 2 || klev < 2 || !(block == 0 || block == 64 || block == 128 || block == 256 || block == 512))
    return cloudsc_capi::fail(CLOUDSC_EDIMS, "cloudsc_tune_set_block: variant 0..2, KLEV >= 2, block 0/64/128/256/512");
  if (block == 0) {
    cloudsc_gpu::set_tuned_launch(variant, klev, nullptr);
  } else {
    cloudsc_gpu::LaunchCfg l;
    l.block = block;
    cloudsc_gpu::set_tuned_launch(variant, klev, &l);
  }
  return CLOUDSC_OK;
}

extern "C" int cloudsc_tune_get_block(int variant, int klev) {
  return cloudsc_gpu::tuned_launch(variant, klev).block;
}

extern "C" int cloudsc_tune_set_launch(int variant, int klev, int block, int waves_per_simd, int klev_compiled) {
  if (variant < 0 || variant > 2 || klev < 2)
    return cloudsc_capi::fail(CLOUDSC_EDIMS, "cloudsc_tune_set_launch: variant 0..2, KLEV >= 2");
  if (block == 0) {
    cloudsc_gpu::set_tuned_launch(variant, klev, nullptr);
    return CLOUDSC_OK;
  }
  cloudsc_gpu::LaunchCfg l;
  l.block = block;
  l.wpe = waves_per_simd;
  l.kspec = klev_compiled;
  if (!cloudsc_gpu::valid_launch(variant, l))
    return cloudsc_capi::fail(CLOUDSC_EDIMS,
                              "cloudsc_tune_set_launch: block 64/128/256/512; waves per SIMD 0, or 2 (dp), 4 (sp), "
                              "3 or 4 (mixed); klev_compiled 0 or 1; waves per SIMD and klev_compiled only with "
                              "block 128 or 256");
  cloudsc_gpu::set_tuned_launch(variant, klev, &l);
  return CLOUDSC_OK;
}

extern "C" int cloudsc_tune_get_launch(int variant, int klev, int* block, int* waves_per_simd, int* klev_compiled) {
  if (variant < 0 || variant > 2 || !block || !waves_per_simd || !klev_compiled)
    return cloudsc_capi::fail(CLOUDSC_ENULL, "cloudsc_tune_get_launch: variant 0..2 and three output pointers");
  const cloudsc_gpu::LaunchCfg l = cloudsc_gpu::tuned_launch(variant, klev);
  *block = l.block;
  *waves_per_simd = l.wpe;
  *klev_compiled = l.kspec;
  return CLOUDSC_OK;
}
