// Device-resident blocked fields with the field API's ownership model, and the physics step on them
// (include/cloudsc_field.h).  Status handling follows ifs-source/contrib/field_api/field_RANKSUFF_module.fypp
// (line numbers below); the data movement is HIP instead of OpenACC.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/cloudsc_field.h"
#include "../../include/cloudsc_hip.h"
#include "../../include/cloudsc_satadj.h"
#include "capi_util.h"

using cloudsc_capi::fail;

namespace {
constexpr int UNINITIALIZED = CLOUDSC_FIELD_DEVFRESH | CLOUDSC_FIELD_HSTFRESH;  // :28, both copies "fresh"
}

struct cloudsc_field {
  bool owner = false;
  size_t esz = 0;
  int rank = 0;
  int64_t lb[5] = {1, 1, 1, 1, 1}, ub[5] = {0, 0, 0, 0, 0};
  size_t nelem = 0;
  void* host = nullptr;
  void* dev = nullptr;
  int status = 0;
  bool has_init = false;
  alignas(16) unsigned char init[16] = {};
  int64_t h2d = 0, d2h = 0;

  size_t bytes() const { return nelem * esz; }
  int64_t extent(int i) const { return ub[i] - lb[i] + 1; }
};

namespace {

int set_shape(cloudsc_field* f, size_t esz, int rank, const int64_t* lb, const int64_t* ub, const char* who) {
  if (rank < 1 || rank > 5 || (esz != 1 && esz != 2 && esz != 4 && esz != 8 && esz != 16))
    return fail(CLOUDSC_EDIMS, who);
  f->esz = esz;
  f->rank = rank;
  f->nelem = 1;
  for (int i = 0; i < rank; ++i) {
    f->lb[i] = lb ? lb[i] : 1;
    f->ub[i] = ub[i];
    if (f->extent(i) < 1) return fail(CLOUDSC_EDIMS, who);
    f->nelem *= (size_t)f->extent(i);
  }
  return CLOUDSC_OK;
}

// OWNER ALLOCATE (:185-191) and the INIT_VALUE fill of OWNER_INIT (:175-181)
int owner_allocate(cloudsc_field* f) {
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  const hipError_t e = hipHostMalloc(&f->host, f->bytes(), hipHostMallocDefault);
  if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloudsc_field: hipHostMalloc", e);
  f->status = UNINITIALIZED;
  if (f->has_init) {
    unsigned char* h = (unsigned char*)f->host;
    for (size_t i = 0; i < f->nelem; ++i) std::memcpy(h + i * f->esz, f->init, f->esz);
    f->status = CLOUDSC_FIELD_HSTFRESH;
  }
  return CLOUDSC_OK;
}

int copy(cloudsc_field* f, bool to_device, void* stream) {
  const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  void* dst = to_device ? f->dev : f->host;
  const void* src = to_device ? f->host : f->dev;
  const hipError_t e = stream ? hipMemcpyAsync(dst, src, f->bytes(), kind, (hipStream_t)stream)
                              : hipMemcpy(dst, src, f->bytes(), kind);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_field: copy", e);
  if (to_device)
    ++f->h2d;
  else
    ++f->d2h;
  return CLOUDSC_OK;
}

}  // namespace

extern "C" {

int cloudsc_field_wrap(void* host, size_t elem_size, int rank, const int64_t* lbounds, const int64_t* ubounds,
                       cloudsc_field** out) {
  if (!host || !ubounds || !out) return fail(CLOUDSC_ENULL, "cloudsc_field_wrap: NULL argument");
  cloudsc_field* f = new (std::nothrow) cloudsc_field;
  if (!f) return fail(CLOUDSC_ENOMEM, "cloudsc_field_wrap: out of host memory");
  const int rc = set_shape(f, elem_size, rank, lbounds, ubounds, "cloudsc_field_wrap: bad rank, bounds or size");
  if (rc) {
    delete f;
    return rc;
  }
  f->host = host;
  f->status = CLOUDSC_FIELD_HSTFRESH;  // WRAP (:123)
  *out = f;
  return CLOUDSC_OK;
}

int cloudsc_field_owner(size_t elem_size, int rank, const int64_t* lbounds, const int64_t* ubounds, int delayed,
                        const void* init_value, cloudsc_field** out) {
  if (!ubounds || !out) return fail(CLOUDSC_ENULL, "cloudsc_field_owner: NULL argument");
  cloudsc_field* f = new (std::nothrow) cloudsc_field;
  if (!f) return fail(CLOUDSC_ENOMEM, "cloudsc_field_owner: out of host memory");
  int rc = set_shape(f, elem_size, rank, lbounds, ubounds, "cloudsc_field_owner: bad rank, bounds or size");
  if (rc) {
    delete f;
    return rc;
  }
  f->owner = true;
  f->status = CLOUDSC_FIELD_UNALLOCATED;  // OWNER_INIT (:168)
  if (init_value) {
    f->has_init = true;
    std::memcpy(f->init, init_value, elem_size);
  }
  if (!delayed) {
    rc = owner_allocate(f);
    if (rc) {
      delete f;
      return rc;
    }
  }
  *out = f;
  return CLOUDSC_OK;
}

// GET_DEVICE_DATA (:456-478) with OWNER_GET_DEVICE_DATA's allocation on first use (:480-497)
int cloudsc_field_get_device_data(cloudsc_field* f, int mode, void** dev, void* stream) {
  if (!f) return fail(CLOUDSC_ENULL, "cloudsc_field_get_device_data: NULL field");
  if (mode < 1 || mode > 3) return fail(CLOUDSC_EDIMS, "cloudsc_field_get_device_data: mode must be RD, WR or RDWR");
  int rc = cloudsc_capi::check_device();
  if (rc) return rc;
  if (f->status == CLOUDSC_FIELD_UNALLOCATED) {
    rc = owner_allocate(f);
    if (rc) return rc;
  }
  if (!f->dev) {
    const hipError_t e = hipMalloc(&f->dev, f->bytes());
    if (e != hipSuccess) {
      f->dev = nullptr;
      return fail(CLOUDSC_ENOMEM, "cloudsc_field_get_device_data: hipMalloc", e);
    }
  }
  if (!(f->status & CLOUDSC_FIELD_DEVFRESH)) {
    rc = copy(f, true, stream);
    if (rc) return rc;
    f->status |= CLOUDSC_FIELD_DEVFRESH;
  }
  if (mode & CLOUDSC_FIELD_WR) f->status &= ~CLOUDSC_FIELD_HSTFRESH;
  if (dev) *dev = f->dev;
  return CLOUDSC_OK;
}

// GET_HOST_DATA (:377-396) with OWNER_GET_HOST_DATA's allocation on first use (:398-413)
int cloudsc_field_get_host_data(cloudsc_field* f, int mode, void** host, void* stream) {
  if (!f) return fail(CLOUDSC_ENULL, "cloudsc_field_get_host_data: NULL field");
  if (mode < 1 || mode > 3) return fail(CLOUDSC_EDIMS, "cloudsc_field_get_host_data: mode must be RD, WR or RDWR");
  int rc;
  if (f->status == CLOUDSC_FIELD_UNALLOCATED) {
    rc = owner_allocate(f);
    if (rc) return rc;
  }
  if (!(f->status & CLOUDSC_FIELD_HSTFRESH)) {
    if (!f->dev) return fail(CLOUDSC_ESTATE, "cloudsc_field_get_host_data: no fresh copy anywhere");
    rc = copy(f, false, stream);
    if (rc) return rc;
    f->status |= CLOUDSC_FIELD_HSTFRESH;
  }
  if (mode & CLOUDSC_FIELD_WR) f->status &= ~CLOUDSC_FIELD_DEVFRESH;
  if (host) *host = f->host;
  return CLOUDSC_OK;
}

int cloudsc_field_sync_device(cloudsc_field* f, int mode, void* stream) {
  return cloudsc_field_get_device_data(f, mode, nullptr, stream);
}

int cloudsc_field_sync_host(cloudsc_field* f, int mode, void* stream) {
  return cloudsc_field_get_host_data(f, mode, nullptr, stream);
}

// GET_VIEW (:193-220), persistent fields: the slice of one last-dimension index
int cloudsc_field_get_view(cloudsc_field* f, int64_t block, void** view) {
  if (!f || !view) return fail(CLOUDSC_ENULL, "cloudsc_field_get_view: NULL argument");
  if (f->status == CLOUDSC_FIELD_UNALLOCATED || !(f->status & CLOUDSC_FIELD_HSTFRESH))
    return fail(CLOUDSC_ESTATE, "cloudsc_field_get_view: data is not present on the host");
  const int last = f->rank - 1;
  if (block < f->lb[last] || block > f->ub[last]) return fail(CLOUDSC_EDIMS, "cloudsc_field_get_view: block index");
  const size_t slice = f->nelem / (size_t)f->extent(last);
  *view = (unsigned char*)f->host + (size_t)(block - f->lb[last]) * slice * f->esz;
  f->status &= ~CLOUDSC_FIELD_DEVFRESH;  // the view may be written (:216-219)
  return CLOUDSC_OK;
}

// DELETE_DEVICE (:222-229).  A host copy that is stale is brought back first, and the device copy loses its
// fresh bit, so that no later call reads a freed or re-created device array as current.
int cloudsc_field_delete_device(cloudsc_field* f) {
  if (!f) return fail(CLOUDSC_ENULL, "cloudsc_field_delete_device: NULL field");
  if (!f->dev) return CLOUDSC_OK;
  if (f->status != CLOUDSC_FIELD_UNALLOCATED && !(f->status & CLOUDSC_FIELD_HSTFRESH)) {
    const int rc = cloudsc_field_get_host_data(f, CLOUDSC_FIELD_RD, nullptr, nullptr);
    if (rc) return rc;
  }
  const hipError_t e = hipFree(f->dev);
  f->dev = nullptr;
  f->status &= ~CLOUDSC_FIELD_DEVFRESH;
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_field_delete_device: hipFree", e);
  return CLOUDSC_OK;
}

int cloudsc_field_status(const cloudsc_field* f, int* status) {
  if (!f || !status) return fail(CLOUDSC_ENULL, "cloudsc_field_status: NULL argument");
  *status = f->status;
  return CLOUDSC_OK;
}

int cloudsc_field_stats(const cloudsc_field* f, int64_t* h2d, int64_t* d2h) {
  if (!f || !h2d || !d2h) return fail(CLOUDSC_ENULL, "cloudsc_field_stats: NULL argument");
  *h2d = f->h2d;
  *d2h = f->d2h;
  return CLOUDSC_OK;
}

// FIELD_DELETE: WRAPPER_FINAL (:238-244) syncs the wrapped array; OWNER_FINAL (:246-253) frees the host data
int cloudsc_field_delete(cloudsc_field* f) {
  if (!f) return CLOUDSC_OK;
  int rc = CLOUDSC_OK;
  if (!f->owner && f->dev && !(f->status & CLOUDSC_FIELD_HSTFRESH))
    rc = cloudsc_field_get_host_data(f, CLOUDSC_FIELD_RD, nullptr, nullptr);
  if (f->dev) {
    const hipError_t e = hipFree(f->dev);
    if (e != hipSuccess && !rc) rc = fail(CLOUDSC_EHIP, "cloudsc_field_delete: hipFree", e);
  }
  if (f->owner && f->host) {
    const hipError_t e = hipHostFree(f->host);
    if (e != hipSuccess && !rc) rc = fail(CLOUDSC_EHIP, "cloudsc_field_delete: hipHostFree", e);
  }
  delete f;
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// physics step on fields
// ---------------------------------------------------------------------------
namespace {

enum Kind { K2D, KH, KC5, KC4, KC2, K1D };

struct Want {
  cloudsc_field* f;
  Kind kind;
  int mode;
  bool integer;
};

int check_shape(const cloudsc_field* f, Kind k, bool integer, size_t esz, int64_t np, int64_t kl, int64_t nb) {
  const size_t want = integer ? 4 : esz;
  if (f->esz != want) return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: field element size");
  int64_t ext[4];
  int rank = 0;
  switch (k) {
    case K2D: ext[0] = np, ext[1] = kl, ext[2] = nb, rank = 3; break;
    case KH: ext[0] = np, ext[1] = kl + 1, ext[2] = nb, rank = 3; break;
    case KC5: ext[0] = np, ext[1] = kl, ext[2] = 5, ext[3] = nb, rank = 4; break;
    case KC4: ext[0] = np, ext[1] = kl, ext[2] = 4, ext[3] = nb, rank = 4; break;
    case KC2: ext[0] = np, ext[1] = kl, ext[2] = 2, ext[3] = nb, rank = 4; break;
    case K1D: ext[0] = np, ext[1] = nb, rank = 2; break;
  }
  if (f->rank != rank) return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: field rank");
  for (int i = 0; i < rank; ++i)
    if (f->extent(i) != ext[i]) return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: field extent");
  return CLOUDSC_OK;
}

template <typename R>
int run_step(const cloudsc_params_t* q, int precision, int ngptot, double ptsphy, const cloudsc_step_fields* s,
             void* stream) {
  const int RD = CLOUDSC_FIELD_RD, RW = CLOUDSC_FIELD_RDWR;
  const bool sat = s->phrsw || s->phrlw || s->ptendency_vdf_t || s->ptendency_vdf_q || s->pvervel;
  const int cml = sat ? RW : RD;
  Want w[] = {{s->pt, K2D, RD, false},
              {s->pq, K2D, RD, false},
              {s->ptendency_cml_t, K2D, cml, false},
              {s->ptendency_cml_q, K2D, cml, false},
              {s->ptendency_cml_a, K2D, cml, false},
              {s->ptendency_cml_cld, KC5, cml, false},
              {s->ptendency_loc_t, K2D, RW, false},
              {s->ptendency_loc_q, K2D, RW, false},
              {s->ptendency_loc_a, K2D, RW, false},
              {s->ptendency_loc_cld, KC5, RW, false},
              {s->pvfl, K2D, RD, false},
              {s->pvfi, K2D, RD, false},
              {s->pap, K2D, RD, false},
              {s->paph, KH, RD, false},
              {s->plsm, K1D, RD, false},
              {s->ldcum, K1D, RD, true},
              {s->ktype, K1D, RD, true},
              {s->peis, K1D, RW, false},
              {s->plu, K2D, RD, false},
              {s->plude, K2D, RW, false},
              {s->pludeli, KC4, RD, false},
              {s->psnde, KC2, RD, false},
              {s->pmfu, K2D, RD, false},
              {s->pmfd, K2D, RD, false},
              {s->pa, K2D, RD, false},
              {s->pclv, KC5, RD, false},
              {s->pcovptot, K2D, RW, false},
              {s->prainfrac_toprfz, K1D, RW, false},
              {s->pfsqlf, KH, RW, false},
              {s->pfsqif, KH, RW, false},
              {s->pfcqnng, KH, RW, false},
              {s->pfcqlng, KH, RW, false},
              {s->pfsqrf, KH, RW, false},
              {s->pfsqsf, KH, RW, false},
              {s->pfcqrng, KH, RW, false},
              {s->pfcqsng, KH, RW, false},
              {s->pfsqltur, KH, RW, false},
              {s->pfsqitur, KH, RW, false},
              {s->pfplsl, KH, RW, false},
              {s->pfplsn, KH, RW, false},
              {s->pfhpsl, KH, RW, false},
              {s->pfhpsn, KH, RW, false},
              {s->phrsw, K2D, RD, false},
              {s->phrlw, K2D, RD, false},
              {s->ptendency_vdf_t, K2D, RD, false},
              {s->ptendency_vdf_q, K2D, RD, false},
              {s->pvervel, K2D, RD, false},
              {s->ptendency_satadj_t, K2D, RW, false},
              {s->ptendency_satadj_q, K2D, RW, false},
              {s->ptendency_satadj_a, K2D, RW, false},
              {s->ptendency_satadj_cld, KC5, RW, false}};
  constexpr int NW = sizeof(w) / sizeof(w[0]);
  constexpr int NCLOUDSC = 42, NSATIN = 5;
  if (!s->pt) return fail(CLOUDSC_ENULL, "cloudsc_step_fields_run: PT field is NULL");
  if (s->pt->rank != 3) return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: PT must be (NPROMA,KLEV,NGPBLKS)");
  const int64_t np = s->pt->extent(0), kl = s->pt->extent(1), nb = s->pt->extent(2);
  if (nb != ((int64_t)ngptot + np - 1) / np) return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: NGPBLKS");
  for (int i = 0; i < NW; ++i) {
    if (!w[i].f) {
      if (i < NCLOUDSC) return fail(CLOUDSC_ENULL, "cloudsc_step_fields_run: a CLOUDSC field is NULL");
      if (i < NCLOUDSC + NSATIN && sat) return fail(CLOUDSC_ENULL, "cloudsc_step_fields_run: a CLOUD_SATADJ field is NULL");
      continue;
    }
    const int rc = check_shape(w[i].f, w[i].kind, w[i].integer, sizeof(R), np, kl, nb);
    if (rc) return rc;
  }
  void* d[NW] = {};
  for (int i = 0; i < NW; ++i) {
    if (!w[i].f) continue;
    const int rc = cloudsc_field_get_device_data(w[i].f, w[i].mode, &d[i], stream);
    if (rc) return rc;
  }
#define A(i) ((R*)d[i])
#define I(i) ((const int32_t*)d[i])
  const int n = (int)np, k = (int)kl, b = (int)nb;
  const R dt = (R)ptsphy;
  int rc;
  if (precision == 0) {
    rc = sat ? cloudsc_satadj_cloudsc_blocked_dp_c(
                   q, ngptot, n, k, b, (double)dt, (double*)A(0), (double*)A(1), (double*)A(2), (double*)A(3),
                   (double*)A(4), (double*)A(5), (double*)A(6), (double*)A(7), (double*)A(8), (double*)A(9),
                   (double*)A(10), (double*)A(11), (double*)A(12), (double*)A(13), (double*)A(14), I(15), I(16),
                   (double*)A(17), (double*)A(18), (double*)A(19), (double*)A(20), (double*)A(21), (double*)A(22),
                   (double*)A(23), (double*)A(24), (double*)A(25), (double*)A(26), (double*)A(27), (double*)A(28),
                   (double*)A(29), (double*)A(30), (double*)A(31), (double*)A(32), (double*)A(33), (double*)A(34),
                   (double*)A(35), (double*)A(36), (double*)A(37), (double*)A(38), (double*)A(39), (double*)A(40),
                   (double*)A(41), (double*)A(42), (double*)A(43), (double*)A(44), (double*)A(45), (double*)A(46),
                   (double*)A(47), (double*)A(48), (double*)A(49), (double*)A(50), stream)
             : cloudsc_blocked_dp_c(q, ngptot, n, k, b, (double)dt, (double*)A(0), (double*)A(1), (double*)A(2),
                                    (double*)A(3), (double*)A(4), (double*)A(5), (double*)A(6), (double*)A(7),
                                    (double*)A(8), (double*)A(9), (double*)A(10), (double*)A(11), (double*)A(12),
                                    (double*)A(13), (double*)A(14), I(15), I(16), (double*)A(17), (double*)A(18),
                                    (double*)A(19), (double*)A(20), (double*)A(21), (double*)A(22), (double*)A(23),
                                    (double*)A(24), (double*)A(25), (double*)A(26), (double*)A(27), (double*)A(28),
                                    (double*)A(29), (double*)A(30), (double*)A(31), (double*)A(32), (double*)A(33),
                                    (double*)A(34), (double*)A(35), (double*)A(36), (double*)A(37), (double*)A(38),
                                    (double*)A(39), (double*)A(40), (double*)A(41), stream);
  } else {
    auto fused = precision == 1 ? cloudsc_satadj_cloudsc_blocked_sp_c : cloudsc_satadj_cloudsc_blocked_mixed_c;
    auto plain = precision == 1 ? cloudsc_blocked_sp_c : cloudsc_blocked_mixed_c;
    rc = sat ? fused(q, ngptot, n, k, b, (float)dt, (float*)A(0), (float*)A(1), (float*)A(2), (float*)A(3),
                     (float*)A(4), (float*)A(5), (float*)A(6), (float*)A(7), (float*)A(8), (float*)A(9),
                     (float*)A(10), (float*)A(11), (float*)A(12), (float*)A(13), (float*)A(14), I(15), I(16),
                     (float*)A(17), (float*)A(18), (float*)A(19), (float*)A(20), (float*)A(21), (float*)A(22),
                     (float*)A(23), (float*)A(24), (float*)A(25), (float*)A(26), (float*)A(27), (float*)A(28),
                     (float*)A(29), (float*)A(30), (float*)A(31), (float*)A(32), (float*)A(33), (float*)A(34),
                     (float*)A(35), (float*)A(36), (float*)A(37), (float*)A(38), (float*)A(39), (float*)A(40),
                     (float*)A(41), (float*)A(42), (float*)A(43), (float*)A(44), (float*)A(45), (float*)A(46),
                     (float*)A(47), (float*)A(48), (float*)A(49), (float*)A(50), stream)
             : plain(q, ngptot, n, k, b, (float)dt, (float*)A(0), (float*)A(1), (float*)A(2), (float*)A(3),
                     (float*)A(4), (float*)A(5), (float*)A(6), (float*)A(7), (float*)A(8), (float*)A(9),
                     (float*)A(10), (float*)A(11), (float*)A(12), (float*)A(13), (float*)A(14), I(15), I(16),
                     (float*)A(17), (float*)A(18), (float*)A(19), (float*)A(20), (float*)A(21), (float*)A(22),
                     (float*)A(23), (float*)A(24), (float*)A(25), (float*)A(26), (float*)A(27), (float*)A(28),
                     (float*)A(29), (float*)A(30), (float*)A(31), (float*)A(32), (float*)A(33), (float*)A(34),
                     (float*)A(35), (float*)A(36), (float*)A(37), (float*)A(38), (float*)A(39), (float*)A(40),
                     (float*)A(41), stream);
  }
#undef A
#undef I
  return rc;
}

}  // namespace

extern "C" int cloudsc_step_fields_run(const cloudsc_params_t* ydparams, int precision, int ngptot, double ptsphy,
                                       const cloudsc_step_fields* fields, void* stream) {
  if (!ydparams || !fields) return fail(CLOUDSC_ENULL, "cloudsc_step_fields_run: NULL argument");
  if (precision == 0) return run_step<double>(ydparams, 0, ngptot, ptsphy, fields, stream);
  if (precision == 1 || precision == 2) return run_step<float>(ydparams, precision, ngptot, ptsphy, fields, stream);
  return fail(CLOUDSC_EDIMS, "cloudsc_step_fields_run: precision must be 0 (FP64), 1 (FP32) or 2 (mixed)");
}
