// Synthetic-atmosphere generator: device kernel and host loop over the same
// column function (synth.h).
#include <hip/hip_runtime.h>

#include "../../include/cloudsc_synth.h"
#include "synth.h"

namespace {

synth::Thermo make_thermo(const cloudsc_params_t& p) {
  synth::Thermo t;
  t.rd = p.rd;
  t.rg = p.rg;
  t.retv = p.retv;
  t.rtt = p.rtt;
  t.r2es = p.r2es;
  t.r3les = p.r3les;
  t.r3ies = p.r3ies;
  t.r4les = p.r4les;
  t.r4ies = p.r4ies;
  t.rtice = p.rtice;
  t.rtwat = p.rtwat;
  t.rtwat_rtice_r = p.rtwat_rtice_r;
  return t;
}

template <typename R>
__global__ __launch_bounds__(256) void synth_kernel(uint64_t seed, synth::Thermo th, int klev, int nproma,
                                                    const double* ahalf, const double* bhalf, synth::Fields<R> f,
                                                    int64_t ncols) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ncols) return;
  synth::fill_column<R>(seed, th, klev, nproma, ahalf, bhalf, f, g);
}

bool bad_desc(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d) {
  return !p || !d || !d->ahalf || !d->bhalf || d->nproma < 1 || d->klev < 2 || d->ngptot < 1 ||
         d->ngpblks != (d->ngptot + d->nproma - 1) / d->nproma;
}

#define SYN_FIELDS_INIT                                                                                      \
  f.pt = pt; f.pq = pq; f.tcml_t = ptendency_cml_t; f.tcml_q = ptendency_cml_q; f.tcml_a = ptendency_cml_a; \
  f.tcml_cld = ptendency_cml_cld; f.pvfl = pvfl; f.pvfi = pvfi; f.pap = pap; f.paph = paph; f.plsm = plsm;  \
  f.ldcum = ldcum; f.ktype = ktype; f.peis = peis; f.plu = plu; f.plude = plude; f.pludeli = pludeli;       \
  f.psnde = psnde; f.pmfu = pmfu; f.pmfd = pmfd; f.pa = pa; f.pclv = pclv;

template <typename R>
int synth_device(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, const synth::Fields<R>& f, void* stream) {
  if (bad_desc(p, d)) return CLOUDSC_EDIMS;
  const int64_t ncols = (int64_t)d->ngpblks * d->nproma;
  const int64_t nblk = (ncols + 255) / 256;
  hipLaunchKernelGGL(synth_kernel<R>, dim3((unsigned)nblk), dim3(256), 0, (hipStream_t)stream, d->seed,
                     make_thermo(*p), d->klev, d->nproma, d->ahalf, d->bhalf, f, ncols);
  return hipGetLastError() == hipSuccess ? CLOUDSC_OK : CLOUDSC_EHIP;
}

template <typename R>
int synth_host(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, int64_t col0, int64_t ncol,
               const synth::Fields<R>& f) {
  if (bad_desc(p, d)) return CLOUDSC_EDIMS;
  const int64_t ncols = (int64_t)d->ngpblks * d->nproma;
  if (col0 < 0 || ncol < 0 || col0 + ncol > ncols) return CLOUDSC_EDIMS;
  const synth::Thermo th = make_thermo(*p);
  for (int64_t g = col0; g < col0 + ncol; ++g)
    synth::fill_column<R>(d->seed, th, d->klev, d->nproma, d->ahalf, d->bhalf, f, g);
  return CLOUDSC_OK;
}

}  // namespace

extern "C" {

int cloudsc_synth_sp(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, CLOUDSC_SYNTH_FIELDS(float),
                     void* stream) {
  synth::Fields<float> f;
  SYN_FIELDS_INIT
  return synth_device<float>(p, d, f, stream);
}
int cloudsc_synth_dp(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, CLOUDSC_SYNTH_FIELDS(double),
                     void* stream) {
  synth::Fields<double> f;
  SYN_FIELDS_INIT
  return synth_device<double>(p, d, f, stream);
}
int cloudsc_synth_host_sp(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, int64_t col0, int64_t ncol,
                          CLOUDSC_SYNTH_FIELDS(float)) {
  synth::Fields<float> f;
  SYN_FIELDS_INIT
  return synth_host<float>(p, d, col0, ncol, f);
}
int cloudsc_synth_host_dp(const cloudsc_params_t* p, const cloudsc_synth_desc_t* d, int64_t col0, int64_t ncol,
                          CLOUDSC_SYNTH_FIELDS(double)) {
  synth::Fields<double> f;
  SYN_FIELDS_INIT
  return synth_host<double>(p, d, col0, ncol, f);
}
void cloudsc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  synth::philox4x32_10(ctr, key, out);
}

}  // extern "C"
