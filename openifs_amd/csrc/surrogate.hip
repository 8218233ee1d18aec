// Learned CLOUDSC surrogate and the hybrid dispatcher (declared in include/cloudsc_surrogate.h).
//
// The surrogate maps one column's CLOUDSC inputs (14 features x KLEV) to its main outputs (6 targets x KLEV) with
// the BF16 MLP of mlp.hip.  Feature and target matrices are built directly from the IFS blocked fields
// (NPROMA, KLEV[,n], NGPBLKS) through LDS-tiled transposes, so that both the blocked reads and the row-major
// matrix writes are coalesced.  The hybrid entry reruns a seeded random subset of columns, and every column its
// gate flags, through the plain kernel: the flagged columns are gathered into a compact blocked field set,
// run by cloudsc_blocked_sp_c's kernel and scattered back.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/cloudsc_surrogate.h"
#include "capi_util.h"
#include "cloudsc_kernel.h"
#include "mlp_internal.h"
#include "synth.h"

namespace cloudsc_gpu {
}

using cloudsc_capi::fail;
using mlp::bf16;
using cloudsc_gpu::KDims;
using cloudsc_gpu::KPtrs;

namespace surr {

constexpr int NF = CLOUDSC_SURR_NFEAT, NT = CLOUDSC_SURR_NTARG;
constexpr int TILE = 64;  // columns x matrix entries per transpose tile
// Standardised features and training targets are clipped to +-CLIP standard deviations: the heavy tails of the
// tendencies (rare columns with values thousands of standard deviations out) would otherwise dominate the loss.
constexpr float CLIP = 10.0f;

inline int pad128(int n) { return (n + 127) / 128 * 128; }
inline int pad256(int n) { return (n + 255) / 256 * 256; }

// Saturation constants (TTHF, fcttre.func.h) for the relative-humidity feature and the gate.
struct Sat {
  float r2es, r3les, r3ies, r4les, r4ies, rtt, rtice, rtwat_rtice_r, retv;
};

struct Dims {
  int ngptot, nproma, klev, ngpblks;
  Sat th;
};

// Saturation specific humidity over the mixed phase, FOEEWM / p limited to 0.5 (fcttre.func.h, cloudsc.F90:1087).
__device__ __forceinline__ float qsat_mix(const Sat& c, float t, float pap) {
  const float x = (fmaxf(c.rtice, fminf(c.rtt, t)) - c.rtice) * c.rtwat_rtice_r;
  const float alfa = fminf(1.0f, x * x);
  const float es = c.r2es * (alfa * expf(c.r3les * (t - c.rtt) / (t - c.r4les)) +
                             (1.0f - alfa) * expf(c.r3ies * (t - c.rtt) / (t - c.r4ies)));
  const float e = fminf(es / pap, 0.5f);
  return e / (1.0f - c.retv * e);
}

// Element (column c, level k) of a (NPROMA, L, n, NGPBLKS) field, species s.
__device__ __forceinline__ size_t fidx(int c, int k, int s, int n, int L, int nproma) {
  const int b = c / nproma, jl = c - b * nproma;
  return (((size_t)b * n + s) * L + k) * nproma + jl;
}

__device__ __forceinline__ float slog(float x) { return log10f(fmaxf(x, 0.0f) + 1e-10f); }

// Raw feature f at level k of column c.
__device__ __forceinline__ float feature(const KPtrs<float>& p, const Dims& d, int c, int f, int k) {
  const int np = d.nproma, L = d.klev;
  switch (f) {
    case 0: return p.pt[fidx(c, k, 0, 1, L, np)];
    case 1: return p.pq[fidx(c, k, 0, 1, L, np)];
    case 2: return p.pa[fidx(c, k, 0, 1, L, np)];
    case 3: case 4: case 5: case 6: return slog(p.pclv[fidx(c, k, f - 3, 5, L, np)]);
    case 7: return p.tcml_t[fidx(c, k, 0, 1, L, np)];
    case 8: return p.tcml_q[fidx(c, k, 0, 1, L, np)];
    case 9: return p.tcml_a[fidx(c, k, 0, 1, L, np)];
    case 10: case 11: return p.tcml_cld[fidx(c, k, f - 10, 5, L, np)];
    case 12: {
      const size_t i = fidx(c, k, 0, 1, L, np);
      return p.pq[i] / qsat_mix(d.th, p.pt[i], p.pap[i]);  // relative humidity over the mixed phase
    }
    default: return slog(p.plude[fidx(c, k, 0, 1, L, np)]);
  }
}

// Raw target t at level k of column c (outputs of the plain kernel).
__device__ __forceinline__ float target(const KPtrs<float>& p, const Dims& d, int c, int t, int k) {
  const int np = d.nproma, L = d.klev;
  switch (t) {
    case 0: return p.tloc_t[fidx(c, k, 0, 1, L, np)];
    case 1: return p.tloc_q[fidx(c, k, 0, 1, L, np)];
    case 2: return p.tloc_a[fidx(c, k, 0, 1, L, np)];
    case 3: case 4: return p.tloc_cld[fidx(c, k, t - 3, 5, L, np)];
    default: return p.pfplsl[fidx(c, k + 1, 0, 1, L + 1, np)] + p.pfplsn[fidx(c, k + 1, 0, 1, L + 1, np)];
  }
}

// Feature matrix X [M x nin] (bf16, standardised) and optionally X^T [nin x M] for columns col0 .. col0+M-1.
// Block: 256 threads, tile of 64 columns x 64 matrix entries; columns beyond NGPTOT and padding entries are 0.
__global__ __launch_bounds__(256) void features(KPtrs<float> p, Dims d, int col0, int M, int nin,
                                                const float* __restrict__ mu, const float* __restrict__ rsd,
                                                bf16* __restrict__ X, bf16* __restrict__ XT) {
  __shared__ float tile[TILE][TILE + 1];
  const int c0 = blockIdx.x * TILE, e0 = blockIdx.y * TILE;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int nfk = NF * d.klev;
  for (int i = ty; i < TILE; i += 4) {  // read: lanes over columns
    const int e = e0 + i, c = c0 + tx, gc = col0 + c;
    float v = 0.0f;
    if (e < nfk && c < M && gc < d.ngptot) {
      const int f = e / d.klev, k = e - f * d.klev;
      v = fminf(fmaxf((feature(p, d, gc, f, k) - mu[e]) * rsd[e], -CLIP), CLIP);
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < TILE; i += 4) {  // write X rows: lanes over entries
    const int c = c0 + i, e = e0 + tx;
    if (c < M && e < nin) X[(size_t)c * nin + e] = (bf16)tile[tx][i];
  }
  if (XT)
    for (int i = ty; i < TILE; i += 4) {  // write X^T rows: lanes over columns
      const int e = e0 + i, c = c0 + tx;
      if (c < M && e < nin) XT[(size_t)e * M + c] = (bf16)tile[i][tx];
    }
}

// Target matrix T [M x nout] (fp32, standardised) for columns col0 .. col0+M-1.
__global__ __launch_bounds__(256) void targets(KPtrs<float> p, Dims d, int col0, int M, int nout,
                                               const float* __restrict__ mu, const float* __restrict__ rsd,
                                               float* __restrict__ T) {
  __shared__ float tile[TILE][TILE + 1];
  const int c0 = blockIdx.x * TILE, e0 = blockIdx.y * TILE;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int ntk = NT * d.klev;
  for (int i = ty; i < TILE; i += 4) {
    const int e = e0 + i, c = c0 + tx, gc = col0 + c;
    float v = 0.0f;
    if (e < ntk && c < M && gc < d.ngptot) {
      const int t = e / d.klev, k = e - t * d.klev;
      v = fminf(fmaxf((target(p, d, gc, t, k) - mu[e]) * rsd[e], -CLIP), CLIP);
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < TILE; i += 4) {
    const int c = c0 + i, e = e0 + tx;
    if (c < M && e < nout) T[(size_t)c * nout + e] = tile[tx][i];
  }
}

// Network output Y [M x nout] (standardised) -> CLOUDSC outputs of columns col0 .. col0+M-1.
__global__ __launch_bounds__(256) void outputs(KPtrs<float> p, Dims d, int col0, int M, int nout,
                                               const float* __restrict__ mu, const float* __restrict__ sd,
                                               const float* __restrict__ Y, float rtt, float rlvtt, float rlstt,
                                               int ncldtop) {
  __shared__ float tile[TILE][TILE + 1];
  const int c0 = blockIdx.x * TILE, e0 = blockIdx.y * TILE;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int ntk = NT * d.klev, L = d.klev, np = d.nproma;
  for (int i = ty; i < TILE; i += 4) {  // read Y rows: lanes over entries
    const int c = c0 + i, e = e0 + tx;
    tile[tx][i] = (c < M && e < ntk) ? Y[(size_t)c * nout + e] * sd[e] + mu[e] : 0.0f;
  }
  __syncthreads();
  for (int i = ty; i < TILE; i += 4) {  // write fields: lanes over columns
    const int e = e0 + i, c = c0 + tx, gc = col0 + c;
    if (e >= ntk || c >= M || gc >= d.ngptot) continue;
    const int t = e / L, k = e - t * L;
    const float v = tile[i][tx];
    switch (t) {
      case 0: p.tloc_t[fidx(gc, k, 0, 1, L, np)] = v; break;
      case 1: p.tloc_q[fidx(gc, k, 0, 1, L, np)] = v; break;
      case 2: p.tloc_a[fidx(gc, k, 0, 1, L, np)] = v; break;
      case 3: case 4:
        p.tloc_cld[fidx(gc, k, t - 3, 5, L, np)] = v;
        p.tloc_cld[fidx(gc, k, t - 1, 5, L, np)] = 0.0f;  // rain and snow tendencies are not predicted
        break;
      default: {
        const float pr = fmaxf(v, 0.0f);
        const bool liquid = p.pt[fidx(gc, k, 0, 1, L, np)] > rtt;
        const float fl = liquid ? pr : 0.0f, fs = liquid ? 0.0f : pr;
        const size_t h = fidx(gc, k + 1, 0, 1, L + 1, np);
        p.pfplsl[h] = fl;
        p.pfplsn[h] = fs;
        p.pfhpsl[h] = -rlvtt * fl;
        p.pfhpsn[h] = -rlstt * fs;
        if (k >= ncldtop - 1) p.pcovptot[fidx(gc, k, 0, 1, L, np)] = 0.0f;  // the plain kernel's write set
        float* ddh[10] = {p.pfsqlf, p.pfsqif, p.pfcqnng, p.pfcqlng, p.pfsqrf,
                          p.pfsqsf, p.pfcqrng, p.pfcqsng, p.pfsqltur, p.pfsqitur};
        for (int j = 0; j < 10; ++j) ddh[j][h] = 0.0f;
        if (k == 0) {
          const size_t h0 = fidx(gc, 0, 0, 1, L + 1, np);
          p.pfplsl[h0] = 0.0f;
          p.pfplsn[h0] = 0.0f;
          p.pfhpsl[h0] = 0.0f;
          p.pfhpsn[h0] = 0.0f;
          for (int j = 0; j < 10; ++j) ddh[j][h0] = 0.0f;
          p.prainfrac_toprfz[fidx(gc, 0, 0, 1, 1, np)] = 0.0f;
        }
      }
    }
  }
}

// Mean and standard deviation over all columns of raw feature (kind 0) or target (kind 1) entry e = blockIdx.x.
__global__ __launch_bounds__(256) void moments(KPtrs<float> p, Dims d, int kind, float* __restrict__ mu,
                                               float* __restrict__ sd) {
  const int e = blockIdx.x, k = e % d.klev, f = e / d.klev;
  double s = 0.0, s2 = 0.0;
  for (int c = threadIdx.x; c < d.ngptot; c += 256) {
    const double v = kind == 0 ? feature(p, d, c, f, k) : target(p, d, c, f, k);
    s += v;
    s2 += v * v;
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = s;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r1[threadIdx.x] += r1[threadIdx.x + h];
      r2[threadIdx.x] += r2[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = r1[0] / d.ngptot;
    const double var = std::max(r2[0] / d.ngptot - m * m, 0.0);
    mu[e] = (float)m;
    sd[e] = (float)std::sqrt(var);  // raw; floors are applied on the host (cloudsc_surrogate_fit_norm)
  }
}

__global__ void reciprocal(const float* __restrict__ a, float* __restrict__ r, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = 1.0f / a[i];
}

__global__ void add_scalar(float* __restrict__ acc, const float* __restrict__ x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) acc[0] += x[0];
}

// PRAINFRAC_TOPRFZ is zeroed for all KLON columns of every block, the padding of the last block included
// (cloudsc.F90:1010); the network path writes the active columns in `outputs`.
__global__ void zero_padding_columns(float* __restrict__ prainfrac, int ngptot, int ntotal) {
  const int c = ngptot + blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ntotal) prainfrac[c] = 0.0f;
}

__global__ void fill_f32(float* p, float v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// Gate score and rerun selection.  Score: max over levels with T < 238 K of q / qsat_mix (FOEEWM with FOEALFA,
// fcttre.func.h).  With flags, each column's selection (seeded 1 % draw or score above the threshold) is stored and
// each group of 256 columns writes how many it selected, for the compaction below.
__global__ __launch_bounds__(256) void select_columns(KPtrs<float> p, Dims d, float frac, uint64_t seed, float thr,
                                                      float* __restrict__ score, uint8_t* __restrict__ flags,
                                                      int* __restrict__ group_counts) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool sel = false;
  if (c < d.ngptot) {
    float sc = 0.0f;
    for (int k = 0; k < d.klev; ++k) {
      const size_t i = fidx(c, k, 0, 1, d.klev, d.nproma);
      const float t = p.pt[i];
      if (t >= 238.0f) continue;
      sc = fmaxf(sc, p.pq[i] / qsat_mix(d.th, t, p.pap[i]));
    }
    if (score) score[c] = sc;
    if (flags) {
      const synth::Draw u = synth::draw(seed, 0x52455255u /* "RERU" */, (uint32_t)c, 0u);
      sel = u.u0 < (double)frac || sc > thr;
      flags[c] = sel ? 1 : 0;
    }
  }
  if (!flags) return;
  __shared__ int wcount[4];
  const uint64_t b = __ballot(sel);
  if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) group_counts[blockIdx.x] = wcount[0] + wcount[1] + wcount[2] + wcount[3];
}

// Exclusive prefix sum of the n group counts in place (one group of 1024 lanes); *total = their sum.
__global__ __launch_bounds__(1024) void scan_counts(int* __restrict__ counts, int n, int* __restrict__ total) {
  __shared__ int part[1024];
  const int t = threadIdx.x, per = (n + 1023) / 1024, lo = min(n, t * per), hi = min(n, lo + per);
  int sum = 0;
  for (int i = lo; i < hi; ++i) sum += counts[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the per-lane sums
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;  // exclusive start of this lane's range
  for (int i = lo; i < hi; ++i) {
    const int ci = counts[i];
    counts[i] = run;
    run += ci;
  }
  if (t == 1023) *total = part[1023];
}

// The selected columns in ascending order: list[offset of the group + rank within it] = column.  Ascending order
// keeps neighbouring lanes of the gather and scatter on neighbouring blocks of the fields.
__global__ __launch_bounds__(256) void compact_columns(const uint8_t* __restrict__ flags, const int* __restrict__ offsets,
                                                       int ngptot, int* __restrict__ list) {
  const int c = blockIdx.x * 256 + threadIdx.x, w = threadIdx.x >> 6;
  const bool sel = c < ngptot && flags[c];
  const uint64_t b = __ballot(sel);
  __shared__ int wcount[4];
  if ((threadIdx.x & 63) == 0) wcount[w] = __popcll(b);
  __syncthreads();
  int base = offsets[blockIdx.x];
  for (int i = 0; i < w; ++i) base += wcount[i];
  const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
  if (sel) list[base + rank] = c;
}

// Column copies between a blocked field set (NPROMA = np_full) and a compact one (NPROMA = np_c), for one
// 2-D slice (NPROMA, L) of each array: element (column, level) at blk*stride + off + level*NPROMA + jl.
struct Slice {
  uint32_t* full;
  uint32_t* comp;
  long full_stride, comp_stride;  // elements per block
  long full_off, comp_off;
  int L;
};
constexpr int MAX_SLICES = 64;
struct Slices {
  Slice s[MAX_SLICES];
  int n;
};

__global__ __launch_bounds__(256) void copy_columns(Slices sl, const int* __restrict__ list, int nsel, int np_full,
                                                    int np_c, int to_compact) {
  const Slice S = sl.s[blockIdx.y];
  const long total = (long)nsel * S.L;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const int j = (int)(i % nsel), lev = (int)(i / nsel);
    const int c = list[j];
    const int bf = c / np_full, jf = c - bf * np_full;
    const int bc = j / np_c, jc = j - bc * np_c;
    uint32_t* f = S.full + bf * S.full_stride + S.full_off + (long)lev * np_full + jf;
    uint32_t* q = S.comp + bc * S.comp_stride + S.comp_off + (long)lev * np_c + jc;
    if (to_compact)
      *q = *f;
    else
      *f = *q;
  }
}

}  // namespace surr

// ---------------------------------------------------------------------------------------------------------
struct cloudsc_surrogate {
  cloudsc_mlp* net = nullptr;
  int klev = 0, nin = 0, nout = 0, max_batch = 0;
  float *mu_x = nullptr, *sd_x = nullptr, *rsd_x = nullptr, *mu_y = nullptr, *sd_y = nullptr, *rsd_y = nullptr;
  float* wgt = nullptr;  // per output entry: 1 for the 6*KLEV targets, 0 for padding
  bf16 *X = nullptr, *XT = nullptr;
  float *T = nullptr, *Y = nullptr, *loss_acc = nullptr;
  // hybrid: selection list and the compact field set
  int* list = nullptr;
  int* count = nullptr;
  int list_cap = 0;
  void* comp = nullptr;
  size_t comp_bytes = 0;
  std::vector<void*> allocs;
};

namespace {

constexpr int NP_C = 128;  // NPROMA of the compact field set

int salloc(cloudsc_surrogate* s, size_t bytes, void** p) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloudsc_surrogate: hipMalloc", e);
  s->allocs.push_back(*p);
  return CLOUDSC_OK;
}

int check_fields(const cloudsc_surrogate* s, const cloudsc_params_t* q, int ngptot, int nproma, int klev,
                 int ngpblks, const KPtrs<float>& p) {
  if (!s || !q) return fail(CLOUDSC_ENULL, "cloudsc_surrogate: NULL object or parameters");
  const int rc = cloudsc_capi::check_config(q);  // the plain kernel reruns columns: the same switches apply
  if (rc) return rc;
  if (klev != s->klev) return fail(CLOUDSC_EDIMS, "cloudsc_surrogate: KLEV differs from the surrogate's");
  if (nproma < 1 || ngptot < 1 || ngpblks != (ngptot + nproma - 1) / nproma)
    return fail(CLOUDSC_EDIMS, "cloudsc_surrogate: need NGPBLKS = ceil(NGPTOT/NPROMA)");
  const void* req[] = {p.pt, p.pq, p.tcml_t, p.tcml_q, p.tcml_a, p.tcml_cld, p.tloc_t, p.tloc_q, p.tloc_a,
                       p.tloc_cld, p.pap, p.paph, p.plude, p.pa, p.pclv, p.pcovptot, p.prainfrac_toprfz,
                       p.pfplsl, p.pfplsn, p.pfhpsl, p.pfhpsn, p.pfsqlf, p.pfsqif, p.pfcqnng, p.pfcqlng,
                       p.pfsqrf, p.pfsqsf, p.pfcqrng, p.pfcqsng, p.pfsqltur, p.pfsqitur};
  for (const void* x : req)
    if (!x) return fail(CLOUDSC_ENULL, "cloudsc_surrogate: NULL array");
  return CLOUDSC_OK;
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

surr::Sat sat_consts(const cloudsc_params_t* q) {
  if (!q) return surr::Sat{};
  return surr::Sat{(float)q->r2es, (float)q->r3les, (float)q->r3ies, (float)q->r4les, (float)q->r4ies,
                   (float)q->rtt,  (float)q->rtice, (float)q->rtwat_rtice_r, (float)q->retv};
}

// Network prediction for all columns, in batches of max_batch.
hipError_t predict(cloudsc_surrogate* s, const cloudsc_params_t* q, const KPtrs<float>& p, const surr::Dims& d,
                   hipStream_t st) {
  hipError_t e = hipSuccess;
  const int ntotal = d.ngpblks * d.nproma;
  if (ntotal > d.ngptot)
    hipLaunchKernelGGL(surr::zero_padding_columns, dim3(nblk(ntotal - d.ngptot, 256)), dim3(256), 0, st,
                       p.prainfrac_toprfz, d.ngptot, ntotal);
  for (int c0 = 0; c0 < d.ngptot && e == hipSuccess; c0 += s->max_batch) {
    const int n = std::min(s->max_batch, d.ngptot - c0);
    const int M = surr::pad128(n);
    hipLaunchKernelGGL(surr::features, dim3(nblk(M, surr::TILE), nblk(s->nin, surr::TILE)), dim3(256), 0, st, p, d,
                       c0, M, s->nin, s->mu_x, s->rsd_x, s->X, (bf16*)nullptr);
    e = mlp::forward(s->net, s->X, M, s->Y, false, st);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(surr::outputs, dim3(nblk(M, surr::TILE), nblk(surr::NT * d.klev, surr::TILE)), dim3(256), 0,
                       st, p, d, c0, n, s->nout, s->mu_y, s->sd_y, s->Y, (float)q->rtt, (float)q->rlvtt,
                       (float)q->rlstt, q->ncldtop);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace

#define SURR_PTRS                                                                                                 \
  KPtrs<float> p;                                                                                                 \
  p.pt = pt; p.pq = pq; p.tcml_t = ptendency_cml_t; p.tcml_q = ptendency_cml_q; p.tcml_a = ptendency_cml_a;       \
  p.tcml_cld = ptendency_cml_cld; p.tloc_t = ptendency_loc_t; p.tloc_q = ptendency_loc_q;                         \
  p.tloc_a = ptendency_loc_a; p.tloc_cld = ptendency_loc_cld; p.pvfl = pvfl; p.pvfi = pvfi; p.pap = pap;          \
  p.paph = paph; p.plsm = plsm; p.ldcum = ldcum; p.ktype = ktype; p.peis = peis; p.plu = plu; p.plude = plude;    \
  p.pludeli = pludeli; p.psnde = psnde; p.pmfu = pmfu; p.pmfd = pmfd; p.pa = pa; p.pclv = pclv;                   \
  p.pcovptot = pcovptot; p.prainfrac_toprfz = prainfrac_toprfz; p.pfsqlf = pfsqlf; p.pfsqif = pfsqif;             \
  p.pfcqnng = pfcqnng; p.pfcqlng = pfcqlng; p.pfsqrf = pfsqrf; p.pfsqsf = pfsqsf; p.pfcqrng = pfcqrng;             \
  p.pfcqsng = pfcqsng; p.pfsqltur = pfsqltur; p.pfsqitur = pfsqitur; p.pfplsl = pfplsl; p.pfplsn = pfplsn;         \
  p.pfhpsl = pfhpsl; p.pfhpsn = pfhpsn;                                                                           \
  const surr::Dims d{ngptot, nproma, klev, ngpblks, sat_consts(ydparams)};                                                              \
  hipStream_t st = (hipStream_t)stream;                                                                           \
  (void)ptsphy

extern "C" {

int cloudsc_surrogate_create(cloudsc_surrogate_t** out, int klev, int n_hidden, int n_hidden_layers, int max_batch,
                             uint64_t seed) {
  if (!out) return fail(CLOUDSC_ENULL, "cloudsc_surrogate_create: NULL output");
  *out = nullptr;
  if (klev < 2 || max_batch < 128 || max_batch % 128)
    return fail(CLOUDSC_EDIMS, "cloudsc_surrogate_create: KLEV >= 2 and max_batch a positive multiple of 128");
  cloudsc_surrogate* s = new cloudsc_surrogate();
  s->klev = klev;
  s->nin = surr::pad128(surr::NF * klev);
  s->nout = surr::pad256(surr::NT * klev);  // the output product then runs on the 256 x 256-tile GEMM
  s->max_batch = max_batch;
  int rc = cloudsc_mlp_create(&s->net, s->nin, n_hidden, n_hidden_layers, s->nout, max_batch, seed);
  auto F = [&](size_t n, float** p) { return salloc(s, n * sizeof(float), (void**)p); };
  if (!rc)
    rc = F(s->nin, &s->mu_x) || F(s->nin, &s->sd_x) || F(s->nin, &s->rsd_x) || F(s->nout, &s->mu_y) ||
         F(s->nout, &s->sd_y) || F(s->nout, &s->rsd_y) || F(s->nout, &s->wgt) ||
         salloc(s, (size_t)max_batch * s->nin * sizeof(bf16), (void**)&s->X) ||
         salloc(s, (size_t)max_batch * s->nin * sizeof(bf16), (void**)&s->XT) ||
         F((size_t)max_batch * s->nout, &s->T) || F((size_t)max_batch * s->nout, &s->Y) || F(1, &s->loss_acc) ||
         salloc(s, sizeof(int), (void**)&s->count);
  if (rc) {
    cloudsc_surrogate_destroy(s);
    return rc < 0 ? rc : CLOUDSC_ENOMEM;
  }
  // identity statistics until fitted; unit weights on the real targets
  std::vector<float> zx(s->nin, 0.0f), ox(s->nin, 1.0f), zy(s->nout, 0.0f), oy(s->nout, 1.0f), w(s->nout, 0.0f);
  for (int i = 0; i < surr::NT * klev; ++i) w[i] = 1.0f;
  hipError_t e = hipMemcpy(s->mu_x, zx.data(), s->nin * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->sd_x, ox.data(), s->nin * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->rsd_x, ox.data(), s->nin * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->mu_y, zy.data(), s->nout * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->sd_y, oy.data(), s->nout * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->rsd_y, oy.data(), s->nout * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->wgt, w.data(), s->nout * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cloudsc_surrogate_destroy(s);
    return fail(CLOUDSC_EHIP, "cloudsc_surrogate_create: initialisation", e);
  }
  *out = s;
  return CLOUDSC_OK;
}

int cloudsc_surrogate_destroy(cloudsc_surrogate_t* s) {
  if (!s) return CLOUDSC_OK;
  cloudsc_mlp_destroy(s->net);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->list) (void)hipFree(s->list);
  if (s->comp) (void)hipFree(s->comp);
  delete s;
  return CLOUDSC_OK;
}

cloudsc_mlp_t* cloudsc_surrogate_mlp(cloudsc_surrogate_t* s) { return s ? s->net : nullptr; }

int cloudsc_surrogate_dims(const cloudsc_surrogate_t* s, int* klev, int* n_in, int* n_out, int* max_batch) {
  if (!s) return fail(CLOUDSC_ENULL, "cloudsc_surrogate_dims: NULL object");
  if (klev) *klev = s->klev;
  if (n_in) *n_in = s->nin;
  if (n_out) *n_out = s->nout;
  if (max_batch) *max_batch = s->max_batch;
  return CLOUDSC_OK;
}

int cloudsc_surrogate_norm(cloudsc_surrogate_t* s, float* mu_x, float* sd_x, float* mu_y, float* sd_y, float* wgt_y,
                           int get) {
  if (!s || !mu_x || !sd_x || !mu_y || !sd_y) return fail(CLOUDSC_ENULL, "cloudsc_surrogate_norm: NULL argument");
  const size_t nx = (size_t)surr::NF * s->klev * 4, ny = (size_t)surr::NT * s->klev * 4;
  hipError_t e = hipDeviceSynchronize();
  if (get) {
    if (e == hipSuccess) e = hipMemcpy(mu_x, s->mu_x, nx, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sd_x, s->sd_x, nx, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(mu_y, s->mu_y, ny, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(sd_y, s->sd_y, ny, hipMemcpyDeviceToHost);
    if (e == hipSuccess && wgt_y) e = hipMemcpy(wgt_y, s->wgt, ny, hipMemcpyDeviceToHost);
  } else {
    if (e == hipSuccess) e = hipMemcpy(s->mu_x, mu_x, nx, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->sd_x, sd_x, nx, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->mu_y, mu_y, ny, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->sd_y, sd_y, ny, hipMemcpyHostToDevice);
    if (e == hipSuccess && wgt_y) e = hipMemcpy(s->wgt, wgt_y, ny, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(surr::reciprocal, dim3(nblk(s->nin, 256)), dim3(256), 0, 0, s->sd_x, s->rsd_x,
                         surr::NF * s->klev);
      hipLaunchKernelGGL(surr::reciprocal, dim3(nblk(s->nout, 256)), dim3(256), 0, 0, s->sd_y, s->rsd_y,
                         surr::NT * s->klev);
      e = hipDeviceSynchronize();
    }
  }
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_norm", e);
  return CLOUDSC_OK;
}

int cloudsc_surrogate_fit_norm(cloudsc_surrogate_t* s, CLOUDSC_BLOCKED_ARGS(float)) {
  SURR_PTRS;
  int rc = check_fields(s, ydparams, ngptot, nproma, klev, ngpblks, p);
  if (rc) return rc;
  hipLaunchKernelGGL(surr::moments, dim3(surr::NF * klev), dim3(256), 0, st, p, d, 0, s->mu_x, s->sd_x);
  hipLaunchKernelGGL(surr::moments, dim3(surr::NT * klev), dim3(256), 0, st, p, d, 1, s->mu_y, s->sd_y);
  const int nx = surr::NF * klev, ny = surr::NT * klev;
  std::vector<float> mx(nx), sx(nx), my(ny), sy(ny), w(s->nout, 0.0f);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(mx.data(), s->mu_x, nx * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(sx.data(), s->sd_x, nx * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(my.data(), s->mu_y, ny * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(sy.data(), s->sd_y, ny * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_fit_norm", e);
  // Features: constant entries (the model-top pressure, levels without a quantity) are not scaled; the others
  // get at least 1e-3 of their feature's largest level spread.
  for (int f = 0; f < surr::NF; ++f) {
    float top = 0.0f;
    for (int k = 0; k < klev; ++k) top = std::max(top, sx[f * klev + k]);
    for (int k = 0; k < klev; ++k) {
      float& v = sx[f * klev + k];
      v = (v <= 1e-6f * std::fabs(mx[f * klev + k]) || v <= 1e-30f) ? 1.0f : std::max(v, 1e-3f * top);
    }
  }
  // Targets: entries whose spread is below 1e-3 of their target's largest level spread (tendencies above
  // NCLDTOP, levels where a process never acts) are not trained (weight 0) and are predicted at their mean.
  for (int t = 0; t < surr::NT; ++t) {
    float top = 0.0f;
    for (int k = 0; k < klev; ++k) top = std::max(top, sy[t * klev + k]);
    for (int k = 0; k < klev; ++k) {
      float& v = sy[t * klev + k];
      const bool live = top > 0.0f && v > 1e-3f * top;
      w[t * klev + k] = live ? 1.0f : 0.0f;
      v = live ? v : (top > 0.0f ? 1e-3f * top : 1.0f);
    }
  }
  e = hipMemcpyAsync(s->sd_x, sx.data(), nx * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->sd_y, sy.data(), ny * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->wgt, w.data(), s->nout * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(surr::reciprocal, dim3(nblk(s->nin, 256)), dim3(256), 0, st, s->sd_x, s->rsd_x, nx);
    hipLaunchKernelGGL(surr::reciprocal, dim3(nblk(s->nout, 256)), dim3(256), 0, st, s->sd_y, s->rsd_y, ny);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_fit_norm", e);
  return CLOUDSC_OK;
}

int cloudsc_surrogate_train(cloudsc_surrogate_t* s, float lr, int batch, float* loss, CLOUDSC_BLOCKED_ARGS(float)) {
  SURR_PTRS;
  int rc = check_fields(s, ydparams, ngptot, nproma, klev, ngpblks, p);
  if (rc) return rc;
  hipLaunchKernelGGL(surr::fill_f32, dim3(1), dim3(64), 0, st, s->loss_acc, 0.0f);
  if (batch < 128 || batch % 128 || batch > s->max_batch)
    return fail(CLOUDSC_EDIMS, "cloudsc_surrogate_train: batch must be a multiple of 128 and <= max_batch");
  int nb = 0;
  const int M = batch;
  hipError_t e = hipGetLastError();
  for (int c0 = 0; c0 + M <= ngptot && e == hipSuccess; c0 += M, ++nb) {
    hipLaunchKernelGGL(surr::features, dim3(nblk(M, surr::TILE), nblk(s->nin, surr::TILE)), dim3(256), 0, st, p, d,
                       c0, M, s->nin, s->mu_x, s->rsd_x, s->X, s->XT);
    hipLaunchKernelGGL(surr::targets, dim3(nblk(M, surr::TILE), nblk(s->nout, surr::TILE)), dim3(256), 0, st, p, d,
                       c0, M, s->nout, s->mu_y, s->rsd_y, s->T);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    rc = cloudsc_mlp_train_step(s->net, (const uint16_t*)s->X, (const uint16_t*)s->XT, s->T, s->wgt, M, lr, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(surr::add_scalar, dim3(1), dim3(64), 0, st, s->loss_acc, s->net->loss);
    e = hipGetLastError();
  }
  float h = 0.0f;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, s->loss_acc, sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_train", e);
  if (loss) *loss = nb ? h / nb : 0.0f;
  return CLOUDSC_OK;
}

int cloudsc_surrogate_predict_sp(cloudsc_surrogate_t* s, CLOUDSC_BLOCKED_ARGS(float)) {
  SURR_PTRS;
  int rc = check_fields(s, ydparams, ngptot, nproma, klev, ngpblks, p);
  if (rc) return rc;
  hipError_t e = predict(s, ydparams, p, d, st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_predict_sp", e);
  return CLOUDSC_OK;
}

int cloudsc_surrogate_gate_scores(cloudsc_surrogate_t* s, float* score, CLOUDSC_BLOCKED_ARGS(float)) {
  SURR_PTRS;
  int rc = check_fields(s, ydparams, ngptot, nproma, klev, ngpblks, p);
  if (rc) return rc;
  if (!score) return fail(CLOUDSC_ENULL, "cloudsc_surrogate_gate_scores: NULL score");
  hipLaunchKernelGGL(surr::select_columns, dim3(nblk(ngptot, 256)), dim3(256), 0, st, p, d, 0.0f, (uint64_t)0, 0.0f,
                     score, (uint8_t*)nullptr, (int*)nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_surrogate_gate_scores", e);
  return CLOUDSC_OK;
}

int cloudsc_hybrid_sp_c(cloudsc_surrogate_t* s, float rerun_fraction, uint64_t rerun_seed, float gate_threshold,
                        int* n_plain, CLOUDSC_BLOCKED_ARGS(float)) {
  SURR_PTRS;
  int rc = check_fields(s, ydparams, ngptot, nproma, klev, ngpblks, p);
  if (rc) return rc;
  if (!p.pvfl || !p.pvfi || !p.plsm || !p.ldcum || !p.ktype || !p.peis || !p.plu || !p.pludeli || !p.psnde ||
      !p.pmfu || !p.pmfd)
    return fail(CLOUDSC_ENULL, "cloudsc_hybrid_sp_c: NULL array on the default path");
  const cloudsc_params_t* q = ydparams;
  // 1. network on every column
  hipError_t e = predict(s, q, p, d, st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_hybrid_sp_c: prediction", e);
  // 2. columns for the plain kernel
  // selection list (ngptot ints), per-group counts and offsets, per-column flags: one allocation
  const int ngroups = (int)nblk(ngptot, 256);
  if (s->list_cap < ngptot) {
    if (s->list) (void)hipFree(s->list);
    s->list = nullptr;
    s->list_cap = 0;
    e = hipMalloc((void**)&s->list, (size_t)(ngptot + ngroups) * sizeof(int) + (size_t)ngptot);
    if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloudsc_hybrid_sp_c: hipMalloc", e);
    s->list_cap = ngptot;
  }
  int* group_offsets = s->list + ngptot;
  uint8_t* flags = reinterpret_cast<uint8_t*>(group_offsets + ngroups);
  hipLaunchKernelGGL(surr::select_columns, dim3(ngroups), dim3(256), 0, st, p, d, rerun_fraction, rerun_seed,
                     gate_threshold, (float*)nullptr, flags, group_offsets);
  hipLaunchKernelGGL(surr::scan_counts, dim3(1), dim3(1024), 0, st, group_offsets, ngroups, s->count);
  hipLaunchKernelGGL(surr::compact_columns, dim3(ngroups), dim3(256), 0, st, flags, group_offsets, ngptot, s->list);
  int nsel = 0;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&nsel, s->count, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_hybrid_sp_c: selection", e);
  if (n_plain) *n_plain = nsel;
  if (nsel == 0) return CLOUDSC_OK;
  // 3. compact field set of the selected columns
  const int ncb = (nsel + NP_C - 1) / NP_C;
  const long L = klev, c2 = (long)NP_C * L, ch = (long)NP_C * (L + 1), c1 = NP_C;
  struct Arr {
    void* full;
    int n, L;   // species, levels per species
    bool in, out;
    int nout;   // species written (tloc_cld: 4 of 5)
  };
  Arr arrs[] = {{(void*)p.pt, 1, (int)L, true, false, 0},        {(void*)p.pq, 1, (int)L, true, false, 0},
                {(void*)p.tcml_t, 1, (int)L, true, false, 0},    {(void*)p.tcml_q, 1, (int)L, true, false, 0},
                {(void*)p.tcml_a, 1, (int)L, true, false, 0},    {(void*)p.tcml_cld, 5, (int)L, true, false, 0},
                {(void*)p.tloc_t, 1, (int)L, false, true, 1},    {(void*)p.tloc_q, 1, (int)L, false, true, 1},
                {(void*)p.tloc_a, 1, (int)L, false, true, 1},    {(void*)p.tloc_cld, 5, (int)L, false, true, 4},
                {(void*)p.pvfl, 1, (int)L, true, false, 0},      {(void*)p.pvfi, 1, (int)L, true, false, 0},
                {(void*)p.pap, 1, (int)L, true, false, 0},       {(void*)p.paph, 1, (int)L + 1, true, false, 0},
                {(void*)p.plsm, 1, 1, true, false, 0},           {(void*)p.ldcum, 1, 1, true, false, 0},
                {(void*)p.ktype, 1, 1, true, false, 0},          {(void*)p.peis, 1, 1, true, false, 0},
                {(void*)p.plu, 1, (int)L, true, false, 0},       {(void*)p.plude, 1, (int)L, true, true, 1},
                {(void*)p.pludeli, 4, (int)L, true, false, 0},   {(void*)p.psnde, 2, (int)L, true, false, 0},
                {(void*)p.pmfu, 1, (int)L, true, false, 0},      {(void*)p.pmfd, 1, (int)L, true, false, 0},
                {(void*)p.pa, 1, (int)L, true, false, 0},        {(void*)p.pclv, 5, (int)L, true, false, 0},
                {(void*)p.pcovptot, 1, (int)L, true, true, 1},  {(void*)p.prainfrac_toprfz, 1, 1, false, true, 1},
                {(void*)p.pfsqlf, 1, (int)L + 1, false, true, 1}, {(void*)p.pfsqif, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfcqnng, 1, (int)L + 1, false, true, 1}, {(void*)p.pfcqlng, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfsqrf, 1, (int)L + 1, false, true, 1}, {(void*)p.pfsqsf, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfcqrng, 1, (int)L + 1, false, true, 1}, {(void*)p.pfcqsng, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfsqltur, 1, (int)L + 1, false, true, 1}, {(void*)p.pfsqitur, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfplsl, 1, (int)L + 1, false, true, 1}, {(void*)p.pfplsn, 1, (int)L + 1, false, true, 1},
                {(void*)p.pfhpsl, 1, (int)L + 1, false, true, 1}, {(void*)p.pfhpsn, 1, (int)L + 1, false, true, 1}};
  const int narr = sizeof(arrs) / sizeof(arrs[0]);
  size_t offs[sizeof(arrs) / sizeof(arrs[0])], need = 0;
  for (int i = 0; i < narr; ++i) {
    offs[i] = need;
    need += (((size_t)ncb * NP_C * arrs[i].n * arrs[i].L * 4) + 255) & ~(size_t)255;
  }
  if (s->comp_bytes < need) {
    if (s->comp) (void)hipFree(s->comp);
    s->comp = nullptr;
    s->comp_bytes = 0;
    e = hipMalloc(&s->comp, need);
    if (e != hipSuccess) return fail(CLOUDSC_ENOMEM, "cloudsc_hybrid_sp_c: hipMalloc", e);
    s->comp_bytes = need;
  }
  char* cb = (char*)s->comp;
  auto slices = [&](bool inputs, surr::Slices& sl) {
    sl.n = 0;
    for (int i = 0; i < narr; ++i) {
      const Arr& a = arrs[i];
      if (inputs ? !a.in : !a.out) continue;
      const int ns = inputs ? a.n : a.nout;
      for (int sp = 0; sp < ns; ++sp) {
        surr::Slice& x = sl.s[sl.n++];
        x.full = (uint32_t*)a.full;
        x.comp = (uint32_t*)(cb + offs[i]);
        x.full_stride = (long)nproma * a.n * a.L;
        x.comp_stride = (long)NP_C * a.n * a.L;
        x.full_off = (long)nproma * sp * a.L;
        x.comp_off = (long)NP_C * sp * a.L;
        x.L = a.L;
      }
    }
  };
  surr::Slices sl;
  slices(true, sl);
  const unsigned gx = std::min<unsigned>(nblk((long)nsel * (L + 1), 256), 4096u);
  hipLaunchKernelGGL(surr::copy_columns, dim3(gx, sl.n), dim3(256), 0, st, sl, s->list, nsel, nproma, NP_C, 1);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_hybrid_sp_c: gather", e);
  // 4. plain kernel on the compact set
  auto C = [&](int i) { return (float*)(cb + offs[i]); };
  KPtrs<float> pc;
  pc.pt = C(0); pc.pq = C(1); pc.tcml_t = C(2); pc.tcml_q = C(3); pc.tcml_a = C(4); pc.tcml_cld = C(5);
  pc.tloc_t = C(6); pc.tloc_q = C(7); pc.tloc_a = C(8); pc.tloc_cld = C(9); pc.pvfl = C(10); pc.pvfi = C(11);
  pc.pap = C(12); pc.paph = C(13); pc.plsm = C(14); pc.ldcum = (const int32_t*)C(15);
  pc.ktype = (const int32_t*)C(16); pc.peis = C(17); pc.plu = C(18); pc.plude = C(19); pc.pludeli = C(20);
  pc.psnde = C(21); pc.pmfu = C(22); pc.pmfd = C(23); pc.pa = C(24); pc.pclv = C(25); pc.pcovptot = C(26);
  pc.prainfrac_toprfz = C(27); pc.pfsqlf = C(28); pc.pfsqif = C(29); pc.pfcqnng = C(30); pc.pfcqlng = C(31);
  pc.pfsqrf = C(32); pc.pfsqsf = C(33); pc.pfcqrng = C(34); pc.pfcqsng = C(35); pc.pfsqltur = C(36);
  pc.pfsqitur = C(37); pc.pfplsl = C(38); pc.pfplsn = C(39); pc.pfhpsl = C(40); pc.pfhpsn = C(41);
  const KDims kd{nsel, NP_C, klev, ncb, 1, NP_C};
  e = cloudsc_gpu::launch_cloudsc<float, float>(*q, (double)ptsphy, pc, kd, 0, st);
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_hybrid_sp_c: plain kernel", e);
  // 5. scatter the plain results over the network's
  slices(false, sl);
  hipLaunchKernelGGL(surr::copy_columns, dim3(gx, sl.n), dim3(256), 0, st, sl, s->list, nsel, nproma, NP_C, 0);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(CLOUDSC_EHIP, "cloudsc_hybrid_sp_c: scatter", e);
  return CLOUDSC_OK;
}

}  // extern "C"
