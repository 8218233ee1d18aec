// Device helpers shared by the CLOUDSC and CLOUD_SATADJ kernels: FP32/FP64 math on the gfx950 instructions,
// the Tetens saturation functions, CLOUD_SUPERSATCHECK for one point, and 32-bit-offset loads and stores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cloudsc_gpu {

// CLOUDSC_FAST_TRANSCENDENTALS selects all three hardware forms below (exp/pow/log, division, square root);
// CLOUDSC_FAST_EXP, CLOUDSC_FAST_DIV and CLOUDSC_FAST_SQRT select them one by one.
#ifdef CLOUDSC_FAST_TRANSCENDENTALS
#define CLOUDSC_FAST_EXP
#define CLOUDSC_FAST_DIV
#define CLOUDSC_FAST_SQRT
#endif
#ifdef CLOUDSC_FAST_EXP
// FP32 exp and pow on the hardware base-2 instructions (v_exp_f32, v_log_f32):
// exp(x) = 2^(x log2 e), pow(x,y) = 2^(y log2 x).  Relative error ~ (1 + |y log2 x|) ulp.
__device__ __forceinline__ float m_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float m_pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ float m_log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
#else
__device__ __forceinline__ float m_exp(float x) { return expf(x); }
__device__ __forceinline__ float m_pow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ float m_log(float x) { return logf(x); }
#endif
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ double m_log(double x) { return log(x); }
#ifdef CLOUDSC_FAST_DIV
// FP32 division as reciprocal (v_rcp_f32, 1 ulp) times numerator instead of the ~10-instruction
// IEEE sequence.  Measured against the FP64 oracle on 8192 L137 columns, the FP32 kernel's
// relative RMS error per output field stays within 7% of the IEEE build's (both are dominated
// by FP32 threshold flips, not by the last bit of a quotient).  FP64 keeps IEEE division.
__device__ __forceinline__ float m_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#else
__device__ __forceinline__ float m_div(float a, float b) { return a / b; }
#endif
__device__ __forceinline__ double m_div(double a, double b) { return a / b; }
// Quotients that are exactly 1, or exactly the numerator's factor, when the denominator is the numerator's own
// sum or factor (the liquid share of the condensate, the rain share of the precipitation, the precipitation
// cover's overlap factor when the cloud cover does not grow): IEEE division.  a * rcp(a) can be 1 - 2^-24
// instead, which leaves a share of ~6e-8 to the other phase and switches on its processes.
template <typename T>
__device__ __forceinline__ T m_share(T a, T b) { return a / b; }
#ifdef CLOUDSC_FAST_SQRT
// square roots occur only inside rate expressions: v_sqrt_f32 (1 ulp)
__device__ __forceinline__ float m_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
__device__ __forceinline__ float m_sqrt(float x) { return sqrtf(x); }
#endif
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
// MAX/MIN on v_max/v_min (identical to the Fortran intrinsics for non-NaN arguments)
__device__ __forceinline__ float m_max(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float m_min(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double m_max(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double m_min(double a, double b) { return __builtin_fmin(a, b); }
template <typename T>
__device__ __forceinline__ T m_abs(T a) { return a < T(0) ? -a : a; }

// Tetens saturation vapour pressures (fcttre.func.h:84-93,165-166).
template <typename T>
struct Sat {
  T alfa, el, ei;  // FOEALFA, EXP(R3LES*(T-RTT)/(T-R4LES)), EXP(R3IES*(T-RTT)/(T-R4IES))
};
template <typename T, typename KCT>
__device__ __forceinline__ Sat<T> saturation(const KCT& c, T t) {
  Sat<T> s;
  T x = (m_max(c.rtice, m_min(c.rtwat, t)) - c.rtice) * c.rtwat_rtice_r;
  s.alfa = m_min(T(1), x * x);
  s.el = m_exp(m_div(c.r3les * (t - c.rtt), t - c.r4les));
  s.ei = m_exp(m_div(c.r3ies * (t - c.rtt), t - c.r4ies));
  return s;
}

// CLOUD_SUPERSATCHECK for one point, KFTLIQICE = 1 (cloud_supersatcheck.F90:111-226).
template <typename T, typename KCT>
__device__ __forceinline__ void supersatcheck(const KCT& c, T pt, T pq, T pa, T pp, T& t_adj, T& q_adj,
                                              T& a_adj, T& l_adj, T& i_adj) {
  const T zepsec = T(1.e-14);
  T zda = T(0);
  T zt = m_max(pt, T(160));
  T za = m_max(m_min(pa, T(1)), T(0));
  const T zp_r = m_div(T(1), pp);
  const Sat<T> s = saturation<T, KCT>(c, zt);
  const T foeeliq = c.r2es * s.el;
  const T zfoeeliq = m_min(foeeliq * zp_r, T(0.5));
  // FOEEW: water above RTT, ice below (FOEDELTA = MAX(0,SIGN(1,T-RTT)))
  const T foeew = (zt - c.rtt >= T(0)) ? c.r2es * s.el : c.r2es * s.ei;
  const T zfoeeice = m_min(foeew * zp_r, T(0.5));
  const T zfokoop = m_min(zfoeeice * (c.rkoop1 - c.rkoop2 * zt), zfoeeliq);
  T zqsliqk = m_div(zfokoop, T(1) - c.retv * zfokoop);
  const T zfoeemix = m_min(c.r2es * (s.alfa * s.el + (T(1) - s.alfa) * s.ei) * zp_r, T(0.5));
  const T zcor = m_div(T(1), T(1) - c.retv * zfoeemix);
  const T zqsmix = zfoeemix * zcor;
  const T dl = zt - c.r4les, di = zt - c.r4ies;
  const T foedem = s.alfa * c.r5alvcp * m_div(T(1), dl * dl) + (T(1) - s.alfa) * c.r5alscp * m_div(T(1), di * di);
  const T zcorqsmix = T(1) + zqsmix * zcor * foedem;
  zqsliqk = T(0.5) * (zqsliqk + zqsmix);
  T zqslim = za * zqsmix + (T(1) - za) * zqsliqk;
  T zcond = m_max(pq - zqslim, T(0));
  if (zcond > zepsec) {
    const T zfaci = s.alfa + (T(1) - s.alfa) * c.rssicefactor;
    zda = (T(1) - za) * zfaci;
    const T za_upd = za + zda;
    zqslim = za_upd * zqsmix + (T(1) - za_upd) * zqsliqk;
    zcond = m_div(pq - zqslim, zcorqsmix);
  }
  const T foeldcpm = s.alfa * c.ralvdcp + (T(1) - s.alfa) * c.ralsdcp;
  t_adj = foeldcpm * zcond;
  q_adj = -zcond;
  a_adj = zda;
  l_adj = zcond * s.alfa;
  i_adj = zcond * (T(1) - s.alfa);
}

// Loads and stores at 32-bit byte offsets from an array base: the form the hardware
// addresses directly (64-bit base in SGPRs, 32-bit offset in one VGPR).  The launcher
// splits the grid so that every array of one launch spans less than 4 GiB.
// u: byte offset common to the wave (block and level); v: the lane's byte offset within the level.  The sum is
// formed in 32 bits, so every array access is one vector offset on a scalar base; with wave-uniform blocks
// (NPROMA a multiple of 64) u is a scalar and the block-dependent offsets cost no vector registers.
template <typename T>
__device__ __forceinline__ T LD(const T* base, uint32_t u, uint32_t v) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (u + v));
}
template <typename T>
__device__ __forceinline__ void ST(T* base, uint32_t u, uint32_t v, T x) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (u + v)) = x;
}

}  // namespace cloudsc_gpu
