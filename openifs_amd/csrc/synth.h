// Deterministic synthetic atmospheres for CLOUDSC: one column at a time,
// identical code on host and device.  Every random number is a function of
// (seed, field, global column, level) through Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), so any
// blocking, thread count or device produces the same fields.
//
// Distributions (column c, level k):
//   ps ~ N(100000, 1500) Pa; half levels PAPH = A + B*ps; full levels PAP by
//     Simmons-Burridge, p_k = exp((p+ ln p+ - p- ln p-)/(p+ - p-) - 1), with
//     p_1 = PAPH(2)/2 at the model top.
//   T = max(216.65, 288.15 (PAP/ps)^(RD*0.0065/RG)) + N(0, 2 K): a 6.5 K/km
//     lapse rate to a 216.65 K tropopause, written in pressure coordinates.
//   RH ~ U(0.3, 1.05); q = RH * qsat_mix(T, PAP) (FOEEWM, capped at 0.5).
//   a ~ U(0,1) where RH > 0.8, else 0.
//   l, i, r, s ~ 1e-5 exp(1.5 N(0,1)) kg/kg; i and s are 0 where T >= 273 K.
//   Tendencies: T ~ N(0,1e-4) K/s; a ~ N(0,1e-4) /s; q and each hydrometeor
//     x ~ x * N(0,1e-4) /s (relative, so dt*tendency stays a fraction of x).
//   PMFU, PMFD ~ U(0, 0.05) kg m-2 s-1.
//   KTYPE in {0,1,2,3} with probability 0.5/0.25/0.15/0.10; LDCUM = KTYPE > 0.
//   Where LDCUM: PLU ~ 1e-4 exp(1.5N) kg/kg, PLUDE ~ 1e-6 exp(1.5N) kg m-2 s-1,
//     PLUDELI(:,:,1:2) = PLUDE*(alpha, 1-alpha) with alpha = FOEALFA(T),
//     PSNDE(:,:,1:2) ~ 1e-7 exp(1.5N); all four are 0 where not LDCUM.
//   PVFL, PVFI ~ N(0, 1e-9) /s; PLSM = 1 with probability 0.3, else 0;
//   PEIS ~ U(0, 15) K.  All other CLOUDSC inputs are zero.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define SYN_HD __host__ __device__
#else
#define SYN_HD
#endif

#include <math.h>

namespace synth {

struct Philox {
  uint32_t c[4];
};

SYN_HD inline void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds (Random123 constants).
SYN_HD inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Two uniforms in (0,1) and one standard normal (Box-Muller) per draw.
struct Draw {
  double u0, u1, n;
};

SYN_HD inline Draw draw(uint64_t seed, uint32_t field, uint32_t col, uint32_t lev) {
  const uint32_t ctr[4] = {col, lev, field, 0x434C4453u /* "CLDS" */};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox4x32_10(ctr, key, o);
  const double s = 2.3283064365386963e-10;  // 2^-32
  Draw d;
  d.u0 = ((double)o[0] + 0.5) * s;
  d.u1 = ((double)o[1] + 0.5) * s;
  const double v0 = ((double)o[2] + 0.5) * s, v1 = ((double)o[3] + 0.5) * s;
  d.n = sqrt(-2.0 * log(v0)) * cos(6.283185307179586 * v1);
  return d;
}

enum Field : uint32_t {
  F_PS = 0, F_T, F_RH, F_A, F_L, F_I, F_R, F_S, F_TT, F_TQ, F_TA, F_TL, F_TI, F_TR, F_TS,
  F_MFU, F_MFD, F_KTYPE, F_LU, F_LUDE, F_SNDE1, F_SNDE2, F_VFL, F_VFI, F_LSM, F_EIS
};

// Constants the generator needs (the CLOUDSC defaults of TCST/TTHF).
struct Thermo {
  double rd, rg, retv, rtt, r2es, r3les, r3ies, r4les, r4ies, rtice, rtwat, rtwat_rtice_r;
};

SYN_HD inline double foealfa(const Thermo& t, double x) {
  double y = (fmax(t.rtice, fmin(t.rtwat, x)) - t.rtice) * t.rtwat_rtice_r;
  return fmin(1.0, y * y);
}
SYN_HD inline double qsat_mix(const Thermo& t, double temp, double p) {
  const double a = foealfa(t, temp);
  const double e = t.r2es * (a * exp(t.r3les * (temp - t.rtt) / (temp - t.r4les)) +
                             (1.0 - a) * exp(t.r3ies * (temp - t.rtt) / (temp - t.r4ies)));
  const double es = fmin(e / p, 0.5);
  return es / (1.0 - t.retv * es);
}

// Output pointers in the blocked layout (NPROMA, KLEV[+1], n, NGPBLKS).
template <typename R>
struct Fields {
  R *pt, *pq, *tcml_t, *tcml_q, *tcml_a, *tcml_cld, *pvfl, *pvfi, *pap, *paph, *plsm;
  int32_t *ldcum, *ktype;
  R *peis, *plu, *plude, *pludeli, *psnde, *pmfu, *pmfd, *pa, *pclv;
};

// Fill global column g (0-based) of a blocked field set.
template <typename R>
SYN_HD inline void fill_column(uint64_t seed, const Thermo& th, int klev, int nproma, const double* ahalf,
                               const double* bhalf, const Fields<R>& f, int64_t g) {
  const int64_t b = g / nproma;
  const int jl = (int)(g - b * nproma);
  const uint32_t col = (uint32_t)g;
  const int64_t s2 = (int64_t)nproma * klev, sh = (int64_t)nproma * (klev + 1);
  auto I2 = [&](int k) { return b * s2 + (int64_t)k * nproma + jl; };           // (NPROMA,KLEV,NGPBLKS)
  auto IH = [&](int k) { return b * sh + (int64_t)k * nproma + jl; };           // (NPROMA,KLEV+1,NGPBLKS)
  auto I3 = [&](int k, int m, int n) { return (b * n + m) * s2 + (int64_t)k * nproma + jl; };  // (NPROMA,KLEV,n,NGPBLKS)
  auto I1 = [&]() { return b * nproma + jl; };

  const double ps = 100000.0 + 1500.0 * draw(seed, F_PS, col, 0).n;
  const Draw dk = draw(seed, F_KTYPE, col, 0);
  const int ktype = dk.u0 < 0.5 ? 0 : (dk.u0 < 0.75 ? 1 : (dk.u0 < 0.9 ? 2 : 3));
  const bool cum = ktype > 0;
  f.ktype[I1()] = ktype;
  f.ldcum[I1()] = cum ? 1 : 0;
  f.plsm[I1()] = (R)(draw(seed, F_LSM, col, 0).u0 < 0.3 ? 1.0 : 0.0);
  f.peis[I1()] = (R)(15.0 * draw(seed, F_EIS, col, 0).u0);
  const double kappa = th.rd * 0.0065 / th.rg;

  for (int k = 0; k <= klev; ++k) f.paph[IH(k)] = (R)(ahalf[k] + bhalf[k] * ps);
  for (int k = 0; k < klev; ++k) {
    const double pm = ahalf[k] + bhalf[k] * ps, pp = ahalf[k + 1] + bhalf[k + 1] * ps;
    const double p = (k == 0) ? 0.5 * pp : exp((pp * log(pp) - pm * log(pm)) / (pp - pm) - 1.0);
    f.pap[I2(k)] = (R)p;
    const double t = fmax(216.65, 288.15 * pow(p / ps, kappa)) + 2.0 * draw(seed, F_T, col, k).n;
    f.pt[I2(k)] = (R)t;
    const double rh = 0.3 + 0.75 * draw(seed, F_RH, col, k).u0;
    const double q = rh * qsat_mix(th, t, p);
    f.pq[I2(k)] = (R)q;
    f.pa[I2(k)] = (R)(rh > 0.8 ? draw(seed, F_A, col, k).u0 : 0.0);
    const bool warm = t >= 273.0;
    const double ql = 1e-5 * exp(1.5 * draw(seed, F_L, col, k).n);
    const double qi = warm ? 0.0 : 1e-5 * exp(1.5 * draw(seed, F_I, col, k).n);
    const double qr = 1e-5 * exp(1.5 * draw(seed, F_R, col, k).n);
    const double qs = warm ? 0.0 : 1e-5 * exp(1.5 * draw(seed, F_S, col, k).n);
    f.pclv[I3(k, 0, 5)] = (R)ql;
    f.pclv[I3(k, 1, 5)] = (R)qi;
    f.pclv[I3(k, 2, 5)] = (R)qr;
    f.pclv[I3(k, 3, 5)] = (R)qs;
    f.pclv[I3(k, 4, 5)] = (R)0;
    f.tcml_t[I2(k)] = (R)(1e-4 * draw(seed, F_TT, col, k).n);
    f.tcml_q[I2(k)] = (R)(q * 1e-4 * draw(seed, F_TQ, col, k).n);
    f.tcml_a[I2(k)] = (R)(1e-4 * draw(seed, F_TA, col, k).n);
    f.tcml_cld[I3(k, 0, 5)] = (R)(ql * 1e-4 * draw(seed, F_TL, col, k).n);
    f.tcml_cld[I3(k, 1, 5)] = (R)(qi * 1e-4 * draw(seed, F_TI, col, k).n);
    f.tcml_cld[I3(k, 2, 5)] = (R)(qr * 1e-4 * draw(seed, F_TR, col, k).n);
    f.tcml_cld[I3(k, 3, 5)] = (R)(qs * 1e-4 * draw(seed, F_TS, col, k).n);
    f.tcml_cld[I3(k, 4, 5)] = (R)0;
    f.pmfu[I2(k)] = (R)(0.05 * draw(seed, F_MFU, col, k).u0);
    f.pmfd[I2(k)] = (R)(0.05 * draw(seed, F_MFD, col, k).u0);
    const double lu = cum ? 1e-4 * exp(1.5 * draw(seed, F_LU, col, k).n) : 0.0;
    const double lude = cum ? 1e-6 * exp(1.5 * draw(seed, F_LUDE, col, k).n) : 0.0;
    const double alpha = foealfa(th, t);
    f.plu[I2(k)] = (R)lu;
    f.plude[I2(k)] = (R)lude;
    f.pludeli[I3(k, 0, 4)] = (R)(lude * alpha);
    f.pludeli[I3(k, 1, 4)] = (R)(lude * (1.0 - alpha));
    f.pludeli[I3(k, 2, 4)] = (R)0;
    f.pludeli[I3(k, 3, 4)] = (R)0;
    f.psnde[I3(k, 0, 2)] = (R)(cum ? 1e-7 * exp(1.5 * draw(seed, F_SNDE1, col, k).n) : 0.0);
    f.psnde[I3(k, 1, 2)] = (R)(cum ? 1e-7 * exp(1.5 * draw(seed, F_SNDE2, col, k).n) : 0.0);
    f.pvfl[I2(k)] = (R)(1e-9 * draw(seed, F_VFL, col, k).n);
    f.pvfi[I2(k)] = (R)(1e-9 * draw(seed, F_VFI, col, k).n);
  }
}

}  // namespace synth
