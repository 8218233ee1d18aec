// CLOUDSC process budgets (LCLDBUD*) for the budget-instrumented instantiation of the column kernel.
//
// The reference keeps per-level budget arrays ZBUDCC(KLON,12), ZBUDL(KLON,22), ZBUDI(KLON,18)
// (ifs-source/arpifs/phys_ec/cloudsc.F90:557-559, zeroed every level :1271-1273), sets single entries where the
// processes run, and writes them into PEXTRA at the end of the level (:3826-4101).  Here each entry the default
// path can set is one register of Bud<R, true>; Bud<R, false> is empty, so the plain kernels are unchanged.
#pragma once

#include <stdint.h>

#include "cloudsc_device.h"

namespace cloudsc_gpu {

// Switch bits of cloudsc_params_t::lcldbud (include/cloudsc_hip.h).
enum : int32_t {
  BUD_C = 1, BUD_L = 2, BUD_I = 4, BUD_T = 8, BUD_VERTINT = 16, BUD_TIMEINT = 32, BUD_LBUD23 = 64, BUD_CHECK = 128,
  BUD_OUT = BUD_C | BUD_L | BUD_I | BUD_T | BUD_VERTINT  // the switches that write PEXTRA
};

// Budget entries the default path sets (names: array and reference index).  L12/L13 (rain freezing from
// autoconversion) and the condensation entries 1, 8, 9, 10 stay zero on the default path (IP_SNOW_ACCRETES_RAIN = 1,
// cloudsc.F90:771) and are not stored.
enum BudIdx {
  CC2, CC3, CC4, CC5, CC6, CC7,
  L2, L3, L4, L5, L6, L7, L11, L14, L15, L17, L18, L19, L20, L21,
  I2, I3, I4, I5, I6, I7, I11, I12, I13, I14, I15, I16, I17,
  PSDEP, PSACR, PIACR,
  NBUD
};

template <typename R, bool ON>
struct Bud;
template <typename R>
struct Bud<R, false> {
  __device__ __forceinline__ Bud(R*, uint32_t) {}
  __device__ __forceinline__ void set(int, R) {}
  __device__ __forceinline__ R get(int) const { return R(0); }
};
// The entries live in LDS (this lane's NBUD words at stride `n`, conflict-free across a wave): the process sites
// store into them and section 7 reads them back, so they do not hold registers across the level.
template <typename R>
struct Bud<R, true> {
  R* s;
  uint32_t n;
  __device__ __forceinline__ Bud(R* lds, uint32_t stride) : s(lds), n(stride) {
#pragma unroll
    for (int i = 0; i < NBUD; ++i) s[i * n] = R(0);
  }
  __device__ __forceinline__ void set(int i, R x) { s[i * n] = x; }
  __device__ __forceinline__ R get(int i) const { return s[i * n]; }
};

// Device pointers of the diagnostics, blocked layout: the budget inputs and PEXTRA (NPROMA, KLEV, KFLDX, NGPBLKS)
// when flags has output bits; PGAW (NPROMA, NGPBLKS) and PFSD (NPROMA, KLEV, NGPBLKS) when fsd
// (LRAD_CLOUD_INHOMOG).
template <typename R>
struct KBudPtrs {
  const R *pvfa, *pdyna, *pdynl, *pdyni, *pdyns, *phrsw, *phrlw, *pvfl, *pvfi;
  const int32_t *kpbltype, *kctop;
  R* pextra;
  int32_t kfldx, flags;
  R reisthsc, ralfdcp;
  const R* pgaw;
  R* pfsd;
  int32_t fsd;
  R ra, rpi, rcloud_frac_std;
  int32_t inhomog;  // LCLOUD_INHOMOG: autoconversion and accretion enhancement from the condensate FSD
};

// Cloud condensate FSD for the radiation (LRAD_CLOUD_INHOMOG; the ZLFSD/ZIFSD/ZFSD of cloudsc.F90).
// Representative grid-box length in km from the normalised Gaussian weight PGAW (:1944, :4313).
template <typename R>
__device__ __forceinline__ R fsd_gridlen(R ra, R rpi, R pgaw) {
  return R(2) * ra * m_sqrt(rpi * pgaw) * R(0.001);
}
// LCLOUD_INHOMOG (:2075-2098): enhancement of the KK2000 autoconversion and accretion rates for sub-grid
// lognormal cloud and rain water (Boutle et al. 2013), E[q_c^b]/E[q_c]^b = (1+f_c^2)^(b(b-1)/2) and, for accretion,
// the bivariate moment with the cloud-rain correlation 1 - 0.8 a.  The rain FSD is Boutle et al.'s fit in the grid
// length and cloud fraction.
template <typename R>
struct InhomogFactors {
  R aut, acc;
};
template <typename R>
__device__ __forceinline__ InhomogFactors<R> inhomog_factors(R fsdc, R za, R zgridlen, R kkbauq, R kkbac) {
  const R fc2 = R(1) + fsdc * fsdc;
  InhomogFactors<R> e;
  e.aut = m_pow(fc2, (kkbauq - R(1)) * kkbauq / R(2));
  const R ga = zgridlen * za;
  const R zphir = m_pow(ga, R(1) / R(3)) * m_pow(m_pow(R(0.11) * ga, R(1.14)) + R(1), R(-0.22));
  R zfracsdr = (R(1.1) - R(0.8) * za) * zphir;
  if (za > R(0.95)) zfracsdr = R(0.3) * zphir;
  const R fr2 = R(1) + zfracsdr * zfracsdr;
  const R ex = (kkbac - R(1)) * kkbac / R(2);
  const R corr = R(1) - R(0.8) * za;  // ZCLDRAINCORR (:1947)
  e.acc = m_pow(fc2, ex) * m_pow(fr2, ex) * m_exp(corr * kkbac * kkbac * m_sqrt(m_log(fc2) * m_log(fr2)));
  return e;
}

// Liquid FSD where the in-cloud liquid exceeds ZEPSEC (:1937-1967): Boutle et al. (2013) with the IFS parameters,
// from the tidied vapour and liquid and the tidied cloud fraction.
template <typename R>
__device__ __forceinline__ R liquid_fsd(R qv, R ql, R za, R zgridlen) {
  const R zqtot = m_min((qv + ql) * R(1000.0), R(30.0));
  const R zphip1 = R(0.2) + R(0.01) * zqtot + R(0.0027) * (zqtot * zqtot) - R(0.00008) * (zqtot * zqtot * zqtot);
  const R zphip2 = zphip1 - R(0.2);
  const R zphip3 = R(0.123) * m_exp(-(zphip1 * R(0.55)));
  const R zanew2 = m_max(za, R(0.1));
  const R zphic = m_pow(zgridlen * zanew2, R(1) / R(3)) *
                  m_pow(m_pow(zphip3 * zgridlen * zanew2, R(1.5)) + R(3) * zphip3, R(-0.17));
  R zfracsdc = (zphip1 - zphip2 * zanew2) * zphic;
  if (za > R(0.95)) zfracsdc = R(0.17) * zphic;
  return R(1.3) * zfracsdc;  // ZR12, the 1D-to-2D enhancement
}
// PFSD of one level (:4307-4359).  qv, qi: tidied state; zanewp: ZANEWP (:3818-3822); zratfsd: ZRATFSD
// (:4112-4121); zicef, zliqf: final ice and liquid (:4123-4124); zlfsd: the liquid FSD of the level.  The
// literals .5, .001, .24, .11, 10. and .03 of :4320-4322 and 0.00001 of :4315 are default real (FP32) in the
// reference and are kept so.
template <typename R>
__device__ __forceinline__ R cloud_fsd(R rcloud_frac_std, R zgridlen, R qv, R qi, R zdp, R zrho, R rg, R zanewp,
                                       R zratfsd, R zicef, R zliqf, R zlfsd) {
  const R zifsdback = R(0.446585) * m_pow(zgridlen, R(1) / R(3)) *
                      m_pow(m_pow(R(0.308061) * zgridlen, R(0.395736)) + R(1), R(-0.744527));
  const R zqtot = m_max(m_min((qv + qi) * R(1000.0), R(30.0)), R(0.00001f));
  R zdelz = m_div(R(1), rg * zrho) * zdp * R(0.001);
  zdelz = m_max(m_min(zdelz, R(0.5f)), R(0.001f));
  const R zphi = m_pow(m_div(zdelz, R(0.24f)), R(0.11f)) * m_pow(m_div(zqtot, R(10.0f)), R(0.03f));
  R zifsd = rcloud_frac_std;
  if (zanewp > R(0.95)) {  // overcast
    zifsd = R(1.3) * (zifsdback * zphi + zratfsd * R(1.5));
  } else if (zanewp > R(0.001) && zanewp <= R(0.95)) {
    zifsd = R(1.3) * (zifsdback * zphi * R(1.5) + zratfsd * R(1.5));
  }
  R zfsd = rcloud_frac_std;
  if (zicef > R(1.e-14)) zfsd = zifsd;
  if (zliqf > R(0) && zicef <= R(1.e-14)) zfsd = zlfsd;
  return m_min(m_max(R(0.1), zfsd), R(3.575));
}

// The vertical-integral field (IS+1) adds into slots 4-12, 19-39, 44-58 (process terms) and 70-82 (counters and
// PEIS) at every level of the column and assigns slots 60 and 83.  The 42 process sums live in LDS for the whole
// column: loaded from PEXTRA before the level loop, added to in the reference's order, stored after it, so each
// level adds old + x exactly as PEXTRA(JL,IK,IS+1) = PEXTRA(JL,IK,IS+1) + x does.  The counter slots add a column
// constant (1 or PEIS) at every level: vertint_end adds it once per level, in sequence, to the loaded value.
constexpr int NVI = 42;
__host__ __device__ constexpr int vi_slot(int j) {  // slot of accumulator j
  return j < 8 ? (j == 0 ? 4 : j + 5) : j < 28 ? (j == 8 ? 19 : j + 12) : (j == 28 ? 44 : j + 17);
}
__host__ __device__ constexpr int vi_index(int slot) {  // accumulator of a slot
  return slot == 4 ? 0 : slot <= 12 ? slot - 5 : slot == 19 ? 8 : slot <= 39 ? slot - 12 : slot == 44 ? 28
       : slot - 17;
}
static_assert(vi_slot(vi_index(4)) == 4 && vi_slot(vi_index(39)) == 39 && vi_slot(vi_index(46)) == 46 &&
              vi_slot(vi_index(58)) == 58 && vi_index(58) == NVI - 1, "vertical-integral slot map");

template <typename R>
struct ViAcc {
  R* s;  // this lane's NVI sums at stride n
  uint32_t n;
  __device__ __forceinline__ void add(int slot, R x) const {
    R& a = s[vi_index(slot) * n];
    a = a + x;
  }
};

// Load (before the level loop) and store (after it) the accumulating slots; ext(f, k) as for budget_write.
template <typename R, typename EXT>
__device__ __forceinline__ void vertint_begin(const ViAcc<R>& acc, const KBudPtrs<R>& p, EXT ext) {
  const int F = ((p.flags & BUD_LBUD23) ? 26 : 0) + 1;
  const char* base = reinterpret_cast<const char*>(p.pextra);
#pragma unroll
  for (int j = 0; j < NVI; ++j) acc.s[j * acc.n] = *reinterpret_cast<const R*>(base + ext(F, vi_slot(j)));
}
// nlev: the levels the scheme ran (NCLDTOP..KLEV), each of which added the counters once (:3924-3939).
template <typename R, typename EXT>
__device__ __forceinline__ void vertint_end(const ViAcc<R>& acc, const KBudPtrs<R>& p, EXT ext, R prainfrac,
                                            int32_t ktype, R peis, uint32_t o1, int nlev) {
  const int is = (p.flags & BUD_LBUD23) ? 26 : 0, F = is + 1;
  char* base = reinterpret_cast<char*>(p.pextra);
#pragma unroll
  for (int j = 0; j < NVI; ++j) *reinterpret_cast<R*>(base + ext(F, vi_slot(j))) = acc.s[j * acc.n];
  // slot 60 holds PRAINFRAC_TOPRFZ as of the last level, slot 83 KCTOP where KTYPE > 0 (:3922, :3939)
  *reinterpret_cast<R*>(base + ext(F, 60)) = prainfrac;
  if (ktype > 0) *reinterpret_cast<R*>(base + ext(F, 83)) = R(LD(p.kctop, o1, 0u));
  const int32_t kpbl = LD(p.kpbltype, o1, 0u);
  const bool on[13] = {ktype == 0, ktype == is + 1,  // the reference compares KTYPE and KPBLTYPE with IS+1 (:3926)
                       ktype == 2, ktype == 3, kpbl == 0, kpbl == is + 1, kpbl == 2, kpbl == 3, true,
                       peis > R(6.0), peis > R(8.0), peis > R(10.0), ktype >= 2 && peis < p.reisthsc};
  R old[13];
#pragma unroll
  for (int j = 0; j < 13; ++j) old[j] = *reinterpret_cast<const R*>(base + ext(F, 70 + j));
#pragma unroll
  for (int j = 0; j < 13; ++j) {
    if (!on[j]) continue;
    const R x = j == 8 ? peis : R(1);
    R a = old[j];
    for (int l = 0; l < nlev; ++l) a = a + x;
    *reinterpret_cast<R*>(base + ext(F, 70 + j)) = a;
  }
}

// The 3D budgets: IS of each block for the switches in force, and the fields each block updates, in the order
// budget_write produces their values.
struct BudBlocks {
  int c, l, i, t;
};
__device__ __forceinline__ BudBlocks bud_blocks(int32_t f) {
  BudBlocks o;
  int is = ((f & BUD_LBUD23) ? 26 : 0) + ((f & BUD_VERTINT) ? 1 : 0);
  o.c = is;
  is += (f & BUD_C) ? 12 : 0;
  o.l = is;
  is += (f & BUD_L) ? 22 : 0;
  o.i = is;
  is += (f & BUD_I) ? 20 : 0;
  o.t = is;
  return o;
}
constexpr int NC3 = 7, NL3 = 17, NI3 = 15, NT3 = 13;
__host__ __device__ constexpr int c_field(int j) { return j == 0 ? 4 : j < 5 ? j + 5 : j + 6; }  // 4, 6-9, 11, 12
__host__ __device__ constexpr int l_field(int j) { return j == 0 ? 5 : j + 6; }                 // 5, 7-22
__host__ __device__ constexpr int i_field(int j) { return j == 0 ? 5 : j + 6; }                 // 5, 7-20
__host__ __device__ constexpr int t_field(int j) { return j + 1; }                              // 1-13
static_assert(c_field(NC3 - 1) == 12 && l_field(NL3 - 1) == 22 && i_field(NI3 - 1) == 20, "3D budget fields");

// The PEXTRA words of this level that the 3D budgets update, loaded at the start of the level so that the loads
// overlap the level's computation; budget_write adds to these values and stores.  No other write touches these
// words in between (the vertical integrals are in field IS+1, the 3D fields are distinct per level).
template <typename R>
struct Pre3D {
  R c[NC3], l[NL3], i[NI3], t[NT3];
};
template <typename R, typename EXT>
__device__ __forceinline__ void budget_prefetch(Pre3D<R>& pre, const KBudPtrs<R>& p, EXT ext, uint32_t lev) {
  const int32_t f = p.flags;
  const BudBlocks o = bud_blocks(f);
  const char* base = reinterpret_cast<const char*>(p.pextra);
  auto ld = [&](int fld) { return *reinterpret_cast<const R*>(base + ext(fld, lev)); };
  if (f & BUD_C) {
#pragma unroll
    for (int j = 0; j < NC3; ++j) pre.c[j] = ld(o.c + c_field(j));
  }
  if (f & BUD_L) {
#pragma unroll
    for (int j = 0; j < NL3; ++j) pre.l[j] = ld(o.l + l_field(j));
  }
  if (f & BUD_I) {
#pragma unroll
    for (int j = 0; j < NI3; ++j) pre.i[j] = ld(o.i + i_field(j));
  }
  if (f & BUD_T) {
#pragma unroll
    for (int j = 0; j < NT3; ++j) pre.t[j] = ld(o.t + t_field(j));
  }
}

// Section 7 of CLOUDSC for one column and level (:3831-4101).  ext(f, k) is the byte offset of PEXTRA(JL,k,f)
// (1-based f and k); lev = JK (1-based); zdz = layer depth, as ZDP/(ZRHO*RG); acc: the vertical-integral sums;
// pre: this level's 3D PEXTRA words as loaded by budget_prefetch.  Rates the default path never sets (ZBUDCC 1,
// 10-12; ZBUDL 1, 8-10, 12-13, 16, 22; ZBUDI 1, 8-10, 18; ZPRACS, ZPRACI, ZPIEVAP) are zero and enter as such, so
// every word the reference touches is touched here with the same value.
template <typename R, typename EXT>
__device__ __forceinline__ void budget_write(const Bud<R, true>& b, const ViAcc<R>& acc, const Pre3D<R>& pre,
                                             const KBudPtrs<R>& p, EXT ext, uint32_t lev, uint32_t o2, R zqtmst,
                                             R ptsphy, R zdz, R qxn_l, R qxn_i, R ralvdcp, R ralsdcp, R ralfdcp) {
  const int32_t f = p.flags;
  const R zdt = (f & BUD_TIMEINT) ? ptsphy : R(1);
  const R pvfa = LD(p.pvfa, o2, 0u), pdyna = LD(p.pdyna, o2, 0u), pvfl = LD(p.pvfl, o2, 0u);
  const R pdynl = LD(p.pdynl, o2, 0u), pvfi = LD(p.pvfi, o2, 0u), pdyni = LD(p.pdyni, o2, 0u);
  if (f & BUD_VERTINT) {  // vertical integrals into field IS+1, "levels" used as slots (:3849-3943)
    auto vi = [&](R x) { return x * zdz * zdt; };  // X*ZDZ*ZDT, in the reference's order
    acc.add(4, vi(b.get(CC2)));
    acc.add(6, vi(b.get(CC3)));
    acc.add(7, vi((b.get(CC4) + b.get(CC6))));
    acc.add(8, vi(b.get(CC5)));
    acc.add(9, vi(b.get(CC7)));
    acc.add(10, vi(R(0)));  // ZBUDCC(JL,12), the tidy-up, is never set
    acc.add(11, vi(pvfa));
    acc.add(12, vi(pdyna));
    acc.add(19, vi(b.get(L2)));
    acc.add(21, vi(b.get(L3)));
    acc.add(22, vi((b.get(L4) + b.get(L6))));
    acc.add(23, vi(b.get(L5)));
    acc.add(24, vi(b.get(L7)));
    acc.add(25, vi(b.get(L11)));
    acc.add(26, vi(b.get(L14)));
    acc.add(27, vi(b.get(L15)));
    acc.add(28, vi(R(0)));
    acc.add(29, vi(R(0)));
    acc.add(30, vi(b.get(L17)));
    acc.add(31, vi((b.get(L18) + b.get(L19))));
    acc.add(32, vi(b.get(L20)));
    acc.add(33, vi(b.get(L21)));
    acc.add(34, vi(R(0)) * zqtmst);  // ZPRACS: rain accreting snow is off on the default path
    acc.add(35, vi(b.get(PSACR)) * zqtmst);
    acc.add(36, vi(R(0)) * zqtmst);  // ZPRACI: off on the default path
    acc.add(37, vi(b.get(PIACR)) * zqtmst);
    acc.add(38, vi(pvfl));
    acc.add(39, vi(pdynl));
    acc.add(44, vi(b.get(I2)));
    acc.add(46, vi(b.get(I3)));
    acc.add(47, vi((b.get(I4) + b.get(I6))));
    acc.add(48, vi(b.get(I5)));
    acc.add(49, vi(b.get(I7)));
    acc.add(50, vi(b.get(I11)));
    acc.add(51, vi((b.get(I12) + b.get(I13))));
    acc.add(52, vi(b.get(I14)));
    acc.add(53, vi(b.get(I16)));
    acc.add(54, vi(b.get(I17)));
    acc.add(55, vi(b.get(PSDEP) * zqtmst));
    acc.add(56, vi(pvfi));
    acc.add(57, vi(pdyni));
    acc.add(58, vi(LD(p.pdyns, o2, 0u)));
  }
  const BudBlocks o = bud_blocks(f);
  char* base = reinterpret_cast<char*>(p.pextra);
  auto st = [&](int fld, R x) { *reinterpret_cast<R*>(base + ext(fld, lev)) = x; };
  if (f & BUD_C) {  // cloud fraction (:3948-3966)
    const R v[NC3] = {b.get(CC2) * zdt, b.get(CC3) * zdt, (b.get(CC4) + b.get(CC6)) * zdt, b.get(CC5) * zdt,
                      b.get(CC7) * zdt, pvfa * zdt, pdyna * zdt};
#pragma unroll
    for (int j = 0; j < NC3; ++j) st(o.c + c_field(j), pre.c[j] + v[j]);
  }
  if (f & BUD_L) {  // cloud liquid (:3971-4003)
    const R v[NL3] = {b.get(L2) * zdt,  b.get(L3) * zdt,  (b.get(L4) + b.get(L6)) * zdt, b.get(L5) * zdt,
                      b.get(L7) * zdt,  b.get(L11) * zdt, R(0) * zdt,  R(0) * zdt,  b.get(L14) * zdt,
                      b.get(L15) * zdt, b.get(L17) * zdt, (b.get(L18) + b.get(L19)) * zdt, b.get(L20) * zdt,
                      b.get(L21) * zdt, pvfl * zdt,       pdynl * zdt, qxn_l};
#pragma unroll
    for (int j = 0; j < NL3; ++j) st(o.l + l_field(j), pre.l[j] + v[j]);
  }
  if (f & BUD_I) {  // cloud ice (:4008-4041); field 18 is ZPIEVAP: ice sublimation (ISUBLICE) is off by default
    const R v[NI3] = {b.get(I2) * zdt,  b.get(I3) * zdt,  (b.get(I4) + b.get(I6)) * zdt, b.get(I5) * zdt,
                      b.get(I7) * zdt,  b.get(I11) * zdt, (b.get(I12) + b.get(I13)) * zdt, b.get(I14) * zdt,
                      (b.get(I15) + b.get(I16)) * zdt, b.get(PSDEP) * zqtmst * zdt, b.get(I17) * zdt,
                      pvfi * zdt, R(0) * zqtmst * zdt, pdyni * zdt, qxn_i};
#pragma unroll
    for (int j = 0; j < NI3; ++j) st(o.i + i_field(j), pre.i[j] + v[j]);
  }
  if (f & BUD_T) {  // temperature (:4052-4101)
    const R v[NT3] = {LD(p.phrsw, o2, 0u),
                      LD(p.phrlw, o2, 0u),
                      (b.get(L2) * ralvdcp - b.get(I11) * ralvdcp) * zdt,
                      (b.get(I2) + b.get(I11)) * ralsdcp * zdt,
                      b.get(I11) * ralfdcp * zdt,
                      (b.get(L7) + b.get(L5)) * ralvdcp * zdt,
                      (b.get(I7) + b.get(I5)) * ralsdcp * zdt,
                      b.get(L20) * ralvdcp * zdt,
                      b.get(I17) * ralsdcp * zdt,
                      b.get(I15) * ralfdcp * zdt,
                      b.get(I16) * ralfdcp * zdt,
                      R(0) - (b.get(L18) + b.get(L19)) * ralfdcp * zdt,
                      b.get(L21) * ralfdcp * zdt};
#pragma unroll
    for (int j = 0; j < NT3; ++j) st(o.t + t_field(j), pre.t[j] + v[j]);
  }
}

// Number of PEXTRA fields the switches need (IS after the last block), checked against KFLDX on the host.
inline int budget_fields(int32_t f) {
  int is = (f & BUD_LBUD23) ? 26 : 0;
  if (f & BUD_VERTINT) is += 1;
  if (f & BUD_C) is += 12;
  if (f & BUD_L) is += 22;
  if (f & BUD_I) is += 20;
  if (f & BUD_T) is += 13;
  return is;
}

}  // namespace cloudsc_gpu
