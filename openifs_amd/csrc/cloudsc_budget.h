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
  BUD_C = 1, BUD_L = 2, BUD_I = 4, BUD_T = 8, BUD_VERTINT = 16, BUD_TIMEINT = 32, BUD_LBUD23 = 64, BUD_CHECK = 128
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
  __device__ __forceinline__ void set(int, R) {}
  __device__ __forceinline__ R get(int) const { return R(0); }
};
template <typename R>
struct Bud<R, true> {
  R v[NBUD];
  __device__ __forceinline__ Bud() {
#pragma unroll
    for (int i = 0; i < NBUD; ++i) v[i] = R(0);
  }
  __device__ __forceinline__ void set(int i, R x) { v[i] = x; }
  __device__ __forceinline__ R get(int i) const { return v[i]; }
};

// Device pointers of the budget inputs and PEXTRA, blocked layout; PEXTRA is (NPROMA, KLEV, KFLDX, NGPBLKS).
template <typename R>
struct KBudPtrs {
  const R *pvfa, *pdyna, *pdynl, *pdyni, *pdyns, *phrsw, *phrlw, *pvfl, *pvfi;
  const int32_t *kpbltype, *kctop;
  R* pextra;
  int32_t kfldx, flags;
  R reisthsc, ralfdcp;
};

// Section 7 of CLOUDSC for one column and level (:3831-4101).  ext(f, k) is the byte offset of PEXTRA(JL,k,f)
// (1-based f and k); lev = JK (1-based); zdz = layer depth, as ZDP/(ZRHO*RG).
template <typename R, typename EXT>
__device__ __forceinline__ void budget_write(const Bud<R, true>& b, const KBudPtrs<R>& p, EXT ext, uint32_t lev,
                                             uint32_t o2, uint32_t o1, R zqtmst, R ptsphy, R zdz, R qxn_l,
                                             R qxn_i, R prainfrac, int32_t ktype, R peis, R ralvdcp, R ralsdcp,
                                             R ralfdcp) {
  const int32_t f = p.flags;
  int is = (f & BUD_LBUD23) ? 26 : 0;
  const R zdt = (f & BUD_TIMEINT) ? ptsphy : R(1);
  auto add = [&](int fld, uint32_t k, R x) {
    R* a = reinterpret_cast<R*>(reinterpret_cast<char*>(p.pextra) + ext(fld, k));
    *a = *a + x;
  };
  auto put = [&](int fld, uint32_t k, R x) { *reinterpret_cast<R*>(reinterpret_cast<char*>(p.pextra) + ext(fld, k)) = x; };
  const R pvfa = LD(p.pvfa, o2, 0u), pdyna = LD(p.pdyna, o2, 0u), pvfl = LD(p.pvfl, o2, 0u);
  const R pdynl = LD(p.pdynl, o2, 0u), pvfi = LD(p.pvfi, o2, 0u), pdyni = LD(p.pdyni, o2, 0u);
  if (f & BUD_VERTINT) {  // vertical integrals into field IS+1, "levels" used as slots (:3849-3943)
    const int F = is + 1;
    auto vi = [&](R x) { return x * zdz * zdt; };  // X*ZDZ*ZDT, in the reference's order
    add(F, 4, vi(b.get(CC2)));
    add(F, 6, vi(b.get(CC3)));
    add(F, 7, vi((b.get(CC4) + b.get(CC6))));
    add(F, 8, vi(b.get(CC5)));
    add(F, 9, vi(b.get(CC7)));
    add(F, 10, vi(R(0)));  // ZBUDCC(JL,12), the tidy-up, is never set
    add(F, 11, vi(pvfa));
    add(F, 12, vi(pdyna));
    add(F, 19, vi(b.get(L2)));
    add(F, 21, vi(b.get(L3)));
    add(F, 22, vi((b.get(L4) + b.get(L6))));
    add(F, 23, vi(b.get(L5)));
    add(F, 24, vi(b.get(L7)));
    add(F, 25, vi(b.get(L11)));
    add(F, 26, vi(b.get(L14)));
    add(F, 27, vi(b.get(L15)));
    add(F, 28, vi(R(0)));
    add(F, 29, vi(R(0)));
    add(F, 30, vi(b.get(L17)));
    add(F, 31, vi((b.get(L18) + b.get(L19))));
    add(F, 32, vi(b.get(L20)));
    add(F, 33, vi(b.get(L21)));
    add(F, 34, vi(R(0)) * zqtmst);  // ZPRACS: rain accreting snow is off on the default path
    add(F, 35, vi(b.get(PSACR)) * zqtmst);
    add(F, 36, vi(R(0)) * zqtmst);  // ZPRACI: off on the default path
    add(F, 37, vi(b.get(PIACR)) * zqtmst);
    add(F, 38, vi(pvfl));
    add(F, 39, vi(pdynl));
    add(F, 44, vi(b.get(I2)));
    add(F, 46, vi(b.get(I3)));
    add(F, 47, vi((b.get(I4) + b.get(I6))));
    add(F, 48, vi(b.get(I5)));
    add(F, 49, vi(b.get(I7)));
    add(F, 50, vi(b.get(I11)));
    add(F, 51, vi((b.get(I12) + b.get(I13))));
    add(F, 52, vi(b.get(I14)));
    add(F, 53, vi(b.get(I16)));
    add(F, 54, vi(b.get(I17)));
    add(F, 55, vi(b.get(PSDEP) * zqtmst));
    add(F, 56, vi(pvfi));
    add(F, 57, vi(pdyni));
    add(F, 58, vi(LD(p.pdyns, o2, 0u)));
    put(F, 60, prainfrac);
    const int32_t kpbl = LD(p.kpbltype, o1, 0u);
    if (ktype == 0) add(F, 70, R(1));
    if (ktype == is + 1) add(F, 71, R(1));  // the reference compares KTYPE with IS+1 here (:3926)
    if (ktype == 2) add(F, 72, R(1));
    if (ktype == 3) add(F, 73, R(1));
    if (kpbl == 0) add(F, 74, R(1));
    if (kpbl == is + 1) add(F, 75, R(1));  // and KPBLTYPE with IS+1 (:3930)
    if (kpbl == 2) add(F, 76, R(1));
    if (kpbl == 3) add(F, 77, R(1));
    add(F, 78, peis);
    if (peis > R(6.0)) add(F, 79, R(1));
    if (peis > R(8.0)) add(F, 80, R(1));
    if (peis > R(10.0)) add(F, 81, R(1));
    if (ktype >= 2 && peis < p.reisthsc) add(F, 82, R(1));
    if (ktype > 0) put(F, 83, R(LD(p.kctop, o1, 0u)));
    is = is + 1;
  }
  if (f & BUD_C) {  // cloud fraction (:3948-3966)
    add(is + 4, lev, b.get(CC2) * zdt);
    add(is + 6, lev, b.get(CC3) * zdt);
    add(is + 7, lev, (b.get(CC4) + b.get(CC6)) * zdt);
    add(is + 8, lev, b.get(CC5) * zdt);
    add(is + 9, lev, b.get(CC7) * zdt);
    add(is + 11, lev, pvfa * zdt);
    add(is + 12, lev, pdyna * zdt);
    is = is + 12;
  }
  if (f & BUD_L) {  // cloud liquid (:3971-4003)
    add(is + 5, lev, b.get(L2) * zdt);
    add(is + 7, lev, b.get(L3) * zdt);
    add(is + 8, lev, (b.get(L4) + b.get(L6)) * zdt);
    add(is + 9, lev, b.get(L5) * zdt);
    add(is + 10, lev, b.get(L7) * zdt);
    add(is + 11, lev, b.get(L11) * zdt);
    add(is + 12, lev, R(0) * zdt);
    add(is + 13, lev, R(0) * zdt);
    add(is + 14, lev, b.get(L14) * zdt);
    add(is + 15, lev, b.get(L15) * zdt);
    add(is + 16, lev, b.get(L17) * zdt);
    add(is + 17, lev, (b.get(L18) + b.get(L19)) * zdt);
    add(is + 18, lev, b.get(L20) * zdt);
    add(is + 19, lev, b.get(L21) * zdt);
    add(is + 20, lev, pvfl * zdt);
    add(is + 21, lev, pdynl * zdt);
    add(is + 22, lev, qxn_l);
    is = is + 22;
  }
  if (f & BUD_I) {  // cloud ice (:4008-4041)
    add(is + 5, lev, b.get(I2) * zdt);
    add(is + 7, lev, b.get(I3) * zdt);
    add(is + 8, lev, (b.get(I4) + b.get(I6)) * zdt);
    add(is + 9, lev, b.get(I5) * zdt);
    add(is + 10, lev, b.get(I7) * zdt);
    add(is + 11, lev, b.get(I11) * zdt);
    add(is + 12, lev, (b.get(I12) + b.get(I13)) * zdt);
    add(is + 13, lev, b.get(I14) * zdt);
    add(is + 14, lev, (b.get(I15) + b.get(I16)) * zdt);
    add(is + 15, lev, b.get(PSDEP) * zqtmst * zdt);
    add(is + 16, lev, b.get(I17) * zdt);
    add(is + 17, lev, pvfi * zdt);
    add(is + 18, lev, R(0) * zqtmst * zdt);  // ZPIEVAP: ice sublimation (ISUBLICE) is off on the default path
    add(is + 19, lev, pdyni * zdt);
    add(is + 20, lev, qxn_i);
    is = is + 20;
  }
  if (f & BUD_T) {  // temperature (:4052-4101)
    add(is + 1, lev, LD(p.phrsw, o2, 0u));
    add(is + 2, lev, LD(p.phrlw, o2, 0u));
    add(is + 3, lev, (b.get(L2) * ralvdcp - b.get(I11) * ralvdcp) * zdt);
    add(is + 4, lev, (b.get(I2) + b.get(I11)) * ralsdcp * zdt);
    add(is + 5, lev, b.get(I11) * ralfdcp * zdt);
    add(is + 6, lev, (b.get(L7) + b.get(L5)) * ralvdcp * zdt);
    add(is + 7, lev, (b.get(I7) + b.get(I5)) * ralsdcp * zdt);
    add(is + 8, lev, b.get(L20) * ralvdcp * zdt);
    add(is + 9, lev, b.get(I17) * ralsdcp * zdt);
    add(is + 10, lev, b.get(I15) * ralfdcp * zdt);
    add(is + 11, lev, b.get(I16) * ralfdcp * zdt);
    add(is + 12, lev, R(0) - (b.get(L18) + b.get(L19)) * ralfdcp * zdt);
    add(is + 13, lev, b.get(L21) * ralfdcp * zdt);
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
