// CLOUDSC for gfx950 -- the column kernel.  See cloudsc_kernel.h for the
// data layout and the register-level reformulation.  Line citations refer to
// ifs-source/arpifs/phys_ec/cloudsc.F90 unless another file is named.
#include <hip/hip_runtime.h>

#include <map>
#include <type_traits>
#include <mutex>
#include <utility>

#include "cloudsc_device.h"
#include "cloudsc_kernel.h"
#include "cloudsc_budget.h"
#include "cloudsc_truncation.h"
#include "satadj_device.h"

namespace cloudsc_gpu {

// The reference's negative-condensate warnings (cloudsc.F90:3794 liquid, :3802 ice) as counts: one atomic add per
// wave and counter, by the wave's first active lane, only when some lane of the wave warns.
__device__ __forceinline__ void count_warnings(uint32_t* n, bool wl, bool wi) {
  const uint64_t bl = __ballot(wl), bi = __ballot(wi);
  if ((bl | bi) == 0) return;
  const uint64_t act = __ballot(1);
  if (__lane_id() == (unsigned)(__ffsll((long long)act) - 1)) {
    if (bl) atomicAdd(n, (uint32_t)__popcll(bl));
    if (bi) atomicAdd(n + 1, (uint32_t)__popcll(bi));
  }
}

// KC / KCD: constants of the working precision R and of the water-budget precision RD, either the run-time
// structure KConst or the compile-time default configuration KConstDefault (same member names).
// UNI: every wave lies in one block (NPROMA a multiple of 64), so block-dependent offsets are wave-uniform.
// SAT: the fused CALLPAR sequence (callpar.F90:1253-1293): on every level at or below NCLDTOP the column first runs
// CLOUD_SATADJ for the point (satadj_point), adds its tendencies to TENDENCY_CML in place (STATE_INCREMENT,
// :1278-1279), stores them as TENDENCY_SATADJ (STATE_COPY, :1289), and CLOUDSC then reads the incremented CML
// tendencies from registers.
// BUD: the process budgets (LCLDBUD*) are recorded per level and written into PEXTRA (cloudsc_budget.h).
// WPE: minimum waves per SIMD the register allocation must allow (0: the compiler's choice); fewer registers per
// lane for more waves in flight, at the price of spills when the live state does not fit.
// KL: number of levels fixed at compile time (0: d.klev at run time), so that the level loop and every level
// offset are constants of the instantiation.
template <typename R, typename RD, typename KC, typename KCD, int BLOCK, bool UNI, bool SAT, bool BUD, int WPE = 0,
          int KL = 0>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1))) void
cloudsc_column_kernel(const KC c, const KCD cd, const KPtrs<R> p, const KDims d, int ssc_off, const KSatPtrs<R> s,
                      const KBudPtrs<R> bp) {
  const uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t nproma = (uint32_t)d.nproma, klev = KL > 0 ? (uint32_t)KL : (uint32_t)d.klev;
  if (g >= (uint32_t)d.ngpblks * nproma) return;
  uint32_t b = g / nproma;
  if (UNI) b = __builtin_amdgcn_readfirstlane(b);
  const uint32_t jl = g - b * nproma;  // 0-based JL-1
  constexpr uint32_t E = sizeof(R);
  const uint32_t jo = jl * E, joi = jl * 4u;           // lane offsets (R and int32 arrays)
  const uint32_t u1 = b * nproma * E, u1i = b * nproma * 4u;
  // PRAINFRAC_TOPRFZ is zeroed for all KLON columns of the block (:1010)
  ST(p.prainfrac_toprfz, u1, jo, R(0));
  const int kfdia = min(d.kfdia, d.ngptot - (int)(b * nproma));
  if ((int)jl < d.kidia - 1 || (int)jl >= kfdia) return;

  const uint32_t s2 = nproma * klev;
  const uint32_t lstep = nproma * E;                  // one level
  const uint32_t sstep = s2 * E;                      // one species of a (NPROMA,KLEV,n) array
  const uint32_t b2 = b * s2 * E;                     // (NPROMA,KLEV) level 0
  const uint32_t bh = b * (s2 + nproma) * E;          // (NPROMA,KLEV+1) level 0
  const uint32_t b5 = b * s2 * 5 * E, b4 = b * s2 * 4 * E, b2x = b * s2 * 2 * E;

  const R ptsphy = c.ptsphy, zqtmst = c.zqtmst, zepsec = R(1.e-14);
  const R rlmin = c.rlmin;
  R* bud_lds = nullptr;  // this lane's budget entries and vertical-integral sums (BUD only)
  ViAcc<R> vacc{nullptr, BLOCK};
  const uint32_t kfx = (uint32_t)bp.kfldx;
  auto ext = [&](int fld, uint32_t kk) {  // byte offset of PEXTRA(JL,kk,fld) of this block
    return (((b * kfx + (uint32_t)(fld - 1)) * klev + (kk - 1)) * nproma) * E + jo;
  };
  if constexpr (BUD) {
    __shared__ R s_bud[(NBUD + NVI) * BLOCK];
    bud_lds = s_bud + threadIdx.x;
    vacc.s = s_bud + NBUD * BLOCK + threadIdx.x;
    if (bp.flags & BUD_VERTINT) vertint_begin(vacc, bp, ext);
  }
  // LRAD_CLOUD_INHOMOG / LCLOUD_INHOMOG: grid length of the column (diagnostic instantiation only)
  const R zgridlen = (BUD && (bp.fsd || bp.inhomog)) ? fsd_gridlen(bp.ra, bp.rpi, LD(bp.pgaw, u1, jo)) : R(0);

  // Per-column inputs
  const R plsm = LD(p.plsm, u1, jo);
  const int32_t ktype = LD(p.ktype, u1i, joi);
  const bool ldcum = LD(p.ldcum, u1i, joi) != 0;
  const R peis = LD(p.peis, u1, jo);
  const R paph_surf = LD(p.paph, bh + klev * lstep, jo);
  const R zadvw = (ktype == 1 && c.rmfadvw > R(0)) ? R(1) - c.rmfadvw : R(1);                  // :696-703
  const R zadvwd = (ktype == 1 && c.rmfadvw > R(0)) ? R(1) - c.rmfadvw * c.rmfadvwdd : R(1);
  const R zconst = plsm > R(0.5) ? c.rcl_kk_cloud_num_land : c.rcl_kk_cloud_num_sea;          // :2063-2069
  const R zconst_pow = m_pow(zconst, c.rcl_kkbaun);
  const int ncldtop0 = c.ncldtop - 1;  // 0-based first level of the cloud scheme

  // Level-carried state (:1205-1211 and :1011)
  R qxnm1_l = R(0), qxnm1_i = R(0), qxnm1_r = R(0), qxnm1_s = R(0);
  R zanewm1 = R(0), zcovptot = R(0), zcovpmax = R(0), zcldtopdist = R(0);
  bool llrainliq = true;
  R prainfrac = R(0);
  R pf_i = R(0), pf_r = R(0), pf_s = R(0);  // ZPFPLSX(JK) of the falling species
  R za_prev = R(0), qr_prev = R(0), qs_prev = R(0), t_prev = R(0), pap_prev = R(0);
  // DDH flux running sums at the upper half level (:4231-4242)
  R sqlf = R(0), sqif = R(0), cqlng = R(0), cqnng = R(0), sqltur = R(0), sqitur = R(0);
  ST(p.pfsqlf, bh, jo, R(0));
  ST(p.pfsqif, bh, jo, R(0));
  ST(p.pfsqrf, bh, jo, R(0));
  ST(p.pfsqsf, bh, jo, R(0));
  ST(p.pfcqlng, bh, jo, R(0));
  ST(p.pfcqnng, bh, jo, R(0));
  ST(p.pfcqrng, bh, jo, R(0));
  ST(p.pfcqsng, bh, jo, R(0));
  ST(p.pfsqltur, bh, jo, R(0));
  ST(p.pfsqitur, bh, jo, R(0));

  R paph_k = LD(p.paph, bh, jo);
  R pmfu_k = LD(p.pmfu, b2, jo), pmfd_k = LD(p.pmfd, b2, jo);

  for (uint32_t k = 0; k < klev; ++k) {
    const uint32_t i2 = b2 + k * lstep;        // level k of a (NPROMA,KLEV) array
    const uint32_t ih0 = bh + k * lstep;       // upper half level of layer k
    const uint32_t ih1 = ih0 + lstep;          // lower half level
    const uint32_t i5 = b5 + k * lstep;
    const uint32_t i4 = b4 + k * lstep;        // (all wave-uniform with UNI; the lane adds jo)
    const bool has_next = k + 1 < klev;
    const R pmfu_n = has_next ? LD(p.pmfu, i2 + lstep, jo) : R(0);
    const R pmfd_n = has_next ? LD(p.pmfd, i2 + lstep, jo) : R(0);
    const R plu_n = has_next ? LD(p.plu, i2 + lstep, jo) : R(0);
    const R paph_k1 = LD(p.paph, ih1, jo);
    const R pap = LD(p.pap, i2, jo);
    const R pt = LD(p.pt, i2, jo), pq = LD(p.pq, i2, jo), pa = LD(p.pa, i2, jo);
    R tcml_t = LD(p.tcml_t, i2, jo), tcml_q = LD(p.tcml_q, i2, jo), tcml_a = LD(p.tcml_a, i2, jo);
    const R pvfl = LD(p.pvfl, i2, jo), pvfi = LD(p.pvfi, i2, jo);
    const R pludeli1 = LD(p.pludeli, i4, jo);
    const R pludeli2 = LD(p.pludeli, i4 + sstep, jo);
    const R pcl_l = LD(p.pclv, i5, jo), pcl_i = LD(p.pclv, i5 + sstep, jo);
    R tcml_l = LD(p.tcml_cld, i5, jo), tcml_i = LD(p.tcml_cld, i5 + sstep, jo);

    if constexpr (SAT) {
      if ((int)k >= ncldtop0) {  // CLOUD_SATADJ for this point, then TENDENCY_CML += its tendencies
        SatIn<R> x;
        x.pt = pt, x.pq = pq, x.pa = pa, x.pl = pcl_l, x.pi = pcl_i, x.pap = pap;
        x.paph0 = paph_k, x.paph1 = paph_k1, x.psurf = paph_surf;
        x.tct = tcml_t, x.tcq = tcml_q, x.tca = tcml_a, x.tcl = tcml_l, x.tci = tcml_i;
        x.hrsw = LD(s.phrsw, i2, jo), x.hrlw = LD(s.phrlw, i2, jo);
        x.vdft = LD(s.tvdf_t, i2, jo), x.vdfq = LD(s.tvdf_q, i2, jo);
        x.lude3 = LD(p.pludeli, i4 + 2 * sstep, jo), x.lude4 = LD(p.pludeli, i4 + 3 * sstep, jo);
        x.vervel = LD(s.pvervel, i2, jo);
        const SatOut<R> y = satadj_point<R, RD, KC, KCD>(c, cd, x);
        tcml_t = tcml_t + y.t;
        tcml_q = tcml_q + y.q;
        tcml_a = tcml_a + y.a;
        tcml_l = tcml_l + y.l;
        tcml_i = tcml_i + y.i;
        ST(s.tcml_t, i2, jo, tcml_t);
        ST(s.tcml_q, i2, jo, tcml_q);
        ST(s.tcml_a, i2, jo, tcml_a);
        ST(s.tcml_cld, i5, jo, tcml_l);
        ST(s.tcml_cld, i5 + sstep, jo, tcml_i);
        if (s.tsat_t) {
          ST(s.tsat_t, i2, jo, y.t);
          ST(s.tsat_q, i2, jo, y.q);
          ST(s.tsat_a, i2, jo, y.a);
          ST(s.tsat_cld, i5, jo, y.l);
          ST(s.tsat_cld, i5 + sstep, jo, y.i);
        }
      } else if (s.tsat_t) {  // zero above NCLDTOP (cloud_satadj.F90:338-348); TENDENCY_CML unchanged
        ST(s.tsat_t, i2, jo, R(0));
        ST(s.tsat_q, i2, jo, R(0));
        ST(s.tsat_a, i2, jo, R(0));
        ST(s.tsat_cld, i5, jo, R(0));
        ST(s.tsat_cld, i5 + sstep, jo, R(0));
      }
    }

    // ---- 1. initial values (:982-1002) ----
    const R ztp1 = pt + ptsphy * tcml_t;
    R qv = pq + ptsphy * tcml_q;
    R za = pa + ptsphy * tcml_a;
    const R zaorig = za;
    const R ql0 = pcl_l + ptsphy * tcml_l;
    const R qi0 = pcl_i + ptsphy * tcml_i;
    const R qr0 = LD(p.pclv, i5 + 2 * sstep, jo) + ptsphy * LD(p.tcml_cld, i5 + 2 * sstep, jo);
    const R qs0 = LD(p.pclv, i5 + 3 * sstep, jo) + ptsphy * LD(p.tcml_cld, i5 + 3 * sstep, jo);
    R ql = ql0, qi = qi0, qr = qr0, qs = qs0;

    // saturation values (:1087-1115)
    const Sat<R> sat = saturation<R, KC>(c, ztp1);
    const R foeeliq = c.r2es * sat.el, foeeice = c.r2es * sat.ei;
    const R zfoeewm = m_min(m_div(c.r2es * (sat.alfa * sat.el + (R(1) - sat.alfa) * sat.ei), pap), R(0.5));
    const R zqsmix = m_div(zfoeewm, R(1) - c.retv * zfoeewm);
    const R zdelta = (ztp1 - c.rtt >= R(0)) ? R(1) : R(0);
    const R zfoeew = m_min(m_div(zdelta * foeeliq + (R(1) - zdelta) * foeeice, pap), R(0.5));
    const R zqsice = m_div(zfoeew, R(1) - c.retv * zfoeew);
    const R zfoeeliqt = m_min(m_div(foeeliq, pap), R(0.5));
    const R zqsliq = m_div(zfoeeliqt, R(1) - c.retv * zfoeeliqt);

    // ---- 2. tidy up (:1127-1196) ----
    R tt = R(0), tq = R(0);
    R zlneg_l = R(0), zlneg_i = R(0), zlneg_r = R(0), zlneg_s = R(0);
    if (ql + qi < rlmin || za < c.ramin) {
      zlneg_l = zlneg_l + ql;
      R zqadj = ql * zqtmst;
      tq = tq + zqadj;
      tt = tt - c.ralvdcp * zqadj;
      qv = qv + ql;
      ql = R(0);
      zlneg_i = zlneg_i + qi;
      zqadj = qi * zqtmst;
      tq = tq + zqadj;
      tt = tt - c.ralsdcp * zqadj;
      qv = qv + qi;
      qi = R(0);
      za = R(0);
    }
#define CLOUDSC_TIDY(x, zlneg, rlat)   \
  if (x < rlmin) {                     \
    zlneg = zlneg + x;                 \
    const R zqadj = x * zqtmst;        \
    tq = tq + zqadj;                   \
    tt = tt - rlat * zqadj;            \
    qv = qv + x;                       \
    x = R(0);                          \
  }
    CLOUDSC_TIDY(ql, zlneg_l, c.ralvdcp)
    CLOUDSC_TIDY(qi, zlneg_i, c.ralsdcp)
    CLOUDSC_TIDY(qr, zlneg_r, c.ralvdcp)
    CLOUDSC_TIDY(qs, zlneg_s, c.ralsdcp)
#undef CLOUDSC_TIDY
    za = m_max(R(0), m_min(R(1), za));
    const R zli = ql + qi;
    R zliqfrac = R(0), zicefrac = R(0);
    if (zli > rlmin) {
      zliqfrac = m_share(ql, zli);
      zicefrac = R(1) - zliqfrac;
    }

    // Outputs of this level; defaults are those of levels above NCLDTOP.
    R qxn_l = R(0), qxn_i = R(0), qxn_r = R(0), qxn_s = R(0);
    R tcld_l = R(0), tcld_i = R(0), tcld_r = R(0), tcld_s = R(0), ta = R(0);
    // PFPLSL/PFPLSN/PFHPSL/PFHPSN at the upper half level (:4218-4224,4297-4302)
    {
      const R fl = pf_r + R(0), fn = pf_s + pf_i;
      ST(p.pfplsl, ih0, jo, fl);
      ST(p.pfplsn, ih0, jo, fn);
      ST(p.pfhpsl, ih0, jo, -c.rlvtt * fl);
      ST(p.pfhpsn, ih0, jo, -c.rlstt * fn);
    }

    // ZLFSD, ZANEWP, ZRATFSD of this level (LRAD_CLOUD_INHOMOG); above NCLDTOP they keep these values (:944-951)
    R fsd_liq = BUD ? bp.rcloud_frac_std : R(0), fsd_anewp = R(0), fsd_rat = R(0);

    if ((int)k >= ncldtop0) {
      const bool above_top = (int)k > ncldtop0;  // JK > NCLDTOP
      Bud<R, BUD> bud(bud_lds, BLOCK);  // ZBUDCC/ZBUDL/ZBUDI of this level, zeroed (:1271-1273)
      // FP32: the 3D budget words are loaded here and arrive during the level; FP64 loads them at section 7
      // (holding them across the level would spill)
      Pre3D<R> bud_pre;
      if constexpr (BUD && sizeof(R) == 4) budget_prefetch(bud_pre, bp, ext, k + 1);
      R zrainaut = R(0), zrainacc = R(0), zsnowaut = R(0), zsnowrime = R(0);  // (:1239-1242)
      const R plude_in = LD(p.plude, i2, jo);
      // derived variables (:1281-1324)
      const R zdp = paph_k1 - paph_k;
      const R zgdp = m_div(c.rg, zdp);
      const R zrho = m_div(pap, c.rd * ztp1);
      const R zdtgdp = ptsphy * zgdp;
      const R zrdtgdp = zdp * c.rdtg_r;
      const R tl = ztp1 - c.r4les, ti = ztp1 - c.r4ies;
      const R zfacw = m_div(c.r5les, tl * tl);
      R zcor = m_div(R(1), R(1) - c.retv * zfoeeliqt);
      const R zdqsliqdt = zfacw * zcor * zqsliq;
      const R zcorqsliq = R(1) + c.ralvdcp * zdqsliqdt;
      const R zfaci = m_div(c.r5ies, ti * ti);
      zcor = m_div(R(1), R(1) - c.retv * zfoeew);
      const R zdqsicedt = zfaci * zcor * zqsice;
      const R zcorqsice = R(1) + c.ralsdcp * zdqsicedt;
      const R zfac = sat.alfa * zfacw + (R(1) - sat.alfa) * zfaci;
      zcor = m_div(R(1), R(1) - c.retv * zfoeewm);
      const R zdqsmixdt = zfac * zcor * zqsmix;
      const R foeldcpm = sat.alfa * c.ralvdcp + (R(1) - sat.alfa) * c.ralsdcp;
      const R zcorqsmix = R(1) + foeldcpm * zdqsmixdt;
      const R zevaplimmix = m_max(m_div(zqsmix - qv, zcorqsmix), R(0));
      const R ztmpa = m_div(R(1), m_max(za, R(0.01)));
      R zliqcld = ql * ztmpa;
      R zicecld = qi * ztmpa;
      const R zlicld = zliqcld + zicecld;

      // Explicit (ZSOLQA) entries: vX = ZSOLQA(V,X) = -ZSOLQA(X,V); il = ZSOLQA(I,L);
      // ri = ZSOLQA(R,I); rs = ZSOLQA(R,S); dX = ZSOLQA(X,X).  Implicit (ZSOLQB) entries qbXY.
      // The level's water budget -- ZSOLQA, the first guesses ZQXFG and the sinks limited by them -- is kept in
      // RD: FP64 in the mixed variant, so that a species a limited sink removes completely leaves an FP64-sized
      // residue that the ZEPSEC clip after the solve removes, as in the FP64 build (R in the dp and sp variants).
      RD vl = RD(0), vi = RD(0), vr = RD(0), vs = RD(0), il = RD(0), ri = RD(0), rs = RD(0);
      RD dl = RD(0), di = RD(0), dr = RD(0), ds = RD(0);
      R qbll = R(0), qbii = R(0), qbrl = R(0), qbsi = R(0), qbsl = R(0);
      R zsolab = R(0), zsolac = R(0);
      R csrc_l = R(0), csrc_i = R(0), csrc_r = R(0), csrc_s = R(0), csink = R(0);
      R fsrc_i = R(0), fsrc_r = R(0), fsrc_s = R(0);
      RD qxfg_l = ql, qxfg_i = qi, qxfg_r = qr, qxfg_s = qs;
      R zqpretot = R(0);

      if (ql < rlmin) vl = ql;  // :1334-1342
      if (qi < rlmin) vi = qi;

      // 3.2 detrainment (:1362-1398)
      R plude = plude_in;
      if (has_next) {
        plude = plude_in * zdtgdp;
        if (ldcum && plude > zepsec && plu_n > zepsec) {
          zsolac = zsolac + m_div(plude, plu_n);
          csrc_l = pludeli1 * zdtgdp;
          csrc_i = pludeli2 * zdtgdp;
          dl = dl + csrc_l;
          di = di + csrc_i;
          bud.set(L3, csrc_l * zqtmst);  // (:1378-1380)
          bud.set(I3, csrc_i * zqtmst);
          bud.set(CC3, m_div(zqtmst * plude, plu_n));
        } else {
          plude = R(0);
        }
        if (c.lmfdsnow) {
          csrc_r = LD(p.psnde, b2x + k * lstep, jo) * zdtgdp;
          csrc_s = LD(p.psnde, b2x + k * lstep + sstep, jo) * zdtgdp;
          dr = dr + csrc_r;
          ds = ds + csrc_s;
        }
        ST(p.plude, i2, jo, plude);
      }

      // 3.3 subsidence from the level above (:1419-1492)
      if (above_top) {
        const R zmf = m_max(R(0), (pmfu_k * zadvw + pmfd_k * zadvwd) * zdtgdp);
        const R zacust = zmf * zanewm1;
        const R lcust_l = zmf * qxnm1_l, lcust_i = zmf * qxnm1_i;
        csrc_l = csrc_l + lcust_l;
        csrc_i = csrc_i + lcust_i;
        const R zdtdp = m_div(c.zrdcp * R(0.5) * (t_prev + ztp1), paph_k);
        const R zwtot = m_min(c.rg * (pmfu_k * zadvw + pmfd_k * zadvwd), m_div(pap - pap_prev, ptsphy));
        const R zdtforc = zdtdp * ptsphy * zwtot;
        const R zdqs = zanewm1 * zdtforc * zdqsmixdt;
        R zlfinalsum = R(0);
        {
          R zlfinal = m_max(R(0), lcust_l - zdqs);
          const R zevap = m_min(lcust_l - zlfinal, zevaplimmix);
          zlfinal = lcust_l - zevap;
          zlfinalsum = zlfinalsum + zlfinal;
          dl = dl + lcust_l;
          vl = vl + zevap;
          bud.set(L4, lcust_l * zqtmst);  // (:1459-1460)
          bud.set(L5, -zevap * zqtmst);
        }
        {
          R zlfinal = m_max(R(0), lcust_i - zdqs);
          const R zevap = m_min(lcust_i - zlfinal, zevaplimmix);
          zlfinal = lcust_i - zevap;
          zlfinalsum = zlfinalsum + zlfinal;
          di = di + lcust_i;
          vi = vi + zevap;
          bud.set(I4, lcust_i * zqtmst);  // (:1475-1476)
          bud.set(I5, -zevap * zqtmst);
        }
        zsolac = zsolac + zacust;
        bud.set(CC4, zacust * zqtmst);  // (:1484-1488)
        if (zlfinalsum < zepsec) {
          zsolac = zsolac - zacust;
          bud.set(CC5, -zacust * zqtmst);
        }
      }
      // subsidence sink to the level below, implicit (:1499-1514)
      if (has_next) {
        const R zmfdn = m_max(R(0), (pmfu_n * zadvw + pmfd_n * zadvwd) * zdtgdp);
        zsolab = zsolab + zmfdn;
        qbll = qbll + zmfdn;
        qbii = qbii + zmfdn;
        csink = zmfdn;
      }

      // 3.4 turbulent erosion, ITURBEROSION = 3 (:1689-1754)
      {
        R zldifdt = c.rcldiff * ptsphy;
        if (ktype >= 2 && plude > zepsec && peis < R(10.0)) {
          zldifdt = R(20.0) * c.rcldiff_convi * zldifdt;
        } else if (ktype == 1 && plude > zepsec) {
          zldifdt = R(2.0) * c.rcldiff_convi * zldifdt;
        }
        if (zli > zepsec) {
          const R zqe = m_div(qv - za * zqsmix, m_max(zepsec, R(1) - za));
          const R ze = m_max(zqsmix - zqe, R(0));
          R zleros = R(0.333) * za * (R(1) - za) * zldifdt * ze;
          zleros = m_min(zleros, zevaplimmix);
          zleros = m_min(zleros, zli);
          const R zaeros = m_div(R(0.5) * zleros, zlicld);
          zsolac = zsolac - zaeros;
          vl = vl + zliqfrac * zleros;
          vi = vi + zicefrac * zleros;
          bud.set(L7, -zliqfrac * zleros * zqtmst);  // (:1749-1751)
          bud.set(I7, -zicefrac * zleros * zqtmst);
          bud.set(CC7, -zaeros * zqtmst);
        }
      }

      // 4.1 sedimentation (:1783-1877)
      const R zfallcorr = m_pow(m_div(c.rdensref, zrho), R(0.4));
      const R ztemp_r = R(0.5) * (qr + qxnm1_r);
      const R zrlambda = m_pow(m_div(zrho * ztemp_r, c.rcl_lam1r), c.rcl_lam2r);
      const R vfall_r = c.rxfallspeedrain * zfallcorr * c.rcl_const7r * m_pow(zrlambda, c.rcl_dr);
      const R vfall_i = c.rvice, vfall_s = c.rvsnow;
      if (above_top) {
        fsrc_i = pf_i * zdtgdp;
        di = di + fsrc_i;
        bud.set(I12, fsrc_i * zqtmst);  // (:1864-1866)
        qxfg_i = qxfg_i + fsrc_i;
        zqpretot = R(zqpretot + qxfg_i);
        fsrc_r = pf_r * zdtgdp;
        dr = dr + fsrc_r;
        qxfg_r = qxfg_r + fsrc_r;
        zqpretot = R(zqpretot + qxfg_r);
        fsrc_s = pf_s * zdtgdp;
        ds = ds + fsrc_s;
        qxfg_s = qxfg_s + fsrc_s;
        zqpretot = R(zqpretot + qxfg_s);
      }
      const R fsink_i = zdtgdp * zrho * vfall_i;
      const R fsink_r = zdtgdp * zrho * vfall_r;
      const R fsink_s = zdtgdp * zrho * vfall_s;

      // 4.2 precipitation fraction (:1898-1919)
      R zcovpclr = R(0), zraincld = R(0), zraincldm1 = R(0), zsnowcld = R(0);
      if (zqpretot > zepsec) {
        zcovptot = R(1) - m_share((R(1) - zcovptot) * (R(1) - m_max(za, za_prev)),
                                   R(1) - m_min(za_prev, R(1) - R(1.E-06)));
        zcovptot = m_max(zcovptot, c.rcovpmin);
        zcovpclr = m_max(R(0), zcovptot - za);
        zraincld = m_div(R(0.5) * (qr + qr_prev), zcovptot);
        zraincldm1 = m_div(R(0.5) * (qr + qxnm1_r), zcovptot);
        zsnowcld = m_div(R(0.5) * (qs + qs_prev), zcovptot);
        zcovpmax = m_max(zcovptot, zcovpmax);
      } else {
        zcovptot = R(0);
        zcovpmax = R(0);
      }

      // 4.3 warm rain, implicit KK2000 (:2050-2254)
      if (zliqcld > zepsec) {
        R c_aut = c.c_aut, c_acc = c.c_acc;
        if constexpr (BUD) {
          if (bp.fsd || bp.inhomog) fsd_liq = liquid_fsd(qv, ql, za, zgridlen);  // (:1937-1971)
          if (bp.inhomog) {  // LCLOUD_INHOMOG: enhancement factors from the condensate FSD (:2075-2098)
            const InhomogFactors<R> e = inhomog_factors(fsd_liq, za, zgridlen, c.rcl_kkbauq, c.rcl_kkbac);
            c_aut = e.aut * ptsphy * c.rcl_kkaau;
            c_acc = e.acc * ptsphy * c.rcl_kkaac;
          }
        }
        zrainaut = c_aut * m_pow(zliqcld, c.rcl_kkbauq - R(1)) * zconst_pow;
        qbrl = qbrl + zrainaut;
        if (zraincld > zepsec)
          zrainacc = c_acc * m_pow(zliqcld, c.rcl_kkbac - R(1)) * m_pow(zraincld, c.rcl_kkbac);
        qbrl = qbrl + zrainacc;
      }

      // 4.4 rain evaporation (:2271-2378)
      {
        R zzrh = c.rprecrhmax + m_div((R(1) - c.rprecrhmax) * zcovpmax, m_max(zepsec, R(1) - za));
        zzrh = m_min(m_max(zzrh, c.rprecrhmax), R(1));
        R zrhc = c.ramid;
        const R zsigk = m_div(pap, paph_surf);
        if (zsigk > R(0.8)) {
          const R s = m_div(zsigk - R(0.8), R(0.2));
          zrhc = c.ramid + (R(1) - c.ramid) * (s * s);
        }
        zzrh = m_min(zrhc, zzrh);
        zzrh = m_min(R(0.9), zzrh);
        const R zqe = m_max(R(0), m_min(qv, zqsliq));
        if (zcovpclr > zepsec && zraincldm1 > zepsec && zqe < zzrh * zqsliq) {
          const R zpreclr = zraincldm1;
          const R zfallcorr2 = m_pow(m_div(c.rdensref, zrho), R(0.4f));
          const R zesatliq = c.rv_rd * foeeliq;
          const R zlambda = m_pow(m_div(c.rcl_lam1r, zrho * zpreclr), c.rcl_lam2r);
          const R t = ztp1;
          const R zevap_denom = c.rcl_cdenom1 * zesatliq - c.rcl_cdenom2 * t * zesatliq + c.rcl_cdenom3 * (t * t * t) * pap;
          const R zcorr2 = m_div(m_pow(m_div(t, R(273.)), R(1.5)) * R(393.), t + R(120.));
          const R zsubsat = m_max(zzrh * zqsliq - zqe, R(0));
          const R zbeta = m_div(R(0.15), zqsliq) * (t * t) * zesatliq * c.rcl_const1r * m_div(zcorr2, zevap_denom) *
                          (m_div(R(0.78), m_pow(zlambda, c.rcl_const4r)) +
                           m_div(c.rcl_const2r * m_sqrt(zrho * zfallcorr2), m_sqrt(zcorr2) * m_pow(zlambda, c.rcl_const3r)));
          const R zdenom = R(1) + zbeta * ptsphy * zcorqsliq;
          const R zdpevap = m_div(zcovpclr * zbeta * ptsphy * zsubsat, zdenom);
          const RD zevap = m_min(RD(zdpevap), qxfg_r);
          vr = vr + zevap;
          qxfg_r = qxfg_r - zevap;
          bud.set(L20, -zevap * zqtmst);  // (:2363)
        }
      }

      // 4.5 ice deposition, IDEPICE = 1 (:2404-2499)
      const R zvpice = m_div(foeeice * c.rv, c.rd);
      {
        if (za_prev < c.rcldtopcf && za >= c.rcldtopcf) {
          zcldtopdist = R(0);
        } else {
          zcldtopdist = zcldtopdist + m_div(zdp, zrho * c.rg);
        }
        if (ztp1 < c.rtt - R(5.) && qxfg_l > rlmin) {
          const R zsupersatice = m_div(zqsliq - zqsice, zqsice);
          const R zicenuclei = R(1000.0) * m_exp(R(12.96) * zsupersatice - R(0.639));
          const R t = ztp1;
          const R zadd = m_div(c.rlstt * (m_div(c.rlstt, c.rv * t) - R(1)), R(2.4E-2) * t);
          const R zbdd = m_div(c.rv * t * pap, R(2.21) * zvpice);
          const R zcvds = m_div(R(7.8) * m_pow(m_div(zicenuclei, zrho), R(0.666)) * zsupersatice, R(8.87) * (zadd + zbdd));
          const R zice0 = m_max(zicecld, m_div(zicenuclei * c.riceinit, zrho));
          const R zinew = m_pow(R(0.666) * zcvds * ptsphy + m_pow(zice0, R(0.666)), R(1.5));
          RD zdepos = m_max(c.rcl_overlapliqice * za * (zinew - zice0), R(0));
          zdepos = m_min(zdepos, qxfg_l);
          const R zinfactor = m_min(m_div(zicenuclei, R(15000.)), R(1));
          zdepos = zdepos * m_min(zinfactor + (R(1) - zinfactor) * (c.rdepliqrefrate + m_div(zcldtopdist, c.rdepliqrefdepth)), R(1));
          il = il + zdepos;
          qxfg_i = qxfg_i + zdepos;
          qxfg_l = qxfg_l - zdepos;
          bud.set(L11, -zdepos * zqtmst);  // (:2495-2496)
          bud.set(I11, zdepos * zqtmst);
        }
      }

      // 4.7 deposition onto snow, IDEPSNOW = 1 (:2700-2745)
      {
        const R zsupersatice = m_max(m_div(qv - zqsmix, zqsmix), R(0));
        if (ztp1 < c.rtt && zsnowcld > zepsec && zsupersatice > R(0)) {
          const R t = ztp1;
          const R zaplusb = c.rcl_apb1 * zvpice - c.rcl_apb2 * zvpice * t + pap * c.rcl_apb3 * (t * t * t);
          const R zcorrfac = m_sqrt(m_div(R(1), zrho));
          const R zcorrfac2 = m_pow(m_div(t, R(273.0)), R(1.5)) * m_div(R(393.0), t + R(120.0));
          const R zpr02 = zrho * zsnowcld * c.rcl_const1s;
          const R zterm1 = m_div(zsupersatice * (t * t) * zvpice * zcorrfac2 * c.rcl_const2s, zrho * zaplusb);
          const R zterm2 = R(0.65) * c.rcl_const6s * m_pow(zpr02, c.rcl_const4s) +
                           m_div(c.rcl_const3s * m_sqrt(zcorrfac) * m_sqrt(zrho) * m_pow(zpr02, c.rcl_const5s), m_sqrt(zcorrfac2));
          R zpsdep = m_max(zcovptot * zterm1 * zterm2 * ptsphy, R(0));
          zpsdep = m_div(m_min(zpsdep, qv - zqsmix), zcorqsmix);
          vs = vs - zpsdep;
          bud.set(PSDEP, zpsdep);  // ZSOLQA(S,V) += ZPSDEP, ZSOLQA(V,S) -= ZPSDEP
        }
      }

      // revised in-cloud condensate (:2750-2754)
      zliqcld = R(qxfg_l * ztmpa);
      zicecld = R(qxfg_i * ztmpa);

      // 4.8 ice -> snow autoconversion (:2769-2800)
      if (ztp1 <= c.rtt && zicecld > zepsec) {
        const R zzco = ptsphy * c.rsnowlin1 * m_exp(c.rsnowlin2 * (ztp1 - c.rtt));
        const R r = m_div(zicecld, c.rlcritsnow);
        const R ztemp = R(-1.) * (r * r);
        zsnowaut = (ztemp < R(50.0)) ? zzco * (R(1) - m_exp(ztemp)) : zzco;
        qbsi = qbsi + zsnowaut;
      }

      // 4.12 melting of ice and snow (:2815-2862)
      {
        const R zicetot = R(qxfg_i + qxfg_s);
        R zmeltmax = R(0);
        if (zicetot > zepsec && ztp1 > c.rtt) {
          const R zsubsat = m_max(zqsice - qv, R(0));
          const R ztdmtw0 = ztp1 - c.rtt - zsubsat * (R(1329.31) + R(0.0074615) * (pap - R(0.85E5)) - R(40.637) * (ztp1 - R(275.0)));
          const R zcons1 = m_abs(m_div(ptsphy * (R(1) + R(0.5) * ztdmtw0), c.rtaumel));
          zmeltmax = m_max(ztdmtw0 * zcons1 * c.zrldcp, R(0));
        }
        if (zmeltmax > zepsec && zicetot > zepsec) {
          R zalfa = R(m_div(qxfg_i, RD(zicetot)));
          RD zmelt = m_min(qxfg_i, RD(zalfa * zmeltmax));
          qxfg_i = qxfg_i - zmelt;
          qxfg_r = qxfg_r + zmelt;
          ri = ri + zmelt;  // ZSOLQA(R,I) += melt
          bud.set(I15, -zmelt * zqtmst);  // (:2856-2858)
          bud.set(L17, zmelt * zqtmst);
          zalfa = R(m_div(qxfg_s, RD(zicetot)));
          zmelt = m_min(qxfg_s, RD(zalfa * zmeltmax));
          qxfg_s = qxfg_s - zmelt;
          qxfg_r = qxfg_r + zmelt;
          rs = rs + zmelt;  // ZSOLQA(R,S) += melt
          bud.set(I16, -zmelt * zqtmst);
        }
      }

      // 4.13 freezing of rain (:2877-2933)
      if (qr > zepsec) {
        if (ztp1 <= c.rtt && t_prev > c.rtt) {
          const R zqp = m_max(qs + qr, zepsec);
          prainfrac = m_share(qr, zqp);
          llrainliq = prainfrac > R(0.8f);
        }
        if (ztp1 < c.rtt) {
          R zfrzmax;
          if (llrainliq) {
            const R zlambda = m_pow(m_div(c.rcl_lam1r, zrho * qr), c.rcl_lam2r);
            const R ztemp = m_min(c.rcl_fzrab * (ztp1 - c.rtt), R(50.));
            const R zfrz = ptsphy * m_div(c.rcl_const5r, zrho) * (m_exp(ztemp) - R(1.)) * m_pow(zlambda, c.rcl_const6r);
            zfrzmax = m_max(zfrz, R(0));
          } else {
            const R zcons1 = m_abs(m_div(ptsphy * (R(1) + R(0.5) * (c.rtt - ztp1)), c.rtaumel));
            zfrzmax = m_max((c.rtt - ztp1) * zcons1 * c.zrldcp, R(0));
          }
          if (zfrzmax > zepsec) {
            const R zfrz = m_min(qr, zfrzmax);
            rs = rs - zfrz;  // ZSOLQA(S,R) += frz
            bud.set(L18, -zfrz * zqtmst);  // (:2927)
          }
        }
      }

      // 4.14 homogeneous freezing of cloud liquid (:2948-2960)
      {
        const R zfrzmax = m_max((c.rthomo - ztp1) * c.zrldcp, R(0));
        if (zfrzmax > zepsec && qxfg_l > zepsec) {
          const RD zfrz = m_min(qxfg_l, RD(zfrzmax));
          il = il + zfrz;
          bud.set(L19, -zfrz * zqtmst);  // (:2958)
        }
      }

      // 4.9 riming of snow (:2975-3017)
      if (ztp1 <= c.rtt && zliqcld > zepsec) {
        if (zsnowcld > zepsec && zcovptot > R(0.01)) {
          const R zfallcorr3 = m_pow(m_div(c.rdensref, zrho), R(0.4f));
          zsnowrime = c.rcl_effrime * zcovptot * ptsphy * c.rcl_const7s * zfallcorr3 *
                        m_pow(zrho * zsnowcld * c.rcl_const1s, c.rcl_const8s);
          zsnowrime = m_min(zsnowrime, R(1));
          qbsl = qbsl + zsnowrime;
        }
      }

      // PSD slopes, PSACR and PIACR (:3029-3198)
      if (ztp1 <= c.rtt && prainfrac < zepsec && zraincld > zepsec) {
        const R zlamr = m_pow(m_div(c.rcl_lam1r, zrho * zraincld), c.rcl_lam2r);
        const R zn0r = c.rcl_x1r * m_pow(zlamr, c.rcl_x2r);
        const R lr3 = zlamr * zlamr * zlamr;
        if (zsnowcld > zepsec) {
          const R zlams = m_pow(m_div(c.rcl_lam1s, zrho * zsnowcld), c.rcl_lam2s);
          const R zfalldiff = m_max(m_abs(vfall_r - vfall_s), (vfall_r + vfall_s) / R(8.));
          const R zfac = m_div(c.rpi * c.rpi * c.rdenswat * R(1.0) * zrho * zn0r * c.rcl_x1s, zlams * lr3);
          const R zrat = m_div(zlamr, zlams);
          const R zpsacr = m_div(zcovptot * ptsphy * zfac * zfalldiff * ((R(0.5) * zrat + R(2.)) * zrat + R(5.)), lr3);
          rs = rs - zpsacr;  // ZSOLQA(S,R) += psacr
          bud.set(PSACR, zpsacr);
        }
        if (zicecld > zepsec) {
          const R zlami = m_pow(m_div(c.rcl_lam1s, zrho * zicecld), c.rcl_lam2s);
          const R zfalldiff = m_max(m_abs(vfall_r - vfall_i), (vfall_r + vfall_i) / R(8.));
          const R zfac = m_div(c.rpi * c.rpi * c.rdenswat * R(0.1) * zrho * zn0r * c.rcl_x1s, zlami * lr3);
          const R zrat = m_div(zlamr, zlami);
          const R zpiacr = m_div(zcovptot * ptsphy * zfac * zfalldiff * ((R(0.5) * zrat + R(2.)) * zrat + R(5.)), lr3);
          ri = ri - zpiacr;  // ZSOLQA(I,R) += piacr
          bud.set(PIACR, zpiacr);
        }
      }

      // 4.15 snow sublimation, ISUBLSNOW = 1 (:3254-3318)
      {
        R zzrh = c.rprecrhmax + m_div((R(1) - c.rprecrhmax) * zcovpmax, m_max(zepsec, R(1) - za));
        zzrh = m_min(m_max(zzrh, c.rprecrhmax), R(1));
        R zqe = m_div(qv - za * zqsice, m_max(zepsec, R(1) - za));
        zqe = m_max(R(0), m_min(zqe, zqsice));
        if (zcovpclr > zepsec && qxfg_s > zepsec && zqe < zzrh * zqsice) {
          const R dd = zcovptot * zdtgdp;
          const R den = m_max(m_abs(dd), c.zepsilon);
          const R zpreclr = R(m_div(qxfg_s * RD(zcovpclr), RD(dd < R(0) ? -den : den)));
          const R zbeta1 = m_div(m_div(m_sqrt(m_div(pap, paph_surf)), c.rvrfactor) * zpreclr, m_max(zcovpclr, zepsec));
          const R zbeta = c.rg * c.rpecons * m_pow(zbeta1, R(0.5777));
          const R zdenom = R(1) + zbeta * ptsphy * zcorqsice;
          const R zdpr = m_div(zcovpclr * zbeta * (zqsice - zqe), zdenom) * zdp * c.zrg_r;
          const R zdpevap = zdpr * zdtgdp;
          const RD zevap = m_min(RD(zdpevap), qxfg_s);
          vs = vs + zevap;
          bud.set(I17, -zevap * zqtmst);  // (:3304)
          zcovptot = m_max(c.rcovpmin, zcovptot - m_max(R(0), R(m_div(RD(zcovptot - za) * zevap, qxfg_s))));
          qxfg_s = qxfg_s - zevap;
        }
      }

      // evaporate small precipitation amounts (:3402-3412)
      if (qxfg_i < rlmin) vi = vi + qxfg_i;
      if (qxfg_r < rlmin) vr = vr + qxfg_r;
      if (qxfg_s < rlmin) vs = vs + qxfg_s;

      // 5.1 cloud cover (:3429-3438)
      R zanew = m_div(za + zsolac, R(1) + zsolab);
      zanew = m_min(zanew, R(1));
      if (zanew < c.ramin) zanew = R(0);
      const R zda = zanew - zaorig;
      zanewm1 = zanew;

      // 5.2.1 truncation of explicit sinks (:3455-3536)
      {
        const RD qxs[5] = {ql, qi, qr, qs, qv};
        truncate_sinks(dl, di, dr, ds, il, ri, rs, vl, vi, vr, vs, qxs, RD(zepsec),
                       [](RD a, RD b) { return m_div(a, b); }, [](RD a, RD b) { return m_max(a, b); });
      }

      // 5.2.2 implicit solve (:3546-3628); ZQLHS is lower triangular here.  The right-hand side -- the level's
      // water budget -- is summed in RD and the elimination runs in R: a species removed completely by a limited
      // sink has a right-hand side of RD size, which the ZEPSEC clip below removes.
      R qxn_v_solve;
      {
        const R one = R(1);
        const R dL = ((one + qbll) + qbrl) + qbsl;
        const R dI = ((one + fsink_i) + qbii) + qbsi;
        const R dR = one + fsink_r;
        const R dS = one + fsink_s;
        const R xL = R(RD(ql) + ((dl - il) - vl));
        const R xI = R(RD(qi) + (((il + di) - ri) - vi));
        const R xR = R(RD(qr) + (((ri + dr) + rs) - vr));
        const R xS = R(RD(qs) + ((-rs + ds) - vs));
        const R xV = R(RD(qv) + (((vl + vi) + vr) + vs));
        const R lRL = m_div(-qbrl, dL), lSL = m_div(-qbsl, dL), lSI = m_div(-qbsi, dI);
        const R yR = xR - lRL * xL;
        const R yS = (xS - lSL * xL) - lSI * xI;
        qxn_l = m_div(xL, dL);
        qxn_i = m_div(xI, dI);
        qxn_r = m_div(yR, dR);
        qxn_s = m_div(yS, dS);
        qxn_v_solve = xV;  // row V of ZQLHS is the identity
      }
      R qxn_v = qxn_v_solve;
      // clip small values (:3633-3640)
      if (qxn_l < zepsec) { qxn_v = qxn_v + qxn_l; qxn_l = R(0); }
      if (qxn_i < zepsec) { qxn_v = qxn_v + qxn_i; qxn_i = R(0); }
      if (qxn_r < zepsec) { qxn_v = qxn_v + qxn_r; qxn_r = R(0); }
      if (qxn_s < zepsec) { qxn_v = qxn_v + qxn_s; qxn_s = R(0); }
      qxnm1_l = qxn_l;
      qxnm1_i = qxn_i;
      qxnm1_r = qxn_r;
      qxnm1_s = qxn_s;
      // diagnosed rates of the implicit processes (:3674-3698), IWARMRAIN = 3 without rain freezing
      bud.set(L14, -zrainaut * qxn_l * zqtmst);
      bud.set(L15, -zrainacc * qxn_l * zqtmst);
      bud.set(CC6, -csink * zanewm1 * zqtmst);
      bud.set(L6, -csink * qxn_l * zqtmst);
      bud.set(I6, -csink * qxn_i * zqtmst);
      bud.set(I13, -fsink_i * qxn_i * zqtmst);
      bud.set(L21, -zsnowrime * qxn_l * zqtmst);
      bud.set(I14, -zsnowaut * qxn_i * zqtmst);

      // 5.3 fluxes to the next level (:3658-3672)
      pf_i = fsink_i * qxn_i * zrdtgdp;
      pf_r = fsink_r * qxn_r * zrdtgdp;
      pf_s = fsink_s * qxn_s * zrdtgdp;
      if (pf_s + pf_r < zepsec) zcovptot = R(0);

      // 6.1-6.3 tendencies (:3713-3758)
      {
        const R fq_l = csrc_l - csink * qxn_l;
        const R fq_i = (csrc_i + fsrc_i) - (fsink_i + csink) * qxn_i;
        const R fq_r = (csrc_r + fsrc_r) - fsink_r * qxn_r;
        const R fq_s = (csrc_s + fsrc_s) - fsink_s * qxn_s;
        tt = tt + c.ralvdcp * (qxn_l - ql - fq_l) * zqtmst;
        tcld_l = (qxn_l - ql0) * zqtmst;
        tt = tt + c.ralsdcp * (qxn_i - qi - fq_i) * zqtmst;
        tcld_i = (qxn_i - qi0) * zqtmst;
        tt = tt + c.ralvdcp * (qxn_r - qr - fq_r) * zqtmst;
        tcld_r = (qxn_r - qr0) * zqtmst;
        tt = tt + c.ralsdcp * (qxn_s - qs - fq_s) * zqtmst;
        tcld_s = (qxn_s - qs0) * zqtmst;
        tq = tq + (qxn_v - qv) * zqtmst;
        ta = zda * zqtmst;
      }

      // 6.4 CLOUD_SUPERSATCHECK on the updated state (:3765-3777), in R
      {
        const R zq_upd = pq + ptsphy * (tq + tcml_q);
        const R zt_upd = pt + ptsphy * (tt + tcml_t);
        const R za_upd = pa + ptsphy * (ta + tcml_a);
        const R pp = ssc_off ? LD(p.pap, i2 + (uint32_t)ssc_off * E, jo) : pap;
        R t_adj, q_adj, a_adj, l_adj, i_adj;
        supersatcheck<R, KC>(c, zt_upd, zq_upd, za_upd, pp, t_adj, q_adj, a_adj, l_adj, i_adj);
        tq = tq + R(q_adj) * zqtmst;
        tt = tt + R(t_adj) * zqtmst;
        ta = ta + R(a_adj) * zqtmst;
        tcld_l = tcld_l + R(l_adj) * zqtmst;
        tcld_i = tcld_i + R(i_adj) * zqtmst;
        bud.set(CC2, R(a_adj) * zqtmst);  // (:3787-3789)
        bud.set(L2, R(l_adj) * zqtmst);
        bud.set(I2, R(i_adj) * zqtmst);
      }
      // negative clean-up (:3793-3807)
      {
        const R zl = ql0 + tcld_l * ptsphy;
        const R zi0 = qi0 + tcld_i * ptsphy;
        if (d.nwarn) count_warnings(d.nwarn, zl < -rlmin, zi0 < -rlmin);
        if (zl < R(0)) {
          tcld_l = tcld_l - zl * zqtmst;
          tq = tq + zl * zqtmst;
          tt = tt - c.ralvdcp * zl * zqtmst;
        }
        const R zi = qi0 + tcld_i * ptsphy;
        if (zi < R(0)) {
          tcld_i = tcld_i - zi * zqtmst;
          tq = tq + zi * zqtmst;
          tt = tt - c.ralsdcp * zi * zqtmst;
        }
      }
      ST(p.pcovptot, i2, jo, zcovptot);  // (:3814-3816)
      if constexpr (BUD) {
        if (bp.fsd) {
          // cloud fraction at the end of the step with the precipitation fraction (:3818-3822); ZDA = ZANEW - ZAORIG
          const R a_end = pa + (zanewm1 - zaorig);
          if (a_end > R(0.001)) fsd_anewp = m_max(m_max(a_end, zcovptot), R(0.1));
          // detrained share of the condensate (:4112-4121), from PLUDE as rescaled above
          if (ql0 + qi0 + plude > rlmin) fsd_rat = m_max(R(0), m_min(R(1), m_share(plude, plude + ql0 + qi0)));
        }
      }
      // section 7 (:3826-4101), only when a budget writes PEXTRA: otherwise the budget inputs may be NULL (a run
      // with only LRAD_CLOUD_INHOMOG)
      if constexpr (BUD) {
        if (bp.flags & BUD_OUT) {
          const R zdz = m_div(zdp, zrho * c.rg);
          if constexpr (sizeof(R) == 8) budget_prefetch(bud_pre, bp, ext, k + 1);
          budget_write<R>(bud, vacc, bud_pre, bp, ext, k + 1, i2 + jo, zqtmst, ptsphy, zdz, qxn_l, qxn_i, c.ralvdcp,
                          c.ralsdcp, bp.ralfdcp);
        }
      }
    }

    if constexpr (BUD) {
      if (bp.fsd) {  // PFSD of every level (:4307-4359); ZQX and ZTP1 are the tidied start-of-step state
        const R zrho_f = m_div(pap, c.rd * ztp1);
        ST(bp.pfsd, i2, jo, cloud_fsd(bp.rcloud_frac_std, zgridlen, qv, qi, paph_k1 - paph_k, zrho_f, c.rg,
                                      fsd_anewp, fsd_rat, qxn_i, qxn_l, fsd_liq));
      }
    }

    // store tendencies of this level (levels above NCLDTOP keep the tidy-up terms only)
    ST(p.tloc_t, i2, jo, tt);
    ST(p.tloc_q, i2, jo, tq);
    ST(p.tloc_a, i2, jo, ta);
    ST(p.tloc_cld, i5, jo, tcld_l);
    ST(p.tloc_cld, i5 + sstep, jo, tcld_i);
    ST(p.tloc_cld, i5 + 2 * sstep, jo, tcld_r);
    ST(p.tloc_cld, i5 + 3 * sstep, jo, tcld_s);

    // DDH fluxes at the lower half level (:4244-4291)
    {
      const R zgdph_r = -c.zrg_r * (paph_k1 - paph_k) * zqtmst;
      const R sqrf = sqlf + (qxn_r - qr0) * zgdph_r;
      const R sqsf = sqif + (qxn_s - qs0) * zgdph_r;
      const R cqrng = cqlng + zlneg_r * zgdph_r;
      const R cqsng = cqnng + zlneg_s * zgdph_r;
      sqlf = sqlf + (qxn_l - ql0 + pvfl * ptsphy) * zgdph_r + pludeli1;
      cqlng = cqlng + zlneg_l * zgdph_r;
      sqltur = sqltur + pvfl * ptsphy * zgdph_r;
      sqif = sqif + (qxn_i - qi0 + pvfi * ptsphy) * zgdph_r + pludeli2;
      cqnng = cqnng + zlneg_i * zgdph_r;
      sqitur = sqitur + pvfi * ptsphy * zgdph_r;
      ST(p.pfsqlf, ih1, jo, sqlf);
      ST(p.pfsqif, ih1, jo, sqif);
      ST(p.pfsqrf, ih1, jo, sqrf);
      ST(p.pfsqsf, ih1, jo, sqsf);
      ST(p.pfcqlng, ih1, jo, cqlng);
      ST(p.pfcqnng, ih1, jo, cqnng);
      ST(p.pfcqrng, ih1, jo, cqrng);
      ST(p.pfcqsng, ih1, jo, cqsng);
      ST(p.pfsqltur, ih1, jo, sqltur);
      ST(p.pfsqitur, ih1, jo, sqitur);
    }

    // carry to the next level
    za_prev = za;
    qr_prev = qr;
    qs_prev = qs;
    t_prev = ztp1;
    pap_prev = pap;
    paph_k = paph_k1;
    pmfu_k = pmfu_n;
    pmfd_k = pmfd_n;
  }
  if constexpr (BUD) {
    if (bp.flags & BUD_VERTINT) vertint_end(vacc, bp, ext, prainfrac, ktype, peis, u1i + joi, (int)klev - ncldtop0);
  }
  // precipitation fluxes at the surface half level
  {
    const uint32_t ih = bh + klev * lstep;
    const R fl = pf_r + R(0), fn = pf_s + pf_i;
    ST(p.pfplsl, ih, jo, fl);
    ST(p.pfplsn, ih, jo, fn);
    ST(p.pfhpsl, ih, jo, -c.rlvtt * fl);
    ST(p.pfhpsn, ih, jo, -c.rlstt * fn);
  }
  ST(p.prainfrac_toprfz, u1, jo, prainfrac);
}

// ---------------------------------------------------------------------------
// Host-side launchers (declared in cloudsc_launch.h)
// ---------------------------------------------------------------------------
template <typename T>
KConst<T> make_kconst(const cloudsc_params_t& q, double ptsphy_in) {
  KConst<T> c{};
  const T ptsphy = T(ptsphy_in);
  c.rg = T(q.rg); c.rd = T(q.rd); c.rv = T(q.rv); c.retv = T(q.retv); c.rtt = T(q.rtt);
  c.rlvtt = T(q.rlvtt); c.rlstt = T(q.rlstt); c.rpi = T(q.rpi);
  c.rv_rd = c.rv / c.rd;
  c.r2es = T(q.r2es); c.r3les = T(q.r3les); c.r3ies = T(q.r3ies); c.r4les = T(q.r4les); c.r4ies = T(q.r4ies);
  c.r5les = T(q.r5les); c.r5ies = T(q.r5ies); c.r5alvcp = T(q.r5alvcp); c.r5alscp = T(q.r5alscp);
  c.ralvdcp = T(q.ralvdcp); c.ralsdcp = T(q.ralsdcp); c.rtwat = T(q.rtwat); c.rtice = T(q.rtice);
  c.rtwat_rtice_r = T(q.rtwat_rtice_r); c.rkoop1 = T(q.rkoop1); c.rkoop2 = T(q.rkoop2);
  c.ramid = T(q.ramid); c.ramin = T(q.ramin); c.rlmin = T(q.rlmin); c.rcldiff = T(q.rcldiff);
  c.rcldiff_convi = T(q.rcldiff_convi); c.rkconv = T(q.rkconv);
  c.rsnowlin1 = T(q.rsnowlin1); c.rsnowlin2 = T(q.rsnowlin2); c.rlcritsnow = T(q.rlcritsnow);
  c.rvice = T(q.rvice); c.rvrain = T(q.rvrain); c.rvsnow = T(q.rvsnow); c.rxfallspeedrain = T(q.rxfallspeedrain);
  c.rcldtopcf = T(q.rcldtopcf); c.rdepliqrefrate = T(q.rdepliqrefrate); c.rdepliqrefdepth = T(q.rdepliqrefdepth);
  c.rcl_overlapliqice = T(q.rcl_overlapliqice); c.rcl_effrime = T(q.rcl_effrime); c.riceinit = T(q.riceinit);
  c.rthomo = T(q.rthomo); c.rprecrhmax = T(q.rprecrhmax); c.rpecons = T(q.rpecons); c.rvrfactor = T(q.rvrfactor);
  c.rcovpmin = T(q.rcovpmin); c.rtaumel = T(q.rtaumel); c.rssicefactor = T(q.rssicefactor);
  c.rdenswat = T(q.rdenswat); c.rdensref = T(q.rdensref);
  c.rcl_dr = T(q.rcl_dr); c.rcl_x1r = T(q.rcl_x1r); c.rcl_x2r = T(q.rcl_x2r); c.rcl_x1s = T(q.rcl_x1s);
  c.rcl_lam1r = T(q.rcl_lam1r); c.rcl_lam2r = T(q.rcl_lam2r); c.rcl_lam1s = T(q.rcl_lam1s); c.rcl_lam2s = T(q.rcl_lam2s);
  c.rcl_cdenom1 = T(q.rcl_cdenom1); c.rcl_cdenom2 = T(q.rcl_cdenom2); c.rcl_cdenom3 = T(q.rcl_cdenom3);
  c.rcl_const1r = T(q.rcl_const1r); c.rcl_const2r = T(q.rcl_const2r); c.rcl_const3r = T(q.rcl_const3r);
  c.rcl_const4r = T(q.rcl_const4r); c.rcl_const5r = T(q.rcl_const5r); c.rcl_const6r = T(q.rcl_const6r);
  c.rcl_const7r = T(q.rcl_const7r); c.rcl_kkaac = T(q.rcl_kkaac); c.rcl_kkbac = T(q.rcl_kkbac);
  c.rcl_kkaau = T(q.rcl_kkaau);
  c.rcl_kkbauq = T(q.rcl_kkbauq); c.rcl_kkbaun = T(q.rcl_kkbaun);
  c.rcl_kk_cloud_num_sea = T(q.rcl_kk_cloud_num_sea); c.rcl_kk_cloud_num_land = T(q.rcl_kk_cloud_num_land);
  c.rcl_fzrab = T(q.rcl_fzrab); c.rcl_apb1 = T(q.rcl_apb1); c.rcl_apb2 = T(q.rcl_apb2); c.rcl_apb3 = T(q.rcl_apb3);
  c.rcl_const1s = T(q.rcl_const1s); c.rcl_const2s = T(q.rcl_const2s); c.rcl_const3s = T(q.rcl_const3s);
  c.rcl_const4s = T(q.rcl_const4s); c.rcl_const5s = T(q.rcl_const5s); c.rcl_const6s = T(q.rcl_const6s);
  c.rcl_const7s = T(q.rcl_const7s); c.rcl_const8s = T(q.rcl_const8s);
  c.rmfadvw = T(q.rmfadvw); c.rmfadvwdd = T(q.rmfadvwdd);
  c.ptsphy = ptsphy;
  c.zqtmst = T(1) / ptsphy;                   // :798
  c.zrdcp = c.rd / T(q.rcpd);                 // :800
  c.zrg_r = T(1) / c.rg;                      // :803
  c.zrldcp = T(1) / (c.ralsdcp - c.ralvdcp);  // :804
  c.zepsilon = T(100) * std::numeric_limits<T>::epsilon();  // :693
  c.rdtg_r = T(1) / (ptsphy * c.rg);          // :1285
  c.c_aut = T(q.rcl_inhomogaut) * ptsphy * T(q.rcl_kkaau);
  c.c_acc = T(q.rcl_inhomogacc) * ptsphy * T(q.rcl_kkaac);
  c.ncldtop = q.ncldtop;
  c.lmfdsnow = q.lmfdsnow;
  return c;
}

template <typename T>
KConstDefault<T> make_kconst_default(const cloudsc_params_t& q, double ptsphy_in) {
  KConstDefault<T> c{};
  const T ptsphy = T(ptsphy_in);
  c.ptsphy = ptsphy;
  c.zqtmst = T(1) / ptsphy;
  c.rdtg_r = T(1) / (ptsphy * KConstDefault<T>::rg);
  c.c_aut = T(cloudsc_defaults::rcl_inhomogaut) * ptsphy * T(cloudsc_defaults::rcl_kkaau);
  c.c_acc = T(cloudsc_defaults::rcl_inhomogacc) * ptsphy * T(cloudsc_defaults::rcl_kkaac);
  c.rssicefactor = T(q.rssicefactor);
  c.ncldtop = q.ncldtop;
  return c;
}

// True when every configuration value except RSSICEFACTOR (time step) and NCLDTOP (levels)
// equals the default configuration the specialised kernels are compiled with.
bool is_default_config(const cloudsc_params_t& q) {
  using namespace cloudsc_defaults;
#define CLOUDSC_SAME(name) (q.name == name) &&
  return CLOUDSC_SAME(rg) CLOUDSC_SAME(rd) CLOUDSC_SAME(rv) CLOUDSC_SAME(rcpd) CLOUDSC_SAME(retv) CLOUDSC_SAME(rtt)
      CLOUDSC_SAME(rlvtt) CLOUDSC_SAME(rlstt) CLOUDSC_SAME(rpi) CLOUDSC_SAME(r2es) CLOUDSC_SAME(r3les)
      CLOUDSC_SAME(r3ies) CLOUDSC_SAME(r4les) CLOUDSC_SAME(r4ies) CLOUDSC_SAME(r5les) CLOUDSC_SAME(r5ies)
      CLOUDSC_SAME(r5alvcp) CLOUDSC_SAME(r5alscp) CLOUDSC_SAME(ralvdcp) CLOUDSC_SAME(ralsdcp) CLOUDSC_SAME(rtwat)
      CLOUDSC_SAME(rtice) CLOUDSC_SAME(rtwat_rtice_r) CLOUDSC_SAME(rkoop1) CLOUDSC_SAME(rkoop2) CLOUDSC_SAME(ramid)
      CLOUDSC_SAME(ramin) CLOUDSC_SAME(rlmin) CLOUDSC_SAME(rcldiff) CLOUDSC_SAME(rcldiff_convi)
      CLOUDSC_SAME(rcl_inhomogaut) CLOUDSC_SAME(rcl_inhomogacc) CLOUDSC_SAME(rsnowlin1) CLOUDSC_SAME(rsnowlin2)
      CLOUDSC_SAME(rlcritsnow) CLOUDSC_SAME(rvice) CLOUDSC_SAME(rvrain) CLOUDSC_SAME(rvsnow)
      CLOUDSC_SAME(rxfallspeedrain) CLOUDSC_SAME(rcldtopcf) CLOUDSC_SAME(rdepliqrefrate)
      CLOUDSC_SAME(rdepliqrefdepth) CLOUDSC_SAME(rcl_overlapliqice) CLOUDSC_SAME(rcl_effrime) CLOUDSC_SAME(riceinit)
      CLOUDSC_SAME(rthomo) CLOUDSC_SAME(rprecrhmax) CLOUDSC_SAME(rpecons) CLOUDSC_SAME(rvrfactor)
      CLOUDSC_SAME(rcovpmin) CLOUDSC_SAME(rtaumel) CLOUDSC_SAME(rdenswat) CLOUDSC_SAME(rdensref) CLOUDSC_SAME(rcl_dr)
      CLOUDSC_SAME(rcl_x1r) CLOUDSC_SAME(rcl_x2r) CLOUDSC_SAME(rcl_x1s) CLOUDSC_SAME(rcl_lam1r)
      CLOUDSC_SAME(rcl_lam2r) CLOUDSC_SAME(rcl_lam1s) CLOUDSC_SAME(rcl_lam2s) CLOUDSC_SAME(rcl_cdenom1)
      CLOUDSC_SAME(rcl_cdenom2) CLOUDSC_SAME(rcl_cdenom3) CLOUDSC_SAME(rcl_const1r) CLOUDSC_SAME(rcl_const2r)
      CLOUDSC_SAME(rcl_const3r) CLOUDSC_SAME(rcl_const4r) CLOUDSC_SAME(rcl_const5r) CLOUDSC_SAME(rcl_const6r)
      CLOUDSC_SAME(rcl_const7r) CLOUDSC_SAME(rcl_kkaac) CLOUDSC_SAME(rcl_kkbac) CLOUDSC_SAME(rcl_kkaau)
      CLOUDSC_SAME(rcl_kkbauq) CLOUDSC_SAME(rcl_kkbaun) CLOUDSC_SAME(rcl_kk_cloud_num_sea)
      CLOUDSC_SAME(rcl_kk_cloud_num_land) CLOUDSC_SAME(rcl_fzrab) CLOUDSC_SAME(rcl_apb1) CLOUDSC_SAME(rcl_apb2)
      CLOUDSC_SAME(rcl_apb3) CLOUDSC_SAME(rcl_const1s) CLOUDSC_SAME(rcl_const2s) CLOUDSC_SAME(rcl_const3s)
      CLOUDSC_SAME(rcl_const4s) CLOUDSC_SAME(rcl_const5s) CLOUDSC_SAME(rcl_const6s) CLOUDSC_SAME(rcl_const7s)
      CLOUDSC_SAME(rcl_const8s) CLOUDSC_SAME(rmfadvw) CLOUDSC_SAME(rmfadvwdd) (q.lmfdsnow == lmfdsnow);
#undef CLOUDSC_SAME
}
template <typename R>
KPtrs<R> offset_ptrs(const KPtrs<R>& p, int64_t b0, int64_t nproma, int64_t klev) {
  const int64_t s1 = nproma, s2 = nproma * klev, sh = nproma * (klev + 1);
  KPtrs<R> q = p;
  q.pt += b0 * s2; q.pq += b0 * s2; q.tcml_t += b0 * s2; q.tcml_q += b0 * s2; q.tcml_a += b0 * s2;
  q.tcml_cld += b0 * s2 * 5; q.tloc_t += b0 * s2; q.tloc_q += b0 * s2; q.tloc_a += b0 * s2;
  q.tloc_cld += b0 * s2 * 5; q.pvfl += b0 * s2; q.pvfi += b0 * s2; q.pap += b0 * s2; q.paph += b0 * sh;
  q.plsm += b0 * s1; q.ldcum += b0 * s1; q.ktype += b0 * s1; q.peis += b0 * s1; q.plu += b0 * s2;
  q.plude += b0 * s2; q.pludeli += b0 * s2 * 4; q.psnde += b0 * s2 * 2; q.pmfu += b0 * s2; q.pmfd += b0 * s2;
  q.pa += b0 * s2; q.pclv += b0 * s2 * 5; q.pcovptot += b0 * s2; q.prainfrac_toprfz += b0 * s1;
  R** h[] = {&q.pfsqlf, &q.pfsqif, &q.pfcqnng, &q.pfcqlng, &q.pfsqrf, &q.pfsqsf, &q.pfcqrng,
             &q.pfcqsng, &q.pfsqltur, &q.pfsqitur, &q.pfplsl, &q.pfplsn, &q.pfhpsl, &q.pfhpsn};
  for (R** x : h) *x += b0 * sh;
  return q;
}

// Launch configuration per (arithmetic variant, KLEV), set by the autotuner (cloudsc_tune_set_launch): work-group
// size, minimum waves per SIMD of the register allocation, and whether the level count is compiled in.  Default:
// 256 columns per group, the compiler's register allocation, run-time KLEV.
namespace {
std::mutex g_tune_mutex;
std::map<std::pair<int, int>, LaunchCfg> g_tuned;
}  // namespace

template <typename R, typename RD>
constexpr int variant_id() {
  return sizeof(R) == 8 ? 0 : (sizeof(RD) == 8 ? 2 : 1);  // 0 dp, 1 sp, 2 mixed
}

// Waves-per-SIMD bounds instantiated for each variant besides 0: dp holds its state in 256 registers (1 wave),
// sp in 137 (3 waves), mixed in 177 (2 waves); each gets the next occupancy steps its state fits with few spills.
bool valid_wpe(int variant, int wpe) {
  if (wpe == 0) return true;
  switch (variant) {
    case 0: return wpe == 2;
    case 1: return wpe == 4;
    case 2: return wpe == 3 || wpe == 4;
    default: return false;
  }
}

bool valid_launch(int variant, const LaunchCfg& l) {
  const bool blk = l.block == 64 || l.block == 128 || l.block == 256 || l.block == 512;
  if (!blk || !valid_wpe(variant, l.wpe) || (l.kspec != 0 && l.kspec != 1)) return false;
  // register bounds and compiled-in level counts are instantiated for 128- and 256-column groups only
  return (l.wpe == 0 && l.kspec == 0) || l.block == 128 || l.block == 256;
}

LaunchCfg tuned_launch(int variant, int klev) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  const auto it = g_tuned.find({variant, klev});
  return it == g_tuned.end() ? LaunchCfg{} : it->second;
}

void set_tuned_launch(int variant, int klev, const LaunchCfg* l) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  if (!l)
    g_tuned.erase({variant, klev});
  else
    g_tuned[{variant, klev}] = *l;
}

template <int BLOCK, int WPE, int KL, bool SAT, typename R, typename RD, typename KC, typename KCD>
hipError_t launch_one(unsigned ncols, const KC& c, const KCD& cd, const KPtrs<R>& p, const KDims& d, int ssc_off,
                      const KSatPtrs<R>& s, hipStream_t stream) {
  const KBudPtrs<R> nb{};
  hipLaunchKernelGGL((cloudsc_column_kernel<R, RD, KC, KCD, BLOCK, true, SAT, false, WPE, KL>),
                     dim3((ncols + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, c, cd, p, d, ssc_off, s, nb);
  return hipGetLastError();
}

// Level count of the tuned instantiation: the compiled-in 137 or 91 when asked for and matching, else run time.
template <int BLOCK, int WPE, bool SAT, typename R, typename RD, typename KC, typename KCD>
hipError_t launch_kl(int kspec, unsigned ncols, const KC& c, const KCD& cd, const KPtrs<R>& p, const KDims& d,
                     int ssc_off, const KSatPtrs<R>& s, hipStream_t stream) {
  if (kspec && d.klev == 137) return launch_one<BLOCK, WPE, 137, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
  if (kspec && d.klev == 91) return launch_one<BLOCK, WPE, 91, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
  return launch_one<BLOCK, WPE, 0, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
}

template <int BLOCK, bool SAT, typename R, typename RD, typename KC, typename KCD>
hipError_t launch_wpe(const LaunchCfg& l, unsigned ncols, const KC& c, const KCD& cd, const KPtrs<R>& p,
                      const KDims& d, int ssc_off, const KSatPtrs<R>& s, hipStream_t stream) {
  constexpr int v = variant_id<R, RD>();
  constexpr int w1 = v == 0 ? 2 : (v == 1 ? 4 : 3);
  constexpr int w2 = 4;  // mixed only
  if (l.wpe == w1) return launch_kl<BLOCK, w1, SAT, R, RD>(l.kspec, ncols, c, cd, p, d, ssc_off, s, stream);
  if (v == 2 && l.wpe == w2) return launch_kl<BLOCK, w2, SAT, R, RD>(l.kspec, ncols, c, cd, p, d, ssc_off, s, stream);
  return launch_kl<BLOCK, 0, SAT, R, RD>(l.kspec, ncols, c, cd, p, d, ssc_off, s, stream);
}

template <bool SAT, typename R, typename RD, typename KC, typename KCD>
hipError_t launch_variant(const LaunchCfg& l, bool uni, unsigned ncols, const KC& c, const KCD& cd,
                          const KPtrs<R>& p, const KDims& d, int ssc_off, const KSatPtrs<R>& s, hipStream_t stream) {
  const KBudPtrs<R> nb{};
  if (!uni) {  // blocks not aligned to waves: per-lane block offsets, 256-lane groups
    hipLaunchKernelGGL((cloudsc_column_kernel<R, RD, KC, KCD, 256, false, SAT, false>), dim3((ncols + 255) / 256),
                       dim3(256), 0, stream, c, cd, p, d, ssc_off, s, nb);
    return hipGetLastError();
  }
  // the tuned instantiations (register bound, compiled-in KLEV) exist for the default configuration of CLOUDSC
  constexpr bool tuned = !SAT && std::is_same<KC, KConstDefault<R>>::value;
  switch (l.block) {
    case 64:
      return launch_one<64, 0, 0, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
    case 128:
      if constexpr (tuned) return launch_wpe<128, SAT, R, RD>(l, ncols, c, cd, p, d, ssc_off, s, stream);
      return launch_one<128, 0, 0, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
    case 512:
      return launch_one<512, 0, 0, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
    default:
      if constexpr (tuned) return launch_wpe<256, SAT, R, RD>(l, ncols, c, cd, p, d, ssc_off, s, stream);
      return launch_one<256, 0, 0, SAT, R, RD>(ncols, c, cd, p, d, ssc_off, s, stream);
  }
}

template <typename R, typename RD, typename KC, typename KCD>
hipError_t launch_either(bool sat, const LaunchCfg& l, bool uni, unsigned ncols, const KC& c, const KCD& cd,
                         const KPtrs<R>& p, const KDims& d, int ssc_off, const KSatPtrs<R>& s,
                         const KBudPtrs<R>* bud, hipStream_t stream) {
  if (bud) {  // budget-instrumented kernel: one instantiation per variant (diagnostic runs); one wave per group
    hipLaunchKernelGGL((cloudsc_column_kernel<R, RD, KC, KCD, 64, false, false, true>), dim3((ncols + 63) / 64),
                       dim3(64), 0, stream, c, cd, p, d, ssc_off, s, *bud);
    return hipGetLastError();
  }
  return sat ? launch_variant<true, R, RD>(l, uni, ncols, c, cd, p, d, ssc_off, s, stream)
             : launch_variant<false, R, RD>(l, uni, ncols, c, cd, p, d, ssc_off, s, stream);
}

template <typename R>
KSatPtrs<R> offset_sat_ptrs(const KSatPtrs<R>& p, int64_t b0, int64_t nproma, int64_t klev) {
  const int64_t s2 = nproma * klev;
  KSatPtrs<R> q = p;
  q.phrsw += b0 * s2; q.phrlw += b0 * s2; q.tvdf_t += b0 * s2; q.tvdf_q += b0 * s2; q.pvervel += b0 * s2;
  q.tcml_t += b0 * s2; q.tcml_q += b0 * s2; q.tcml_a += b0 * s2; q.tcml_cld += b0 * s2 * 5;
  if (q.tsat_t) {
    q.tsat_t += b0 * s2; q.tsat_q += b0 * s2; q.tsat_a += b0 * s2; q.tsat_cld += b0 * s2 * 5;
  }
  return q;
}

template <typename R>
KBudPtrs<R> offset_bud_ptrs(const KBudPtrs<R>& p, int64_t b0, int64_t nproma, int64_t klev) {
  const int64_t s2 = nproma * klev;
  KBudPtrs<R> q = p;
  auto adv = [](auto*& x, int64_t n) {  // arrays of a diagnostic that is off are NULL
    if (x) x += n;
  };
  adv(q.pvfa, b0 * s2); adv(q.pdyna, b0 * s2); adv(q.pdynl, b0 * s2); adv(q.pdyni, b0 * s2);
  adv(q.pdyns, b0 * s2); adv(q.phrsw, b0 * s2); adv(q.phrlw, b0 * s2); adv(q.pvfl, b0 * s2); adv(q.pvfi, b0 * s2);
  adv(q.kpbltype, b0 * nproma); adv(q.kctop, b0 * nproma);
  adv(q.pextra, b0 * s2 * q.kfldx);
  adv(q.pgaw, b0 * nproma); adv(q.pfsd, b0 * s2);
  return q;
}

template <typename R, typename RD>
hipError_t launch_cloudsc(const cloudsc_params_t& q, double ptsphy, const KPtrs<R>& p, const KDims& d, int ssc_off,
                          hipStream_t stream, const KSatPtrs<R>* sat, const KBudPtrs<R>* bud) {
  const bool def = is_default_config(q);
  const LaunchCfg tl = def ? tuned_launch(variant_id<R, RD>(), d.klev) : LaunchCfg{};
  // every wave inside one block (64-lane waves, 64 | NPROMA): block offsets are uniform across the wave
  const bool uni = d.nproma % 64 == 0;
  // blocks per launch: the largest array, (NPROMA,KLEV,5) or PEXTRA (NPROMA,KLEV,KFLDX), stays below 4 GiB
  const int64_t nfld = bud ? std::max<int64_t>(5, bud->kfldx) : 5;
  const int64_t per_block = (int64_t)d.nproma * d.klev * nfld * (int64_t)sizeof(R);
  const int64_t max_blocks = std::max<int64_t>(1, ((int64_t)1 << 32) / per_block - 1);
  // launches of equal size: a short last launch would leave most of the GPU idle while it runs
  const int64_t nlaunch = (d.ngpblks + max_blocks - 1) / max_blocks;
  const int64_t step = (d.ngpblks + nlaunch - 1) / nlaunch;
  for (int64_t b0 = 0; b0 < d.ngpblks; b0 += step) {
    const int64_t nb = std::min<int64_t>(step, d.ngpblks - b0);
    KDims dc = d;
    dc.ngpblks = (int32_t)nb;
    dc.ngptot = (int32_t)(d.ngptot - b0 * d.nproma);
    const KPtrs<R> pc = offset_ptrs(p, b0, d.nproma, d.klev);
    const KSatPtrs<R> sc = sat ? offset_sat_ptrs(*sat, b0, d.nproma, d.klev) : KSatPtrs<R>{};
    const KBudPtrs<R> bc = bud ? offset_bud_ptrs(*bud, b0, d.nproma, d.klev) : KBudPtrs<R>{};
    const unsigned ncols = (unsigned)(nb * d.nproma);
    hipError_t e;
    if (def) {
      const KConstDefault<R> c = make_kconst_default<R>(q, ptsphy);
      const KConstDefault<RD> cd = make_kconst_default<RD>(q, ptsphy);
      e = launch_either<R, RD>(sat != nullptr, tl, uni, ncols, c, cd, pc, dc, ssc_off, sc, bud ? &bc : nullptr,
                               stream);
    } else {
      const KConst<R> c = make_kconst<R>(q, ptsphy);
      const KConst<RD> cd = make_kconst<RD>(q, ptsphy);
      e = launch_either<R, RD>(sat != nullptr, LaunchCfg{}, uni, ncols, c, cd, pc, dc, ssc_off, sc,
                               bud ? &bc : nullptr, stream);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template KConst<float> make_kconst<float>(const cloudsc_params_t&, double);
template KConst<double> make_kconst<double>(const cloudsc_params_t&, double);
template KConstDefault<float> make_kconst_default<float>(const cloudsc_params_t&, double);
template KConstDefault<double> make_kconst_default<double>(const cloudsc_params_t&, double);

template hipError_t launch_cloudsc<double, double>(const cloudsc_params_t&, double, const KPtrs<double>&,
                                                   const KDims&, int, hipStream_t, const KSatPtrs<double>*,
                                                   const KBudPtrs<double>*);
template hipError_t launch_cloudsc<float, float>(const cloudsc_params_t&, double, const KPtrs<float>&, const KDims&,
                                                 int, hipStream_t, const KSatPtrs<float>*, const KBudPtrs<float>*);
template hipError_t launch_cloudsc<float, double>(const cloudsc_params_t&, double, const KPtrs<float>&, const KDims&,
                                                  int, hipStream_t, const KSatPtrs<float>*, const KBudPtrs<float>*);

}  // namespace cloudsc_gpu
