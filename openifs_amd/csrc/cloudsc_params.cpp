// Default configuration of CLOUDSC: the constants SUPHEC, SUCLDP and their
// neighbours put into TCST/TTHF/TECLDP/TECUMF/TERAD/TVDF before the first
// physics call.  Host code; compiled into libcloudsc_hip.so.
//
// Sources restated (formulas, not code):
//   ifs-source/arpifs/module/yomcst.F90:508-629     fundamental constants
//   ifs-source/arpifs/phys_ec/suphec.F90:231-254    TTHF thermodynamic constants
//   ifs-source/arpifs/phys_ec/sucldp.F90:243-655    TECLDP defaults and PSD constants
//   ifs-source/arpifs/utility/fcgeneralized_gamma.F90:99-122  Lanczos Gamma
//   ifs-source/arpifs/phys_ec/sucumf.F90:203,228-229, suvdf.F90:135,
//   ifs-source/arpifs/phys_radi/suecrad.F90:745, setup/su0phy.F90:1298,1336,1343,
//   ifs-source/arpifs/setup/sudyncore.F90:86-90 (RPLRG = RPLDARE = 1)
#include <cmath>
#include <cstring>

#include "../../include/cloudsc_hip.h"

namespace {

// Gamma function by the 6-term Lanczos series of Numerical Recipes (2nd ed.,
// gammln), as FCGENERALIZED_GAMMA evaluates it for x > 0
// (fcgeneralized_gamma.F90:99-122).  The reference declares the series
// coefficients, the sqrt(2 pi) factor and the leading 1.000000000190015 with
// default-real (single precision) literals, so they enter its FP64 build
// rounded to FP32; the same rounding is applied here to keep the derived
// PSD constants identical to the reference's.
double lanczos_gamma(double px) {
  const double stp = (double)2.5066282746310005f;
  const double c1 = (double)76.18009172947146f, c2 = (double)-86.50532032941677f,
               c3 = (double)24.01409824083091f, c4 = (double)-1.231739572450155f,
               c5 = (double)0.1208650973866179E-2f, c6 = (double)-0.5395239384953E-5f;
  const double sercst = (double)1.000000000190015f;
  const double x = px;  // only positive arguments occur in SUCLDP
  const double t = x + 5.5;
  const double lt = (x + 0.5) * std::log(t) - t;
  const double ser = sercst + c1 / (1. + x) + c2 / (2. + x) + c3 / (3. + x) + c4 / (4. + x) +
                     c5 / (5. + x) + c6 / (6. + x);
  return std::exp(lt + std::log(stp * ser / x));
}

}  // namespace

extern "C" int cloudsc_params_default(cloudsc_params_t* p, double ptsphy, int nflevg,
                                      const double* stpre) {
  if (!p || !stpre || nflevg < 2 || !(ptsphy > 0.0)) return CLOUDSC_EDIMS;
  std::memset(p, 0, sizeof(*p));

  // ---- TCST: yomcst.F90:508-591 -------------------------------------------
  const double rpi = 2.0 * std::asin(1.0);
  const double rkbol = 1.380658e-23, rnavo = 6.0221367e+23;
  const double r = rnavo * rkbol;
  const double rmd = 28.9644, rmv = 18.0153;
  p->rpi = rpi;
  p->ra = 6371229.0;  // RA_EARTH * RPLRADI, RPLRADI = 1
  p->rg = 9.80665;
  p->rd = 1000.0 * r / rmd;
  p->rv = 1000.0 * r / rmv;
  p->rcpd = 3.5 * p->rd;
  p->retv = p->rv / p->rd - 1.0;
  p->rtt = 273.16;
  p->rlvtt = 2.5008e+6;
  p->rlstt = 2.8345e+6;
  p->rlmlt = p->rlstt - p->rlvtt;

  // ---- TTHF: suphec.F90:231-254 -------------------------------------------
  p->r2es = 611.21 * p->rd / p->rv;
  p->r3les = 17.502;
  p->r3ies = 22.587;
  p->r4les = 32.19;
  p->r4ies = -0.7;
  p->r5les = p->r3les * (p->rtt - p->r4les);
  p->r5ies = p->r3ies * (p->rtt - p->r4ies);
  p->r5alvcp = p->r5les * p->rlvtt / p->rcpd;
  p->r5alscp = p->r5ies * p->rlstt / p->rcpd;
  p->ralvdcp = p->rlvtt / p->rcpd;
  p->ralsdcp = p->rlstt / p->rcpd;
  p->ralfdcp = p->rlmlt / p->rcpd;
  p->rtwat = p->rtt;
  p->rtice = p->rtt - 23.0;
  p->rtwat_rtice_r = 1.0 / (p->rtwat - p->rtice);
  p->rkoop1 = 2.583;
  p->rkoop2 = 0.48116e-2;

  // ---- TECLDP: sucldp.F90 (RPLRG = RPLDARE = 1, so ZPLRG = 1) ---------------
  const double rplrg = 1.0, rpldare = 1.0;
  const double zplrg = rpldare * std::sqrt(rplrg);
  p->ramid = 0.8;
  p->rprc1 = 100.0;
  p->ramin = 1.e-8;
  p->rlmin = 1.e-8;
  p->rcldiff = 6.0e-6 * rplrg * rpldare;
  p->rcldiff_convi = 10.0;
  p->rclcrit_sea = 2.5e-4;
  p->rclcrit_land = 5.5e-4;
  p->rkconv = rplrg * rpldare / 6000.0;
  p->rcl_inhomogaut = 1.5;
  p->rcl_inhomogacc = 3.0;
  p->rsnowlin1 = rplrg * rpldare / 1000.0;
  p->rsnowlin2 = 0.030;
  p->rlcritsnow = 2.0e-5;
  p->rvice = 0.13 * zplrg;
  p->rvrain = 4.0 * zplrg;
  p->rvsnow = 1.0 * zplrg;
  p->rxfallspeedrain = 1.0;
  p->rxfallspeedice = 1.0;
  p->rnice = 0.027;
  p->rccn = 125.0;
  p->rcldtopcf = 0.01;
  p->rdepliqrefrate = 0.5;
  p->rdepliqrefdepth = 500.0 / rplrg;
  p->rcl_overlapliqice = 0.65;
  p->rcl_effrime = 1.0;
  p->riceinit = 1.e-12;
  p->rthomo = p->rtt - 38.0;
  p->rprecrhmax = 0.7;
  p->rpecons = 5.44e-4 * rplrg * rpldare / p->rg;
  p->rvrfactor = 5.09e-3 * zplrg;
  p->rcovpmin = 0.1;
  p->rtaumel = 2.0 * 3.6e3 / (rplrg * rpldare);
  p->rkooptau = 3.0 * 3.6e3 / (rplrg * rpldare);
  p->rdenswat = 1000.0;
  p->rdensref = 1.0;

  // Rain PSD (sucldp.F90:362-443)
  const double rcl_cr = 386.8 * zplrg;
  const double rcl_dr = 0.67;
  const double rcl_ar = p->rdenswat * rpi / 6.0;
  const double rcl_br = 3.0;
  const double rcl_x1r = 0.22, rcl_x2r = 2.20, rcl_x4r = 0.0;
  p->rcl_dr = rcl_dr;
  p->rcl_x1r = rcl_x1r;
  p->rcl_x2r = rcl_x2r;
  p->rcl_lam1r = rcl_ar * rcl_x1r * lanczos_gamma(rcl_br + 1.0 + rcl_x4r);
  p->rcl_lam2r = 1.0 / (rcl_br + 1.0 + rcl_x4r - rcl_x2r);
  p->rcl_cdenom1 = 5.57e11;
  p->rcl_cdenom2 = 1.03e8;
  p->rcl_cdenom3 = 2.04e2;
  p->rcl_ka273 = 2.4e-2;
  const double rcl_schmidt = 0.6, rcl_dynvisc = 1.717e-5;
  const double g1r = lanczos_gamma((rcl_dr + 5.0) / 2.0);
  const double g3r = lanczos_gamma(rcl_br + 1.0 + rcl_x4r);
  const double g4r = lanczos_gamma(rcl_dr + rcl_br + 1.0 + rcl_x4r);
  const double g5r = lanczos_gamma(3.0 + rcl_dr + rcl_x4r);
  const double g6r = lanczos_gamma(4.0 + rcl_br + rcl_x4r);
  p->rcl_const1r = 2.0 * rpi * rcl_x1r * rplrg * rpldare;
  p->rcl_const2r = 0.31 * std::pow(rcl_schmidt, 0.3333) * g1r * std::pow(rcl_cr / rcl_dynvisc, 0.5);
  p->rcl_const3r = (rcl_dr + 5.0) / 2.0 - rcl_x2r;
  p->rcl_const4r = 2.0 - rcl_x2r;
  p->rcl_const7r = rcl_cr * g4r / g3r;
  p->rcl_const9r = rpi * 0.25 * rcl_x1r * rcl_cr * g5r;
  p->rcl_const10r = 3.0 + rcl_dr - rcl_x2r;
  p->rcl_eff_racw = 0.7;
  p->rcl_kkaac = 67.0;
  p->rcl_kkbac = 1.15;
  p->rcl_kkaau = 1350.0;
  p->rcl_kkbauq = 2.47;
  p->rcl_kkbaun = -1.79;
  p->rcl_kk_cloud_num_sea = 50.0;
  p->rcl_kk_cloud_num_land = 300.0;
  p->rcl_fzrab = -0.66;
  const double rcl_fzrbb = 2.e2;
  p->rcl_const5r = rpi / 6.0 * rcl_fzrbb * rcl_ar * rcl_x1r * g6r;
  p->rcl_const6r = -1.0 * (4.0 + rcl_br + rcl_x4r - rcl_x2r);

  // Ice PSD (sucldp.F90:448-489); RCL_CI, RCL_DI overridden to 100, 1 (:456-457)
  const double rcl_ai = 0.069, rcl_bi = 2.0, rcl_ci = 100.0, rcl_di = 1.0;
  const double rcl_x1i = 2.0e6, rcl_x2i = 0.0, rcl_x4i = 0.0;
  const double g1i = lanczos_gamma(rcl_bi + 1.0 + rcl_x4i);
  const double g2i = lanczos_gamma(rcl_x4i + 2.0);
  const double g3i = lanczos_gamma((rcl_di + 5.0 + 2.0 * rcl_x4i) / 2.0);
  const double g4i = lanczos_gamma(rcl_di + rcl_bi + 1.0 + rcl_x4i);
  p->rcl_const1i = 1.0 / (rcl_ai * rcl_x1i * g1i);
  p->rcl_const2i = 1.0 * rpi * rcl_x1i;
  p->rcl_const3i = 0.44 * std::pow(0.6, 1.0 / 3.0) * std::pow(rcl_ci, 0.5) * g3i / std::pow(1.717e-5, 0.5);
  p->rcl_const4i = (2.0 + rcl_x4i - rcl_x2i) / (rcl_bi + 1.0 + rcl_x4i - rcl_x2i);
  p->rcl_const5i = (5.0 + rcl_di + 2.0 * rcl_x4i - 2.0 * rcl_x2i) / (2.0 * (rcl_bi + 1.0 + rcl_x4i - rcl_x2i));
  p->rcl_const6i = g2i;
  p->rcl_const7i = rcl_ci * g4i / g1i;
  p->rcl_apb1 = 7.14e11;
  p->rcl_apb2 = 1.16e8;
  p->rcl_apb3 = 2.416e2;

  // Snow PSD (sucldp.F90:495-530)
  const double rcl_as = 0.069, rcl_bs = 2.0, rcl_cs = 16.8 * zplrg, rcl_ds = 0.527;
  const double rcl_x1s = 2.0e6, rcl_x2s = 0.0, rcl_x4s = 0.0;
  const double g1s = lanczos_gamma(rcl_bs + 1.0 + rcl_x4s);
  const double g2s = lanczos_gamma(3.0 + rcl_ds + rcl_x4s);
  const double g3s = lanczos_gamma((rcl_ds + 5.0 + 2.0 * rcl_x4s) / 2.0);
  const double g4s = lanczos_gamma(rcl_x4s + 2.0);
  p->rcl_x1s = rcl_x1s;
  p->rcl_lam1s = rcl_as * rcl_x1s * lanczos_gamma(rcl_bs + 1.0 + rcl_x4s);
  p->rcl_lam2s = 1.0 / (rcl_bs + 1.0 + rcl_x4s - rcl_x2s);
  p->rcl_const1s = 1.0 / (rcl_as * rcl_x1s * g1s);
  p->rcl_const2s = 1.0 * rpi * rcl_x1s;
  p->rcl_const3s = 0.44 * std::pow(0.6, 1.0 / 3.0) * std::pow(rcl_cs, 0.5) * g3s / std::pow(1.717e-5, 0.5);
  p->rcl_const4s = (2.0 + rcl_x4s - rcl_x2s) / (rcl_bs + 1.0 + rcl_x4s - rcl_x2s);
  p->rcl_const5s = (5.0 + rcl_ds + 2.0 * rcl_x4s - 2.0 * rcl_x2s) / (2.0 * (rcl_bs + 1.0 + rcl_x4s - rcl_x2s));
  p->rcl_const6s = g4s;
  p->rcl_const7s = rpi * 0.25 * rcl_x1s * rcl_cs * g2s;
  p->rcl_const8s = (3.0 + rcl_ds + rcl_x4s - rcl_x2s) / (rcl_bs + 1.0 + rcl_x4s - rcl_x2s);

  // NCLDTOP from the standard atmosphere (sucldp.F90:631-634), RCLDTOPP = 100 Pa
  const double rcldtopp = 100.0;
  int ncldtop = 2;
  for (int jlev = nflevg; jlev >= 2; --jlev)
    if (stpre[jlev - 1] > rcldtopp) ncldtop = jlev;
  p->ncldtop = ncldtop;

  // RSSICEFACTOR (sucldp.F90:655)
  p->rssicefactor = std::fmin(ptsphy / p->rkooptau, 1.0);

  // ---- Other types --------------------------------------------------------
  p->lmfdsnow = 1;        // sucumf.F90:203
  p->rmfadvw = 0.0;       // sucumf.F90:228
  p->rmfadvwdd = 0.0;     // sucumf.F90:229
  p->rcloud_frac_std = 1.0;  // suecrad.F90:745
  p->reisthsc = 5.0;      // suvdf.F90:135
  p->lcloud_inhomog = 0;  // sucldp.F90:432
  p->lrad_cloud_inhomog = 0;  // su0phy.F90:1336
  p->laeraware = 0;       // NAERCLD = 0 (sucldp.F90)
  p->lspp = 0;            // spp_mod.F90:63
  p->lcldbud = 0;         // NCLDDIAG = 0, LCLDBUD* = .F., LBUD23 = .F.
  return CLOUDSC_OK;
}
