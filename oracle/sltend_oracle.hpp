// CPU oracle of SLTEND's supersaturation adjustment -- TEST INFRASTRUCTURE ONLY.
//
// Restates section 3 of ifs-source/arpifs/phys_ec/sltend.F90 (:370-409): on the levels NCLDTOP..KLEV the state
// updated with the tendencies so far, T + TSPHY*(PCENT+PTENT) etc., goes through CLOUD_SUPERSATCHECK with
// KFTLIQICE = 1, and the adjustments divided by TSPHY are added to PTENT, PTENQ, PTENA and PTENCLD(:,:,NCLDQL/QI),
// and kept for the budget diagnostics (ZADJA/ZADJL/ZADJI).  One block, Fortran order (KLON,KLEV) arrays,
// PRS1 (KLON,KLEV+1), PTENCLD (KLON,KLEV,5).  R = JPRB, RD = type of CLOUD_SUPERSATCHECK.
#pragma once

#include <vector>

#include "cloudsc_oracle.hpp"

namespace oracle {

template <typename R>
struct SltendArgs {
  int kidia, kfdia, klon, klev;
  R tsphy;
  const R *pt, *pq, *pa, *prsf1, *prs1, *pcent, *pcenq, *pcena;
  R *ptent, *ptenq, *ptena, *ptencld;
  R *adja, *adjl, *adji;  // may be null
};

template <typename R, typename RD>
void sltend_supersat(const cloudsc_params_t& p, const SltendArgs<R>& a) {
  const int klon = a.klon, klev = a.klev;
  auto I2 = [klon](int jl, int jk) { return (size_t)(jk - 1) * klon + (jl - 1); };
  const R zrtsphy = R(1) / a.tsphy;  // :271
  const int n = a.kfdia - a.kidia + 1;
  std::vector<RD> zt(n), zq(n), za(n), t_adj(n), q_adj(n), a_adj(n), l_adj(n), i_adj(n);
  std::vector<R> pp(n), ps(n);
  for (int jk = p.ncldtop; jk <= klev; ++jk) {
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {  // :376-380
      const int i = jl - a.kidia;
      zt[i] = RD(a.pt[I2(jl, jk)] + a.tsphy * (a.pcent[I2(jl, jk)] + a.ptent[I2(jl, jk)]));
      zq[i] = RD(a.pq[I2(jl, jk)] + a.tsphy * (a.pcenq[I2(jl, jk)] + a.ptenq[I2(jl, jk)]));
      za[i] = RD(a.pa[I2(jl, jk)] + a.tsphy * (a.pcena[I2(jl, jk)] + a.ptena[I2(jl, jk)]));
      pp[i] = a.prsf1[I2(jl, jk)];  // PRSF1(KIDIA:KFDIA,JK): element JL-KIDIA+1 of the section
      ps[i] = a.prs1[(size_t)klev * klon + (jl - 1)];
    }
    // :384-387.  The reference passes the sections PRSF1(KIDIA:KFDIA,JK) and PRS1(KIDIA:KFDIA,KLEV+1), which
    // the callee indexes from 1 while it indexes ZT_UPD etc. by JL; the two agree for KIDIA = 1, the only case
    // restated here (full blocks, as EC_PHYS_DRV calls it).
    supersatcheck<R, RD>(p, 1, n, zt.data(), zq.data(), za.data(), pp.data(), ps.data(), t_adj.data(),
                         q_adj.data(), a_adj.data(), l_adj.data(), i_adj.data(), 1);
    for (int jl = a.kidia; jl <= a.kfdia; ++jl) {  // :389-401
      const int i = jl - a.kidia;
      const size_t o = I2(jl, jk);
      a.ptent[o] = a.ptent[o] + R(t_adj[i]) * zrtsphy;
      a.ptenq[o] = a.ptenq[o] + R(q_adj[i]) * zrtsphy;
      a.ptena[o] = a.ptena[o] + R(a_adj[i]) * zrtsphy;
      a.ptencld[o] = a.ptencld[o] + R(l_adj[i]) * zrtsphy;
      a.ptencld[o + (size_t)klev * klon] = a.ptencld[o + (size_t)klev * klon] + R(i_adj[i]) * zrtsphy;
      if (a.adja) a.adja[o] = R(a_adj[i]) * zrtsphy;
      if (a.adjl) a.adjl[o] = R(l_adj[i]) * zrtsphy;
      if (a.adji) a.adji[o] = R(i_adj[i]) * zrtsphy;
    }
  }
}

}  // namespace oracle
