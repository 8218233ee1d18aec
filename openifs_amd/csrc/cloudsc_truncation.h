// Truncation of the explicit sinks of one level (cloudsc.F90:3455-3536) on the CLOUDSC kernel's registers.
//
// The kernel keeps ZSOLQA of one point as its 4 diagonal entries dl, di, dr, ds (ZSOLQA(L,L) .. ZSOLQA(S,S); the
// vapour diagonal is zero on the default path) and 7 antisymmetric pairs, each stored once as x = ZSOLQA(a,b) =
// -ZSOLQA(b,a): il = (I,L), ri = (R,I), rs = (R,S), vl = (V,L), vi = (V,I), vr = (V,R), vs = (V,S); the other
// off-diagonal entries are zero.  Scaling both halves of a pair by the same factor keeps it antisymmetric exactly,
// so one stored value per pair gives the reference's results bit for bit.
//
// Plain C++ apart from the attribute macro, so that the host test (tests/test_truncation.py) compiles it with the
// oracle's truncate_sinks and compares them case by case.
#pragma once

#ifdef __HIPCC__
#define CLOUDSC_HD __host__ __device__
#else
#define CLOUDSC_HD
#endif

namespace cloudsc_gpu {

// Species that share a pair of ZSOLQA ("partners"): L-I, L-V, I-R, I-V, R-S, R-V, S-V (bits L=1 I=2 R=4 S=8 V=16).
// Bit a of the result is set when the species set a holds a partner pair.
constexpr unsigned truncation_partner_sets() {
  const unsigned pairs[7] = {1u | 2u, 1u | 16u, 2u | 4u, 2u | 16u, 4u | 8u, 4u | 16u, 8u | 16u};
  unsigned sets = 0;
  for (unsigned a = 0; a < 32; ++a)
    for (unsigned p : pairs)
      if ((a & p) == p) sets |= 1u << a;
  return sets;
}

// qx: ZQX of the level (L, I, R, S, V); div(a, b) and vmax(a, b): the division and MAX of the precision T.
template <typename T, typename Div, typename Max>
CLOUDSC_HD inline void truncate_sinks(T& dl, T& di, T& dr, T& ds, T& il, T& ri, T& rs, T& vl, T& vi, T& vr, T& vs,
                                      const T qx[5], T zepsec, Div div, Max vmax) {
  // Sink sums of the rows of ZSOLQA (minus sign: sinks positive), each in the reference's column order L,I,R,S,V.
  const auto sum_l = [&] { return ((T(0) - dl) + il) + vl; };             // (L,L)=dl (L,I)=-il (L,V)=-vl
  const auto sum_i = [&] { return (((T(0) - il) - di) + ri) + vi; };      // il, di, (I,R)=-ri, (I,V)=-vi
  const auto sum_r = [&] { return (((T(0) - ri) - dr) - rs) + vr; };      // ri, dr, rs, (R,V)=-vr
  const auto sum_s = [&] { return ((T(0) + rs) - ds) + vs; };             // (S,R)=-rs, ds, (S,V)=-vs
  const auto sum_v = [&] { return (((T(0) - vl) - vi) - vr) - vs; };      // vl, vi, vr, vs
  // ZZRATIO applied to the sinks ZSOLQA(JO,JN) < 0 of one species and their mirror entries ZSOLQA(JN,JO)
  // (:3528-3533).  A stored pair x = ZSOLQA(a,b) = -ZSOLQA(b,a) is a sink of a when x < 0 and of b when x > 0; a
  // negative diagonal entry is scaled twice (:3531-3532).  Each returns the rows that hold the mirrors.
  const auto scale_l = [&](T zz) {
    if (il > T(0)) il = il * zz;
    if (dl < T(0)) dl = (dl * zz) * zz;
    if (vl > T(0)) vl = vl * zz;
    return 2 | 16;
  };
  const auto scale_i = [&](T zz) {
    if (il < T(0)) il = il * zz;
    if (di < T(0)) di = (di * zz) * zz;
    if (ri > T(0)) ri = ri * zz;
    if (vi > T(0)) vi = vi * zz;
    return 1 | 4 | 16;
  };
  const auto scale_r = [&](T zz) {
    if (ri < T(0)) ri = ri * zz;
    if (dr < T(0)) dr = (dr * zz) * zz;
    if (rs < T(0)) rs = rs * zz;
    if (vr > T(0)) vr = vr * zz;
    return 2 | 8 | 16;
  };
  const auto scale_s = [&](T zz) {
    if (rs > T(0)) rs = rs * zz;
    if (ds < T(0)) ds = (ds * zz) * zz;
    if (vs > T(0)) vs = vs * zz;
    return 4 | 16;
  };
  const auto scale_v = [&](T zz) {
    if (vl < T(0)) vl = vl * zz;
    if (vi < T(0)) vi = vi * zz;
    if (vr < T(0)) vr = vr * zz;
    if (vs < T(0)) vs = vs * zz;
    return 1 | 2 | 4 | 8;
  };

  const T rsum[5] = {sum_l(), sum_i(), sum_r(), sum_s(), sum_v()};
  // ZRATIO = ZMAX/MAX(sinksum,ZMAX) is exactly 1 unless the sinks overshoot; with no
  // overshoot anywhere every scale factor below is 1 and the step is a no-op.
  T zmax[5];
  bool any_trunc = false;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    zmax[m] = vmax(qx[m], zepsec);
    any_trunc = any_trunc || (rsum[m] > zmax[m]);
  }
  if (!any_trunc) return;  // lanes without an overshoot are no-ops: every scale factor would be 1

  // ZRATIO: a quotient only for the species that overshoot (liquid and rain nearly always).  `act` marks the
  // species whose step can scale: those that overshoot, and those whose row a scaling has changed since.  Any
  // other species has its first sink sum, at most ZMAX, so its step scales by 1: skipped.  `first` marks the
  // species with ZRATIO < 1, which the sort (:3479-3495) puts ahead of all species at ratio 1.
  T zratio[5];
  int act = 0, first = 0;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    zratio[m] = T(1);
    if (rsum[m] > zmax[m]) {
      zratio[m] = div(zmax[m], rsum[m]);
      act |= 1 << m;
      if (zratio[m] < T(1)) first |= 1 << m;
    }
  }

  if (!((truncation_partner_sets() >> first) & 1u)) {
    // No two species of `first` are partners: no step of theirs changes another's row or entries, so their steps
    // commute and each finds its row as it was, i.e. scales by its ZRATIO.  They run in species order, then the
    // species at ratio 1 in species order (the order of the stable sort), each only if its row has changed.
    if (first & 1) act |= scale_l(zratio[0]);
    if (first & 2) act |= scale_i(zratio[1]);
    if (first & 4) act |= scale_r(zratio[2]);
    if (first & 8) act |= scale_s(zratio[3]);
    if (first & 16) act |= scale_v(zratio[4]);
    if ((act & ~first) & 1) {
      const T s = sum_l();
      if (s > zmax[0]) act |= scale_l(div(zmax[0], s));
    }
    if ((act & ~first) & 2) {
      const T s = sum_i();
      if (s > zmax[1]) act |= scale_i(div(zmax[1], s));
    }
    if ((act & ~first) & 4) {
      const T s = sum_r();
      if (s > zmax[2]) act |= scale_r(div(zmax[2], s));
    }
    if ((act & ~first) & 8) {
      const T s = sum_s();
      if (s > zmax[3]) act |= scale_s(div(zmax[3], s));
    }
    if ((act & ~first) & 16) {
      const T s = sum_v();
      if (s > zmax[4]) act |= scale_v(div(zmax[4], s));
    }
    return;
  }

  // Partners among the species at ratio < 1: the full sort.  Stable ascending order of zratio (insertion sort of
  // :3479-3495): rank[m] = position of species m = #{n < m: zratio[n] <= zratio[m]} + #{n > m: zratio[n] <
  // zratio[m]}, one comparison per pair.
  int rank[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int m = 1; m < 5; ++m) {
#pragma unroll
    for (int n = 0; n < m; ++n) {
      const bool le = zratio[n] <= zratio[m];
      rank[m] += le ? 1 : 0;
      rank[n] += le ? 0 : 1;
    }
  }
  // Step s processes the species of rank s (:3500-3536) from its current entries.  Each species has its own
  // branch with only its own entries, so a wave executes only the species its lanes process.
#pragma unroll
  for (int s = 0; s < 5; ++s) {
    if (rank[0] == s && (act & 1)) {
      const T ss = sum_l();
      if (ss > zmax[0]) act |= scale_l(div(zmax[0], ss));
    }
    if (rank[1] == s && (act & 2)) {
      const T ss = sum_i();
      if (ss > zmax[1]) act |= scale_i(div(zmax[1], ss));
    }
    if (rank[2] == s && (act & 4)) {
      const T ss = sum_r();
      if (ss > zmax[2]) act |= scale_r(div(zmax[2], ss));
    }
    if (rank[3] == s && (act & 8)) {
      const T ss = sum_s();
      if (ss > zmax[3]) act |= scale_s(div(zmax[3], ss));
    }
    if (rank[4] == s && (act & 16)) {
      const T ss = sum_v();
      if (ss > zmax[4]) act |= scale_v(div(zmax[4], ss));
    }
  }
}

}  // namespace cloudsc_gpu
